"""semilocal_amd -- MI355X-native semi-local LCS engine.

The product is libsemilocal_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/semilocal_hip.h, plus the C++ drop-in strategies in
include/semilocal_hip.hpp).  This package builds it and exposes a thin
Python binding used by the tests and the benchmark.
"""
from ._lib import (AMatch, BITPAR_CARRY, BITPAR_COMBING, Batch, BitParallel, BraidIndex, Hybrid, HyyroBatch, SemiLocalError, braid_multiply,
                   braid_multiply_batch, braid_multiply_time, comb_batch, device_info, duplicate_search,
                   hybrid_autotune_blocks, hybrid_bucket_blocks, hybrid_iterative, lib, llcs_bitparallel, llcs_hyyro, plan_queue, semi_local_comb,
                   staggered_multiply, Server)

__all__ = ["AMatch", "BITPAR_CARRY", "BITPAR_COMBING", "Batch", "BitParallel", "BraidIndex", "Hybrid", "HyyroBatch", "SemiLocalError",
           "braid_multiply", "braid_multiply_batch", "braid_multiply_time", "comb_batch", "device_info",
           "duplicate_search", "hybrid_autotune_blocks", "hybrid_bucket_blocks", "hybrid_iterative", "lib", "llcs_bitparallel", "llcs_hyyro", "plan_queue",
           "semi_local_comb", "staggered_multiply", "Server", "llcs_from_kernel"]


def llcs_from_kernel(perm, m: int, n: int) -> int:
    """LLCS = m - #{s < m : perm[s] >= n}  (reference include/LCS/LCS.h:223-233)."""
    import numpy as np

    perm = np.asarray(perm)
    return int(m - np.count_nonzero(perm[:m] >= n))
