"""Re-entrancy of the C ABI: the reference's hybrids call their block solver
from concurrent OpenMP tasks (SemiLocalLCS.h:674-697, semi_local.h:596-623),
so every one-shot entry point must be safe to call from many host threads at
once.  Eight threads each run combing, a ragged batch, a braid product, a
hybrid, Hyyro LLCS and a duplicate search on their own seeded inputs (ctypes
releases the GIL during each call, so the calls overlap on the device); every
result must equal the oracle's.  Bar: exact."""
import threading

import numpy as np
import pytest

from oracle.binding import AMatchOracle

pytestmark = pytest.mark.gpu


def _work(gpu, oracle, seed, errors):
    try:
        g = np.random.default_rng(seed)
        for it in range(3):
            a = g.integers(0, 4, int(g.integers(200, 3000))).astype(np.uint8)
            b = g.integers(0, 4, int(g.integers(200, 3000))).astype(np.uint8)
            perm = gpu.semi_local_comb(a, b)
            if not np.array_equal(perm, oracle.comb(a, b)):
                errors.append(f"seed {seed} it {it}: semi_local_comb")
            pairs = [(g.integers(0, 20, int(g.integers(20, 500))).astype(np.uint8),
                      g.integers(0, 20, int(g.integers(20, 500))).astype(np.uint8)) for _ in range(12)]
            for (x, y), r in zip(pairs, gpu.comb_batch(pairs)):
                if not np.array_equal(r, oracle.comb(x, y)):
                    errors.append(f"seed {seed} it {it}: comb_batch")
                    break
            p, q = g.permutation(2000).astype(np.int32), g.permutation(2000).astype(np.int32)
            if not np.array_equal(gpu.braid_multiply(p, q), oracle.steady_ant(p, q)):
                errors.append(f"seed {seed} it {it}: braid_multiply")
            if not np.array_equal(gpu.hybrid_iterative(a, b, 2, 2), perm):
                errors.append(f"seed {seed} it {it}: hybrid_iterative")
            if gpu.llcs_hyyro(a, b) != gpu.llcs_from_kernel(perm, len(a), len(b)):
                errors.append(f"seed {seed} it {it}: llcs_hyyro")
            pat = g.integers(0, 4, 20).astype(np.uint8)
            text = g.integers(0, 4, 600).astype(np.uint8)
            text[100:120] = pat
            if gpu.duplicate_search(pat, text, 0.8, 1) != AMatchOracle().dup_search(pat, text, 0.8, 1):
                errors.append(f"seed {seed} it {it}: duplicate_search")
    except Exception as e:   # reported by the main thread
        errors.append(f"seed {seed}: {type(e).__name__}: {e}")


def test_entry_points_from_eight_threads_at_once(gpu, oracle):
    errors = []
    threads = [threading.Thread(target=_work, args=(gpu, oracle, 100 + k, errors)) for k in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a worker thread did not finish"
    assert errors == []
