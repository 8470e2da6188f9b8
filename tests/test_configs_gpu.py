"""Full-size workloads on the GPU against the reference's own outputs
(SHA-256 of the row->col permutation, LLCS, sampled entries) and
size-independent properties."""
import numpy as np
import pytest

import golden_io as G
from semilocal_amd import workloads as W

pytestmark = pytest.mark.gpu


def _check(perm, rec):
    assert W.sha256(perm) == rec["perm_sha256"]
    idx = np.array(rec["sample_idx"])
    assert np.array_equal(perm[idx], np.array(rec["sample_val"]))


def test_binary_pair_2p20_permutation_and_score(gpu):
    rec = G.configs()["binary_pair_2p20"]
    a, b = W.binary_pair()
    assert W.sha256(a, b) == rec["input_sha256"]
    perm = gpu.semi_local_comb(a, b)
    assert np.array_equal(np.sort(perm), np.arange(len(perm)))      # a permutation
    _check(perm, rec)
    llcs = gpu.llcs_from_kernel(perm, len(a), len(b))
    assert llcs == rec["llcs"] == rec["llcs_bitparallel"]
    assert gpu.llcs_bitparallel(a, b) == rec["llcs_bitparallel"]
    # the hybrid's block grids the bench chooses among: the same permutation
    for sm, sn in [(1, 2), (2, 1)]:
        h = gpu.Hybrid(a, b, sm, sn)
        _check(h.run(), rec)
        h.close()


def test_similar_pair_1e6_direct_combing_equals_reference_hybrid(gpu):
    # the kernel is unique: combing the whole pair equals the reference's
    # 8x8-block hybrid with steady-ant merges
    rec = G.configs()["similar_pair_1e6"]
    a, b = W.similar_pair()
    assert W.sha256(a, b) == rec["input_sha256"]
    perm = gpu.semi_local_comb(a, b)
    _check(perm, rec)
    assert gpu.llcs_from_kernel(perm, len(a), len(b)) == rec["llcs"]


def test_similar_pair_1e6_hybrid_8x8_matches_reference_hybrid(gpu):
    # the configuration's own path: 64 blocks combed in one batched launch,
    # merged level by level with staggered steady-ant multiplications
    rec = G.configs()["similar_pair_1e6"]
    a, b = W.similar_pair()
    perm = gpu.hybrid_iterative(a, b, 8, 8)
    _check(perm, rec)
    assert gpu.llcs_from_kernel(perm, len(a), len(b)) == rec["llcs"]


def test_similar_pair_blocks_every_tile_shape_gives_the_same_braids(gpu):
    # The hybrid's combing launch on this configuration: 64 blocks of ~125k x
    # 125k combed as one batch.  Every shape of the autotuner's grid (queue
    # order x rows per lane, fixed or per pair, x waves per CU) must give the braids of the
    # default shape, whose hybrid result matches the reference's fixture.
    a, b = W.similar_pair()
    ba = [a[i * len(a) // 8:(i + 1) * len(a) // 8] for i in range(8)]
    bb = [b[j * len(b) // 8:(j + 1) * len(b) // 8] for j in range(8)]
    bt = gpu.Batch([(x, y) for x in ba for y in bb])
    want = bt.run().results()
    bad = []
    shapes = [(sched, R, w) for sched in (0, 1, 2, 3) for R in (4, 8, 16) for w in (4, 8, 12, 16, 20, 24, 32)]
    shapes += [(sched, 0, w) for sched in (2, 4, 6) for w in (8, 12)]   # a height per pair, longer lags
    for sched, R, w in shapes:
        bt.set_schedule(sched)
        bt.set_tiles(R, w)
        try:
            got = bt.run().results()
            ok = all(np.array_equal(x, y) for x, y in zip(got, want))
        except gpu.SemiLocalError as e:
            ok = False
            print(f"shape sched={sched} R={R} wpc={w}: {e}", flush=True)
        if not ok:
            bad.append((sched, R, w))
    bt.close()
    assert not bad, bad
