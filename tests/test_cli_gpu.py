"""GPU parity of the command-line drivers (semilocal_amd/bin/): each driver
runs on input files in the reference's format and its stdout, parsed the way
the reference's runners parse it (utils/runners.py:16-74; any stderr output
is a failure there and here), carries the oracle's result: the permutation
hash of the oracle's kernel, the DP score, the oracle's braid product.
Bar: exact integers."""
import numpy as np
import pytest

from semilocal_amd import harness

pytestmark = pytest.mark.gpu


def _files(tmp_path, a, b):
    fa, fb = str(tmp_path / "a.fna"), str(tmp_path / "b.fna")
    harness.write_sequence(fa, "seq-a", a)
    harness.write_sequence(fb, "seq-b", b)
    return fa, fb


@pytest.mark.parametrize("driver,threads", [("semi_local_hip_iterative", 4), ("semi_local_hip_hybrid_iterative", 1),
                                            ("semi_local_hip_hybrid_iterative", 4)])
@pytest.mark.parametrize("m,n,sigma", [(3000, 2500, 4), (700, 1900, 26), (1, 40, 2)])
def test_combing_drivers_print_the_oracle_hash(tmp_path, oracle, driver, threads, m, n, sigma):
    g = np.random.default_rng(m + 3 * n + sigma)
    a = g.integers(0, sigma, m).astype(np.int32)
    b = g.integers(0, sigma, n).astype(np.int32)
    fa, fb = _files(tmp_path, a, b)
    print(f"driver {driver} threads {threads} m {m} n {n}", flush=True)
    r = harness.run_combing(driver, threads, fa, fb, timeout=120)
    assert (r.a_size, r.b_size, r.a_name, r.b_name) == (m, n, "seq-a", "seq-b")
    assert r.value == harness.permutation_hash(oracle.comb(a, b))
    assert r.preprocess_ms >= 0 and r.algo_ms >= 0


@pytest.mark.parametrize("driver,m,n,sigma", [("prefix_lcs_hip", 3000, 2500, 4), ("prefix_lcs_hip", 500, 64, 20),
                                              ("prefix_lcs_hip", 400, 900, 300),     # > 256 symbols: combing path
                                              ("prefix_lcs_hip_bit_parallel", 1024, 2048, 2),
                                              ("prefix_lcs_hip_bit_parallel", 4096, 640, 2),
                                              ("prefix_lcs_hip_bit_parallel", 1000, 777, 2)])
def test_lcs_drivers_print_the_dp_score(tmp_path, oracle, driver, m, n, sigma):
    g = np.random.default_rng(m * 5 + n)
    a = g.integers(0, sigma, m).astype(np.int32)
    b = g.integers(0, sigma, n).astype(np.int32)
    fa, fb = _files(tmp_path, a, b)
    print(f"driver {driver} m {m} n {n}", flush=True)
    r = harness.run_combing(driver, 2, fa, fb, timeout=120)
    assert (r.a_size, r.b_size) == (m, n)
    assert r.value == oracle.llcs_dp(a, b)


@pytest.mark.parametrize("n,seed", [(1, 3), (5000, 42), (1 << 16, 7)])
def test_braid_driver_prints_the_oracle_product_hash(tmp_path, oracle, n, seed):
    dump = str(tmp_path / "pqr.txt")
    print(f"braid driver n {n} seed {seed}", flush=True)
    r = harness.run_braid(2, n, seed, dump=dump, timeout=120)
    assert (r.n, r.seed) == (n, seed)
    p, q, prod = (np.array(line.split(), np.int32) for line in open(dump).read().splitlines())
    assert np.array_equal(np.sort(p), np.arange(n)) and np.array_equal(np.sort(q), np.arange(n))
    assert np.array_equal(prod, oracle.steady_ant(p, q))
    assert r.hash == harness.permutation_hash(prod)
    # the inputs depend on the seed alone
    again = harness.run_braid(0, n, seed, timeout=120)
    assert again.hash == r.hash
