"""Request loop on the GPU: continuous batching with overlapping
micro-batches, admission by device memory, resident braids and follow-up
queries.  Every result must equal the one-shot entry point's on the same
input (bit-exact: the braid and the LCS values are integers)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def _H(perm, m, n, i, j):
    # H(i, j) = j - i - #{r >= i + m : P[r] < j}  (approximate_matching/utils.h:281-293)
    return j - i - int(np.count_nonzero(perm[i + m:] < j))


def _pairs(g, count, median, lo, hi, sigma):
    out = []
    for _ in range(count):
        m, n = np.clip(np.round(np.exp(g.normal(np.log(median), 1.0, size=2))), lo, hi).astype(int)
        out.append((g.integers(0, sigma, size=m).astype(np.uint8), g.integers(0, sigma, size=n).astype(np.uint8)))
    return out


@pytest.mark.parametrize("index_min", [0, 2000])
def test_results_in_the_mix_equal_results_alone(gpu, oracle, index_min):
    # index_min 0: every query scans its braid; 2000: braids of m+n >= 2000
    # get a wavelet-matrix index and their queries take the index path
    g = _rng(101)
    pairs = _pairs(g, 96, 1500, 20, 30000, 4)
    srv = gpu.Server(pool_bytes=1 << 30, max_inflight=4, max_batch=12)
    srv.set_index(index_min)
    ids = [srv.submit_compare(a, b) for a, b in pairs]   # a burst: micro-batches overlap
    big = sum(len(a) + len(b) >= index_min for a, b in pairs)
    if index_min > 0:
        # index builds run after the results: wait until every large braid has one
        for rid in ids:
            srv.result(rid)
        t0 = time.time()
        while srv.stats()["indexed_braids"] < big and time.time() - t0 < 30:
            time.sleep(0.001)
        assert srv.stats()["indexed_braids"] == big > 0
    queries = []
    for k in range(80):
        t = int(g.integers(0, len(pairs)))
        a, b = pairs[t]
        if k % 2:
            i = int(g.integers(0, len(b) + 1)); j = int(g.integers(i, len(b) + 1))
            kind = srv.STRING_SUBSTRING
        else:
            i = int(g.integers(0, len(a) + 1)); j = int(g.integers(i, len(a) + 1))
            kind = srv.SUBSTRING_STRING
        queries.append((srv.submit_query(ids[t], kind, i, j), t, kind, i, j))
    for rid, (a, b) in zip(ids, pairs):
        alone = gpu.semi_local_comb(a, b)
        val, lat = srv.result(rid)
        assert val == gpu.llcs_from_kernel(alone, len(a), len(b))
        assert lat >= 0
        assert np.array_equal(srv.braid(rid, len(a) + len(b)), alone)
    for rid, t, kind, i, j in queries:
        a, b = pairs[t]
        perm = gpu.semi_local_comb(a, b)
        m, n = len(a), len(b)
        want = _H(perm, m, n, i, j) if kind == srv.STRING_SUBSTRING else _H(perm, m, n, -i, n + m - j) - i - (m - j)
        assert srv.result(rid)[0] == want
    # a few windows against dynamic programming directly
    for rid, t, kind, i, j in queries[:12]:
        a, b = pairs[t]
        if len(a) * len(b) > 4e6:
            continue
        want = oracle.llcs_dp(a, b[i:j]) if kind == srv.STRING_SUBSTRING else oracle.llcs_dp(a[i:j], b)
        assert srv.result(rid)[0] == want
    st = srv.stats()
    assert st["compares_done"] == len(pairs) and st["queries_done"] == len(queries)
    assert st["launches"] > len(pairs) // 12     # several micro-batches
    assert st["peak_inflight_compares"] >= 2      # ... on the device at the same time
    if index_min > 0:
        on_big = sum(len(pairs[t][0]) + len(pairs[t][1]) >= index_min for _, t, _, _, _ in queries)
        assert st["indexed_braids"] == big and st["index_queries"] == on_big > 0
    else:
        assert st["indexed_braids"] == 0 and st["index_queries"] == 0
    srv.close()


def test_query_waits_for_its_comparison(gpu):
    g = _rng(7)
    a = g.integers(0, 4, size=40000).astype(np.uint8)
    b = g.integers(0, 4, size=30000).astype(np.uint8)
    srv = gpu.Server(pool_bytes=256 << 20)
    cid = srv.submit_compare(a, b)
    qid = srv.submit_query(cid, srv.STRING_SUBSTRING, 1000, 25000)   # submitted before the braid exists
    perm = gpu.semi_local_comb(a, b)
    assert srv.result(qid)[0] == _H(perm, len(a), len(b), 1000, 25000)
    assert srv.result(cid)[0] == gpu.llcs_from_kernel(perm, len(a), len(b))
    srv.close()


def test_admission_by_device_memory_and_eviction(gpu):
    # a pool that holds about two of these comparisons at a time (batch
    # buffers ~1 MB plus a 320 KB resident braid each)
    g = _rng(11)
    pairs = [(g.integers(0, 4, size=40000).astype(np.uint8), g.integers(0, 4, size=40000).astype(np.uint8))
             for _ in range(8)]
    srv = gpu.Server(pool_bytes=3 << 20, max_inflight=4, max_batch=4)
    ids = [srv.submit_compare(a, b) for a, b in pairs]
    for rid, (a, b) in zip(ids, pairs):
        perm = gpu.semi_local_comb(a, b)
        assert srv.result(rid)[0] == gpu.llcs_from_kernel(perm, len(a), len(b))
    st = srv.stats()
    assert st["compares_done"] == 8
    assert st["pool_peak"] <= st["pool_bytes"]
    assert st["evictions"] > 0            # resident braids made room for later comparisons
    # the newest braid is still resident; evicted ones are reported as such
    last = ids[-1]
    assert srv.result(srv.submit_query(last, srv.STRING_SUBSTRING, 0, 40000))[0] == srv.result(last)[0]
    gone = 0
    for rid in ids:
        try:
            srv.submit_query(rid, srv.STRING_SUBSTRING, 0, 10)
        except gpu.SemiLocalError:
            gone += 1
    assert gone == st["evictions"] == 8 - st["resident_braids"]
    # a comparison that cannot fit the pool at all fails cleanly
    big = g.integers(0, 4, size=1_000_000).astype(np.uint8)
    rid = srv.submit_compare(big, big)
    with pytest.raises(gpu.SemiLocalError):
        srv.result(rid)
    # and the loop keeps serving afterwards
    a, b = pairs[0]
    rid = srv.submit_compare(a, b)
    assert srv.result(rid)[0] == srv.result(ids[0])[0]
    srv.close()


def test_release_and_argument_errors(gpu):
    g = _rng(13)
    a = g.integers(0, 20, size=3000).astype(np.uint8)
    b = g.integers(0, 20, size=2000).astype(np.uint8)
    srv = gpu.Server(pool_bytes=64 << 20)
    cid = srv.submit_compare(a, b)
    srv.result(cid)
    with pytest.raises(gpu.SemiLocalError):
        srv.submit_query(cid, srv.STRING_SUBSTRING, 5, 2001)      # window past the end
    with pytest.raises(gpu.SemiLocalError):
        srv.submit_query(cid + 100, srv.STRING_SUBSTRING, 0, 1)  # no such comparison
    srv.release(cid)
    with pytest.raises(gpu.SemiLocalError):
        srv.submit_query(cid, srv.STRING_SUBSTRING, 0, 1)
    nid = srv.submit_compare(a, b, keep=False)                   # score only
    perm = gpu.semi_local_comb(a, b)
    assert srv.result(nid)[0] == gpu.llcs_from_kernel(perm, len(a), len(b))
    assert srv.stats()["resident_braids"] == 0
    srv.close()


def test_config5_stream_against_the_reference_fixture(gpu):
    """Config 5's own workload through the request loop: Poisson arrivals
    (2,000/s, replayed in real time), lengths log-normal (median 5,000)
    clipped to [200, 200,000] with two comparisons at the maximum, alphabet 4,
    30% follow-up queries of both kinds.  Every comparison's braid and LLCS
    and every query's answer equal the fixture written from the reference's
    own combing (tests/golden/serving_stream.json), and comparison
    micro-batches were on the device at the same time."""
    import golden_io as G
    from semilocal_amd import workloads as W

    fx = G.serving_stream()
    reqs = W.serving_parity_stream()
    want_cmp = {c["rid"]: c for c in fx["compares"]}
    want_q = {q["rid"]: q for q in fx["queries"]}
    srv = gpu.Server(pool_bytes=4 << 30)
    ids = {}
    t0 = time.perf_counter()
    for r in reqs:
        delay = r.arrival_s - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        if r.kind == "compare":
            assert want_cmp[r.rid]["input_sha256"] == W.sha256(r.a, r.b)
            ids[r.rid] = srv.submit_compare(r.a, r.b)
        else:
            kind = srv.SUBSTRING_STRING if r.qkind == "substring_string" else srv.STRING_SUBSTRING
            ids[r.rid] = srv.submit_query(ids[r.target], kind, *r.window)
    for r in reqs:
        val, lat = srv.result(ids[r.rid])
        assert lat >= 0
        if r.kind == "compare":
            w = want_cmp[r.rid]
            assert val == w["llcs"], r.rid
            assert W.sha256(srv.braid(ids[r.rid], w["m"] + w["n"])) == w["perm_sha256"], r.rid
        else:
            assert val == want_q[r.rid]["lcs"], (r.rid, r.qkind, r.window)
    st = srv.stats()
    assert st["compares_done"] == len(want_cmp) and st["queries_done"] == len(want_q)
    assert st["peak_inflight_compares"] >= 2, st     # micro-batches overlapped on the device
    srv.close()
