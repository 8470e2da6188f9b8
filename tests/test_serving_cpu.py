"""The serving parity fixture (tests/golden/serving_stream.json, written from
the reference's own combing) checked on the CPU: the regenerated stream is
byte-identical to the one the fixture was made from, the stream follows
config 5's law, and every fixture value that dynamic programming can reach
in seconds equals the LCS computed directly -- which pins the H formulas the
fixture generator read the query answers with."""
import numpy as np

import golden_io as G
from semilocal_amd import workloads as W


def test_fixture_matches_the_regenerated_stream():
    fx = G.serving_stream()
    reqs = W.serving_parity_stream()
    assert fx["requests"] == len(reqs)
    by_rid = {c["rid"]: c for c in fx["compares"]}
    for r in reqs:
        if r.kind == "compare":
            assert by_rid[r.rid]["input_sha256"] == W.sha256(r.a, r.b)
    kinds = {q["qkind"] for q in fx["queries"]}
    assert kinds == {"substring_string", "string_substring"}
    lens = [max(c["m"], c["n"]) for c in fx["compares"]]
    assert max(lens) == 200_000 and min(min(c["m"], c["n"]) for c in fx["compares"]) >= 200
    assert 0.2 < len(fx["queries"]) / fx["requests"] < 0.4


def test_fixture_values_equal_dynamic_programming(oracle):
    fx = G.serving_stream()
    reqs = {r.rid: r for r in W.serving_parity_stream()}
    checked = 0
    for c in fx["compares"]:
        r = reqs[c["rid"]]
        if c["m"] * c["n"] <= 3e7:
            assert oracle.llcs_dp(r.a, r.b) == c["llcs"]
            checked += 1
    assert checked >= 60
    qchecked = 0
    for q in fx["queries"]:
        t = reqs[q["target"]]
        s, e = q["start"], q["stop"]
        if len(t.a) * len(t.b) > 3e7:
            continue
        want = oracle.llcs_dp(t.a[s:e], t.b) if q["qkind"] == "substring_string" else oracle.llcs_dp(t.a, t.b[s:e])
        assert want == q["lcs"], q
        qchecked += 1
    assert qchecked >= 20
