"""CPU pinning for the Hyyro LLCS path: the reference's own lllcs_hyyro
(preprocess + hyyro_magic<uint64_t>, transposition_network_binary_alphabet.h
:743-820) compiled in place agrees with plain LCS dynamic programming, the
value the device's bit-vector kernel is checked against."""
import numpy as np
import pytest

from oracle.binding import Oracle, Reference


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref needs the reference tree to build")
def test_reference_hyyro_equals_dp():
    r, o = Reference(), Oracle()
    g = np.random.default_rng(12)
    for _ in range(300):
        m, n = int(g.integers(0, 300)), int(g.integers(0, 700))
        sigma = int(g.choice([2, 4, 20, 256]))
        a, b = g.integers(0, sigma, m), g.integers(0, sigma, n)
        assert r.hyyro(a, b) == o.llcs_dp(a, b), (m, n, sigma)
    # word boundaries
    for m, n in [(64, 64), (65, 63), (1, 128), (129, 1)]:
        a, b = g.integers(0, 4, m), g.integers(0, 4, n)
        assert r.hyyro(a, b) == o.llcs_dp(a, b)
