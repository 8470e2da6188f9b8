"""GPU parity of kernel (b) (bit-parallel LLCS) in both forms -- the
sliding-diagonal combing of formula 2 and Hyyro's carry recurrence with the
packed bits as masks -- against the reference's golden vectors, DP, and the
integer permutation of kernel (a) (0-1 principle: the 1-bit strands are the
integer strands thresholded at m)."""
import numpy as np
import pytest

import golden_io as G

pytestmark = pytest.mark.gpu

FORMS = [0, 1]   # BITPAR_CARRY, BITPAR_COMBING


def llcs_form(gpu, a, b, form):
    plan = gpu.BitParallel(a, b, form=form)
    try:
        return plan.llcs()
    finally:
        plan.close()


def test_bitparallel_golden_vectors(gpu):
    for a, b, llcs in G.bitpar():
        assert gpu.llcs_bitparallel(a, b) == llcs
        for form in FORMS:
            assert llcs_form(gpu, a, b, form) == llcs, form


@pytest.mark.parametrize("form", FORMS)
def test_bitparallel_vs_dp_random(gpu, oracle, form):
    g = np.random.default_rng(21)
    for _ in range(25):
        la, lb = 64 * int(g.integers(1, 40)), 64 * int(g.integers(1, 40))
        a = g.integers(0, 2, la).astype(np.uint8)
        b = g.integers(0, 2, lb).astype(np.uint8)
        assert llcs_form(gpu, a, b, form) == oracle.llcs_dp(a, b), (la, lb)


@pytest.mark.parametrize("form", FORMS)
def test_bitparallel_matches_integer_permutation(gpu, form):
    g = np.random.default_rng(22)
    for la, lb in [(1 << 14, 1 << 15), (1 << 15, 1 << 14), (4160, 4160 * 3)]:
        a = g.integers(0, 2, la).astype(np.uint8)
        b = g.integers(0, 2, lb).astype(np.uint8)
        perm = gpu.semi_local_comb(a, b)
        assert llcs_form(gpu, a, b, form) == gpu.llcs_from_kernel(perm, la, lb), (la, lb)


def test_carry_form_matches_combing_across_buckets(gpu, oracle):
    """The carry form (the Hyyro kernel on the unpacked full strings) over
    its length buckets -- one workgroup per pair up to 256 packed words,
    segments beyond -- with either string as the packed one and lengths that
    are not whole words of 64, against the combing form and, where small,
    DP.  The packed padding (zero symbols up to whole 64-bit words) is part
    of the reference's grid: its score is |a| - 64 words(a) + LLCS(a +
    padding, b + padding)."""
    g = np.random.default_rng(24)
    shapes = [(64, 64), (64, 8192), (8192, 64), (64 * 255, 64 * 257), (64 * 513, 64 * 511),
              (50, 70), (100, 64 * 300), (64 * 300 + 1, 64 * 2), (16384 + 64, 16384 - 64)]
    for la, lb in shapes:
        for p1 in (0.5, 0.9):
            a = (g.random(la) < p1).astype(np.uint8)
            b = (g.random(lb) < 0.5).astype(np.uint8)
            carry, comb = llcs_form(gpu, a, b, 0), llcs_form(gpu, a, b, 1)
            assert carry == comb, (la, lb, p1)
            if la * lb <= 1 << 22:
                ap = np.zeros(-(-la // 64) * 64, np.uint8)
                bp = np.zeros(-(-lb // 64) * 64, np.uint8)
                ap[:la], bp[:lb] = a, b
                assert carry == la - len(ap) + oracle.llcs_dp(ap, bp), (la, lb, p1)


def test_resident_plan_repeats(gpu, oracle):
    g = np.random.default_rng(23)
    a = g.integers(0, 2, 64 * 300).astype(np.uint8)
    b = g.integers(0, 2, 64 * 500).astype(np.uint8)
    plan = gpu.BitParallel(a, b)
    assert plan.form == gpu.BITPAR_CARRY
    want = oracle.bitpar_llcs(a, b)
    assert [plan.llcs() for _ in range(3)] == [want] * 3
    # forms switched between runs of one plan, each run repeated
    for form in (1, 0, 1, 0):
        plan.form = form
        assert plan.form == form
        assert [plan.llcs() for _ in range(2)] == [want] * 2
    plan.close()


def test_carry_form_refused_past_its_length_limit(gpu):
    """A longer string of 2^20 + 64 symbols is past the carry form's 2^20, so
    the plan offers the combing form only (nothing is run)."""
    a = np.zeros((1 << 20) + 64, np.uint8)
    plan = gpu.BitParallel(a[:4096], a)
    assert plan.form == gpu.BITPAR_COMBING
    with pytest.raises(gpu.SemiLocalError):
        plan.form = gpu.BITPAR_CARRY
    with pytest.raises(gpu.SemiLocalError):
        plan.form = 7
    plan.close()


def test_padding_bits_are_symbols_in_both_forms(gpu, oracle):
    """The reference combs every bit of the packed words and returns
    a_total_symbols - (left strands leaving right), so padding bits past
    a_total are symbols too: the score is a_total - 64 words(a) + LLCS of the
    two full packed strings.  Random (non-zero) padding, a_total below 64
    words(a), both forms, through the C ABI."""
    import ctypes as C

    from semilocal_amd._lib import pack_bits, pack_bits_reverse
    lib = gpu.lib()
    g = np.random.default_rng(41)
    for aw, bw, cut in [(1, 1, 1), (1, 3, 40), (3, 1, 100), (5, 7, 257), (7, 5, 300)]:
        A = g.integers(0, 2, 64 * aw).astype(np.uint8)   # full strings, padding included
        B = g.integers(0, 2, 64 * bw).astype(np.uint8)
        a_total = 64 * aw - cut % 64 - 1
        ar, bp = pack_bits_reverse(A), pack_bits(B)
        want = a_total - 64 * aw + oracle.llcs_dp(A, B)
        if aw <= bw:   # the shapes the reference's diagonal loops accept: its restated combing agrees
            assert oracle.L.orc_bitpar_llcs(ar, aw, bp, bw, a_total, 1) == want
        for form in FORMS:
            h = C.c_void_p()
            assert lib.sl_bitpar_create(C.byref(h), ar.ctypes.data_as(C.c_void_p), aw,
                                        bp.ctypes.data_as(C.c_void_p), bw, a_total) == 0
            assert lib.sl_bitpar_set_form(h, form) == 0
            out = C.c_int()
            assert lib.sl_bitpar_run(h) == 0 and lib.sl_bitpar_result(h, C.byref(out)) == 0
            lib.sl_bitpar_destroy(h)
            assert out.value == want, (aw, bw, a_total, form)
        one = C.c_int()
        assert lib.sl_llcs_bitparallel(ar.ctypes.data_as(C.c_void_p), aw, bp.ctypes.data_as(C.c_void_p), bw,
                                       a_total, 1, 0, C.byref(one)) == 0
        assert one.value == want
