"""Pins the oracle (CPU restatement of the reference Ozaki path) before it is trusted.

Each check is an independent route to the same answer: the word-level restatement of
the AIE split vs the closed form, numpy int64 GEMMs vs the triangular bins, the
mechanistic paired-int64 format + host conversion vs exact-T truncation, exact rational
arithmetic vs the MPFR oracle, and known answers that follow from the mathematics
(A = B = ones gives K; small-integer matrices are exact; power-of-two scalings commute).
"""
from fractions import Fraction
import math

import numpy as np
import pytest

import oracle
from conftest import gen_inputs

REF_PRESETS = [(3, 29), (3, 28), (3, 27), (3, 26), (3, 25), (4, 18), (7, 8)]
GPU_PRESETS = [(7, 3), (7, 7), (7, 9), (7, 10), (5, 12), (6, 16), (2, 30), (1, 48)]


def edge_matrix(rng, m, n, phi=2.0):
    X = gen_inputs(rng, m, n, phi)
    X[0, :6] = [0.0, -0.0, 1.0, 5e-324, 2.2250738585072014e-308, -2.2250738585072009e-308]
    X[1, :] = 0.0
    X[2, 3] = 1e-300
    X[3, :] = -1.0
    X[4, 0] = 1.7976931348623157e308
    X[5, 1] = np.nextafter(1.0, 2.0)
    return X


# ---------------------------------------------------------------- exponent max
def test_max_exp_brute_force():
    rng = np.random.default_rng(1)
    X = edge_matrix(rng, 16, 40)
    E = oracle.max_exp(X, 1)
    for r in range(16):
        ref = max(((int(np.float64(x).view(np.uint64)) >> 52) & 0x7FF) for x in X[r])
        assert E[r] == ref
    Ec = oracle.max_exp(X, 0)
    for c in range(40):
        ref = max(((int(np.float64(x).view(np.uint64)) >> 52) & 0x7FF) for x in X[:, c])
        assert Ec[c] == ref


# ---------------------------------------------------------------- split
@pytest.mark.parametrize("b,d", REF_PRESETS + GPU_PRESETS)
def test_split_closed_form_matches_word_level(b, d):
    rng = np.random.default_rng(b * 100 + d)
    for phi in (0.5, 2.0, 8.0):
        X = edge_matrix(rng, 24, 96, phi)
        E = oracle.max_exp(X, 1)
        s_closed = oracle.split(X, E, b, d, 1, mech=False)
        s_mech = oracle.split(X, E, b, d, 1, mech=True)
        assert np.array_equal(s_closed, s_mech)


def test_split_word_level_all_biases():
    """Every exponent distance 0..140 against a fixed row maximum, one element each."""
    for b, d in [(3, 29), (3, 26), (7, 8), (7, 13), (4, 18), (5, 19)]:
        xs = [1.5]
        for bias in range(0, 141):
            xs.append(math.ldexp(1.0 + 0.123456789, -bias))
            xs.append(-math.ldexp(1.0 + 0.987654321, -bias))
        X = np.array([xs])
        E = oracle.max_exp(X, 1)
        assert np.array_equal(oracle.split(X, E, b, d, 1), oracle.split(X, E, b, d, 1, mech=True))


@pytest.mark.parametrize("b,d", [(3, 29), (7, 8), (7, 3), (4, 18)])
def test_split_invariants(b, d):
    rng = np.random.default_rng(7)
    X = gen_inputs(rng, 20, 64, 1.0)
    E = oracle.max_exp(X, 1)
    S = oracle.split(X, E, b, d, 1).astype(np.int64)
    assert np.abs(S).max() <= (1 << b) - 1
    # sign agreement: every nonzero digit carries the sign of its element
    sgn = np.sign(X)[None, :, :]
    assert np.all((S == 0) | (np.sign(S) == sgn))
    # the digits reassemble trunc(|x| * 2^(1022 + b d - E)) exactly (normal inputs)
    for r in range(20):
        for k in range(64):
            q = 0
            for j in range(d):
                q = (q << b) | abs(int(S[j, r, k]))
            x = Fraction(float(X[r, k]))
            ref = math.floor(abs(x) * Fraction(2) ** (1022 + b * d - int(E[r])))
            assert q == ref
    # power-of-two row scaling leaves the digits unchanged
    X2 = X * 2.0 ** 37
    E2 = oracle.max_exp(X2, 1)
    assert np.array_equal(E2, E + 37)
    assert np.array_equal(oracle.split(X2, E2, b, d, 1), S.astype(np.int8))


def test_split_special_values():
    # x = row max = 1.0 -> s_0 = 2^(b-1) (the implicit one), rest zero
    for b, d in [(3, 29), (7, 8)]:
        X = np.array([[1.0, 0.0, -0.0]])
        S = oracle.split(X, oracle.max_exp(X, 1), b, d, 1)
        assert S[0, 0, 0] == 1 << (b - 1)
        assert not S[1:, 0, 0].any()
        assert not S[:, 0, 1:].any()
    # exponent gap beyond the window -> all-zero digits
    X = np.array([[1.0, 2.0 ** -100]])
    S = oracle.split(X, oracle.max_exp(X, 1), 7, 8, 1)
    assert not S[:, 0, 1].any()


def test_split_column_orientation_is_transpose():
    rng = np.random.default_rng(3)
    X = gen_inputs(rng, 48, 20, 1.0)
    E0 = oracle.max_exp(X, 0)
    E1 = oracle.max_exp(np.ascontiguousarray(X.T), 1)
    assert np.array_equal(E0, E1)
    assert np.array_equal(oracle.split(X, E0, 7, 8, 0), oracle.split(np.ascontiguousarray(X.T), E1, 7, 8, 1))


# ---------------------------------------------------------------- bins, pack, convert
@pytest.mark.parametrize("K", [128, 256, 512, 1024, 2048])
@pytest.mark.parametrize("mode", [oracle.ACCURATE, oracle.FAST])
def test_bins_pack_convert_vs_closed_form(K, mode):
    st, b, d = oracle.plan_bd(K, K, K, mode)
    assert st == 0
    rng = np.random.default_rng(K + mode)
    M, N = 24, 20
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
    SA, SB = oracle.split(A, Er, b, d, 1), oracle.split(B, Ec, b, d, 0)
    bins = oracle.bins(SA, SB)
    ref = np.zeros_like(bins)
    for i in range(d):
        for j in range(d - i):
            ref[i + j] += SA[i].astype(np.int64) @ SB[j].astype(np.int64).T
    assert np.array_equal(bins, ref)
    assert np.abs(bins).max() < 2 ** 31
    P = oracle.pack_paired(bins, K, b, d, Er, Ec)
    C_ref_path = oracle.convert_paired(P, K, b, d)
    T = oracle.bins_to_T(bins, b, d)
    C_closed = np.array([[oracle.T_to_double(T[r, c, 0], T[r, c, 1], int(Er[r]) + int(Ec[c]), b, d, 0)
                          for c in range(N)] for r in range(M)])
    assert np.array_equal(C_ref_path.view(np.int64), C_closed.view(np.int64))
    C_e2e, ov = oracle.ozaki_dgemm(A, B, b, d)
    assert ov == 0
    assert np.array_equal(C_e2e.view(np.int64), C_closed.view(np.int64))


def test_groups_table():
    # n0 / n1 of the reference presets (values follow ozaki_base.cc:151-161)
    assert oracle.groups(7, 8, 128) == (5, 3)
    assert oracle.groups(3, 29, 128) == (16, 13)
    assert oracle.groups(3, 28, 256) == (15, 13)
    assert oracle.groups(3, 27, 512) == (15, 12)
    assert oracle.groups(3, 26, 2048) == (14, 12)


def test_T_to_double_known_answers():
    b, d = 3, 29
    # T = 2^95 with E_r + E_c = 2046 -> 2^(95 + 2046 - 2044 - 90) = 128
    assert oracle.T_to_double(1 << 31, 0, 2046, b, d, 0) == 128.0
    assert oracle.T_to_double(0, 0, 2046, b, d, 0) == 0.0
    # truncation vs nearest on a 54-bit odd T
    t = (1 << 53) + 1
    s = 2046 - 2044 - b * (d + 1)
    assert oracle.T_to_double(0, t, 2046, b, d, 0) == math.ldexp(1 << 53, s)
    assert oracle.T_to_double(0, t + 2, 2046, b, d, 1) == math.ldexp((1 << 53) + 4, s)
    # negative T (two's complement 128-bit)
    neg = -(1 << 60)
    hi, lo = (neg >> 64) & ((1 << 64) - 1), neg & ((1 << 64) - 1)
    to_i64 = lambda u: u - (1 << 64) if u >= 1 << 63 else u
    assert oracle.T_to_double(to_i64(hi), to_i64(lo), 2046, b, d, 0) == -math.ldexp(1, 60 + s)


# ---------------------------------------------------------------- end to end
def test_example_ones_known_answer():
    """example/oztest.cpp: A = B = ones, accurate, M=256 N=512 K=128 -> C = 128 exactly."""
    A = np.ones((256, 128))
    B = np.ones((128, 512))
    C, ov = oracle.ozaki_dgemm(A, B, 3, 29)
    assert ov == 0 and np.all(C == 128.0)


@pytest.mark.parametrize("b,d", [(3, 29), (7, 8), (7, 3), (7, 7)])
def test_small_integers_exact(b, d):
    rng = np.random.default_rng(11)
    A = rng.integers(-127, 128, size=(17, 70)).astype(np.float64)
    B = rng.integers(-127, 128, size=(70, 23)).astype(np.float64)
    A[:, 0] = 127.0
    B[0, :] = -127.0
    C, ov = oracle.ozaki_dgemm(A, B, b, d)
    assert ov == 0
    assert np.array_equal(C, A @ B)


def test_kchunk_is_transparent():
    rng = np.random.default_rng(5)
    A = gen_inputs(rng, 12, 700, 1.0)
    B = gen_inputs(rng, 700, 9, 1.0)
    C1, _ = oracle.ozaki_dgemm(A, B, 7, 8)
    C2, _ = oracle.ozaki_dgemm(A, B, 7, 8, kchunk=64)
    assert np.array_equal(C1.view(np.int64), C2.view(np.int64))


def exact_dot(a, b):
    return sum((Fraction(float(x)) * Fraction(float(y)) for x, y in zip(a, b)), Fraction(0))


def test_ozaki_error_bound_vs_exact():
    """|C - exact| is bounded by the dropped digits: per term at most
    2^(e_r + e_c - b d) * (d + 2), which is what the truncated split and the dropped
    triangle can lose, plus one truncation ulp of the result."""
    rng = np.random.default_rng(9)
    A = gen_inputs(rng, 6, 64, 1.0)
    B = gen_inputs(rng, 64, 5, 1.0)
    for b, d in [(7, 8), (7, 3), (3, 29)]:
        C, _ = oracle.ozaki_dgemm(A, B, b, d)
        Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
        for r in range(6):
            for c in range(5):
                ex = exact_dot(A[r], B[:, c])
                scale = Fraction(2) ** (int(Er[r]) + int(Ec[c]) - 2044 - b * d)
                bound = 64 * (d + 2) * scale * (2 ** (2 * b)) + abs(ex) * Fraction(2) ** -52
                assert abs(Fraction(float(C[r, c])) - ex) <= bound


# ---------------------------------------------------------------- MPFR oracle
@pytest.mark.skipif(not oracle.have_mpfr(), reason="MPFR not available")
def test_mpfr_oracle_vs_exact_rationals():
    rng = np.random.default_rng(4)
    A = gen_inputs(rng, 5, 40, 2.0)
    B = gen_inputs(rng, 40, 4, 2.0)
    rows = np.repeat(np.arange(5), 4)
    cols = np.tile(np.arange(4), 5)
    got = oracle.mpfr_sample(A, B, rows, cols, 106)
    for i, (r, c) in enumerate(zip(rows, cols)):
        ex = exact_dot(A[r], B[:, c])
        # 106-bit intermediate rounding: within a few ulps of 2^-106 relative to the terms
        tol = sum(abs(Fraction(float(x)) * Fraction(float(y))) for x, y in zip(A[r], B[:, c])) * Fraction(2) ** -100
        assert abs(Fraction(float(got[i])) - ex) <= tol + abs(ex) * Fraction(2) ** -53
    # with 2000 bits every dot product here is exact before the final rounding
    got_hi = oracle.mpfr_sample(A, B, rows, cols, 2000)
    for i, (r, c) in enumerate(zip(rows, cols)):
        assert float(got_hi[i]) == float(exact_dot(A[r], B[:, c]))


def test_naive_dgemm_matches_sequential_python():
    rng = np.random.default_rng(2)
    A = gen_inputs(rng, 3, 9, 1.0)
    B = gen_inputs(rng, 9, 2, 1.0)
    C = oracle.naive_dgemm(A, B)
    for r in range(3):
        for c in range(2):
            s = 0.0
            for k in range(9):
                s += float(A[r, k]) * float(B[k, c])
            assert C[r, c] == s


def test_plan_table_matches_reference_validation():
    assert oracle.plan_bd(256, 512, 128, 0)[0] == 1          # non-square rejected
    assert oracle.plan_bd(64, 64, 64, 0)[0] == 1             # K < 128 rejected
    assert oracle.plan_bd(4096, 4096, 4096, 1)[0] == 1       # K > 2048 rejected
    assert oracle.plan_bd(128, 128, 128, 0) == (0, 3, 29)
    assert oracle.plan_bd(2048, 2048, 2048, 0) == (0, 3, 26)
    assert oracle.plan_bd(512, 512, 512, 1) == (0, 7, 8)
