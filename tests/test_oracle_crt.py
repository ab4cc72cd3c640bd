"""Pins for the oracle's CRT mode (oracle/crt_oracle.c).

The mode has no reference counterpart, so every pin is mathematics checked with Python
integers and exact rationals: the moduli and reconstruction constants, the bit-count
bound of the max rule, the norm rule's bound on each row's Euclidean norm and its shift
(largest by its own integer condition, and every row and column of the scaled operands
within sqrt(M / 4)), the integer scaling q(x), the residues, the residue products, the
CRT identity, and end-to-end C against the exact integer product of the scaled operands
under both rules."""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import oracle


def _q_exact(x, E, beta):
    """sign(x) * round(|x| * 2^(1022 + beta - E)) with exact rationals, to nearest with ties
    away from zero (floor(v + 1/2) of the magnitude)."""
    if x == 0:
        return 0
    v = Fraction(abs(float(x))) * Fraction(2) ** (1022 + beta - E) + Fraction(1, 2)
    q = v.numerator // v.denominator
    return -q if x < 0 else q


def _field(x):
    return int((np.float64(x).view(np.uint64) >> np.uint64(52)) & np.uint64(0x7FF))


def _sym(q, p):
    r = q % p
    half = 127 if p == 256 else (p - 1) // 2
    return r - p if r > half else r


def test_moduli_and_constants():
    ps = oracle.crt_moduli(20)
    assert len(set(ps)) == 20 and all(2 <= p <= 256 for p in ps)
    for i in range(20):
        for j in range(i):
            assert math.gcd(ps[i], ps[j]) == 1
    assert oracle.crt_moduli(21)[20] == 0
    for n in (2, 5, 16, 17, 20):
        W, f, M = oracle.crt_consts(n)
        assert M == math.prod(ps[:n])
        for i, w in enumerate(W):
            assert 0 <= w < M
            for j in range(n):
                assert w % ps[j] == (1 if i == j else 0)
            assert abs(f[i] - w / M) <= 4 * 2.0 ** -53


@pytest.mark.parametrize("n", [2, 3, 8, 12, 16, 17, 20])
def test_beta_is_the_largest_safe_bit_count(n):
    M = math.prod(oracle.crt_moduli(n))
    L = M.bit_length() - 1
    for K in (1, 2, 3, 100, 128, 129, 4096, 8192, 65536):
        beta = oracle.crt_beta(n, K)
        if beta < 0:
            assert K * 4 ** 1 > 2 ** (L - 2)
            continue
        assert K * 2 ** (2 * beta) <= M // 4                       # exact rebuild
        lk = (K - 1).bit_length()
        if beta < 62:
            assert 2 * (beta + 1) + lk > L - 2                       # largest by the rule
        assert beta <= 62


def test_q_closed_form_and_bound():
    rng = np.random.default_rng(3)
    xs = list((rng.random(400) - 0.5) * np.exp(8 * rng.standard_normal(400)))
    xs += [0.0, -0.0, 5e-324, -5e-324, 2.2250738585072014e-308, 1.0, -1.0, np.nextafter(1.0, 2.0),
           1.7976931348623157e308, 3e-310]
    for beta in (7, 24, 55, 62):
        for x in xs:
            for E in {_field(x), _field(x) + 1, _field(x) + 30, max(_field(x), 1)}:
                if E > 2047 or E < _field(x):
                    continue
                q = oracle.crt_q(x, E, beta)
                assert q == _q_exact(x, E, beta), (x, E, beta)
                assert abs(q) <= 2 ** beta


def test_split_residues_and_products():
    rng = np.random.default_rng(4)
    A = (rng.random((9, 37)) - 0.5) * np.exp(2 * rng.standard_normal((9, 37)))
    B = (rng.random((37, 11)) - 0.5) * np.exp(2 * rng.standard_normal((37, 11)))
    A[0, 3] = 0.0
    B[5, 2] = -0.0
    n, beta = 16, oracle.crt_beta(16, 37)
    Er = np.array([max(_field(v) for v in row) for row in A], dtype=np.int16)
    Ec = np.array([max(_field(v) for v in col) for col in B.T], dtype=np.int16)
    SA = oracle.crt_split(A, Er, beta, n, axis_contract=1)
    SB = oracle.crt_split(B, Ec, beta, n, axis_contract=0)
    ps = oracle.crt_moduli(n)
    for i, p in enumerate(ps):
        for r in range(A.shape[0]):
            for k in range(A.shape[1]):
                assert SA[i, r, k] == _sym(_q_exact(A[r, k], int(Er[r]), beta), p)
        for c in range(B.shape[1]):
            for k in range(B.shape[0]):
                assert SB[i, c, k] == _sym(_q_exact(B[k, c], int(Ec[c]), beta), p)
    G = oracle.crt_products(SA, SB)
    for i, p in enumerate(ps):
        ref = (SA[i].astype(np.int64) @ SB[i].astype(np.int64).T) % p
        assert np.array_equal(G[i].astype(np.int64), ref)


@pytest.mark.parametrize("n", [2, 7, 16, 17, 20])
def test_rebuild_is_the_crt_inverse(n):
    ps = oracle.crt_moduli(n)
    M = math.prod(ps)
    rnd = random.Random(n)
    vals = [0, 1, -1, M // 4, -(M // 4)] + [rnd.randrange(-(M // 4), M // 4 + 1) for _ in range(300)]
    for v in vals:
        g = np.array([v % p for p in ps], dtype=np.uint8)
        assert oracle.crt_rebuild(g) == v


def _q_sigma_exact(x, sigma):
    """sign(x) * round(|x| * 2^(1075 - sigma)), ties away from zero, exact rationals."""
    if x == 0:
        return 0
    v = Fraction(abs(float(x))) * Fraction(2) ** (1075 - sigma) + Fraction(1, 2)
    q = v.numerator // v.denominator
    return -q if x < 0 else q


def _sigmas(A, B, n, bits):
    """Shifts of the rows of A and the columns of B: the max rule from the exponent fields
    (sigma = 53 + E - beta), the norm rule from the oracle (pinned by the tests below)."""
    if bits:
        return ([53 + max(_field(v) for v in row) - bits for row in A],
                [53 + max(_field(v) for v in col) - bits for col in B.T])
    return [int(v) for v in oracle.crt_sigma(A, n, 1)], [int(v) for v in oracle.crt_sigma(B, n, 0)]


def _exact_C(A, B, n, rounding, bits):
    """C from the exact integer product of the scaled operands, by Python integers."""
    sr, sc = _sigmas(A, B, n, bits)
    C = np.zeros((A.shape[0], B.shape[1]))
    for r in range(A.shape[0]):
        qa = [_q_sigma_exact(v, sr[r]) for v in A[r]]
        for c in range(B.shape[1]):
            qb = [_q_sigma_exact(v, sc[c]) for v in B[:, c]]
            Cp = sum(x * y for x, y in zip(qa, qb))
            C[r, c] = oracle.crt_to_double(Cp, sr[r] + sc[c] - 2150, rounding)
    return C


@pytest.mark.parametrize("rule", ["norm", "max"])
@pytest.mark.parametrize("n,phi,shape", [(16, 1.0, (6, 50, 5)), (16, 4.0, (5, 129, 4)), (9, 0.5, (4, 64, 6)),
                                         (3, 1.0, (3, 8, 3)), (17, 2.0, (4, 70, 5)), (20, 1.0, (3, 40, 4))])
def test_crt_dgemm_equals_exact_integer_product(n, phi, shape, rule):
    M_, K, N_ = shape
    rng = np.random.default_rng(int(phi * 10) + n)
    A = (rng.random((M_, K)) - 0.5) * np.exp(phi * rng.standard_normal((M_, K)))
    B = (rng.random((K, N_)) - 0.5) * np.exp(phi * rng.standard_normal((K, N_)))
    bits = 0 if rule == "norm" else oracle.crt_beta(n, K)
    for rounding in (0, 1):
        C = oracle.crt_dgemm(A, B, n, rounding, bits)
        ref = _exact_C(A, B, n, rounding, bits)
        assert np.array_equal(C.view(np.int64), ref.view(np.int64))


def _norm_rows(rng):
    """Rows with the cases the chunked bound must handle: wide exponent ranges, zeros,
    subnormals, a lone huge element, a chunk boundary inside the row, a row of zeros."""
    rows = []
    for phi, K in ((0.1, 64), (1.0, 300), (4.0, 257), (20.0, 90), (1.0, 16), (1.0, 1), (2.0, 1000)):
        rows.append((rng.random(K) - 0.5) * np.exp(phi * rng.standard_normal(K)))
    r = (rng.random(200) - 0.5)
    r[:30] = 5e-324 * rng.integers(-400, 400, size=30)
    r[100] = 1.5e300
    rows.append(r)
    rows.append(5e-324 * rng.integers(-9, 9, size=70).astype(np.float64))
    rows.append(np.zeros(40))
    return rows


def test_norm_bound_valid_and_tight():
    """||x||^2 <= S 4^(E^ - 1038) (valid), and S U^2 <= ||x||^2 + 2 U sum|x| + 17 nch U^2 with
    U = 2^(E^ - 1038) and nch the number of 16-index chunks (each 16-bit ceiling adds at
    most 2 |x| U_c + U_c^2, U_c <= U, and each chunk's rescaling one unit U^2);
    E^ = max over the nonzero elements of max(e, 1)."""
    rng = np.random.default_rng(11)
    for x in _norm_rows(rng):
        Eh, S = oracle.crt_normbound(x[None, :], 1)
        Eh, S = int(Eh[0]), int(S[0])
        nz = [v for v in x if v != 0]
        if not nz:
            assert S == 0
            continue
        assert Eh == max(max(_field(v), 1) for v in nz)
        U = Fraction(2) ** (Eh - 1038)
        nrm = sum(Fraction(float(v)) ** 2 for v in x)
        absum = sum(abs(Fraction(float(v))) for v in x)
        nch = (len(x) + 15) // 16
        assert nrm <= S * U * U
        assert S * U * U <= nrm + 2 * U * absum + 17 * nch * U * U


@pytest.mark.parametrize("n", [2, 5, 9, 16, 17, 20])
def test_norm_m_is_the_largest_by_its_condition(n):
    """m = the largest integer <= 46 with S (2^22 + 4) 4^m <= M 2^20 - K (2^20 + 2^40)."""
    M = math.prod(oracle.crt_moduli(n))
    rnd = random.Random(n)
    for K in (1, 16, 300, 8192, 65536):
        R = M * 2 ** 20 - K * (2 ** 20 + 2 ** 40)
        if R <= 0:
            assert oracle.crt_norm_m(1, n, K) == -1000
            continue
        for S in [1, 2, 3, 2 ** 32, 2 ** 36, 2 ** 48 - 1] + [rnd.randrange(1, 2 ** 48) for _ in range(40)]:
            m = oracle.crt_norm_m(S, n, K)
            fits = lambda mm: (S * (2 ** 22 + 4) * 4 ** mm <= R) if mm >= 0 else (S * (2 ** 22 + 4) <= R * 4 ** -mm)
            assert m <= 46 and fits(m)
            assert m == 46 or not fits(m + 1)


@pytest.mark.parametrize("n,K", [(16, 8192), (16, 300), (9, 100), (17, 5000), (20, 64), (4, 20)])
def test_norm_rule_bounds_every_row_and_column(n, K):
    """With the norm rule's shifts, every row of the scaled integers has ||q||^2 <= M / 4 (so
    |C'| <= ||q_a|| ||q_b|| <= M / 4 by Cauchy-Schwarz, the rebuild's range), every
    |q| <= 2^62, and q is the closed form round(|x| 2^(1075 - sigma))."""
    rng = np.random.default_rng(K + n)
    M = math.prod(oracle.crt_moduli(n))
    rows = [(rng.random(K) - 0.5) * np.exp(phi * rng.standard_normal(K)) for phi in (0.1, 1.0, 4.0)]
    spike = np.full(K, 1e-3)
    spike[K // 2] = 1e5                       # one element carries the norm
    flat = np.full(K, -0.75)                  # every element at the row maximum
    rows += [spike, flat]
    X = np.array(rows)
    sg = oracle.crt_sigma(X, n, 1, K=K)
    for x, s in zip(X, sg):
        q = [oracle.crt_q_sigma(v, int(s)) for v in x]
        idx = rng.integers(0, K, size=min(K, 64))
        for i in idx:
            assert q[i] == _q_sigma_exact(x[i], int(s))
        assert max(abs(v) for v in q) <= 2 ** 62
        assert 4 * sum(v * v for v in q) <= M


def test_few_moduli_fall_back_to_the_max_rule():
    """Where M 2^20 <= K (2^20 + 2^40) the norm rule has no room and bits 0 means the max rule
    with the largest exact bit count."""
    rng = np.random.default_rng(8)
    A = rng.random((3, 20)) - 0.5
    B = rng.random((20, 4)) - 0.5
    assert oracle.crt_norm_m(1, 3, 20) == -1000 and oracle.crt_beta(3, 20) >= 1
    for rounding in (0, 1):
        C0 = oracle.crt_dgemm(A, B, 3, rounding, 0)
        Cm = oracle.crt_dgemm(A, B, 3, rounding, oracle.crt_beta(3, 20))
        assert np.array_equal(C0.view(np.int64), Cm.view(np.int64))


def test_norm_rule_gains_bits_on_wide_rows():
    """On (U - 1/2) exp(N(0, 1)) rows of K = 8192 the norm rule's scale exceeds the max
    rule's by log2(max |x| sqrt(K) / ||x||) - 1/2 to within the two rules' floors (about 4
    bits here); a row whose norm one element carries gains log2 sqrt(K), and on a flat row
    (every element at the maximum, the norm rule's worst case) it is at most one bit below."""
    rng = np.random.default_rng(21)
    K, n = 8192, 16
    X = (rng.random((7, K)) - 0.5) * np.exp(rng.standard_normal((7, K)))
    X[5] = -0.75
    X[6] = 1e-9
    X[6, 7] = 3.0
    beta = oracle.crt_beta(n, K)
    smax = np.array([53 + max(_field(v) for v in row) - beta for row in X])
    gain = smax - oracle.crt_sigma(X, n, 1)                # bits the norm rule adds
    pred = [math.log2(np.abs(row).max() * math.sqrt(K) / np.linalg.norm(row)) - 0.5 for row in X]
    for g, p in zip(gain[:5], pred[:5]):
        assert abs(g - p) <= 1.5 and g >= 3
    assert gain[5] >= -1
    assert gain[6] >= 6
