"""Pins for the oracle's CRT mode (oracle/crt_oracle.c).

The mode has no reference counterpart, so every pin is mathematics checked with Python
integers and exact rationals: the moduli and reconstruction constants, the bit-count
bound, the integer scaling q(x), the residues, the residue products, the CRT identity,
and end-to-end C against the exact integer product of the scaled operands."""
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


def _exact_C(A, B, n, rounding):
    """C from the exact integer product of the scaled operands, by Python integers."""
    beta = oracle.crt_beta(n, A.shape[1])
    Er = [max(_field(v) for v in row) for row in A]
    Ec = [max(_field(v) for v in col) for col in B.T]
    C = np.zeros((A.shape[0], B.shape[1]))
    for r in range(A.shape[0]):
        qa = [_q_exact(v, Er[r], beta) for v in A[r]]
        for c in range(B.shape[1]):
            qb = [_q_exact(v, Ec[c], beta) for v in B[:, c]]
            Cp = sum(x * y for x, y in zip(qa, qb))
            C[r, c] = oracle.crt_to_double(Cp, Er[r] + Ec[c] - 2044 - 2 * beta, rounding)
    return C


@pytest.mark.parametrize("n,phi,shape", [(16, 1.0, (6, 50, 5)), (16, 4.0, (5, 129, 4)), (9, 0.5, (4, 64, 6)),
                                         (3, 1.0, (3, 8, 3)), (17, 2.0, (4, 70, 5)), (20, 1.0, (3, 40, 4))])
def test_crt_dgemm_equals_exact_integer_product(n, phi, shape):
    M_, K, N_ = shape
    rng = np.random.default_rng(int(phi * 10) + n)
    A = (rng.random((M_, K)) - 0.5) * np.exp(phi * rng.standard_normal((M_, K)))
    B = (rng.random((K, N_)) - 0.5) * np.exp(phi * rng.standard_normal((K, N_)))
    for rounding in (0, 1):
        C = oracle.crt_dgemm(A, B, n, rounding)
        ref = _exact_C(A, B, n, rounding)
        assert np.array_equal(C.view(np.int64), ref.view(np.int64))


def test_conversion_rounding_closed_form():
    rnd = random.Random(5)
    for _ in range(600):
        v = rnd.randrange(-2 ** 126, 2 ** 126) if _ % 2 else rnd.randrange(-2 ** 190, 2 ** 190)
        e2 = rnd.randrange(-1100, -900)
        for rounding in (0, 1):
            got = oracle.crt_to_double(v, e2, rounding)
            exact = Fraction(v) * Fraction(2) ** e2
            if rounding == 1:
                assert got == float(exact)                       # Fraction -> float rounds to nearest-even
            else:
                a = abs(v)
                sh = max(0, a.bit_length() - 53)
                t = (a >> sh) << sh
                assert got == float(Fraction(t if v >= 0 else -t) * Fraction(2) ** e2)


def test_crt_error_bound_vs_exact_product():
    """|C - A B| <= sum_k (|a| d_b + |b| d_a + d_a d_b) + one rounding, d = 2^(E - 1022 - beta) / 2:
    each scaled operand is within half a grid step of the exact one.  (The rounding of each
    scaled element itself is pinned by test_q_closed_form_and_bound.)"""
    rng = np.random.default_rng(6)
    A = (rng.random((4, 300)) - 0.5) * np.exp(rng.standard_normal((4, 300)))
    B = (rng.random((300, 3)) - 0.5) * np.exp(rng.standard_normal((300, 3)))
    n = 16
    beta = oracle.crt_beta(n, 300)
    C = oracle.crt_dgemm(A, B, n, 1)
    for r in range(4):
        Er = max(_field(v) for v in A[r])
        da = Fraction(2) ** (Er - 1022 - beta) / 2
        for c in range(3):
            Ec = max(_field(v) for v in B[:, c])
            db = Fraction(2) ** (Ec - 1022 - beta) / 2
            exact = sum(Fraction(float(x)) * Fraction(float(y)) for x, y in zip(A[r], B[:, c]))
            bound = sum(abs(Fraction(float(x))) * db + abs(Fraction(float(y))) * da + da * db
                        for x, y in zip(A[r], B[:, c]))
            bound += abs(exact) * Fraction(2) ** -52
            assert abs(Fraction(float(C[r, c])) - exact) <= bound


@pytest.mark.parametrize("n", [2, 9, 16, 17, 20])
def test_rebuild_from_offset_symmetric_residues(n):
    """The GPU stores each residue product as a signed byte r (r = g mod p, -128 <= r <= 127,
    128 kept as -128 for p = 256) and rebuilds from u = r + 128: because sum_j W_j = 1 (mod M),
    C' = sum u W - nearest(sum u f) M - 128 for every |C'| <= M / 4."""
    ps = oracle.crt_moduli(n)
    M = math.prod(ps)
    W, f, M2 = oracle.crt_consts(n)
    assert M2 == M and sum(W) % M == 1
    rnd = random.Random(100 + n)
    vals = [0, 1, -1, M // 4, -(M // 4), 128, -128] + [rnd.randrange(-(M // 4), M // 4 + 1) for _ in range(400)]
    for v in vals:
        u = []
        for p in ps:
            r = v % p
            r = r - p if r > p // 2 else r            # symmetric, in [-(p-1)/2, (p-1)/2] or [-128, 128]
            r = ((r & 0xFF) ^ 0x80) - 128             # stored as one byte, read back signed
            assert (r - v) % p == 0
            u.append(r + 128)
        t = 0.0
        for uj, fj in zip(u, f):
            t += float(uj) * fj
        k = round(t)
        R = 2 ** (128 if n <= 16 else 192)
        S = (sum(uj * wj for uj, wj in zip(u, W)) - k * M - 128) % R
        assert (S - R if S >= R // 2 else S) == v


@pytest.mark.parametrize("n", [2, 5, 9, 16, 17, 20])
def test_rebuild_byte_limbs_and_fixed_point_quotient(n):
    """The GPU rebuild's arithmetic against the exact CRT inverse: u = r + 128 per modulus,
    L_i = sum_j u_j w_ji over the byte limbs w_ji of W_j, the quotient
    k = floor((sum_j u_j q_j + 2^15) / 2^16) with q_j = floor(2^16 fl(W_j / M)) (< 2^16), and
    C' = (sum_i (L_i + k m_i) 256^i - 128) mod 2^(8 nl) with m_i the bytes of 2^(8 nl) - M:
    nl = 16 limbs (128 bits) for n <= 16 moduli, 24 (192 bits) above."""
    ps = oracle.crt_moduli(n)
    W, f, M = oracle.crt_consts(n)
    nl = 16 if n <= 16 else 24
    wb = [[(w >> (8 * i)) & 0xFF for i in range(nl)] for w in W]
    q = [int(math.floor(fj * 65536.0)) for fj in f]
    assert all(0 <= qj < 65536 for qj in q)
    mneg = [((2 ** (8 * nl) - M) >> (8 * i)) & 0xFF for i in range(nl)]
    rnd = random.Random(200 + n)
    vals = [0, 1, -1, M // 4, -(M // 4), 127, -128] + [rnd.randrange(-(M // 4), M // 4 + 1) for _ in range(400)]
    for v in vals:
        u = []
        for p in ps:
            r = v % p
            r = r - p if r > p // 2 else r
            u.append((((r & 0xFF) ^ 0x80) - 128) + 128)
        t = sum(uj * qj for uj, qj in zip(u, q))
        assert t < 2 ** 29
        k = (t + 32768) >> 16
        L = [sum(u[j] * wb[j][i] for j in range(n)) + k * mneg[i] for i in range(nl)]
        assert max(L) < 2 ** 22
        R = 2 ** (8 * nl)
        S = (sum(Li << (8 * i) for i, Li in enumerate(L)) - 128) % R
        assert (S - R if S >= R // 2 else S) == v
