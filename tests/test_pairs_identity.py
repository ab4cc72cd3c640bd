"""The paired-digit decomposition the library runs for b <= 3 (the reference's accurate
presets) gives the same exact integer T as the reference's triangle of slice products
(src/aie_kernels/aie2/mm_ozaki.cc:304-357, summed as in ozaki_base.cc:143-340).

Adjacent digits s_2I, s_2I+1 (s_0 most significant, a zero digit appended when d is odd)
become one digit S_I = s_2I 2^b + s_2I+1 below 2^(2b) <= 64.  With H = (d-1)/2 for odd d
and (d-2)/2 for even d, P = H (odd) or H + 1 (even) pair planes and SB_k =
sum_{I+J=k} S_I T_J:
    d odd:   T       = sum_{k<P} SB_k 2^(2b(P-1-k)) + sum_{I+J=H} s_2I t_2J
    d even:  T * 2^b = sum_{k<P} SB_k 2^(2b(P-1-k)) - sum_{I+J=H} s_2I+1 t_2J+1
The check takes the oracle's own digits and its own exact T (both pinned independently in
test_oracle.py), so a wrong pair boundary, weight or sign fails here, on the CPU."""
import numpy as np
import pytest

import oracle
from conftest import gen_inputs


def T_int(Tpair):
    """Python ints from the oracle's (hi, lo) int64 pairs."""
    hi = Tpair[..., 0].astype(object)
    lo = Tpair[..., 1].astype(np.uint64).astype(object)
    return hi * (1 << 64) + lo


def pair_T(SA, SB, b, d):
    """T (times 2^b for even d) from the paired decomposition, in Python integers."""
    SA = SA.astype(np.int64)
    SB = SB.astype(np.int64)
    odd = d % 2
    if odd:
        SA = np.concatenate([SA, np.zeros_like(SA[:1])])
        SB = np.concatenate([SB, np.zeros_like(SB[:1])])
    D = SA.shape[0] // 2
    H = (d - 1) // 2 if odd else (d - 2) // 2
    P = H if odd else H + 1
    PA = SA[0::2] * (1 << b) + SA[1::2]     # pair digits S_I, |S_I| < 2^(2b)
    PB = SB[0::2] * (1 << b) + SB[1::2]
    assert np.abs(PA).max() < (1 << (2 * b)) and np.abs(PB).max() < (1 << (2 * b))
    M, N = SA.shape[1], SB.shape[1]
    T = np.zeros((M, N), dtype=object)
    for k in range(P):
        sb = sum(PA[i] @ PB[k - i].T for i in range(k + 1)).astype(object)
        T = T + sb * (1 << (2 * b * (P - 1 - k)))
    if odd:   # high digits of the pairs on the anti-diagonal I + J = H
        x = sum(SA[2 * i] @ SB[2 * (H - i)].T for i in range(H + 1)).astype(object)
        return T + x
    z = sum(SA[2 * i + 1] @ SB[2 * (H - i) + 1].T for i in range(H + 1)).astype(object)
    return T - z


@pytest.mark.parametrize("b,d,K", [(3, 29, 128), (3, 28, 256), (3, 27, 512), (3, 26, 2048), (3, 25, 4096),
                                   (3, 8, 64), (3, 9, 64), (2, 16, 96), (2, 13, 96), (1, 30, 64)])
def test_pair_decomposition_equals_triangle(b, d, K):
    rng = np.random.default_rng(b * 100 + d)
    M, N = 7, 5
    A = gen_inputs(rng, M, K, 2.0)
    B = gen_inputs(rng, K, N, 2.0)
    A[0, :3] = [0.0, -0.0, 5e-324]     # zeros and a subnormal
    B[:2, 1] = [1.0, -1.0]
    Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
    SA = oracle.split(A, Er, b, d, 1)
    SB = oracle.split(B, Ec, b, d, 0)
    T_ref = T_int(oracle.bins_to_T(oracle.bins(SA, SB), b, d))
    got = pair_T(SA, SB, b, d)
    if d % 2 == 0:
        assert all(int(v) % (1 << b) == 0 for v in got.ravel())
        got = np.array([int(v) >> b for v in got.ravel()], dtype=object).reshape(got.shape)
    assert np.array_equal(got, T_ref)


def test_pair_term_bounds_fit_int32():
    """The planner's int32 bounds for the pair path: the anti-diagonal over all of K and each
    pair bin over one contraction chunk stay below 2^31 at the accurate presets' largest K."""
    for lk, (b, d) in zip(range(7, 14), [(3, 29), (3, 28), (3, 27), (3, 27), (3, 26), (3, 25), (3, 25)]):
        K = 1 << lk
        H = (d - 1) // 2 if d % 2 else (d - 2) // 2
        P = H if d % 2 else H + 1
        assert (H + 1) * (2 ** b - 1) ** 2 * K < 2 ** 31
        assert P * (2 ** (2 * b) - 1) ** 2 * min(K, 2 ** 31 // (P * (2 ** (2 * b) - 1) ** 2)) < 2 ** 31
