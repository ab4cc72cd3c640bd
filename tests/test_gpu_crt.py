"""CRT mode on the GPU (through the C ABI) against the oracle's CRT restatement.

Bar: bit-exact shifts and residues of both operands under the norm rule (the default) and
the max rule (explicit bits), bit-exact residue products for every modulus, and
bit-exact FP64 C (truncating and nearest-even), including transposed operands, fewer
moduli (where the norm rule has no room, the max rule), 17-20 moduli (products beyond
2^127, rebuilt in 192 bits), explicit bit counts, the reference-style entry point with
host arrays, and the K range limits; at 1024^3 the mode's error against the exact
product is below native DGEMM's."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz
from conftest import gen_inputs

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


def bits(x):
    return np.ascontiguousarray(x).view(np.int64)


def default_bits(n, K):
    """What bits 0 selects: the norm rule (0), or for a few moduli the max rule's largest
    bit count."""
    return 0 if oracle.crt_norm_m(1, n, K) != -1000 else oracle.crt_beta(n, K)


def ref_residues(X, n, axis, bits, K):
    if bits:
        return oracle.crt_split(X, oracle.max_exp(X, axis), bits, n, axis)
    return oracle.crt_split_sigma(X, oracle.crt_sigma(X, n, axis, K=K), n, axis)


def edge_matrix(rng, m, n, phi=2.0):
    X = gen_inputs(rng, m, n, phi)
    X[0, :6] = [0.0, -0.0, 1.0, 5e-324, 2.2250738585072014e-308, -2.2250738585072009e-308]
    X[1, :] = 0.0
    X[2, 3] = 1e-300
    X[3, :] = -1.0
    X[4, :] = 3e-310            # a row of subnormals only
    X[5, 1] = np.nextafter(1.0, 2.0)
    return X


@pytest.mark.parametrize("rule", ["default", "max"])
@pytest.mark.parametrize("M,N,K,n", [(37, 45, 100, 16), (70, 33, 300, 16), (9, 260, 64, 5), (130, 7, 129, 2),
                                     (33, 40, 200, 17), (20, 21, 70, 20), (40, 36, 2000, 16)])
def test_crt_stages_bit_exact(gpu, M, N, K, n, rule):
    rng = np.random.default_rng(M * 7 + K)
    A = edge_matrix(rng, M, K)
    B = np.ascontiguousarray(edge_matrix(rng, N, K).T)
    bits = default_bits(n, K) if rule == "default" else oracle.crt_beta(n, K)
    SA, SB, G = oz.debug_crt(dev(A), dev(B), n, 0 if rule == "default" else bits)
    SAo = ref_residues(A, n, 1, bits, K)
    SBo = ref_residues(B, n, 0, bits, K)
    assert np.array_equal(SA, SAo)
    assert np.array_equal(SB, SBo)
    assert np.array_equal(G, oracle.crt_products(SAo, SBo))


@pytest.mark.parametrize("M,N,K,n,phi", [(64, 48, 200, 16, 1.0), (300, 260, 513, 16, 2.0), (40, 520, 64, 16, 4.0),
                                         (256, 256, 1024, 9, 1.0), (17, 19, 3000, 3, 0.5), (12, 9, 40000, 16, 2.0),
                                         (40, 30, 3000, 17, 2.0), (70, 66, 600, 18, 1.0), (9, 12, 5000, 20, 0.5)])
@pytest.mark.parametrize("rounding", [oz.ROUND_TRUNCATE, oz.ROUND_NEAREST])
def test_crt_C_bit_exact(gpu, M, N, K, n, phi, rounding):
    import torch
    rng = np.random.default_rng(M + N + K + n)
    A = gen_inputs(rng, M, K, phi)
    B = gen_inputs(rng, K, N, phi)
    A[0, 0] = 0.0
    p = oz.Plan(M, N, K, 0, n, rounding=rounding, crt=True)
    info = p.info()
    assert info["kernel"] == "crt" and info["bits"] == default_bits(n, K) and info["slices"] == n
    C = p(dev(A), dev(B))
    torch.cuda.synchronize()
    ref = oracle.crt_dgemm(A, B, n, rounding)
    assert np.array_equal(bits(C.cpu().numpy()), bits(ref))
    p.close()


@pytest.mark.parametrize("n", [14, 17])
def test_crt_modulus_gemm_forms_agree(gpu, n):
    """Every form of the modulus GEMM -- 256 x 256 tiles (OZAIE_OPT_KERNEL 2), per-wave loads
    (3), 128 x 128 tiles (4) and the size rule (0) -- gives the same C on a shape that is no
    multiple of either tile, equal to the oracle on sampled rows and columns."""
    import torch
    rng = np.random.default_rng(60 + n)
    M, N, K = 700, 900, 1000
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    dA, dB = dev(A), dev(B)
    out = {}
    for kernel in (0, 2, 3, 4):
        oz.set_option(oz.OPT_KERNEL, kernel)
        try:
            p = oz.Plan(M, N, K, 0, n, rounding=oz.ROUND_NEAREST, crt=True)
        finally:
            oz.set_option(oz.OPT_KERNEL, 0)
        out[kernel] = p(dA, dB).cpu().numpy()
        p.close()
    for kernel in (2, 3, 4):
        assert np.array_equal(bits(out[kernel]), bits(out[0])), kernel
    rows = np.array([0, 1, 255, 256, 511, M - 1])
    cols = np.array([0, 127, 128, 255, 256, 640, N - 1])
    ref = oracle.crt_dgemm(A[rows], np.ascontiguousarray(B[:, cols]), n, oz.ROUND_NEAREST)
    assert np.array_equal(bits(out[0][np.ix_(rows, cols)]), bits(ref))


@pytest.mark.parametrize("M,N,K,n", [(130, 67, 300, 16), (64, 128, 100, 16), (33, 36, 90, 20)])
def test_crt_C_into_strided_view(gpu, M, N, K, n):
    """C written through a view that starts 8 bytes past a 16-byte boundary, with a row
    stride above N: equal to the contiguous result, and nothing outside the view written
    (the rebuild stores four columns of a row with 16-byte stores only where they are
    aligned, one double at a time elsewhere)."""
    import torch
    rng = np.random.default_rng(M + 3 * N + K)
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    p = oz.Plan(M, N, K, 0, n, rounding=oz.ROUND_NEAREST, crt=True)
    ref = p(dev(A), dev(B))
    torch.cuda.synchronize()
    big = torch.full((M + 5, N + 7), float("nan"), dtype=torch.float64, device="cuda")
    for r0, c0 in ((3, 1), (2, 2)):
        big.fill_(float("nan"))
        C = big[r0:r0 + M, c0:c0 + N]
        assert C.stride(0) == N + 7
        p(dev(A), dev(B), C)
        torch.cuda.synchronize()
        assert torch.equal(C, ref), (r0, c0)
        outside = big.clone()
        outside[r0:r0 + M, c0:c0 + N] = float("nan")
        assert bool(torch.isnan(outside).all()), (r0, c0)
    assert np.array_equal(bits(ref.cpu().numpy()), bits(oracle.crt_dgemm(A, B, n, oz.ROUND_NEAREST)))
    p.close()


def test_crt_transposes_and_explicit_bits(gpu):
    import torch
    rng = np.random.default_rng(11)
    M, N, K = 90, 70, 333
    A = gen_inputs(rng, M, K, 1.5)
    B = gen_inputs(rng, K, N, 1.5)
    ref = oracle.crt_dgemm(A, B, 16, oz.ROUND_NEAREST)
    p = oz.Plan(M, N, K, 0, 16, rounding=oz.ROUND_NEAREST, crt=True)
    for ta in (False, True):
        for tb in (False, True):
            a = dev(A.T.copy()) if ta else dev(A)
            b = dev(B.T.copy()) if tb else dev(B)
            C = p(a, b, transA=ta, transB=tb)
            torch.cuda.synchronize()
            assert np.array_equal(bits(C.cpu().numpy()), bits(ref)), (ta, tb)
    p.close()
    # a smaller bit count than the largest exact one: the oracle with the same split
    beta = oracle.crt_beta(16, K) - 9
    p = oz.Plan(M, N, K, beta, 16, rounding=oz.ROUND_NEAREST, crt=True)
    C = p(dev(A), dev(B)).cpu().numpy()
    SA, SB, G = oz.debug_crt(dev(A), dev(B), 16, beta)
    assert np.array_equal(SA, oracle.crt_split(A, oracle.max_exp(A, 1), beta, 16, 1))
    assert np.abs(C - A @ B).max() <= 2.0 ** (-beta + 12) * np.abs(A).max() * np.abs(B).max() * K
    p.close()


def test_crt_reference_api_host_arrays(gpu):
    rng = np.random.default_rng(5)
    M, N, K = 128, 96, 256
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    C = np.zeros((M, N))
    assert oz.ozAIE_setup(M, N, K, oz.Mode.crt) == 0
    assert oz.ozAIE_DGEMM(M, N, K, A, B, C) == 0
    assert np.array_equal(bits(C), bits(oracle.crt_dgemm(A, B, 16, oz.get_option(oz.OPT_ROUNDING))))
    oz.release()


def test_crt_limits(gpu):
    with pytest.raises(oz.OzaieError):
        oz.Plan(64, 64, 70000, 0, 16, crt=True)          # K beyond the int32 residue-sum range
    with pytest.raises(oz.OzaieError):
        oz.Plan(64, 64, 64, 0, 21, crt=True)             # more moduli than the table
    with pytest.raises(oz.OzaieError):
        oz.Plan(64, 64, 64, 63, 16, crt=True)            # more bits than the moduli rebuild


def test_crt_error_below_native(gpu):
    import torch
    n = 1024
    rng = np.random.default_rng(12445)
    A = gen_inputs(rng, n, n, 1.0)
    B = gen_inputs(rng, n, n, 1.0)
    rows = rng.integers(0, n, 2000)
    cols = rng.integers(0, n, 2000)
    exact = oz.exact_entries(A, B, rows, cols)
    dA, dB = dev(A), dev(B)
    nat = (dA @ dB).cpu().numpy()[rows, cols]
    p = oz.Plan(n, n, n, 0, 16, rounding=oz.ROUND_NEAREST, crt=True)
    C = p(dA, dB).cpu().numpy()[rows, cols]
    rel = lambda x: np.max(np.abs(x - exact) / np.abs(exact))
    assert rel(C) <= rel(nat)
    p.close()
