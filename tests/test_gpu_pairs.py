"""The paired-digit path (b <= 3 with d >= 8, the reference's accurate presets): adjacent
digits paired into one int8 digit, a triangle of pair products plus one anti-diagonal of
single-digit products (tests/test_pairs_identity.py pins the identity on the CPU).  On
the GPU its C must equal, bit for bit, the oracle's restatement of the reference triangle
and the bin-group kernel that runs the triangle itself, for every accurate preset, odd and
even d, several contraction chunks, transposed operands, ragged shapes and both roundings."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz
from conftest import gen_inputs

pytestmark = pytest.mark.gpu


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def bits(x):
    return np.ascontiguousarray(x).view(np.int64)


def run(M, N, K, b, d, A, B, rounding=oz.ROUND_TRUNCATE, transA=False, transB=False):
    import torch
    p = oz.Plan(M, N, K, b, d, rounding)
    info = p.info()
    C = p(dev(A.T) if transA else dev(A), dev(B.T) if transB else dev(B), transA=transA, transB=transB)
    torch.cuda.synchronize()
    p.close()
    return C.cpu().numpy(), info


CASES = [(128, 3, 29), (256, 3, 28), (512, 3, 27), (1024, 3, 27), (2048, 3, 26), (4096, 3, 25), (8192, 3, 25),
         (200, 3, 8), (200, 3, 9), (300, 2, 16), (150, 2, 13), (100, 3, 16)]


@pytest.mark.parametrize("K,b,d", CASES)
def test_pairs_C_bit_exact_vs_oracle(gpu, K, b, d):
    rng = np.random.default_rng(K + 31 * d)
    M, N = 97, 130
    print(f"pairs b={b} d={d} {M}x{N}x{K}", flush=True)
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    A[3, :5] = [0.0, -0.0, 5e-324, -1e-310, 1.0]
    got, info = run(M, N, K, b, d, A, B)
    assert info["kernel"] == "pairs"
    ref, ov = oracle.ozaki_dgemm(A, B, b, d)
    assert np.array_equal(bits(got), bits(ref))


def test_pairs_equal_bin_group_kernel_and_config1(gpu):
    """The reference example's shape (256 x 512 x 128, phi = 0.5, accurate: b = 3, d = 29)
    through both kernels."""
    import torch
    rng = np.random.default_rng(12445)
    M, N, K = 256, 512, 128
    A = gen_inputs(rng, M, K, 0.5)
    B = gen_inputs(rng, K, N, 0.5)
    got, info = run(M, N, K, 3, 29, A, B)
    assert info["kernel"] == "pairs" and info["passes"] == 1
    oz.set_option(oz.OPT_KERNEL, 1)
    try:
        grp, ginfo = run(M, N, K, 3, 29, A, B)
    finally:
        oz.set_option(oz.OPT_KERNEL, 0)
    assert ginfo["kernel"] == "group" and ginfo["passes"] == 4
    assert np.array_equal(bits(got), bits(grp))
    # the reference API (host pointers, Mode.accurate) takes the paired path too
    assert oz.ozAIE_setup(M, N, K, oz.Mode.accurate) == oz.OK
    assert oz.plan_info()["kernel"] == "pairs"
    C = np.zeros((M, N))
    assert oz.ozAIE_DGEMM(M, N, K, A, B, C) == oz.OK
    oz.release()
    assert np.array_equal(bits(C), bits(got))


@pytest.mark.parametrize("n,b,d", [(2048, 3, 27), (4096, 3, 25), (2048, 3, 9), (2048, 3, 16)])
def test_pairs_large_grids_match_bin_groups_and_oracle(gpu, n, b, d):
    """Grids large enough for the anti-diagonal GEMM's 128 and 256 square workgroup tiles and
    the fused kernel's default wave tiles (4 and 8 pair planes: 64 x 64 and 32 x 64; small
    grids run 64 square and 32 x 32 tiles): C equal, bit for bit, to the bin-group kernel's
    over the whole of C, and to the oracle on a 64 x 64 block (the split scales each row of A
    and each column of B on its own, so the block's rows and columns alone give its C)."""
    K = 256
    rng = np.random.default_rng(n + 7 * d)
    print(f"pairs large grid b={b} d={d} {n}x{n}x{K}", flush=True)
    A = gen_inputs(rng, n, K, 1.0)
    B = gen_inputs(rng, K, n, 1.0)
    got, info = run(n, n, K, b, d, A, B)
    assert info["kernel"] == "pairs"
    oz.set_option(oz.OPT_KERNEL, 1)
    try:
        grp, ginfo = run(n, n, K, b, d, A, B)
    finally:
        oz.set_option(oz.OPT_KERNEL, 0)
    assert ginfo["kernel"] == "group"
    assert np.array_equal(bits(got), bits(grp))
    r0, c0 = (int(x) for x in rng.integers(0, n - 64, 2))
    ref, _ = oracle.ozaki_dgemm(A[r0:r0 + 64], B[:, c0:c0 + 64], b, d)
    assert np.array_equal(bits(got[r0:r0 + 64, c0:c0 + 64]), bits(ref))


@pytest.mark.parametrize("b,d", [(3, 29), (3, 26)])
def test_pairs_several_contraction_chunks(gpu, b, d):
    rng = np.random.default_rng(40 + d)
    M, N, K = 70, 90, 1000
    A = gen_inputs(rng, M, K, 2.0)
    B = gen_inputs(rng, K, N, 2.0)
    oz.set_option(oz.OPT_KCHUNK, 256)
    try:
        got, info = run(M, N, K, b, d, A, B)
    finally:
        oz.set_option(oz.OPT_KCHUNK, 0)
    assert info["kernel"] == "pairs" and info["passes"] == 4
    ref, _ = oracle.ozaki_dgemm(A, B, b, d)
    assert np.array_equal(bits(got), bits(ref))


def test_pairs_transposed_and_nearest(gpu):
    rng = np.random.default_rng(41)
    M, N, K = 65, 300, 512
    A = gen_inputs(rng, M, K, 1.0)
    B = gen_inputs(rng, K, N, 1.0)
    ref, _ = oracle.ozaki_dgemm(A, B, 3, 27)
    for ta in (False, True):
        for tb in (False, True):
            got, _ = run(M, N, K, 3, 27, A, B, transA=ta, transB=tb)
            assert np.array_equal(bits(got), bits(ref)), (ta, tb)
    ref_n, _ = oracle.ozaki_dgemm(A, B, 3, 28, rounding=1)
    got_n, _ = run(M, N, K, 3, 28, A, B, rounding=oz.ROUND_NEAREST)
    assert np.array_equal(bits(got_n), bits(ref_n))


def test_pairs_noalloc(gpu):
    import torch
    rng = np.random.default_rng(42)
    for M, N, K in [(300, 200, 256), (64, 33, 128)]:
        A = gen_inputs(rng, M, K, 1.0)
        B = gen_inputs(rng, K, N, 1.0)
        C = torch.empty(M, N, dtype=torch.float64, device="cuda")
        oz.dgemm_noalloc(dev(A), dev(B), C, 3, 29)
        torch.cuda.synchronize()
        ref, _ = oracle.ozaki_dgemm(A, B, 3, 29)
        assert np.array_equal(bits(C.cpu().numpy()), bits(ref))
