"""ozaie_dgemm_noalloc (the allocation-free entry point) against ozaie_plan_create +
ozaie_dgemm, the entry point it is a variant of: bit-identical C on two calls in a row, the
second with a smaller product, for the fused b = 7 path, the reference's accurate preset
(several passes through 128-bit partial sums) and the CRT mode; and a product too large for
the static workspace is rejected with a clear error."""
import numpy as np
import pytest

import ozaie_amd as oz
from conftest import gen_inputs


def _pair(torch, rng, M, N, K, phi=1.0):
    A = torch.from_numpy(gen_inputs(rng, M, K, phi)).cuda()
    B = torch.from_numpy(gen_inputs(rng, K, N, phi)).cuda()
    return A, B


def _ref(torch, A, B, bits, slices, rounding, crt, transA=False, transB=False):
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    N = B.shape[0] if transB else B.shape[1]
    p = oz.Plan(M, N, K, bits, slices, rounding, crt=crt)
    C = p(A, B, transA=transA, transB=transB)
    torch.cuda.synchronize()
    p.close()
    return C


CASES = [
    # (bits, slices, rounding, crt, first shape, second (smaller) shape)
    (7, 8, oz.ROUND_TRUNCATE, False, (512, 384, 640), (200, 130, 300)),
    (3, 28, oz.ROUND_TRUNCATE, False, (300, 260, 256), (100, 90, 128)),
    (0, 16, oz.ROUND_NEAREST, True, (400, 300, 500), (128, 64, 96)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("bits,slices,rounding,crt,s1,s2", CASES)
def test_noalloc_matches_plan_two_calls(gpu, bits, slices, rounding, crt, s1, s2):
    import torch
    rng = np.random.default_rng(12445 + slices)
    for M, N, K in (s1, s2):
        print(f"noalloc b={bits} d={slices} crt={crt} {M}x{N}x{K}", flush=True)
        A, B = _pair(torch, rng, M, N, K)
        C = torch.full((M, N), float("nan"), dtype=torch.float64, device="cuda")
        oz.dgemm_noalloc(A, B, C, bits, slices, rounding, crt)
        torch.cuda.synchronize()
        ref = _ref(torch, A, B, bits, slices, rounding, crt)
        assert torch.equal(C.view(torch.int64), ref.view(torch.int64)), (M, N, K)


@pytest.mark.gpu
def test_noalloc_transposed_operands(gpu):
    import torch
    rng = np.random.default_rng(7)
    M, N, K = 192, 160, 224
    At = torch.from_numpy(gen_inputs(rng, K, M, 1.0)).cuda()   # stored K x M
    Bt = torch.from_numpy(gen_inputs(rng, N, K, 1.0)).cuda()   # stored N x K
    C = torch.empty(M, N, dtype=torch.float64, device="cuda")
    print("noalloc transposed 192x160x224", flush=True)
    oz.dgemm_noalloc(At, Bt, C, 7, 9, transA=True, transB=True)
    torch.cuda.synchronize()
    ref = _ref(torch, At, Bt, 7, 9, oz.ROUND_TRUNCATE, False, transA=True, transB=True)
    assert torch.equal(C.view(torch.int64), ref.view(torch.int64))


@pytest.mark.gpu
def test_noalloc_rejects_products_beyond_the_workspace(gpu):
    import torch
    # 8192^3 with 8 slices needs 2 x 8 x 8192 x 8192 bytes of slices: beyond 256 MiB; the
    # check happens before any pointer is touched, so small placeholder tensors suffice
    A = torch.zeros(1, 1, dtype=torch.float64, device="cuda")
    rc = oz.lib().ozaie_dgemm_noalloc(8192, 8192, 8192, 7, 8, 0, 0, 0, oz._ptr(A), 8192, oz._ptr(A), 8192,
                                      oz._ptr(A), 8192, None)
    assert rc == oz.ERR_NOMEM
    assert b"static workspace" in oz.lib().ozaie_last_error()
