"""Differentiated emulated GEMM: loss = sum(C * W) with W ~ N(0, 1).

Gradients of the emulated product are checked against the FP64 torch.matmul gradient
(dA = W B^T, dB = A^T W, computed in float64) and against central finite differences
(h = 1e-6) of the emulated forward itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def gen(torch, m, n, phi, g):
    return (torch.rand(m, n, dtype=torch.float64, device="cuda", generator=g) - 0.5) * torch.exp(
        phi * torch.randn(m, n, dtype=torch.float64, device="cuda", generator=g))


@pytest.mark.parametrize("M,N,K,crt", [(2048, 2048, 2048, False), (300, 170, 530, False),
                                       (2048, 2048, 2048, True), (300, 170, 530, True)])
def test_gradients_match_fp64_matmul(gpu, M, N, K, crt):
    import torch
    from ozaie_amd.autograd import ozaki_matmul
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    A = gen(torch, M, K, 0.5, g).requires_grad_(True)
    B = gen(torch, K, N, 0.5, g).requires_grad_(True)
    W = torch.randn(M, N, dtype=torch.float64, device="cuda", generator=g)
    C = ozaki_matmul(A, B, bits=0, slices=16, crt=True) if crt else ozaki_matmul(A, B)
    (C * W).sum().backward()
    with torch.no_grad():
        Cref = A @ B
        dA_ref = W @ B.t()
        dB_ref = A.t() @ W
    def rel(x, y):
        return ((x - y).abs().max() / y.abs().max()).item()
    assert rel(C.detach(), Cref) < 1e-13
    assert rel(A.grad, dA_ref) < 1e-13
    assert rel(B.grad, dB_ref) < 1e-13


def test_central_finite_differences(gpu):
    import torch
    from ozaie_amd.autograd import ozaki_matmul
    g = torch.Generator(device="cuda")
    g.manual_seed(6)
    M, N, K = 96, 80, 112
    A = gen(torch, M, K, 0.5, g).requires_grad_(True)
    B = gen(torch, K, N, 0.5, g).requires_grad_(True)
    W = torch.randn(M, N, dtype=torch.float64, device="cuda", generator=g)
    (ozaki_matmul(A, B) * W).sum().backward()
    h = 1e-6
    rng = np.random.default_rng(0)

    def loss(a, b):
        with torch.no_grad():
            return (ozaki_matmul(a, b) * W).sum().item()
    for _ in range(6):
        i, k = int(rng.integers(M)), int(rng.integers(K))
        Ap = A.detach().clone(); Ap[i, k] += h
        Am = A.detach().clone(); Am[i, k] -= h
        fd = (loss(Ap, B.detach()) - loss(Am, B.detach())) / (2 * h)
        assert abs(fd - A.grad[i, k].item()) <= 1e-5 * max(1.0, abs(fd))
        k2, j = int(rng.integers(K)), int(rng.integers(N))
        Bp = B.detach().clone(); Bp[k2, j] += h
        Bm = B.detach().clone(); Bm[k2, j] -= h
        fd = (loss(A.detach(), Bp) - loss(A.detach(), Bm)) / (2 * h)
        assert abs(fd - B.grad[k2, j].item()) <= 1e-5 * max(1.0, abs(fd))


def test_backward_uses_resplit_transposes(gpu):
    """The gradient of the emulated product equals the emulated product of the
    transposed operands bit for bit (no stored forward slices)."""
    import torch
    import ozaie_amd as oz
    from ozaie_amd.autograd import ozaki_matmul
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    M, N, K = 64, 96, 128
    A = gen(torch, M, K, 1.0, g).requires_grad_(True)
    B = gen(torch, K, N, 1.0, g).requires_grad_(True)
    W = torch.randn(M, N, dtype=torch.float64, device="cuda", generator=g)
    (ozaki_matmul(A, B, 7, 8, oz.ROUND_TRUNCATE) * W).sum().backward()
    dA = oz.Plan(M, K, N, 7, 8)(W, B.detach().t().contiguous())
    dB = oz.Plan(K, N, M, 7, 8)(A.detach().t().contiguous(), W)
    torch.cuda.synchronize()
    assert torch.equal(dA, A.grad)
    assert torch.equal(dB, B.grad)
