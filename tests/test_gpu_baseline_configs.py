"""Parity at the full sizes of the benchmark configurations.

The exponents and slices (or residues) of an element depend only on its row of A, or its
column of B, over the whole contraction axis.  So the oracle run on A[rows, :] and
B[:, cols] reproduces C[rows, cols] of the full product exactly, and a sampled block of a
full-size GPU result can be checked bit for bit at the cost of a small CPU product (the
reference samples large shapes the same way: stochastic verification,
src/programming_examples/basic/matrix_multiplication/test.cpp:479-520, common.h:613-669).

  square 4096^3, phi = 1: b = 7 with 3, 7, 8 slices (truncating) and the CRT mode with 16
    moduli (nearest), a 64 x 64 block;
  square 8192^3, phi = 1 (the headline workload): the CRT mode with 14 and 15 moduli and
    b = 7, d = 9, 4 sampled rows x 64 sampled columns;
  8192 x 8192 x 32768, phi = 2: b = 7, d = 7 (two int32-safe contraction passes) and the
    CRT mode with 15 and 17 moduli (128- and 192-bit rebuilds), 4 sampled rows x 64 sampled
    columns;
  512 problems with M, N, K log-uniform in [32, 2048], phi = 1: the grouped multi-launch
    and persistent modes (b = 7 with 3, 7, 8, 9 slices; the CRT mode with 14 and 16
    moduli), every problem equal to its own single plan, and 8 sampled problems equal to
    the oracle on sampled rows and columns.
"""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz

pytestmark = pytest.mark.gpu

SEED = 12445


def gen(torch, g, rows, cols, phi):
    """(U(0,1) - 0.5) exp(phi N(0,1)) on the GPU (the benchmark's generator)."""
    return (torch.rand(rows, cols, dtype=torch.float64, device="cuda", generator=g) - 0.5) * torch.exp(
        phi * torch.randn(rows, cols, dtype=torch.float64, device="cuda", generator=g))


def bits(x):
    return np.ascontiguousarray(x).view(np.int64)


@pytest.fixture(scope="module")
def square4096(gpu):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED)
    A = gen(torch, g, 4096, 4096, 1.0)
    B = gen(torch, g, 4096, 4096, 1.0)
    rng = np.random.default_rng(SEED + 4096)
    r0, c0 = (int(v) for v in rng.integers(0, 4096 - 64, size=2))
    Ab = A[r0:r0 + 64].cpu().numpy()
    Bb = B[:, c0:c0 + 64].cpu().numpy()
    yield A, B, (r0, c0), Ab, Bb
    del A, B
    torch.cuda.empty_cache()


@pytest.mark.parametrize("d", [3, 7, 8])
def test_square4096_split_block_vs_oracle(square4096, d):
    import torch
    A, B, (r0, c0), Ab, Bb = square4096
    print(f"4096^3 phi=1 b=7 d={d} block ({r0}, {c0})", flush=True)
    p = oz.Plan(4096, 4096, 4096, 7, d)
    assert p.info()["passes"] == 1
    C = p(A, B)
    torch.cuda.synchronize()
    got = C[r0:r0 + 64, c0:c0 + 64].cpu().numpy()
    p.close()
    ref, ov = oracle.ozaki_dgemm(Ab, Bb, 7, d)
    assert ov == 0
    assert np.array_equal(bits(got), bits(ref))


def test_square4096_crt16_block_vs_oracle(square4096):
    import torch
    A, B, (r0, c0), Ab, Bb = square4096
    print(f"4096^3 phi=1 CRT 16 moduli block ({r0}, {c0})", flush=True)
    p = oz.Plan(4096, 4096, 4096, 0, 16, rounding=oz.ROUND_NEAREST, crt=True)
    C = p(A, B)
    torch.cuda.synchronize()
    got = C[r0:r0 + 64, c0:c0 + 64].cpu().numpy()
    C2 = p(A, B)                      # a second run is bit-identical (no state kept)
    torch.cuda.synchronize()
    assert torch.equal(C.view(torch.int64), C2.view(torch.int64))
    p.close()
    ref = oracle.crt_dgemm(Ab, Bb, 16, oz.ROUND_NEAREST)
    assert np.array_equal(bits(got), bits(ref))


@pytest.fixture(scope="module")
def square8192(gpu):
    import torch
    n = 8192
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED + 8)
    A = gen(torch, g, n, n, 1.0)
    B = gen(torch, g, n, n, 1.0)
    rng = np.random.default_rng(SEED + 8)
    rows = np.sort(rng.choice(n, 4, replace=False))
    cols = np.sort(rng.choice(n, 64, replace=False))
    Ar = A[torch.from_numpy(rows).cuda()].cpu().numpy()
    Bc = B[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    yield A, B, rows, cols, Ar, Bc
    del A, B
    torch.cuda.empty_cache()


def _sampled(torch, C, rows, cols):
    return C[torch.from_numpy(rows).cuda()][:, torch.from_numpy(cols).cuda()].cpu().numpy()


@pytest.mark.parametrize("n", [14, 15])
def test_square8192_crt_vs_oracle(square8192, n):
    import torch
    A, B, rows, cols, Ar, Bc = square8192
    print(f"8192^3 phi=1 CRT {n} moduli", flush=True)
    p = oz.Plan(8192, 8192, 8192, 0, n, rounding=oz.ROUND_NEAREST, crt=True)
    assert p.info()["bits"] == 0                                  # the norm rule
    C = p(A, B)
    torch.cuda.synchronize()
    got = _sampled(torch, C, rows, cols)
    del C
    p.close()
    assert np.array_equal(bits(got), bits(oracle.crt_dgemm(Ar, Bc, n, oz.ROUND_NEAREST)))


def test_square8192_split_d9_vs_oracle(square8192):
    import torch
    A, B, rows, cols, Ar, Bc = square8192
    print("8192^3 phi=1 b=7 d=9", flush=True)
    p = oz.Plan(8192, 8192, 8192, 7, 9)
    C = p(A, B)
    torch.cuda.synchronize()
    got = _sampled(torch, C, rows, cols)
    del C
    p.close()
    ref, ov = oracle.ozaki_dgemm(Ar, Bc, 7, 9)
    assert ov == 0
    assert np.array_equal(bits(got), bits(ref))


@pytest.fixture(scope="module")
def long_k(gpu):
    import torch
    M = N = 8192
    K = 32768
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED + 3)
    A = gen(torch, g, M, K, 2.0)
    B = gen(torch, g, K, N, 2.0)
    rng = np.random.default_rng(SEED + 3)
    rows = np.sort(rng.choice(M, 4, replace=False))
    cols = np.sort(rng.choice(N, 64, replace=False))
    Ar = A[torch.from_numpy(rows).cuda()].cpu().numpy()
    Bc = B[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    yield A, B, rows, cols, Ar, Bc
    del A, B
    torch.cuda.empty_cache()


def test_long_k_d7_two_passes_vs_oracle(long_k):
    import torch
    A, B, rows, cols, Ar, Bc = long_k
    print("8192x8192x32768 phi=2 b=7 d=7", flush=True)
    p = oz.Plan(8192, 8192, 32768, 7, 7)
    info = p.info()
    assert info["passes"] == 2
    C = p(A, B)
    torch.cuda.synchronize()
    got = C[torch.from_numpy(rows).cuda()][:, torch.from_numpy(cols).cuda()].cpu().numpy()
    del C
    p.close()
    ref, ov = oracle.ozaki_dgemm(Ar, Bc, 7, 7, kchunk=info["kchunk"])
    assert ov == 0                    # the planner's chunk keeps every int32 bin exact
    assert np.array_equal(bits(got), bits(ref))


@pytest.mark.parametrize("n", [15, 17])
def test_long_k_crt_vs_oracle(long_k, n):
    import torch
    A, B, rows, cols, Ar, Bc = long_k
    print(f"8192x8192x32768 phi=2 CRT {n} moduli", flush=True)
    p = oz.Plan(8192, 8192, 32768, 0, n, rounding=oz.ROUND_NEAREST, crt=True)
    assert p.info()["bits"] == 0                                  # the norm rule
    C = p(A, B)
    torch.cuda.synchronize()
    got = C[torch.from_numpy(rows).cuda()][:, torch.from_numpy(cols).cuda()].cpu().numpy()
    del C
    p.close()
    ref = oracle.crt_dgemm(Ar, Bc, n, oz.ROUND_NEAREST)
    assert np.array_equal(bits(got), bits(ref))


def grouped_shapes(count=512):
    """The benchmark's 512 shapes: M, N, K log-uniform in [32, 2048]."""
    rng = np.random.default_rng(SEED + 4)
    s = np.exp(rng.uniform(np.log(32), np.log(2048), size=(count, 3))).astype(np.int64)
    return [tuple(int(v) for v in row) for row in s]


@pytest.fixture(scope="module")
def group512(gpu):
    import torch
    shapes = grouped_shapes()
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED + 4)
    As, Bs = [], []
    for m, n, k in shapes:
        As.append(gen(torch, g, m, k, 1.0))
        Bs.append(gen(torch, g, k, n, 1.0))
    yield shapes, As, Bs
    del As, Bs
    torch.cuda.empty_cache()


def _single_plans(torch, shapes, As, Bs, bits_, d, crt):
    out = []
    rnd = oz.ROUND_NEAREST if crt else oz.ROUND_TRUNCATE
    for (m, n, k), a, b in zip(shapes, As, Bs):
        p = oz.Plan(m, n, k, bits_, d, rnd, crt=crt)
        out.append(p(a, b))
        torch.cuda.synchronize()
        p.close()
    return out


def _oracle_samples(shapes, As, Bs, Cs, d, crt, seed):
    rng = np.random.default_rng(seed)
    for i in rng.choice(len(shapes), 8, replace=False):
        m, n, k = shapes[i]
        rows = np.sort(rng.choice(m, min(m, 16), replace=False))
        cols = np.sort(rng.choice(n, min(n, 16), replace=False))
        Ar = As[i].cpu().numpy()[rows]
        Bc = Bs[i].cpu().numpy()[:, cols]
        got = Cs[i].cpu().numpy()[np.ix_(rows, cols)]
        if crt:
            ref = oracle.crt_dgemm(Ar, Bc, d, oz.ROUND_NEAREST)
        else:
            ref, ov = oracle.ozaki_dgemm(Ar, Bc, 7, d)
            assert ov == 0
        assert np.array_equal(bits(got), bits(ref)), (int(i), shapes[i])


@pytest.mark.parametrize("d", [3, 7, 8, 9])
def test_group512_modes_vs_single_plans_and_oracle(group512, d):
    import torch
    shapes, As, Bs = group512
    ref = _single_plans(torch, shapes, As, Bs, 7, d, False)
    for persistent in (False, True):
        print(f"512 problems b=7 d={d} persistent={persistent}", flush=True)
        grp = oz.Group(shapes, 7, d, persistent=persistent)
        Cs = [torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda") for m, n, _ in shapes]
        for rep in range(2):          # capture + replay
            grp(As, Bs, Cs)
            torch.cuda.synchronize()
            assert grp.status() == 0
            for i, (c, r) in enumerate(zip(Cs, ref)):
                assert torch.equal(c.view(torch.int64), r.view(torch.int64)), (persistent, rep, i, shapes[i])
        if persistent:
            assert grp.query()["launches"] == 1
        grp.close()
    _oracle_samples(shapes, As, Bs, Cs, d, False, SEED + d)


@pytest.mark.parametrize("nm", [14, 16])
def test_group512_crt_vs_single_plans_and_oracle(group512, nm):
    import torch
    shapes, As, Bs = group512
    ref = _single_plans(torch, shapes, As, Bs, 0, nm, True)
    for persistent in (False, True):
        print(f"512 problems CRT {nm} moduli persistent={persistent}", flush=True)
        grp = oz.Group(shapes, 0, nm, rounding=oz.ROUND_NEAREST, crt=True, persistent=persistent)
        Cs = [torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda") for m, n, _ in shapes]
        for rep in range(2):
            grp(As, Bs, Cs)
            torch.cuda.synchronize()
            assert grp.status() == 0
            for i, (c, r) in enumerate(zip(Cs, ref)):
                assert torch.equal(c.view(torch.int64), r.view(torch.int64)), (persistent, rep, i, shapes[i])
        if persistent:
            assert grp.query()["launches"] == 1
        grp.close()
    _oracle_samples(shapes, As, Bs, Cs, nm, True, SEED + nm)
