"""Grouped products as one persistent launch (split phase, grid barrier, GEMM phase).

Every result must equal, bit for bit, the single-problem plan on the same inputs (which
the parity tests tie to the oracle), on repeated replays with eager work between them,
with new operand pointers, and for every persistent tile configuration; the kernel's
barrier error flag must stay clear."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz

pytestmark = pytest.mark.gpu

from test_gpu_grouped import inputs, shapes, single   # noqa: E402

PERSIST_CFGS = {3: [0, 1, 2], 7: [0, 1], 8: [0], 9: [0]}


@pytest.mark.parametrize("d", [8, 3, 9, 7])
def test_persistent_replays_match_single_plans(gpu, d):
    import torch
    rng = np.random.default_rng(20 + d)
    shp = shapes(rng, 40)
    As, Bs = inputs(torch, shp, 200 + d)
    ref = single(torch, shp, As, Bs, 7, d)
    grp = oz.Group(shp, 7, d, persistent=True)
    q = grp.query()
    assert q["launches"] == 1 and not q["fallback"]
    Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    for rep in range(3):
        for c in Cs:
            c.fill_(float("nan"))                      # eager work between replays
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert grp.status() == 0
        for i, (c, r) in enumerate(zip(Cs, ref)):
            assert torch.equal(c, r), (rep, i, shp[i])
    grp.close()


@pytest.mark.parametrize("d", [3, 7])
def test_every_persistent_configuration_is_exact(gpu, d):
    import torch
    rng = np.random.default_rng(30 + d)
    shp = shapes(rng, 10, 60, 900)
    As, Bs = inputs(torch, shp, 300 + d)
    ref = single(torch, shp, As, Bs, 7, d)
    for c in PERSIST_CFGS[d]:
        grp = oz.Group(shp, 7, d, persistent=True, persistent_cfg=c)
        assert set(grp.configs().tolist()) == {c}
        Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert grp.status() == 0
        for x, r in zip(Cs, ref):
            assert torch.equal(x, r), c
        grp.close()


def test_persistent_vs_oracle_and_nearest(gpu):
    import torch
    rng = np.random.default_rng(8)
    shp = shapes(rng, 6, 20, 150)
    As, Bs = inputs(torch, shp, 17)
    for rounding in (oz.ROUND_TRUNCATE, oz.ROUND_NEAREST):
        Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
        grp = oz.Group(shp, 7, 8, rounding=rounding, persistent=True)
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert grp.status() == 0
        for (m, n, k), a, b, c in zip(shp, As, Bs, Cs):
            ref, _ = oracle.ozaki_dgemm(a.cpu().numpy(), b.cpu().numpy(), 7, 8, rounding=rounding)
            assert np.array_equal(c.cpu().numpy().view(np.int64), ref.view(np.int64))
        grp.close()


def test_persistent_new_pointers_one_problem_and_large(gpu):
    import torch
    # one problem (most workgroups have no items) and a mix with a 2048^3 product
    for shp in ([(300, 200, 500)], [(2048, 2048, 2048), (33, 70, 31), (700, 64, 1500)]):
        As, Bs = inputs(torch, shp, 21)
        ref = single(torch, shp, As, Bs, 7, 9)
        grp = oz.Group(shp, 7, 9, persistent=True)
        C1 = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
        grp(As, Bs, C1)
        As2 = [a.clone() for a in As]
        C2 = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
        grp(As2, Bs, C2)                              # new pointers, same graph
        torch.cuda.synchronize()
        assert grp.status() == 0
        for a, b, r in zip(C1, C2, ref):
            assert torch.equal(a, r) and torch.equal(b, r)
        grp.close()


def test_persistent_eager_mode(gpu):
    import torch
    rng = np.random.default_rng(9)
    shp = shapes(rng, 12)
    As, Bs = inputs(torch, shp, 23)
    ref = single(torch, shp, As, Bs, 7, 3)
    grp = oz.Group(shp, 7, 3, graph=False, persistent=True)
    Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    for _ in range(2):
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert not grp.query()["graph"] and grp.status() == 0
        for c, r in zip(Cs, ref):
            assert torch.equal(c, r)
    grp.close()
