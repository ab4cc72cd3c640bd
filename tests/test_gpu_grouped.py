"""Grouped products: uneven shapes, graph capture and replay, tile configurations.

Every replay of the captured graph (with eager work and synchronizations between the
replays) must equal, bit for bit, the single-problem plan on the same inputs, which the
parity tests tie to the oracle."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz

pytestmark = pytest.mark.gpu


def shapes(rng, n, lo=32, hi=600):
    s = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(n, 3))).astype(np.int64)
    return [tuple(int(v) for v in row) for row in s]


def inputs(torch, shp, seed, phi=1.0):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    As, Bs = [], []
    for m, n, k in shp:
        As.append((torch.rand(m, k, dtype=torch.float64, device="cuda", generator=g) - 0.5)
                  * torch.exp(phi * torch.randn(m, k, dtype=torch.float64, device="cuda", generator=g)))
        Bs.append((torch.rand(k, n, dtype=torch.float64, device="cuda", generator=g) - 0.5)
                  * torch.exp(phi * torch.randn(k, n, dtype=torch.float64, device="cuda", generator=g)))
    return As, Bs


def single(torch, shp, As, Bs, b, d):
    out = []
    for (m, n, k), a, bb in zip(shp, As, Bs):
        p = oz.Plan(m, n, k, b, d)
        out.append(p(a, bb))
        torch.cuda.synchronize()
        p.close()
    return out


@pytest.mark.parametrize("d", [8, 3, 9, 7])
def test_graph_replays_match_single_plans(gpu, d):
    import torch
    rng = np.random.default_rng(d)
    shp = shapes(rng, 40)
    As, Bs = inputs(torch, shp, 100 + d)
    ref = single(torch, shp, As, Bs, 7, d)
    grp = oz.Group(shp, 7, d)
    Cs = [torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    for rep in range(3):
        for c in Cs:
            c.fill_(float("nan"))                      # eager work between replays
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        q = grp.query()
        assert q["graph"] and not q["fallback"] and q["launches"] <= 3 + 5
        for i, (c, r) in enumerate(zip(Cs, ref)):
            assert torch.equal(c, r), (rep, i, shp[i])
    grp.close()


def test_grouped_vs_oracle_small(gpu):
    import torch
    rng = np.random.default_rng(4)
    shp = shapes(rng, 6, 20, 150)
    As, Bs = inputs(torch, shp, 7)
    Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    grp = oz.Group(shp, 7, 8)
    grp(As, Bs, Cs)
    torch.cuda.synchronize()
    for (m, n, k), a, b, c in zip(shp, As, Bs, Cs):
        ref, _ = oracle.ozaki_dgemm(a.cpu().numpy(), b.cpu().numpy(), 7, 8)
        assert np.array_equal(c.cpu().numpy().view(np.int64), ref.view(np.int64))


def test_new_pointers_and_eager_mode(gpu):
    import torch
    rng = np.random.default_rng(5)
    shp = shapes(rng, 12)
    As, Bs = inputs(torch, shp, 9)
    ref = single(torch, shp, As, Bs, 7, 8)
    grp = oz.Group(shp, 7, 8)
    C1 = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    grp(As, Bs, C1)
    As2 = [a.clone() for a in As]
    C2 = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    grp(As2, Bs, C2)                                  # new pointers, same graph
    torch.cuda.synchronize()
    for a, b, r in zip(C1, C2, ref):
        assert torch.equal(a, r) and torch.equal(b, r)
    eager = oz.Group(shp, 7, 8, graph=False)
    C3 = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    eager(As, Bs, C3)
    torch.cuda.synchronize()
    assert not eager.query()["graph"]
    for a, r in zip(C3, ref):
        assert torch.equal(a, r)


def test_fallback_slice_count(gpu):
    import torch
    rng = np.random.default_rng(6)
    shp = shapes(rng, 5, 40, 300)
    As, Bs = inputs(torch, shp, 11)
    ref = single(torch, shp, As, Bs, 6, 5)
    grp = oz.Group(shp, 6, 5)
    Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    for _ in range(2):
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert grp.query()["fallback"]
        for c, r in zip(Cs, ref):
            assert torch.equal(c, r)


@pytest.mark.parametrize("d,cfgs", [(3, [0, 1, 2, 3, 4]), (8, [0, 1, 3, 4]), (9, [0, 3, 4])])
def test_every_tile_configuration_is_exact(gpu, d, cfgs):
    import torch
    rng = np.random.default_rng(7)
    shp = shapes(rng, 8, 60, 700)
    As, Bs = inputs(torch, shp, 13)
    ref = single(torch, shp, As, Bs, 7, d)
    saved = oz.tune_table(d)
    try:
        for c in cfgs:
            oz.set_tune_table(d, [c] * 9)
            grp = oz.Group(shp, 7, d)
            assert set(grp.configs().tolist()) == {c}
            Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
            grp(As, Bs, Cs)
            torch.cuda.synchronize()
            for x, r in zip(Cs, ref):
                assert torch.equal(x, r), c
            grp.close()
    finally:
        oz.set_tune_table(d, saved)


def test_autotune_fills_the_table(gpu):
    table = oz.autotune(7, 8, reps=1)
    assert len(table) == 9
    assert all(t in (0, 1, 3, 4) for t in table)
    assert list(oz.tune_table(8)) == list(table)
    oz.set_tune_table(8, [-1] * 9)
