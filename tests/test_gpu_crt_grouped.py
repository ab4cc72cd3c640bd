"""Grouped products in the CRT mode (OZAIE_GROUP_CRT).

Every replay of the captured graph, for every tile shape, must equal bit for bit the
single-problem CRT plan on the same inputs (tied to the oracle's CRT restatement by
tests/test_gpu_crt.py), and small problems are also checked against the oracle directly.
Covers ragged shapes (M, N, K of 1, not multiples of 4 or 32, K below and above one
contraction ring), both roundings, fixed bit counts, pointer changes and the limits."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz

pytestmark = pytest.mark.gpu


def shapes(rng, n, lo=32, hi=700):
    s = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(n, 3))).astype(np.int64)
    return [tuple(int(v) for v in row) for row in s]


EDGE = [(1, 1, 1), (1, 70, 33), (70, 1, 257), (5, 3, 31), (33, 65, 300), (130, 129, 1), (64, 64, 256),
        (257, 255, 511), (4, 300, 64)]


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


def single(torch, shp, As, Bs, bits, n, rounding):
    out = []
    for (m, nn, k), a, b in zip(shp, As, Bs):
        p = oz.Plan(m, nn, k, bits, n, rounding=rounding, crt=True)
        out.append(p(a, b))
        torch.cuda.synchronize()
        p.close()
    return out


@pytest.mark.parametrize("cfg", [None, 0, 1, 2])
@pytest.mark.parametrize("rounding", [oz.ROUND_TRUNCATE, oz.ROUND_NEAREST])
def test_crt_group_replays_match_single_plans(gpu, cfg, rounding):
    import torch
    rng = np.random.default_rng(31 + (cfg or 0))
    shp = EDGE + shapes(rng, 24)
    As, Bs = inputs(torch, shp, 200 + (cfg or 0))
    ref = single(torch, shp, As, Bs, 0, 16, rounding)
    grp = oz.Group(shp, 0, 16, rounding=rounding, crt=True, crt_cfg=cfg)
    if cfg is not None:
        assert all(int(c) == cfg for c in grp.configs())
    Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    for rep in range(3):
        for c in Cs:
            c.fill_(float("nan"))                      # eager work between replays
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        q = grp.query()
        assert q["graph"] and not q["fallback"] and 5 <= q["launches"] <= 7
        for i, (c, r) in enumerate(zip(Cs, ref)):
            assert torch.equal(c, r), (rep, i, shp[i])
    grp.close()


@pytest.mark.parametrize("n", [16, 9, 3, 17, 20])
def test_crt_group_vs_oracle_small(gpu, n):
    import torch
    rng = np.random.default_rng(40 + n)
    shp = EDGE[:6] + shapes(rng, 5, 20, 150)
    As, Bs = inputs(torch, shp, 9 + n, phi=2.0)
    Cs = [torch.empty((m, nn), dtype=torch.float64, device="cuda") for m, nn, _ in shp]
    grp = oz.Group(shp, 0, n, rounding=oz.ROUND_NEAREST, crt=True)
    grp(As, Bs, Cs)
    torch.cuda.synchronize()
    for (m, nn, k), a, b, c in zip(shp, As, Bs, Cs):
        ref = oracle.crt_dgemm(a.cpu().numpy(), b.cpu().numpy(), n, oz.ROUND_NEAREST)
        assert np.array_equal(c.cpu().numpy().view(np.int64), ref.view(np.int64)), (m, nn, k)
    grp.close()


def test_crt_group_fixed_bits_new_pointers_eager(gpu):
    import torch
    rng = np.random.default_rng(50)
    shp = shapes(rng, 10)
    bits = min(oz.crt_beta(16, k) for _, _, k in shp) - 5
    As, Bs = inputs(torch, shp, 51)
    ref = single(torch, shp, As, Bs, bits, 16, oz.ROUND_TRUNCATE)
    grp = oz.Group(shp, bits, 16, crt=True, graph=False)
    for _ in range(2):                                  # new C tensors: descriptors re-uploaded
        Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert not grp.query()["graph"]
        for i, (c, r) in enumerate(zip(Cs, ref)):
            assert torch.equal(c, r), (i, shp[i])
    grp.close()


def test_crt_group_limits(gpu):
    with pytest.raises(oz.OzaieError):
        oz.Group([(64, 64, 70000)], 0, 16, crt=True)                 # beyond exact int32 residue sums
    with pytest.raises(oz.OzaieError):
        oz.Group([(64, 64, 64)], 0, 21, crt=True)                    # more moduli than the table
    with pytest.raises(oz.OzaieError):
        oz.Group([(64, 64, 64), (64, 64, 4096)], oz.crt_beta(16, 64), 16, crt=True)   # bits too many for K=4096
    with pytest.raises(oz.OzaieError):
        oz.Group([(64, 64, 64)], 0, 16, crt=True, crt_cfg=3)         # no such tile shape


def test_crt_autotune_table(gpu):
    table = oz.autotune(0, 16, reps=1, crt=True)
    assert len(table) == 9 and all(0 <= int(t) <= 2 for t in table)
    assert list(oz.tune_table(16, crt=True)) == list(table)
    oz.set_tune_table(16, [-1] * 9, crt=True)
    assert list(oz.tune_table(16, crt=True)) == [-1] * 9


@pytest.mark.parametrize("mpw", [1, 5, 16])
def test_crt_group_moduli_per_workgroup(gpu, mpw):
    import torch
    rng = np.random.default_rng(60 + mpw)
    shp = EDGE + shapes(rng, 12)
    As, Bs = inputs(torch, shp, 61 + mpw)
    ref = single(torch, shp, As, Bs, 0, 16, oz.ROUND_NEAREST)
    oz.set_option(oz.OPT_CRT_GROUP_MODULI, mpw)
    try:
        for cfg in (0, 1, 2):
            grp = oz.Group(shp, 0, 16, rounding=oz.ROUND_NEAREST, crt=True, crt_cfg=cfg)
            Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
            grp(As, Bs, Cs)
            torch.cuda.synchronize()
            for i, (c, r) in enumerate(zip(Cs, ref)):
                assert torch.equal(c, r), (cfg, i, shp[i])
            grp.close()
    finally:
        oz.set_option(oz.OPT_CRT_GROUP_MODULI, 0)


@pytest.mark.parametrize("n,cfg,rounding", [(16, None, oz.ROUND_NEAREST), (16, 0, oz.ROUND_TRUNCATE),
                                            (16, 2, oz.ROUND_NEAREST), (17, None, oz.ROUND_NEAREST),
                                            (3, 1, oz.ROUND_TRUNCATE)])
def test_crt_persistent_group_matches_single_plans(gpu, n, cfg, rounding):
    """ONE persistent launch (residues, grid barrier, modulus GEMMs, grid barrier, rebuild):
    every replay equal to the single CRT plans, ragged shapes included, per-class and forced
    tile shapes, 17 moduli through the 192-bit rebuild."""
    import torch
    rng = np.random.default_rng(60 + n + (cfg or 0))
    shp = EDGE + shapes(rng, 30)
    As, Bs = inputs(torch, shp, 300 + n)
    ref = single(torch, shp, As, Bs, 0, n, rounding)
    print(f"persistent CRT group n={n} cfg={cfg}", flush=True)
    grp = oz.Group(shp, 0, n, rounding=rounding, crt=True, crt_cfg=cfg, persistent=True)
    q = grp.query()
    assert q["launches"] == 1 and not q["fallback"]
    if cfg is not None:
        assert all(int(c) == cfg for c in grp.configs())
    Cs = [torch.empty((m, nn), dtype=torch.float64, device="cuda") for m, nn, _ in shp]
    for rep in range(3):
        for c in Cs:
            c.fill_(float("nan"))                      # eager work between replays
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert grp.status() == 0
        for i, (c, r) in enumerate(zip(Cs, ref)):
            assert torch.equal(c.view(torch.int64), r.view(torch.int64)), (rep, i, shp[i])
    grp.close()
