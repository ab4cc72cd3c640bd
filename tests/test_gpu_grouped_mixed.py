"""Groups with a slice count per size class (slices=0): every problem takes its class's
count from the slice-count table, in the multi-launch mode (one GEMM launch per slice
count and tile configuration) and in the ONE persistent launch (runs of tiles per slice
count and tile shape).  Each C_i must equal, bit for bit, a single plan of that problem
with its own slice count on every graph replay; the calibration tuner must pick, per
class, a count whose error is no worse than native DGEMM's (or the largest candidate)."""
import numpy as np
import pytest

import ozaie_amd as oz

pytestmark = pytest.mark.gpu

TABLE = [3, 7, 8, 9, 8, 7, 9, 3, 9]


def shapes(rng, n, lo=32, hi=1500):
    s = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(n, 3))).astype(np.int64)
    return [tuple(int(v) for v in row) for row in s]


def inputs(torch, shp, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    As, Bs = [], []
    for m, n, k in shp:
        As.append((torch.rand(m, k, dtype=torch.float64, device="cuda", generator=g) - 0.5)
                  * torch.exp(torch.randn(m, k, dtype=torch.float64, device="cuda", generator=g)))
        Bs.append((torch.rand(k, n, dtype=torch.float64, device="cuda", generator=g) - 0.5)
                  * torch.exp(torch.randn(k, n, dtype=torch.float64, device="cuda", generator=g)))
    return As, Bs


@pytest.fixture
def mixed_table():
    old = oz.slice_table()
    oz.set_slice_table(TABLE)
    yield TABLE
    oz.set_slice_table(old)


@pytest.mark.parametrize("persistent,pcfg", [(False, None), (True, None), (True, 1)])
def test_mixed_slice_counts_match_single_plans(gpu, mixed_table, persistent, pcfg):
    import torch
    rng = np.random.default_rng(77)
    shp = shapes(rng, 48)
    ds = [TABLE[oz.size_class(m, n, k)] for m, n, k in shp]
    assert len(set(ds)) >= 3
    As, Bs = inputs(torch, shp, 78)
    ref = []
    for (m, n, k), a, b, d in zip(shp, As, Bs, ds):
        p = oz.Plan(m, n, k, 7, d)
        ref.append(p(a, b))
        torch.cuda.synchronize()
        p.close()
    print(f"mixed slice counts, persistent={persistent} cfg={pcfg}", flush=True)
    grp = oz.Group(shp, 7, 0, persistent=persistent, persistent_cfg=pcfg)
    q = grp.query()
    assert not q["fallback"] and (q["launches"] == 1) == persistent
    Cs = [torch.empty((m, n), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    for rep in range(3):
        for c in Cs:
            c.fill_(float("nan"))
        grp(As, Bs, Cs)
        torch.cuda.synchronize()
        assert grp.status() == 0
        for i, (c, r) in enumerate(zip(Cs, ref)):
            assert torch.equal(c.view(torch.int64), r.view(torch.int64)), (rep, i, shp[i], ds[i])
    grp.close()


def test_tune_slices_picks_matched_counts(gpu):
    from ozaie_amd import tune
    table, details = tune.tune_slices(phi=1.0, seeds=(0,), sample=16, apply=False)
    assert len(table) == 9
    for t, det in zip(table, details):
        assert t in (3, 7, 8, 9)
        ok = [d for d, e in det["errors"].items() if e <= det["native_max_rel_err"]]
        if ok:
            assert det["errors"][t] <= det["native_max_rel_err"]
            assert t * (t + 1) == min(d * (d + 1) for d in ok)
        else:
            assert t == 9
