"""Grouped products with degenerate shapes: M, N or K of 1, sizes just around the 32-row
slice blocks, the 64 / 128-row tiles and the 64-index contraction padding, mixed in one
group.  Every problem of every launch mode (per-phase graph, one persistent launch, the CRT
mode in both) must equal the oracle's C bit for bit -- the split scheme against the
restated reference rule, the CRT mode against the CRT restatement."""
import numpy as np
import pytest

import oracle
import ozaie_amd as oz

pytestmark = pytest.mark.gpu

EDGE = [(1, 1, 1), (1, 70, 3), (33, 1, 65), (5, 7, 1), (64, 64, 64), (65, 129, 33), (2, 300, 2),
        (31, 33, 95), (128, 1, 129), (1, 128, 64), (97, 63, 31), (129, 65, 200)]


def edge_shapes(seed):
    rng = np.random.default_rng(seed)
    rnd = [tuple(int(v) for v in row) for row in rng.integers(1, 71, size=(12, 3))]
    return EDGE + rnd


def host_inputs(shp, seed):
    rng = np.random.default_rng(seed)
    As = [(rng.random((m, k)) - 0.5) * np.exp(rng.standard_normal((m, k))) for m, n, k in shp]
    Bs = [(rng.random((k, n)) - 0.5) * np.exp(rng.standard_normal((k, n))) for m, n, k in shp]
    return As, Bs


def run_group(torch, grp, shp, As, Bs):
    dA = [torch.from_numpy(a).cuda() for a in As]
    dB = [torch.from_numpy(b).cuda() for b in Bs]
    Cs = [torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda") for m, n, _ in shp]
    grp(dA, dB, Cs)
    torch.cuda.synchronize()
    assert grp.status() == 0
    return [c.cpu().numpy() for c in Cs]


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("d", [3, 9])
def test_split_group_edges_match_oracle(gpu, d, persistent):
    import torch
    shp = edge_shapes(d)
    As, Bs = host_inputs(shp, 10 + d)
    grp = oz.Group(shp, 7, d, persistent=persistent)
    got = run_group(torch, grp, shp, As, Bs)
    grp.close()
    for i, (a, b, c) in enumerate(zip(As, Bs, got)):
        ref, _ = oracle.ozaki_dgemm(a, b, 7, d)
        assert np.array_equal(c.view(np.int64), ref.view(np.int64)), (i, shp[i])


@pytest.mark.parametrize("persistent", [False, True])
def test_crt_group_edges_match_oracle(gpu, persistent):
    import torch
    shp = edge_shapes(5)
    As, Bs = host_inputs(shp, 77)
    grp = oz.Group(shp, 0, 16, rounding=oz.ROUND_NEAREST, crt=True, persistent=persistent)
    got = run_group(torch, grp, shp, As, Bs)
    grp.close()
    for i, (a, b, c) in enumerate(zip(As, Bs, got)):
        ref = oracle.crt_dgemm(a, b, 16, oz.ROUND_NEAREST)
        assert np.array_equal(c.view(np.int64), ref.view(np.int64)), (i, shp[i])
