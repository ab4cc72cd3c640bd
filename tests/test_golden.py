"""Committed fixtures (tests/golden/ozaki_k128.npz, written by make_golden.py from the
oracle): the oracle must keep reproducing them, and the GPU path must equal them."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ozaki_k128.npz")
PRESETS = [(3, 29), (7, 8), (7, 3), (7, 7), (7, 9)]


def load():
    return np.load(GOLD, allow_pickle=False)


@pytest.mark.parametrize("b,d", PRESETS)
def test_oracle_reproduces_fixtures(b, d):
    g = load()
    A, B = g["A"], g["B"]
    K = A.shape[1]
    Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
    assert np.array_equal(Er, g["Er"]) and np.array_equal(Ec, g["Ec"])
    tag = f"b{b}d{d}"
    SA = oracle.split(A, Er, b, d, 1)
    SB = oracle.split(B, Ec, b, d, 0)
    assert np.array_equal(SA, g[f"SA_{tag}"]) and np.array_equal(SB, g[f"SB_{tag}"])
    bins = oracle.bins(SA, SB)
    assert np.array_equal(bins, g[f"bins_{tag}"].astype(np.int64))
    assert np.array_equal(oracle.pack_paired(bins, K, b, d, Er, Ec), g[f"paired_{tag}"])
    C, _ = oracle.ozaki_dgemm(A, B, b, d)
    assert np.array_equal(C.view(np.int64), g[f"C_{tag}"].view(np.int64))


@pytest.mark.gpu
@pytest.mark.parametrize("b,d", PRESETS)
def test_gpu_matches_fixtures(gpu, b, d):
    import torch
    import ozaie_amd as oz
    g = load()
    A, B = g["A"], g["B"]
    tag = f"b{b}d{d}"
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    E, S = oz.debug_split(dA, b, d, 1)
    assert np.array_equal(E, g["Er"].astype(np.int32)) and np.array_equal(S, g[f"SA_{tag}"])
    E, S = oz.debug_split(dB, b, d, 0)
    assert np.array_equal(E, g["Ec"].astype(np.int32)) and np.array_equal(S, g[f"SB_{tag}"])
    assert np.array_equal(oz.debug_bins(dA, dB, b, d), g[f"bins_{tag}"])
    assert np.array_equal(oz.paired_output(dA, dB, b, d), g[f"paired_{tag}"])
    M, K = A.shape
    N = B.shape[1]
    C = oz.Plan(M, N, K, b, d)(dA, dB)
    Cn = oz.Plan(M, N, K, b, d, rounding=oz.ROUND_NEAREST)(dA, dB)
    torch.cuda.synchronize()
    assert np.array_equal(C.cpu().numpy().view(np.int64), g[f"C_{tag}"].view(np.int64))
    assert np.array_equal(Cn.cpu().numpy().view(np.int64), g[f"Cnearest_{tag}"].view(np.int64))
