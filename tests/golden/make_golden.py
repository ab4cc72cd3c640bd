"""Writes tests/golden/ozaki_k128.npz: small fixtures of every intermediate of the
reference Ozaki path, computed by the oracle (oracle/, the C restatement of the
reference rule: word-level split, triangular bins, paired-int64 wire format and the
reference host conversion).  The reference ships no vectors of its own and cannot be
built here, so these fixtures pin the oracle's current output (regression) and let the
GPU tests compare against committed data.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

PRESETS = [(3, 29), (7, 8), (7, 3), (7, 7), (7, 9)]
M, N, K = 16, 24, 128


def main():
    rng = np.random.default_rng(20260101)
    A = (rng.random((M, K)) - 0.5) * np.exp(1.0 * rng.standard_normal((M, K)))
    B = (rng.random((K, N)) - 0.5) * np.exp(1.0 * rng.standard_normal((K, N)))
    A[0, :3] = [0.0, -0.0, 5e-324]
    B[5, 7] = 2.2250738585072014e-308
    out = {"A": A, "B": B}
    Er, Ec = oracle.max_exp(A, 1), oracle.max_exp(B, 0)
    out["Er"], out["Ec"] = Er, Ec
    for b, d in PRESETS:
        SA = oracle.split(A, Er, b, d, 1, mech=True)
        SB = oracle.split(B, Ec, b, d, 0, mech=True)
        bins = oracle.bins(SA, SB)
        paired = oracle.pack_paired(bins, K, b, d, Er, Ec)
        C = oracle.convert_paired(paired, K, b, d)
        Cn, _ = oracle.ozaki_dgemm(A, B, b, d, 0, 1)
        tag = f"b{b}d{d}"
        out[f"SA_{tag}"], out[f"SB_{tag}"] = SA, SB
        out[f"bins_{tag}"] = bins.astype(np.int32)
        out[f"paired_{tag}"] = paired
        out[f"C_{tag}"] = C
        out[f"Cnearest_{tag}"] = Cn
    np.savez_compressed(os.path.join(HERE, "ozaki_k128.npz"), **out)
    print("wrote", os.path.join(HERE, "ozaki_k128.npz"))


if __name__ == "__main__":
    main()
