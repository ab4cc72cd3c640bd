"""Per-size-class slice counts for grouped products.

The reference fixes (b, d) per log2 K from the accuracy each configuration reaches and its
cost (include/oz_aie.h:60-80, derived in Experiments/Program.cs:803-898: "accurate" is
the largest mantissa coverage, "fast" the cheapest configuration by n_muls / throughput at
53 bits).  For grouped products of uneven shape this module makes the same choice per
size class of the grouped autotuner (max(M, N) and K each in {<= 256, <= 1024, larger}):
on calibration products of the class's representative shape, with the workload's input
distribution, it measures the maximum relative error of native FP64 DGEMM (torch.matmul)
and of each candidate slice count against the correctly rounded exact product, and keeps
the cheapest count (fewest slice products, d(d+1)/2) whose maximum error is no worse than
native's.  The table goes to the library (ozaie_tune_set_slices), where groups created
with slices=0 read it, and the tile configurations of every chosen count are tuned by
the library's own autotuner (ozaie_autotune).
"""
import numpy as np

from . import Plan, autotune, exact_entries, set_slice_table

CLASS_MN = (128, 512, 2048)
CLASS_K = (128, 512, 2048)


def _inputs(torch, g, rows, cols, phi):
    return (torch.rand(rows, cols, dtype=torch.float64, device="cuda", generator=g) - 0.5) * torch.exp(
        phi * torch.randn(rows, cols, dtype=torch.float64, device="cuda", generator=g))


def _max_rel(got, ref):
    r = np.abs(got - ref) / np.abs(ref)
    r = r[np.isfinite(r)]
    return float(r.max()) if r.size else 0.0


def tune_slices(phi=1.0, candidates=(3, 7, 8, 9), seeds=(0, 1), sample=32, bits=7, apply=True,
                tune_tiles=True, reps=2):
    """Calibrate the slice count of each of the 9 size classes; returns (table9, details).
    details[c] = {"shape": (mn, mn, k), "native_max_rel_err": e, "errors": {d: e_d}, "slices": d}."""
    import torch
    cands = sorted(set(int(d) for d in candidates), key=lambda d: d * (d + 1))
    table = [-1] * 9
    details = []
    for a, mn in enumerate(CLASS_MN):
        for kb, k in enumerate(CLASS_K):
            cls = 3 * a + kb
            nat_err = 0.0
            errs = {d: 0.0 for d in cands}
            for seed in seeds:
                g = torch.Generator(device="cuda")
                g.manual_seed(1000 * cls + seed)
                A = _inputs(torch, g, mn, k, phi)
                B = _inputs(torch, g, k, mn, phi)
                s = min(sample, mn)
                rows = np.repeat(np.arange(s), s)
                cols = np.tile(np.arange(s), s)
                exact = exact_entries(A[:s].cpu().numpy(), B[:, :s].cpu().numpy(), rows, cols)
                nat = torch.matmul(A, B)[:s, :s].cpu().numpy().ravel()
                nat_err = max(nat_err, _max_rel(nat, exact))
                for d in cands:
                    p = Plan(mn, mn, k, bits, d)
                    C = p(A, B)
                    torch.cuda.synchronize()
                    errs[d] = max(errs[d], _max_rel(C[:s, :s].cpu().numpy().ravel(), exact))
                    p.close()
            pick = next((d for d in cands if errs[d] <= nat_err), cands[-1])
            table[cls] = pick
            details.append({"shape": (mn, mn, k), "native_max_rel_err": nat_err, "errors": errs, "slices": pick})
    if apply:
        set_slice_table(table)
        if tune_tiles:
            for d in sorted(set(table)):
                autotune(bits, d, reps=reps)
    return table, details


__all__ = ["tune_slices", "CLASS_MN", "CLASS_K"]
