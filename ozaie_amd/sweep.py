"""Measurement sweeps: the equivalent of the reference's experiment driver.

The reference drives its NPU runs from a C# SSH orchestrator (Experiments/Program.cs):
accuracy over phi in {0.1, 1, 2, 4} x 10 seeds x K = 2^min..2^max (Program.cs:113-147),
performance with device iterations/warmups 10/10 and medians (Program.cs:658-664), and a
`resume` flag that skips measured configurations (Program.cs:93-103, 158-165).  This
module runs the same sweeps on the local GPU and appends JSON records to a file.

    python -m ozaie_amd.sweep accuracy --out acc.json
    python -m ozaie_amd.sweep summary acc.json
    python -m ozaie_amd.sweep perf --sizes 1024,2048,4096 --out perf.json

Errors are max / median |C - C*| / |C*| over sampled entries against the correctly
rounded exact product (ozaie_exact_entries); native FP64 DGEMM is scored on the same
entries.
"""
import argparse
import json
import os
import sys

MODES = {"accurate": None, "fast": (7, 8), "fast3": (7, 3), "accurate7": (7, 7), "matched": (7, 9), "crt": "crt"}


def mode_bd(mode, K):
    import ozaie_amd as oz
    if mode == "accurate":
        return oz.mode_slices(oz.Mode.accurate, K)
    if mode.startswith("crt"):          # crt (16 moduli) or crtN
        return 0, int(mode[3:] or 16)
    if mode.startswith("b") and "d" in mode:          # e.g. b7d9
        b, d = mode[1:].split("d")
        return int(b), int(d)
    return MODES[mode]


def gen(torch, n, k, phi, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = (torch.rand(n, k, dtype=torch.float64, device="cuda", generator=g) - 0.5) * torch.exp(
        phi * torch.randn(n, k, dtype=torch.float64, device="cuda", generator=g))
    B = (torch.rand(k, n, dtype=torch.float64, device="cuda", generator=g) - 0.5) * torch.exp(
        phi * torch.randn(k, n, dtype=torch.float64, device="cuda", generator=g))
    return A, B


def load_records(path):
    if not path or not os.path.exists(path):
        return []
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def record_key(r):
    return tuple(sorted((k, v) for k, v in r.items() if k in ("kind", "K", "phi", "seed", "mode", "size")))


def append(path, rec):
    if path:
        with open(path, "a") as f:
            f.write(json.dumps(rec) + "\n")
    print(json.dumps(rec), flush=True)


def rel_err(np, got, ref):
    r = np.abs(got - ref) / np.abs(ref)
    r = r[np.isfinite(r)]
    return float(r.max()), float(np.median(r))


def accuracy(args):
    import numpy as np
    import torch
    import ozaie_amd as oz
    done = {record_key(r) for r in load_records(args.out)} if args.resume else set()
    phis = [float(x) for x in args.phis.split(",")]
    ks = [int(x) for x in args.ks.split(",")]
    modes = args.modes.split(",")
    for K in ks:
        for phi in phis:
            for seed in range(args.seeds):
                todo = [m for m in modes if record_key({"kind": "accuracy", "K": K, "phi": phi, "seed": seed,
                                                       "mode": m}) not in done]
                if not todo:
                    continue
                A, B = gen(torch, K, K, phi, 12445 + seed)
                rng = np.random.default_rng(seed)
                ns = min(args.samples, K * K)
                rows = rng.integers(0, K, size=ns)
                cols = rng.integers(0, K, size=ns)
                exact = oz.exact_entries(A.cpu().numpy(), B.cpu().numpy(), rows, cols)
                nat = (A @ B)[rows, cols].cpu().numpy()
                nat_err = rel_err(np, nat, exact)
                for m in todo:
                    b, d = mode_bd(m, K)
                    crt = m.startswith("crt")
                    C = oz.Plan(K, K, K, b, d, rounding=oz.ROUND_NEAREST if crt else oz.ROUND_TRUNCATE, crt=crt)(A, B)
                    torch.cuda.synchronize()
                    err = rel_err(np, C[rows, cols].cpu().numpy(), exact)
                    append(args.out, {"kind": "accuracy", "K": K, "phi": phi, "seed": seed, "mode": m,
                                      "bits": b, "slices": d, "max_rel_err": err[0], "median_rel_err": err[1],
                                      "native_max_rel_err": nat_err[0], "native_median_rel_err": nat_err[1],
                                      "samples": int(ns)})


def int8_products(mode, K):
    """int8 slice or residue products per output element of a preset (its cost measure, as the
    reference ranks configurations by n_muls, Experiments/Program.cs:879-882)."""
    b, d = mode_bd(mode, K)
    if mode.startswith("crt"):
        return d
    if b <= 3 and d >= 8:                 # the paired-digit path the library runs
        H = (d - 1) // 2 if d % 2 else (d - 2) // 2
        P = H if d % 2 else H + 1
        return P * (P + 1) // 2 + H + 1
    return d * (d + 1) // 2


def summary(args):
    """Per (phi, K): native's max / median error, each preset's max and median over seeds, and
    the matched preset: the fewest int8 products whose max error over the seeds is no worse
    than native's max over the same seeds."""
    recs = [r for r in load_records(args.inp) if r.get("kind") == "accuracy"]
    modes = []
    for r in recs:
        if r["mode"] not in modes:
            modes.append(r["mode"])
    keys = sorted({(r["phi"], r["K"]) for r in recs})
    lines = []
    head = f"{'phi':>4s} {'K':>5s} {'seeds':>5s} {'native max':>10s} " + " ".join(f"{m:>11s}" for m in modes) + \
           "   matched (int8 products)"
    lines.append("max relative error over the seeds (median of the medians in the second table)")
    lines.append(head)
    med_lines = [f"{'phi':>4s} {'K':>5s} {'native med':>10s} " + " ".join(f"{m:>11s}" for m in modes)]
    for phi, K in keys:
        rs = [r for r in recs if r["phi"] == phi and r["K"] == K]
        seeds = sorted({r["seed"] for r in rs})
        nat = max(r["native_max_rel_err"] for r in rs)
        natmed = sorted(r["native_median_rel_err"] for r in rs)[len(rs) // 2]
        row, mrow, ok = [], [], []
        for m in modes:
            mr = [r for r in rs if r["mode"] == m]
            if not mr:
                row.append(f"{'-':>11s}")
                mrow.append(f"{'-':>11s}")
                continue
            mx = max(r["max_rel_err"] for r in mr)
            md = sorted(r["median_rel_err"] for r in mr)[len(mr) // 2]
            row.append(f"{mx:11.2e}")
            mrow.append(f"{md:11.2e}")
            if mx <= nat:
                ok.append((int8_products(m, K), m))
        match = f"{min(ok)[1]} ({min(ok)[0]})" if ok else "none"
        lines.append(f"{phi:4g} {K:5d} {len(seeds):5d} {nat:10.2e} " + " ".join(row) + f"   {match}")
        med_lines.append(f"{phi:4g} {K:5d} {natmed:10.2e} " + " ".join(mrow))
    text = "\n".join(lines + ["", "median relative error (median over the seeds)"] + med_lines)
    print(text)
    return text


def perf(args):
    import numpy as np
    import torch
    import ozaie_amd as oz
    done = {record_key(r) for r in load_records(args.out)} if args.resume else set()
    sizes = [int(x) for x in args.sizes.split(",")]
    modes = args.modes.split(",")

    def med_ms(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.iters):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))
    for n in sizes:
        A, B = gen(torch, n, n, args.phi, 12445)
        C = torch.empty(n, n, dtype=torch.float64, device="cuda")
        t_nat = med_ms(lambda: torch.matmul(A, B, out=C))
        for m in modes:
            if record_key({"kind": "perf", "size": n, "mode": m, "phi": args.phi}) in done:
                continue
            b, d = mode_bd(m, n)
            crt = m.startswith("crt")
            p = oz.Plan(n, n, n, b, d, rounding=oz.ROUND_NEAREST if crt else oz.ROUND_TRUNCATE, crt=crt)
            t = med_ms(lambda: p(A, B, C))
            p.close()
            append(args.out, {"kind": "perf", "size": n, "phi": args.phi, "mode": m, "bits": b, "slices": d,
                              "median_ms": t, "tflops": 2.0 * n ** 3 / t / 1e9, "native_ms": t_nat,
                              "native_tflops": 2.0 * n ** 3 / t_nat / 1e9, "speedup": t_nat / t})


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("accuracy")
    a.add_argument("--phis", default="0.1,1,2,4")
    a.add_argument("--seeds", type=int, default=3)
    a.add_argument("--ks", default="128,256,512,1024,2048")
    a.add_argument("--modes", default="accurate,fast,fast3,accurate7,matched,crt")
    a.add_argument("--samples", type=int, default=1000, help="sampled entries of C per run")
    p = sub.add_parser("perf")
    p.add_argument("--sizes", default="1024,2048,4096")
    p.add_argument("--modes", default="fast3,accurate7,fast,matched,crt")
    p.add_argument("--phi", type=float, default=1.0)
    p.add_argument("--iters", type=int, default=10)
    p.add_argument("--warmup", type=int, default=10)
    for s in (a, p):
        s.add_argument("--out", default=None, help="append JSON records to this file")
        s.add_argument("--resume", action="store_true", help="skip configurations already in --out")
    m = sub.add_parser("summary", help="table of an accuracy sweep's records")
    m.add_argument("inp")
    args = ap.parse_args(argv)
    if args.cmd == "accuracy":
        accuracy(args)
    elif args.cmd == "summary":
        summary(args)
    else:
        perf(args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
