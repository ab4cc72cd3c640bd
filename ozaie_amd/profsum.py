"""Text summaries of rocprofv3 output (the files kept under profiles/).

    python -m ozaie_amd.profsum kernels DIR [--top N] [--by-grid]
        per-kernel calls / average / min / max / share from the kernel trace
        (rocprofv3 --kernel-trace --output-format csv -d DIR); --by-grid: one line per
        kernel and grid size (threads in x and y), e.g. the CRT GEMM per modulus count
    python -m ozaie_amd.profsum counters DIR [--simds 1024 --xcds 8]
        per-kernel averages of one --pmc run with SQ_VALU_MFMA_BUSY_CYCLES, SQ_WAVE_CYCLES,
        SQ_WAIT_ANY, SQ_INSTS_LDS, SQ_LDS_BANK_CONFLICT, GRBM_GUI_ACTIVE (with --kernel-trace
        in the same run, also the duration and the clock): cycles per XCD = GRBM_GUI_ACTIVE /
        xcds, MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / simds / cycles per XCD, wait =
        SQ_WAIT_ANY / SQ_WAVE_CYCLES
    python -m ozaie_amd.profsum traffic FETCH_DIR WRITE_DIR [--json OUT]
        HBM bytes per dispatch of every oz:: kernel from two --pmc runs (FETCH_SIZE, then
        WRITE_SIZE, in KiB), scaled by what the same counters read for the 1 GiB device copy
        that `bench.py --profile-run` performs (the copy reads and writes exactly 1 GiB)
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys


def _csv(d, suffix):
    f = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not f:
        raise SystemExit(f"no *{suffix} under {d}")
    return list(csv.DictReader(open(f[0])))


def kernels(d, top, by_grid=False):
    rows = _csv(d, "kernel_trace.csv")
    agg = collections.defaultdict(list)
    for r in rows:
        name = r["Kernel_Name"]
        if by_grid:   # one line per launch shape: e.g. the CRT GEMM's grid.y = moduli
            name = f"[{r['Grid_Size_X']}x{r['Grid_Size_Y']}] " + name
        agg[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    tot = sum(sum(v) for v in agg.values()) or 1.0
    out = [f"{'kernel':62s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    for name, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append(f"{name[:62]:62s} {len(v):6d} {sum(v) / len(v):9.1f} {min(v):9.1f} {max(v):9.1f} "
                   f"{100 * sum(v) / tot:6.2f}")
    return "\n".join(out)


def _pmc(d, counter):
    rows = _csv(d, "counter_collection.csv")
    per = collections.defaultdict(float)
    names = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        did = int(r["Dispatch_Id"])
        per[did] += float(r["Counter_Value"])
        names[did] = r["Kernel_Name"]
    return [(did, names[did], per[did]) for did in sorted(per)]


def counters(d, simds, xcds):
    rows = _csv(d, "counter_collection.csv")
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    for r in rows:
        did = int(r["Dispatch_Id"])
        per[did][r["Counter_Name"]] += float(r["Counter_Value"])
        names[did] = r["Kernel_Name"]
    dur = {}
    try:
        for r in _csv(d, "kernel_trace.csv"):
            dur[int(r["Dispatch_Id"])] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    except SystemExit:
        pass
    agg = collections.OrderedDict()
    for did in sorted(per):
        if "oz::" not in names[did]:
            continue
        name = names[did].split("(")[0].replace("void ", "")
        agg.setdefault(name, []).append((per[did], dur.get(did)))
    out = [f"{'kernel':40s} {'calls':>5s} {'us':>9s} {'cycles/XCD':>11s} {'GHz':>5s} {'MFMA busy':>9s} "
           f"{'wait':>6s} {'LDS insts':>10s} {'LDS confl':>9s}"]
    for name, v in agg.items():
        n = len(v)
        mean = lambda key: sum(c.get(key, 0.0) for c, _ in v) / n
        cyc = mean("GRBM_GUI_ACTIVE") / xcds
        busy = mean("SQ_VALU_MFMA_BUSY_CYCLES") / simds / cyc if cyc else 0.0
        wc = mean("SQ_WAVE_CYCLES")
        wait = mean("SQ_WAIT_ANY") / wc if wc else 0.0
        us = [t for _, t in v if t]
        t = sum(us) / len(us) if us else None
        ghz = cyc / (t * 1e3) if t else None
        out.append(f"{name[:40]:40s} {n:5d} {t if t else 0:9.1f} {cyc:11.0f} {ghz if ghz else 0:5.2f} "
                   f"{100 * busy:8.1f}% {100 * wait:5.1f}% {mean('SQ_INSTS_LDS'):10.0f} "
                   f"{mean('SQ_LDS_BANK_CONFLICT'):9.0f}")
    return "\n".join(out)


def traffic(fetch_dir, write_dir):
    fetch = _pmc(fetch_dir, "FETCH_SIZE")
    write = _pmc(write_dir, "WRITE_SIZE")
    if len(fetch) != len(write):
        raise SystemExit("the two runs dispatched different kernel sequences")
    copy = [i for i, (_, n, _) in enumerate(fetch) if "copyBuffer" in n or "elementwise" in n and "copy" in n.lower()]
    if not copy:
        copy = [i for i, (_, n, _) in enumerate(fetch) if "copy" in n.lower()]
    gib_kib = 1 << 20
    ci = copy[-1]
    ff = gib_kib / fetch[ci][2]
    wf = gib_kib / write[ci][2]
    rec = {"calibration": {"kernel": fetch[ci][1][:60], "bytes": 1 << 30, "fetch_kib": fetch[ci][2],
                           "write_kib": write[ci][2], "fetch_factor": ff, "write_factor": wf},
           "dispatches": []}
    for (did, name, fk), (_, _, wk) in zip(fetch, write):
        if "oz::" not in name:
            continue
        rec["dispatches"].append({"kernel": name.split("(")[0], "fetch_kib": fk, "write_kib": wk,
                                  "hbm_bytes": fk * 1024 * ff + wk * 1024 * wf})
    return rec


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    k.add_argument("--top", type=int, default=25)
    k.add_argument("--by-grid", action="store_true", help="one line per kernel and grid size")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--simds", type=int, default=1024)
    c.add_argument("--xcds", type=int, default=8)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--json", default=None)
    t.add_argument("--n", type=int, default=None, help="square size of the profiled workload (recorded)")
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        print(kernels(a.dir, a.top, a.by_grid))
        return 0
    if a.cmd == "counters":
        print(counters(a.dir, a.simds, a.xcds))
        return 0
    rec = traffic(a.fetch_dir, a.write_dir)
    rec["workload_n"] = a.n
    c = rec["calibration"]
    print(f"calibration: {c['kernel']} of 1 GiB reads FETCH_SIZE {c['fetch_kib']:.0f} KiB, WRITE_SIZE "
          f"{c['write_kib']:.0f} KiB -> scale FETCH by {c['fetch_factor']:.3f}, WRITE by {c['write_factor']:.3f}")
    print(f"{'kernel':48s} {'FETCH KiB':>11s} {'WRITE KiB':>11s} {'HBM bytes':>12s}")
    for x in rec["dispatches"]:
        print(f"{x['kernel'][:48]:48s} {x['fetch_kib']:11.0f} {x['write_kib']:11.0f} {x['hbm_bytes']:12.4g}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
