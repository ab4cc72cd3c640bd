"""Text summaries of rocprofv3 output (the files kept under profiles/).

    python -m ozaie_amd.profsum kernels DIR [--top N]
        per-kernel calls / average / min / max / share from the kernel trace
        (rocprofv3 --kernel-trace --output-format csv -d DIR)
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


def kernels(d, top):
    rows = _csv(d, "kernel_trace.csv")
    agg = collections.defaultdict(list)
    for r in rows:
        agg[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
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
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--json", default=None)
    t.add_argument("--n", type=int, default=None, help="square size of the profiled workload (recorded)")
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        print(kernels(a.dir, a.top))
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
