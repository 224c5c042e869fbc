"""Summaries of rocprofv3 output for the profiles/ folder.

    python -m marching_cube_amd.rocprof_summary stats  <kernel_stats.csv> [--match extract]
    python -m marching_cube_amd.rocprof_summary window <kernel_trace.csv> --first F --count N [--match k ...]
    python -m marching_cube_amd.rocprof_summary counters <pm_counter_collection.csv> [--match k ...]
    python -m marching_cube_amd.rocprof_summary traffic <fetch_dir> <write_dir> --out X.json

`window` averages, per extraction kernel, the dispatches F .. F+N-1 in
launch order (a bench's timed region: it skips the sizing, checking and
warm-up calls before it and the calls after it).

`counters` averages each counter of a ``rocprofv3 --pmc`` run per dispatch,
per kernel (instruction counts: SQ_INSTS_VALU, SQ_INSTS_SALU, SQ_WAVES ...).

`traffic` reads two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs of
the same command (``pm_counter_collection.csv``), scales each counter by what
it reads for the plain copy kernel of the run (torch's copy of a 512 MiB
buffer, ``bench.py --calib-copy``), and sums the extraction kernels (count,
scan, emit) of one call: HBM bytes per extraction.
"""
import argparse
import collections
import csv
import json
import os
import sys

EXTRACT_KERNELS = ("count4_kernel", "count_kernel", "scan_kernel", "emit_kernel",
                   "emit_units_kernel")
COPY_BYTES = 512 << 20


def kernel_stats(path, match=None):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Name"]
            if match and not any(m in name for m in match):
                continue
            rows.append((name, int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                         float(r["TotalDurationNs"]) / 1e3))
    return rows


def window(path, first, count, match=EXTRACT_KERNELS):
    """Per kernel: durations (us) of dispatches first .. first+count-1 in
    start order."""
    per = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if any(m in name for m in match):
                per[name].append((int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    out = []
    for name, ds in per.items():
        ds.sort()
        us = [(e - b) / 1e3 for b, e in ds[first:first + count]]
        if us:
            us_s = sorted(us)
            out.append((name, len(ds), len(us), sum(us) / len(us), us_s[len(us_s) // 2],
                        us_s[0], us_s[-1], ds[first][0]))
    out.sort(key=lambda r: r[-1])
    return out


def _per_kernel(path, counter):
    d = collections.defaultdict(list)
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                d[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return d


def _short(name):
    """Kernel name without return type, namespaces and parameter list."""
    n = name[5:] if name.startswith("void ") else name
    n = n.replace("mcx::(anonymous namespace)::", "").replace("at::native::", "")
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):  # first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return n[:cut][:70]


def counters(path, match=None):
    """Per kernel: dispatches and the mean of every counter per dispatch."""
    d = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if match and not any(m in name for m in match):
                continue
            d[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    out = []
    for name, cs in d.items():
        n = max(len(v) for v in cs.values())
        out.append((name, n, {c: sum(v) / len(v) for c, v in cs.items()}))
    return out


def traffic(fetch_dir, write_dir):
    fetch = _per_kernel(os.path.join(fetch_dir, "pm_counter_collection.csv"), "FETCH_SIZE")
    write = _per_kernel(os.path.join(write_dir, "pm_counter_collection.csv"), "WRITE_SIZE")
    # the copy kernel: the dispatches whose WRITE_SIZE is closest to the copy size
    def copy_scale(d, counter_bytes_expected):
        cands = []
        for name, vals in d.items():
            if "copy" in name.lower():
                for v in vals:
                    cands.append(v * 1024.0)
        best = max(cands) if cands else None
        return (counter_bytes_expected / best) if best else None
    f_scale = copy_scale(fetch, COPY_BYTES)
    w_scale = copy_scale(write, COPY_BYTES)
    out = {"fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    total = 0.0
    for name in sorted(set(fetch) | set(write)):
        short = _short(name)
        if not any(k in name for k in EXTRACT_KERNELS):
            continue
        fv = fetch.get(name, [])
        wv = write.get(name, [])
        fb = sorted(fv)[len(fv) // 2] * 1024.0 * f_scale if fv else 0.0  # median dispatch
        wb = sorted(wv)[len(wv) // 2] * 1024.0 * w_scale if wv else 0.0
        out["kernels"][short] = {"dispatches": len(fv), "fetch_bytes": round(fb),
                                 "write_bytes": round(wb)}
        total += round(fb) + round(wb)  # the sum of the bytes written out
    out["hbm_bytes_per_launch"] = round(total)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--match", nargs="*")
    w = sub.add_parser("window")
    w.add_argument("csv")
    w.add_argument("--first", type=int, required=True)
    w.add_argument("--count", type=int, required=True)
    w.add_argument("--match", nargs="*", help="kernel name parts (default: the extraction kernels)")
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("--match", nargs="*")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--out")
    b.add_argument("--meta", default="{}", help="JSON merged into the output")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print("| kernel | calls | avg us | total us |")
        print("|---|---|---|---|")
        for name, calls, avg, tot in kernel_stats(args.csv, args.match):
            print(f"| {_short(name)} | {calls} | {avg:.2f} | {tot:.1f} |")
        return 0
    if args.cmd == "window":
        rows = window(args.csv, args.first, args.count, tuple(args.match or EXTRACT_KERNELS))
        print("| kernel | dispatches in run | in window | avg us | median us | min us | max us |")
        print("|---|---|---|---|---|---|---|")
        tot = 0.0
        for name, n, k, avg, med, lo, hi, _ in rows:
            print(f"| {_short(name)} | {n} | {k} | {avg:.2f} | {med:.2f} | {lo:.2f} | {hi:.2f} |")
            tot += avg
        print(f"\nsum of the window averages: {tot:.2f} us")
        return 0
    if args.cmd == "counters":
        rows = counters(args.csv, args.match)
        names = sorted({c for _, _, cs in rows for c in cs})
        print("| kernel | dispatches | " + " | ".join(f"{c} per dispatch" for c in names) + " |")
        print("|---|---|" + "---|" * len(names))
        for name, n, cs in rows:
            vals = " | ".join(f"{cs[c]:.0f}" if c in cs else "" for c in names)
            print(f"| {_short(name)} | {n} | {vals} |")
        return 0
    res = traffic(args.fetch_dir, args.write_dir)
    res.update(json.loads(args.meta))
    txt = json.dumps(res, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")
    print(txt)
    return 0


if __name__ == "__main__":
    sys.exit(main())
