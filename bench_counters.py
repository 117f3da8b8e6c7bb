"""Summaries of rocprofv3 counter runs for bench.py (written to profiles/).

bench.py reads two counter summaries and reports them only when they were
taken from the current kernel sources (their "sources_sha16" equals
bench.source_hash()):

* profiles/traffic_2e24.json: per dispatch of each 2^24-multiply kernel, the
  HBM bytes (FETCH_SIZE and WRITE_SIZE, each from its own pass, scaled by
  what the same counter reads for a plain device copy of known size) and
  SQ_INSTS_VALU.
      python bench_counters.py traffic --fetch D --write D --sq D \\
          --copy-fetch D --copy-write D
* profiles/valu_secondary.json: SQ_INSTS_VALU per call of a secondary line,
  summed over the dispatches of `bench.py --only LINE --counter-calls N`
  (the runtime's own copy kernels left out) and divided by its calls.
      python bench_counters.py valu --line LINE --calls FILE --dir D
* a kernel trace (--kernel-trace --stats) as a markdown table:
      python bench_counters.py trace --dir D --out NAME.md --title TITLE

Each D is a rocprofv3 output directory (--output-format csv); the summaries
record the commands given with --cmd.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 1024 * 2.32e9 / 4   # wave-instructions/s: 1024 SIMDs, 4 cycles each, 2.32 GHz
COPY_BYTES = 256 << 20          # bench.py --copy-calibration


def counters(d):
    """{(dispatch id, kernel name): {counter: value}} over every
    *counter_collection.csv under directory d."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    out = collections.defaultdict(dict)
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                key = (f, row["Dispatch_Id"], row["Kernel_Name"])
                name = row["Counter_Name"]
                out[key][name] = out[key].get(name, 0.0) + float(row["Counter_Value"])
    return out


def short(kernel):
    """Kernel name without its argument list (template arguments kept)."""
    depth, end = 0, len(kernel)
    for i, c in enumerate(kernel):
        if c == "<":
            depth += 1
        elif c == ">":
            depth -= 1
        elif c == "(" and depth == 0:
            end = i
            break
    name = kernel[:end].strip()
    if name.startswith("void "):
        name = name[5:]
    return name.replace("cm::", "")


def per_kernel(d, counter):
    """{short kernel name: (mean value per dispatch, dispatches)}."""
    acc = collections.defaultdict(list)
    for (_, _, k), v in counters(d).items():
        if counter in v:
            acc[short(k)].append(v[counter])
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def copy_scale(d, counter):
    """bytes per counter unit (KiB) from the plain copy run."""
    vals = [v for k, v in per_kernel(d, counter).items() if "copy" in k.lower()]
    if not vals:
        raise SystemExit(f"no copy kernel in {d}")
    kib = max(vals, key=lambda x: x[0])[0]
    return COPY_BYTES / (kib * 1024.0), kib


def traffic(args):
    import bench
    fs, fk = copy_scale(args.copy_fetch, "FETCH_SIZE")
    ws, wk = copy_scale(args.copy_write, "WRITE_SIZE")
    fetch = per_kernel(args.fetch, "FETCH_SIZE")
    write = per_kernel(args.write, "WRITE_SIZE")
    sq = per_kernel(args.sq, "SQ_INSTS_VALU")
    n = 1 << 24
    out = {"workload": "2^24 multiply mod 2013265921 (bench.py default)",
           "sources_sha16": bench.source_hash(), "valu_peak_per_s": VALU_PEAK,
           "fetch_scale": fs, "write_scale": ws,
           "copy_calibration": {"bytes": COPY_BYTES, "fetch_kib": fk, "write_kib": wk},
           "commands": args.cmd or [], "kernels": {}}
    for key in ("k_col_fwd", "k_row_mul", "k_col_inv"):
        names = [k for k in fetch if k.startswith(key)]
        if not names:
            continue
        k = names[0]
        fb = fetch[k][0] * 1024 * fs
        wb = write.get(k, (0.0, 0))[0] * 1024 * ws
        algo = bench.algo_bytes(key, n, n // 2, n // 2, n - 1)
        out["kernels"][key] = {"kernel": k, "fetch_bytes": round(fb), "write_bytes": round(wb),
                               "traffic_bytes": round(fb + wb), "algorithmic_bytes": algo,
                               "ratio": round((fb + wb) / algo, 3),
                               "valu_insts": round(sq.get(k, (0.0, 0))[0])}
    path = os.path.join(ROOT, "profiles", "traffic_2e24.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def valu(args):
    import bench
    with open(args.calls) as f:
        info = json.loads(f.read().strip().splitlines()[-1])
    calls = int(info["calls"])
    per = collections.defaultdict(float)
    disp = collections.Counter()
    for (_, _, k), v in counters(args.dir).items():
        if "SQ_INSTS_VALU" not in v or k.startswith("__amd_rocclr"):
            continue
        per[short(k)] += v["SQ_INSTS_VALU"]
        disp[short(k)] += 1
    total = sum(per.values())
    path = os.path.join(ROOT, "profiles", "valu_secondary.json")
    prof = {"lines": {}}
    if os.path.exists(path):
        with open(path) as f:
            prof = json.load(f)
    if prof.get("sources_sha16") != bench.source_hash():
        prof = {"lines": {}}     # every line is re-taken with the current sources
    prof["sources_sha16"] = bench.source_hash()
    prof["valu_peak_per_s"] = VALU_PEAK
    prof["lines"][args.line] = {
        "valu_insts_per_call": total / calls, "calls": calls,
        "dispatches_per_call": sum(disp.values()) / calls,
        "kernels": {k: {"valu_insts_per_call": round(v / calls),
                        "dispatches_per_call": disp[k] / calls}
                    for k, v in sorted(per.items(), key=lambda kv: -kv[1])},
        "commands": args.cmd or []}
    with open(path, "w") as f:
        json.dump(prof, f, indent=1)
    print(json.dumps(prof["lines"][args.line], indent=1))


def trace(args):
    """Markdown table of a rocprofv3 --kernel-trace --stats run (kernel names
    without argument lists; the runtime's own copy kernels included)."""
    files = glob.glob(os.path.join(args.dir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_stats.csv under {args.dir}")
    rows = []
    with open(files[0], newline="") as fh:
        for r in csv.DictReader(fh):
            rows.append((short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e3,
                         float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    rows.sort(key=lambda x: -x[2])
    out = [f"# Kernel trace: {args.title}", "", args.note or "", "",
           "Command (run from /tmp with TMPDIR=/tmp, `R` = repository root):", "", "```",
           *(args.cmd or []), "```", "",
           "| Kernel | Calls | Avg (us) | Total (us) | % |", "|---|---|---|---|---|"]
    out += [f"| `{k}` | {c} | {a:.2f} | {t:.1f} | {p:.1f} |" for k, c, t, a, p in rows[:args.top]]
    with open(os.path.join(ROOT, "profiles", args.out), "w") as f:
        f.write("\n".join(out) + "\n")
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    tr = sub.add_parser("trace")
    tr.add_argument("--dir", required=True)
    tr.add_argument("--out", required=True)
    tr.add_argument("--title", required=True)
    tr.add_argument("--note")
    tr.add_argument("--top", type=int, default=25)
    tr.add_argument("--cmd", action="append")
    t = sub.add_parser("traffic")
    for a in ("--fetch", "--write", "--sq", "--copy-fetch", "--copy-write"):
        t.add_argument(a, required=True)
    t.add_argument("--cmd", action="append")
    v = sub.add_parser("valu")
    v.add_argument("--line", required=True)
    v.add_argument("--calls", required=True)
    v.add_argument("--dir", required=True)
    v.add_argument("--cmd", action="append")
    args = ap.parse_args()
    {"traffic": traffic, "valu": valu, "trace": trace}[args.what](args)


if __name__ == "__main__":
    main()
