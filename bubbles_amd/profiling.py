"""Summaries of rocprofv3 output for the committed profile records.

  python -m bubbles_amd.profiling stats KERNEL_STATS_CSV --command CMD --out OUT.md
  python -m bubbles_amd.profiling traffic FETCH_CSV WRITE_CSV --n N --command CMD \
      --json OUT.json --out OUT.md
  python -m bubbles_amd.profiling sq SQ_CSV --command CMD --out OUT.md

`stats` turns a `rocprofv3 --kernel-trace --stats -f csv` summary into a
markdown table. `traffic` turns two `rocprofv3 --pmc FETCH_SIZE` /
`--pmc WRITE_SIZE` counter collections of the benchmark into HBM bytes per
launch for each kernel.

The two counters are scaled by what they read for `k_stream_copy`, whose
bytes are known: the benchmark's bandwidth probe copies 2^30 bytes, so it
reads 2^30 and writes 2^30 bytes per launch.

The loop kernels are split by role from their own counters: a launch that
writes under 1 B/particle did no work (an earlier iteration had already
integrated); a pressure-force launch that writes over 40 B/particle
integrated; the rest predicted.
"""
from __future__ import annotations

import argparse
import csv
import json
import statistics
from collections import defaultdict

COPY_BYTES = float(2 ** 30)

SHORT = {
    "k_hash": "grid_hash", "k_tile_tables": "tile_tables", "k_scatter": "grid_scatter", "k_rank_move": "grid_rank_move",
    "k_scan_block": "grid_scan", "k_scan_add": "grid_scan_add", "k_scan_add_reduce": "grid_scan_add",
    "k_zero_int2": "grid_clear",
    "k_buckets_density<false, 0>": "buckets_density", "k_buckets_density<true, 0>": "buckets_density_tait",
    "k_buckets_density<false, 1>": "buckets_density_features",
    "k_buckets_density<false, 2>": "buckets_density_warm",
    "k_nonpressure<0>": "nonpressure", "k_nonpressure<1>": "nonpressure_predict",
    "k_nonpressure<2>": "nonpressure_warm",
    "k_mlp_infer": "mlp_infer",
    "k_warm_init": "warm_init", "k_predict_pressure": "predict_pressure",
    "k_pressure_force": "pressure_force", "k_stream_copy": "stream_copy",
    "k_sph_force": "sph_force", "k_integrate": "integrate", "k_xsph": "xsph",
    "k_density": "density", "k_fill_int": "fill",
}


def short_name(kernel: str) -> str:
    base = kernel.split("(")[0].replace("void ", "").strip()
    base = base.split("::")[-1]
    return SHORT.get(base, base)


def loop_kernel_split(trace_csv: str, threshold_us: float = 15.0):
    """Per loop kernel: (working launches, their mean us, early exits, their
    mean us) from a kernel trace.  A launch that finds an earlier iteration
    integrated exits at entry in a few microseconds; a working launch over
    the 1M benchmark takes tens."""
    dur = defaultdict(list)
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            name = short_name(r["Kernel_Name"])
            if name in ("predict_pressure", "pressure_force"):
                dur[name].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = {}
    for name, d in dur.items():
        work = [x for x in d if x >= threshold_us]
        skip = [x for x in d if x < threshold_us]
        out[name] = (len(work), statistics.mean(work) if work else 0.0, len(skip),
                     statistics.mean(skip) if skip else 0.0)
    return out


def stats_markdown(path: str, command: str, trace_csv: str = None) -> str:
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                         float(r["TotalDurationNs"]) / 1e6, float(r["Percentage"])))
    lines = ["# rocprofv3 kernel statistics", "", "Command:", "", "```", command, "```", "",
             "| kernel | calls | avg (us) | total (ms) | % of kernel time |",
             "|---|---|---|---|---|"]
    for name, calls, avg, tot, pct in rows:
        lines.append(f"| {name} | {calls} | {avg:.2f} | {tot:.3f} | {pct:.2f} |")
    lines.append("")
    lines.append("Loop kernels (`predict_pressure`, `pressure_force`) average over all of "
                 "their launches, including those that exit at entry once an earlier "
                 "iteration has integrated.")
    if trace_csv:
        lines += ["", "The same launches split by duration (from the kernel trace, threshold 15 us):", "",
                  "| kernel | working launches | avg (us) | early exits | avg (us) |", "|---|---|---|---|---|"]
        for name, (nw, aw, ns, as_) in sorted(loop_kernel_split(trace_csv).items()):
            lines.append(f"| {name} | {nw} | {aw:.2f} | {ns} | {as_:.2f} |")
    return "\n".join(lines) + "\n"


def _counter_rows(path: str, counter: str):
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            out.append((int(r["Dispatch_Id"]), short_name(r["Kernel_Name"]), float(r["Counter_Value"])))
    out.sort()
    return out


def traffic(fetch_csv: str, write_csv: str, n: int):
    fetch = _counter_rows(fetch_csv, "FETCH_SIZE")
    write = _counter_rows(write_csv, "WRITE_SIZE")
    if len(fetch) != len(write) or any(a[1] != b[1] for a, b in zip(fetch, write)):
        raise ValueError("the two counter runs launched different kernel sequences")
    copy = [(f[2], w[2]) for f, w in zip(fetch, write) if f[1] == "stream_copy"]
    if not copy:
        raise ValueError("no stream_copy dispatch to calibrate against")
    kf = COPY_BYTES / (statistics.median(c[0] for c in copy) * 1024.0)
    kw = COPY_BYTES / (statistics.median(c[1] for c in copy) * 1024.0)
    per = defaultdict(list)
    for (_, name, fv), (_, _, wv) in zip(fetch, write):
        fb, wb = fv * 1024.0 * kf, wv * 1024.0 * kw
        if name in ("predict_pressure", "pressure_force"):
            if wb < n:
                name += "_skipped"
            elif name == "pressure_force":
                name += "_integrate" if wb > 40.0 * n else "_predict"
        per[name].append((fb, wb))
    kernels = {}
    for name, v in per.items():
        fb = statistics.median(x[0] for x in v)
        wb = statistics.median(x[1] for x in v)
        kernels[name] = {"dispatches": len(v), "fetch_bytes": round(fb), "write_bytes": round(wb),
                         "bytes_per_launch": round(fb + wb),
                         "bytes_per_particle": round((fb + wb) / n, 2)}
    return {"n": n, "fetch_scale": round(kf, 4), "write_scale": round(kw, 4), "kernels": kernels}


def traffic_markdown(rec: dict, command: str, algorithmic: dict | None = None) -> str:
    lines = ["# HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE)", "",
             "Commands (one counter per run):", "", "```", command, "```", "",
             f"Calibration: `stream_copy` reads and writes 2^30 bytes per launch; "
             f"FETCH_SIZE x {rec['fetch_scale']}, WRITE_SIZE x {rec['write_scale']} "
             f"(KiB counters scaled to those bytes). n = {rec['n']} particles.", "",
             "| kernel | dispatches | fetch (MB) | write (MB) | B/particle | algorithmic B/particle |",
             "|---|---|---|---|---|---|"]
    for name, k in sorted(rec["kernels"].items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
        alg = (algorithmic or {}).get(name, "")
        lines.append(f"| {name} | {k['dispatches']} | {k['fetch_bytes'] / 1e6:.1f} | "
                     f"{k['write_bytes'] / 1e6:.1f} | {k['bytes_per_particle']} | {alg} |")
    if "buckets_density" in rec["kernels"]:
        lines += ["", "The algorithmic figure of `buckets_density` leaves out the bucket it writes: "
                  "2 B more per bucket entry (the benchmark line's `bytes_per_particle` adds them "
                  "at the measured mean entry count)."]
    return "\n".join(lines) + "\n"


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_VMEM",
               "SQ_WAVE_CYCLES", "SQ_WAIT_ANY", "SQ_BUSY_CYCLES")


def sq_summary(path: str, min_waves: int = 1000):
    """Per kernel: mean duration and the SQ counters per wave, over the
    dispatches with at least `min_waves` waves (launches that exit at entry
    are left out)."""
    per = defaultdict(dict)
    for r in csv.DictReader(open(path)):
        key = (short_name(r["Kernel_Name"]), int(r["Dispatch_Id"]))
        per[key][r["Counter_Name"]] = float(r["Counter_Value"])
        per[key]["_us"] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    agg, cnt = defaultdict(lambda: defaultdict(float)), defaultdict(int)
    for (name, _), c in per.items():
        if c.get("SQ_WAVES", 0) < min_waves:
            continue
        cnt[name] += 1
        for k, v in c.items():
            agg[name][k] += v
    out = {}
    for name, c in agg.items():
        w = c["SQ_WAVES"]
        row = {"dispatches": cnt[name], "us": c["_us"] / cnt[name], "waves": w / cnt[name]}
        for k in ("SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_VMEM"):
            if k in c:
                row[k] = c[k] / w
        if "SQ_WAIT_ANY" in c and c.get("SQ_WAVE_CYCLES"):
            row["wait_frac"] = c["SQ_WAIT_ANY"] / c["SQ_WAVE_CYCLES"]
        out[name] = row
    return out


def sq_markdown(rows: dict, command: str, simds: int = 1024, clock_ghz: float = 2.4) -> str:
    lines = ["# SQ counters per kernel (instructions per wave)", "", "Command:", "", "```", command,
             "```", "",
             "`VALU share` = VALU instructions x 4 cycles (a wave64 instruction on a 16-lane SIMD) "
             f"x waves / ({simds} SIMDs x {clock_ghz} GHz), divided by the kernel's duration under "
             "the profiler: how much of the kernel the vector ALU alone would need.", "",
             "| kernel | dispatches | us | waves | VALU/wave | LDS/wave | SALU/wave | VMEM/wave | wait % | VALU share |",
             "|---|---|---|---|---|---|---|---|---|---|"]
    for name, r in sorted(rows.items(), key=lambda kv: -kv[1]["us"] * kv[1]["dispatches"]):
        valu = r.get("SQ_INSTS_VALU", 0.0)
        share = valu * 4 * r["waves"] / (simds * clock_ghz * 1e3) / r["us"] if r["us"] else 0.0
        lines.append(f"| {name} | {r['dispatches']} | {r['us']:.1f} | {r['waves']:.0f} | {valu:.0f} | "
                     f"{r.get('SQ_INSTS_LDS', 0):.0f} | {r.get('SQ_INSTS_SALU', 0):.0f} | "
                     f"{r.get('SQ_INSTS_VMEM', 0):.0f} | {100 * r.get('wait_frac', 0):.0f} | {share:.2f} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    st = sub.add_parser("stats")
    st.add_argument("csv")
    st.add_argument("--command", required=True)
    st.add_argument("--out", required=True)
    st.add_argument("--trace", default=None, help="kernel trace CSV: split loop-kernel launches")
    tr = sub.add_parser("traffic")
    tr.add_argument("fetch_csv")
    tr.add_argument("write_csv")
    tr.add_argument("--n", type=int, required=True)
    tr.add_argument("--command", required=True)
    tr.add_argument("--json", required=True)
    tr.add_argument("--out", required=True)
    sq = sub.add_parser("sq")
    sq.add_argument("csv")
    sq.add_argument("--command", required=True)
    sq.add_argument("--out", required=True)
    args = ap.parse_args(argv)
    if args.cmd == "sq":
        with open(args.out, "w") as f:
            f.write(sq_markdown(sq_summary(args.csv), args.command))
    elif args.cmd == "stats":
        with open(args.out, "w") as f:
            f.write(stats_markdown(args.csv, args.command, args.trace))
    else:
        rec = traffic(args.fetch_csv, args.write_csv, args.n)
        rec["command"] = args.command
        with open(args.json, "w") as f:
            json.dump(rec, f, indent=1)
        try:
            import sys
            import os

            sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
            from bench import KERNEL_BYTES as alg
        except Exception:  # noqa: BLE001 - the table column is optional
            alg = None
        with open(args.out, "w") as f:
            f.write(traffic_markdown(rec, args.command, alg))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
