"""Summaries of rocprofv3 CSV output for the detection kernels.

    python -m megastation_amd.profiling stats  <dir with *_kernel_stats.csv>
    python -m megastation_amd.profiling traffic <FETCH_SIZE dir> <WRITE_SIZE dir> \
        [--copy-bytes N] [--json out.json]
    python -m megastation_amd.profiling counters <dir with *_counter_collection.csv>
    python -m megastation_amd.profiling sqtable <dir of an SQ counter pass>

`traffic` turns per-dispatch FETCH_SIZE / WRITE_SIZE counters (KiB) into HBM
bytes per launch. Each counter is scaled by what it reads for a plain copy of
known size recorded in the same run (the runtime's __amd_rocclr_copyBuffer of
`--copy-bytes`, bench.py --copy-calibration): factor = copy_bytes / (counter
KiB * 1024) for the copy.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import statistics

MEGA = "mega::detail"


def _short(name: str) -> str:
    if "__amd_rocclr_copyBuffer" in name:
        return "copyBuffer"
    if MEGA in name:
        m = re.search(r"::(\w+(?:<[^>]*>)?)\(", name)
        return m.group(1) if m else name[:40]
    return name[:40]


def kernel_stats(d: str) -> list[dict]:
    path = glob.glob(os.path.join(d, "*kernel_stats.csv"))[0]
    out = []
    for r in csv.DictReader(open(path)):
        out.append({"kernel": _short(r["Name"]), "calls": int(r["Calls"]),
                    "avg_us": float(r["AverageNs"]) / 1e3,
                    "min_us": float(r["MinNs"]) / 1e3, "max_us": float(r["MaxNs"]) / 1e3,
                    "pct": float(r["Percentage"])})
    return out


def counter_by_kernel(d: str) -> dict[str, list[float]]:
    path = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
    agg: dict[str, list[float]] = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        agg[(_short(r["Kernel_Name"]), r["Counter_Name"])].append(float(r["Counter_Value"]))
    return agg


def traffic(fetch_dir: str, write_dir: str, copy_bytes: int) -> dict:
    f = counter_by_kernel(fetch_dir)
    w = counter_by_kernel(write_dir)
    # the calibration copy is the largest copyBuffer dispatch of the run
    copy_f = max(f[("copyBuffer", "FETCH_SIZE")])
    copy_w = max(w[("copyBuffer", "WRITE_SIZE")])
    kf = copy_bytes / (copy_f * 1024)
    kw = copy_bytes / (copy_w * 1024)
    res = {"copy_bytes": copy_bytes, "copy_fetch_kib": copy_f, "copy_write_kib": copy_w,
           "fetch_scale": kf, "write_scale": kw, "kernels": {}}
    for (k, c), vals in f.items():
        if not k.startswith(("equalize", "zf", "detect", "precode", "fft")):
            continue
        wv = w.get((k, "WRITE_SIZE"), [0.0])
        fb = statistics.median(vals) * 1024 * kf
        wb = statistics.median(wv) * 1024 * kw
        res["kernels"][k] = {"fetch_kib": statistics.median(vals),
                             "write_kib": statistics.median(wv),
                             "read_bytes": fb, "write_bytes": wb,
                             "traffic_bytes": fb + wb, "dispatches": len(vals)}
    return res


# MI355X: 8 XCDs, 256 CUs, 4 SIMDs per CU, 2.4 GHz peak engine clock.
XCDS, SIMDS, CLOCK_GHZ = 8, 1024, 2.4


def sq_table(d: str, prefixes: tuple) -> str:
    """Per-kernel medians of one SQ counter pass and the ratios read from them:
      kernel cycles  = GRBM_GUI_ACTIVE / 8 (the counter is summed over the
                       eight XCDs; / 2.4 GHz gives the duration)
      MFMA busy      = SQ_VALU_MFMA_BUSY_CYCLES / (1024 SIMDs * kernel cycles)
                       (the counter sums matrix-core busy cycles over SIMDs)
      VALU active    = SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES (per wave-cycle)
      wait           = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES (share of wave time
                       spent waiting for an operand: memory, LDS, MFMA result)
      waves resident = 4 * SQ_WAVE_CYCLES / kernel cycles / 256 (per CU,
                       average; the SQ wave counters count quad-cycles)
      LDS conflicts  = SQ_LDS_BANK_CONFLICT / SQ_INSTS_LDS (cycles per LDS
                       instruction)"""
    agg = counter_by_kernel(d)
    kernels = sorted({k for (k, _) in agg if k.startswith(prefixes)})
    med = lambda k, c: statistics.median(agg[(k, c)]) if (k, c) in agg else float("nan")
    rows = ["| kernel | dispatches | kernel us | MFMA busy | VALU active | wait | "
            "waves/CU | LDS conflict cyc/instr |", "|---|---|---|---|---|---|---|---|"]
    for k in kernels:
        gui = med(k, "GRBM_GUI_ACTIVE") / XCDS
        wc = med(k, "SQ_WAVE_CYCLES")
        lds = med(k, "SQ_INSTS_LDS")
        conf = med(k, "SQ_LDS_BANK_CONFLICT")
        rows.append(
            f"| {k} | {len(agg[(k, 'SQ_WAVE_CYCLES')])} | {gui / CLOCK_GHZ / 1e3:.1f} "
            f"| {med(k, 'SQ_VALU_MFMA_BUSY_CYCLES') / (SIMDS * gui):.3f} "
            f"| {med(k, 'SQ_ACTIVE_INST_VALU') / wc:.3f} "
            f"| {med(k, 'SQ_WAIT_INST_ANY') / wc:.3f} | {4 * wc / gui / 256:.2f} "
            f"| {(conf / lds) if lds else 0.0:.2f} |")
    return "\n".join(rows)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-bytes", type=int, default=1 << 30)
    t.add_argument("--json")
    t.add_argument("--workload", help="with --update: the workload key to replace")
    t.add_argument("--update", help="traffic file (profiles/traffic.json) whose entry for "
                                    "--workload is replaced by this measurement")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    q = sub.add_parser("sqtable")
    q.add_argument("dir")
    q.add_argument("--prefix", default="equalize,zf,detect,fft,precode,ldpc")
    a = ap.parse_args()
    if a.cmd == "stats":
        print("| kernel | calls | avg us | min us | max us | % time |")
        print("|---|---|---|---|---|---|")
        for r in kernel_stats(a.dir):
            print(f"| {r['kernel']} | {r['calls']} | {r['avg_us']:.1f} | {r['min_us']:.1f} "
                  f"| {r['max_us']:.1f} | {r['pct']:.1f} |")
    elif a.cmd == "traffic":
        res = traffic(a.fetch_dir, a.write_dir, a.copy_bytes)
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(res, fh, indent=1)
        if a.update and a.workload:
            # bench.py reads {workload: {kernel: {traffic_bytes, ...}, _calibration}}
            db = json.load(open(a.update)) if os.path.exists(a.update) else {}
            entry = {k: {"traffic_bytes": v["traffic_bytes"], "read_bytes": v["read_bytes"],
                         "write_bytes": v["write_bytes"], "fetch_kib": v["fetch_kib"],
                         "write_kib": v["write_kib"]} for k, v in res["kernels"].items()}
            entry["_calibration"] = {"copy_bytes": res["copy_bytes"],
                                     "fetch_scale": res["fetch_scale"],
                                     "write_scale": res["write_scale"]}
            db[a.workload] = entry
            with open(a.update, "w") as fh:
                json.dump(db, fh, indent=1)
        print(json.dumps(res, indent=1))
    elif a.cmd == "sqtable":
        print(sq_table(a.dir, tuple(a.prefix.split(","))))
    else:
        for (k, cn), vals in sorted(counter_by_kernel(a.dir).items()):
            print(f"{k:40s} {cn:28s} n={len(vals)} median={statistics.median(vals):.6g}")


if __name__ == "__main__":
    main()
