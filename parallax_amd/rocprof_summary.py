"""Summarise rocprofv3 CSV output of a bench.py --profile run.

    python -m parallax_amd.rocprof_summary --trace DIR [--fetch DIR] [--write DIR]
        [--batch-views B] [--width W] [--height H]

--trace: a `rocprofv3 --kernel-trace --stats --output-format csv` directory
(kernel_stats.csv); --fetch / --write: `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
runs (counter_collection.csv, KiB per dispatch). HBM traffic is calibrated
on kernels of known size in the same run: the device-to-device target copy
(__amd_rocclr_copyBuffer, reads 3*W*H*4 bytes) for FETCH_SIZE and the
parameter-clear fill (k_fill_f32 over the capacity buffer, whose grid size is
its element count) for WRITE_SIZE.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re


def short(name: str) -> str:
    m = re.search(r"(k_\w+|__amd_rocclr_\w+)", name)
    base = m.group(1) if m else name[:40]
    t = re.search(r"k_\w+<(\d)>", name)
    return base + (f"<{t.group(1)}>" if t else "")


def read_stats(d):
    f = glob.glob(os.path.join(d, "*kernel_stats.csv"))[0]
    out = {}
    for row in csv.DictReader(open(f)):
        k = short(row["Name"])
        calls = int(row["Calls"])
        tot = float(row["TotalDurationNs"])
        prev = out.get(k, (0, 0.0))
        out[k] = (prev[0] + calls, prev[1] + tot)
    return out


def read_counters(d):
    f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
    per = {}
    for row in csv.DictReader(open(f)):
        k = short(row["Kernel_Name"])
        per.setdefault(k, []).append((float(row["Counter_Value"]), int(row["Grid_Size"])))
    return per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--batch-views", type=int, default=8)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    a = ap.parse_args()
    stats = read_stats(a.trace)
    fetch = read_counters(a.fetch) if a.fetch else {}
    write = read_counters(a.write) if a.write else {}
    f_scale = w_scale = None
    if "__amd_rocclr_copyBuffer" in fetch:
        known = 3 * a.width * a.height * 4 / 1024.0
        vals = [v for v, g in fetch["__amd_rocclr_copyBuffer"] if v > 0.5 * known]
        if vals:
            f_scale = known / (sum(vals) / len(vals))
    if "k_fill_f32" in write:
        big = max(write["k_fill_f32"], key=lambda vg: vg[1])
        known = big[1] * 4 / 1024.0
        w_scale = known / big[0] if big[0] > 0 else None
    print(f"calibration: FETCH_SIZE x {f_scale:.4f}" if f_scale else "calibration: FETCH_SIZE uncalibrated")
    print(f"calibration: WRITE_SIZE x {w_scale:.4f}" if w_scale else "calibration: WRITE_SIZE uncalibrated")
    total = sum(t for _, t in stats.values())
    print(f"{'kernel':28s} {'calls':>6s} {'avg_us':>10s} {'share':>7s} {'fetch_MB':>10s} {'write_MB':>10s}")
    for k, (calls, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        def avg_mb(src, scale):
            if k not in src:
                return ""
            v = sum(x for x, _ in src[k]) / len(src[k]) / 1024.0
            return f"{v * (scale or 1.0):10.2f}"
        print(f"{k:28s} {calls:6d} {tot / calls / 1e3:10.2f} {100 * tot / total:6.1f}% "
              f"{avg_mb(fetch, f_scale):>10s} {avg_mb(write, w_scale):>10s}")


if __name__ == "__main__":
    main()
