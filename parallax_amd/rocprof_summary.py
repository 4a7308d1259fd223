"""Summarise rocprofv3 CSV output of a bench.py --profile run.

    python -m parallax_amd.rocprof_summary --trace DIR [--fetch DIR] [--write DIR]
        [--calib-bytes BYTES]

--trace: a `rocprofv3 --kernel-trace --stats --output-format csv` directory
(kernel_stats.csv); --fetch / --write: `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE`
runs (counter_collection.csv, KiB per dispatch). HBM traffic is calibrated on
the plain device-to-device copy of known size that `bench.py --profile` makes
after its timed steps (torch's copy, the __amd_rocclr_copyBuffer kernel, which
reads and writes --calib-bytes): the scale of each counter is the copy's true
bytes over what the counter read for it.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re


def short(name: str) -> str:
    """Kernel name without namespaces and parameters, template arguments kept
    (k_backward<64,true> and the autotuner's other variants stay apart)."""
    m = re.search(r"(k_\w+|__amd_rocclr_\w+)(<[^>]*>)?", name)
    if not m:
        return name[:40]
    return m.group(1) + (m.group(2).replace(" ", "") if m.group(2) else "")


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


def copy_scale(src, calib_bytes):
    """True bytes / counter bytes for the calibration copy, or None."""
    known = calib_bytes / 1024.0  # the counters read KiB
    vals = [v for v, _ in src.get("__amd_rocclr_copyBuffer", []) if 0.25 * known < v < 4.0 * known]
    if not vals:
        return None
    return known / max(vals)


def summarise(trace, fetch_dir=None, write_dir=None, calib_bytes=256 << 20):
    """Lines of the summary table: calibration, then one row per kernel."""
    stats = read_stats(trace)
    fetch = read_counters(fetch_dir) if fetch_dir else {}
    write = read_counters(write_dir) if write_dir else {}
    f_scale = copy_scale(fetch, calib_bytes)
    w_scale = copy_scale(write, calib_bytes)
    lines = [f"calibration: FETCH_SIZE x {f_scale:.4f}" if f_scale else "calibration: FETCH_SIZE uncalibrated",
             f"calibration: WRITE_SIZE x {w_scale:.4f}" if w_scale else "calibration: WRITE_SIZE uncalibrated",
             f"(calibration kernel: __amd_rocclr_copyBuffer, {calib_bytes / 2**20:.0f} MiB read + written)"]
    total = sum(t for _, t in stats.values())
    lines.append(f"{'kernel':28s} {'calls':>6s} {'avg_us':>10s} {'share':>7s} {'fetch_MB':>10s} {'write_MB':>10s}")
    for k, (calls, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        def avg_mb(src, scale):
            if k not in src:
                return ""
            v = sum(x for x, _ in src[k]) / len(src[k]) * 1024.0 / 1e6
            return f"{v * (scale or 1.0):10.2f}"
        lines.append(f"{k:28s} {calls:6d} {tot / calls / 1e3:10.2f} {100 * tot / total:6.1f}% "
                     f"{avg_mb(fetch, f_scale):>10s} {avg_mb(write, w_scale):>10s}")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--calib-bytes", type=int, default=256 << 20)
    a = ap.parse_args()
    print("\n".join(summarise(a.trace, a.fetch, a.write, a.calib_bytes)))


if __name__ == "__main__":
    main()
