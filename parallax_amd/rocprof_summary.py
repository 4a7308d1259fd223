"""Summarise rocprofv3 CSV output of a bench.py --profile run.

    python -m parallax_amd.rocprof_summary --trace DIR [--fetch DIR] [--write DIR]
        [--calib-bytes BYTES]
    python -m parallax_amd.rocprof_summary --sq DIR [DIR ...] [--kernels NAME ...]
    python -m parallax_amd.rocprof_summary --trace DIR --split KERNEL

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


def sq_summary(dirs, kernels=None):
    """Per kernel: each SQ/GRBM counter of the `--pmc` passes in `dirs`
    (averaged over the kernel's dispatches), then the derived rates:
    cycles = GRBM_GUI_ACTIVE / 8 XCDs; valu_ipc = SQ_INSTS_VALU / 1024 SIMDs /
    cycles (a wave64 FP32 VALU op holds a gfx950 SIMD for 2 cycles, so 0.5 is
    every cycle used); salu/valu, lds/valu; wait_frac = SQ_WAIT_ANY /
    SQ_WAVE_CYCLES; LDS bank-conflict cycles per LDS-active cycle."""
    per = {}
    for d in dirs:
        f = glob.glob(os.path.join(d, "*counter_collection.csv"))[0]
        acc = {}
        for row in csv.DictReader(open(f)):
            acc.setdefault(short(row["Kernel_Name"]), {}).setdefault(row["Counter_Name"], []).append(
                float(row["Counter_Value"]))
        for k, cs in acc.items():
            for c, v in cs.items():
                per.setdefault(k, {})[c] = sum(v) / len(v)
    lines = []
    for k in kernels or sorted(per):
        c = per.get(k)
        if not c:
            lines.append(f"== {k}: no dispatches")
            continue
        lines.append(f"== {k}")
        lines += [f"   {n:30s} {v:12.4g}" for n, v in sorted(c.items())]
        cyc = c.get("GRBM_GUI_ACTIVE", 0.0) / 8
        if cyc and "SQ_INSTS_VALU" in c:
            valu = c["SQ_INSTS_VALU"]
            parts = [f"cycles {cyc:.4g}", f"valu_ipc {valu / 1024 / cyc:.3f} (of 0.5)"]
            if "SQ_INSTS_SALU" in c:
                parts.append(f"salu/valu {c['SQ_INSTS_SALU'] / max(valu, 1):.3f}")
            if "SQ_INSTS_LDS" in c:
                parts.append(f"lds/valu {c['SQ_INSTS_LDS'] / max(valu, 1):.3f}")
            if "SQ_WAIT_ANY" in c and c.get("SQ_WAVE_CYCLES"):
                parts.append(f"wait_frac {c['SQ_WAIT_ANY'] / c['SQ_WAVE_CYCLES']:.2f}")
            if "SQ_LDS_BANK_CONFLICT" in c and c.get("SQ_ACTIVE_INST_LDS"):
                parts.append(f"lds_bank_conflict/lds_active {c['SQ_LDS_BANK_CONFLICT'] / c['SQ_ACTIVE_INST_LDS']:.3f}")
            lines.append("   -> " + ", ".join(parts))
    return lines


def split_trace(trace, name):
    """Per-kernel dispatch counts and average durations of a kernel trace
    (kernel_trace.csv), in two parts: the dispatches up to and including the
    last dispatch of `name`, and those after it (a run that switches modes
    after its last use of that kernel)."""
    f = glob.glob(os.path.join(trace, "*kernel_trace.csv"))[0]
    rows = sorted(csv.DictReader(open(f)), key=lambda r: int(r["Start_Timestamp"]))
    last = max((i for i, r in enumerate(rows) if short(r["Kernel_Name"]) == name), default=-1)
    lines = []
    for title, part in ((f"up to the last {name}", rows[:last + 1]), (f"after the last {name}", rows[last + 1:])):
        acc = {}
        for r in part:
            k = short(r["Kernel_Name"])
            d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
            c, t = acc.get(k, (0, 0.0))
            acc[k] = (c + 1, t + d)
        lines.append(f"-- {title}")
        lines.append(f"{'kernel':28s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s}")
        for k, (c, t) in sorted(acc.items(), key=lambda kv: -kv[1][1]):
            lines.append(f"{k:28s} {c:6d} {t / c:10.2f} {t / 1e3:10.3f}")
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--calib-bytes", type=int, default=256 << 20)
    ap.add_argument("--sq", nargs="+", help="--pmc pass directories of SQ/GRBM counters")
    ap.add_argument("--kernels", nargs="+", help="with --sq: these kernels, in this order")
    ap.add_argument("--split", help="with --trace: per-dispatch averages before/after the last dispatch of this kernel")
    a = ap.parse_args()
    if a.split:
        print("\n".join(split_trace(a.trace, a.split)))
        return
    if a.sq:
        print("\n".join(sq_summary(a.sq, a.kernels)))
        return
    if not a.trace:
        ap.error("--trace or --sq is required")
    print("\n".join(summarise(a.trace, a.fetch, a.write, a.calib_bytes)))


if __name__ == "__main__":
    main()
