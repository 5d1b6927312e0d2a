"""Condense rocprofv3 output into the per-kernel summary bench.py reports.

Inputs (CSV files written by rocprofv3 with `-f csv`):
  --trace  <run_kernel_trace.csv>   from `--kernel-trace`
  --pmc    <run_counter_collection.csv> (repeatable) from `--pmc ...` runs
  --copy-fetch / --copy-write <run_counter_collection.csv> of
           `bench.py --copy-calibration` under `--pmc FETCH_SIZE` /
           `--pmc WRITE_SIZE`: a device copy of known size (1 GiB read,
           1 GiB written per launch) that gives the scale of each counter.

Output: JSON with, per (kernel, grid size): launches, mean duration (ms), the
mean of every collected counter per dispatch, and for FETCH_SIZE /
WRITE_SIZE the scaled HBM bytes per launch; the same over the "large"
launches only (duration >= 0.6 x the longest of the group), which separates
the several-view launches of a persistent kernel from its one-view launches
at the same grid size. Kernels are keyed by short name and total grid size.

Usage:
  python -m splattar_amd.rocprof_summary --trace T.csv --pmc A.csv --pmc B.csv \\
      --copy-fetch CF.csv --copy-write CW.csv --command "..." --out summary.json
"""

from __future__ import annotations

import argparse
import collections
import csv
import json
import os
import re

COPY_BYTES = 1 << 30  # bench.py --copy-calibration: 1 GiB read + 1 GiB written per launch


def short_name(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"^void ", "", name)
    m = re.match(r"([\w:<>, ]+?)\(", name)
    return (m.group(1) if m else name).strip()


def _grid(row) -> int:
    if "Grid_Size" in row:
        return int(float(row["Grid_Size"]))
    return int(row["Grid_Size_X"]) * int(row["Grid_Size_Y"]) * int(row["Grid_Size_Z"])


def read_trace(path):
    out = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        key = (short_name(r["Kernel_Name"]), _grid(r))
        out[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    return out


LARGE = 0.6  # a launch is "large" at >= this fraction of its group's longest


def read_pmc(path, large_only=False):
    """(kernel, grid) -> counter -> list of per-dispatch values."""
    per_dispatch = collections.defaultdict(dict)
    keys = {}
    dur = {}
    for r in csv.DictReader(open(path)):
        d = r["Dispatch_Id"]
        keys[d] = (short_name(r["Kernel_Name"]), _grid(r))
        dur[d] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        per_dispatch[d][r["Counter_Name"]] = per_dispatch[d].get(r["Counter_Name"], 0.0) + float(
            r["Counter_Value"])
    longest = collections.defaultdict(float)
    for d, k in keys.items():
        longest[k] = max(longest[k], dur[d])
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for d, counters in per_dispatch.items():
        if large_only and dur[d] < LARGE * longest[keys[d]]:
            continue
        for c, v in counters.items():
            out[keys[d]][c].append(v)
    return out


def copy_scale(path, counter):
    """Bytes per counter unit: known copy bytes / measured counter (KiB)."""
    vals = [v for (name, _), cs in read_pmc(path).items() if "copyBuffer" in name
            for v in cs.get(counter, [])]
    if not vals:
        raise SystemExit(f"no copy dispatch with {counter} in {path}")
    mean = sum(vals) / len(vals)
    return COPY_BYTES / mean


VALU_PEAK_WINST = 256 * 2.4e9  # wave instructions / s: 256 CUs, one per SIMD every 4 cycles


def markdown_rows(kernels, top=28):
    """The per-kernel table of a summary: mean duration, HBM bytes, VALU
    issue rate against the non-packed peak, the fraction of wave cycles spent
    waiting, LDS instructions per wave and LDS bank-conflict cycles against
    LDS-active cycles (SQ_LDS_BANK_CONFLICT / SQ_ACTIVE_INST_LDS). A
    persistent kernel whose launches span several views also gets a row over
    its large launches (>= 0.6 x its longest)."""
    head = ("| kernel | grid | launches | mean ms | HBM read MB | HBM write MB | VALU G wave-instr/s "
            "(frac of 614) | wait frac | LDS instr / wave | bank-conflict / LDS-active cycles |\n"
            "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    rows = [head]

    def row(name, grid, launches, k):
        cs = k.get("counters_per_dispatch", {})
        ms = k["mean_ms"]

        def f(v, fmt):
            return fmt.format(v) if v is not None else "-"
        rd = k.get("hbm_read_bytes")
        wr = k.get("hbm_write_bytes")
        valu = cs.get("SQ_INSTS_VALU")
        rate = valu / (ms * 1e-3) if valu is not None else None
        wait = (cs["SQ_WAIT_ANY"] / cs["SQ_WAVE_CYCLES"]) if cs.get("SQ_WAVE_CYCLES") and "SQ_WAIT_ANY" in cs \
            else None
        lds = (cs["SQ_INSTS_LDS"] / cs["SQ_WAVES"]) if cs.get("SQ_WAVES") and "SQ_INSTS_LDS" in cs else None
        bank = (cs["SQ_LDS_BANK_CONFLICT"] / cs["SQ_ACTIVE_INST_LDS"]) \
            if cs.get("SQ_ACTIVE_INST_LDS") and "SQ_LDS_BANK_CONFLICT" in cs else None
        vstr = f"{rate / 1e9:.0f} ({rate / VALU_PEAK_WINST:.2f})" if rate is not None else "-"
        rows.append(f"| {name} | {grid} | {launches} | {ms:.4f} | {f(rd and rd / 1e6, '{:.0f}')} | "
                    f"{f(wr and wr / 1e6, '{:.0f}')} | {vstr} | {f(wait, '{:.2f}')} | {f(lds, '{:.0f}')} | "
                    f"{f(bank, '{:.3f}')} |")

    for k in kernels[:top]:
        row(k["kernel"], k["grid"], k["launches"], k)
        lg = k.get("large")
        if lg and lg["launches"] < k["launches"] and "raster_" in k["kernel"]:
            row(k["kernel"] + " (large launches)", k["grid"], lg["launches"], lg)
    return "\n".join(rows) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--trace", required=True)
    ap.add_argument("--pmc", action="append", default=[])
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--command", default="")
    ap.add_argument("--out", required=True)
    ap.add_argument("--markdown", help="also write the per-kernel table (markdown) here")
    a = ap.parse_args(argv)
    trace = read_trace(a.trace)
    counters = collections.defaultdict(dict)
    counters_large = collections.defaultdict(dict)
    for p in a.pmc:
        if not os.path.exists(p):
            print(f"rocprof_summary: no counter file {p}, skipped")
            continue
        for dst, large in ((counters, False), (counters_large, True)):
            for key, cs in read_pmc(p, large).items():
                for c, vals in cs.items():
                    dst[key][c] = sum(vals) / len(vals)
    fetch_scale = copy_scale(a.copy_fetch, "FETCH_SIZE") if a.copy_fetch else None
    write_scale = copy_scale(a.copy_write, "WRITE_SIZE") if a.copy_write else None
    kernels = []
    for (name, grid), durs in sorted(trace.items(), key=lambda kv: -sum(kv[1])):
        k = {"kernel": name, "grid": grid, "launches": len(durs),
             "mean_ms": sum(durs) / len(durs), "total_ms": sum(durs)}
        cs = counters.get((name, grid), {})
        if cs:
            k["counters_per_dispatch"] = cs
        if fetch_scale and "FETCH_SIZE" in cs:
            k["hbm_read_bytes"] = cs["FETCH_SIZE"] * fetch_scale
        if write_scale and "WRITE_SIZE" in cs:
            k["hbm_write_bytes"] = cs["WRITE_SIZE"] * write_scale
        big = [d for d in durs if d >= LARGE * max(durs)]
        lg = {"launches": len(big), "mean_ms": sum(big) / len(big)}
        cl = counters_large.get((name, grid), {})
        if cl:
            lg["counters_per_dispatch"] = cl
        if fetch_scale and "FETCH_SIZE" in cl:
            lg["hbm_read_bytes"] = cl["FETCH_SIZE"] * fetch_scale
        if write_scale and "WRITE_SIZE" in cl:
            lg["hbm_write_bytes"] = cl["WRITE_SIZE"] * write_scale
        k["large"] = lg
        kernels.append(k)
    out = {"command": a.command, "fetch_bytes_per_unit": fetch_scale,
           "write_bytes_per_unit": write_scale, "kernels": kernels}
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    if a.markdown:
        with open(a.markdown, "w") as f:
            f.write(markdown_rows(kernels))


if __name__ == "__main__":
    main()
