"""HBM traffic per kernel launch from two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE).

Both counters are read per dispatch (KiB). They are converted to bytes with a factor taken
from a plain device copy of known size that bench.py runs inside the same profiled process
(`--calib-copy-mib`): factor = bytes the copy must move / bytes the counter reports for it.

  rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- \\
      python bench.py --steps 4 --warmup 2 --timed-steps 2 --no-cpu --no-graph --calib-copy-mib 512
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- (same)
  python -m morphogenesis_amd.traffic OUT/fetch OUT/write --workload tissue_262k --calib-mib 512

writes profiles/traffic_<kernel>_<workload>.json for each step kernel and a Markdown summary
profiles/traffic_<workload>.md.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys
from collections import defaultdict

STEP_KERNELS = ("k_pressure", "k_force_advance", "k_diffuse_genes_insert", "k_scan_lookback",
                "k_place", "k_sort_gather")


def short_name(name: str) -> str:
    n = name.split("(")[0].strip()
    n = n.split("<")[0]
    return n.split("::")[-1].split(" ")[-1]


def read_counter(run_dir: str, counter: str) -> dict:
    """{kernel full name: [values per dispatch]} for one counter."""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    out = defaultdict(list)
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name", "") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                out[name].append((int(row.get("Grid_Size", 0) or 0), float(row["Counter_Value"])))
    return out


def calib_factor(data: dict, calib_bytes: float) -> tuple:
    """Factor from the calibration copy: the largest dispatches of a kernel named *copy* (the
    runtime's copyBuffer kernel serves torch's device copy; bench.py runs it three times),
    averaged over those within 5% of the largest."""
    cands = [(v, name) for name, vals in data.items() if "copy" in name.lower() for _, v in vals]
    if not cands:
        raise RuntimeError("calibration copy kernel not found in the counter run")
    vmax, name = max(cands)
    top = [v for v, n in cands if v >= 0.95 * vmax]
    mean_kib = sum(top) / len(top)
    return calib_bytes / (mean_kib * 1024.0), name, mean_kib


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_dir")
    ap.add_argument("write_dir")
    ap.add_argument("--workload", default="tissue_262k")
    ap.add_argument("--calib-mib", type=float, default=512.0)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.dirname(
        os.path.abspath(__file__))), "profiles"))
    ap.add_argument("--command", default="", help="the profiled command, recorded in the summary")
    a = ap.parse_args(argv)
    fetch = read_counter(a.fetch_dir, "FETCH_SIZE")
    write = read_counter(a.write_dir, "WRITE_SIZE")
    nbytes = a.calib_mib * 1024 * 1024
    f_fetch, cname, c_fetch = calib_factor(fetch, nbytes)
    f_write, _, c_write = calib_factor(write, nbytes)
    rows = []
    os.makedirs(a.out, exist_ok=True)
    for name in sorted(set(fetch) | set(write)):
        k = short_name(name)
        if k not in STEP_KERNELS:
            continue
        fv = [v for _, v in fetch.get(name, [])]
        wv = [v for _, v in write.get(name, [])]
        if not fv or not wv:
            continue
        fb = sum(fv) / len(fv) * 1024.0 * f_fetch
        wb = sum(wv) / len(wv) * 1024.0 * f_write
        rec = {"kernel": k, "workload": a.workload, "bytes_per_launch": fb + wb,
               "fetch_bytes": fb, "write_bytes": wb, "dispatches": [len(fv), len(wv)],
               "fetch_kib_raw": sum(fv) / len(fv), "write_kib_raw": sum(wv) / len(wv),
               "calibration": {"kernel": cname, "bytes_each_way": nbytes,
                               "fetch_kib": c_fetch, "write_kib": c_write,
                               "fetch_factor": f_fetch, "write_factor": f_write}}
        if k in [r["kernel"] for r in rows]:
            continue
        rows.append(rec)
        with open(os.path.join(a.out, f"traffic_{k}_{a.workload}.json"), "w") as fh:
            json.dump(rec, fh, indent=1)
    md = [f"# HBM traffic per launch, {a.workload}", "",
          "FETCH_SIZE and WRITE_SIZE per dispatch (rocprofv3 --pmc, one counter per run), "
          "scaled by a copy of known size in the same process.", "",
          f"Calibration: `{cname}` moving {nbytes / 2**20:.0f} MiB each way read "
          f"FETCH_SIZE {c_fetch:.0f} KiB (factor {f_fetch:.4f}), WRITE_SIZE {c_write:.0f} KiB "
          f"(factor {f_write:.4f}).", ""]
    if a.command:
        md += ["Commands:", "", "```", a.command, "```", ""]
    md += ["| kernel | fetch MB | write MB | total MB/launch | dispatches |",
           "|---|---|---|---|---|"]
    for r in rows:
        md.append(f"| {r['kernel']} | {r['fetch_bytes'] / 1e6:.2f} | {r['write_bytes'] / 1e6:.2f} | "
                  f"{r['bytes_per_launch'] / 1e6:.2f} | {r['dispatches'][0]} |")
    with open(os.path.join(a.out, f"traffic_{a.workload}.md"), "w") as fh:
        fh.write("\n".join(md) + "\n")
    print("\n".join(md))
    return 0


if __name__ == "__main__":
    sys.exit(main())
