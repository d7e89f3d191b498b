"""HBM traffic per kernel launch from rocprofv3 counter runs.

Reads the counter_collection.csv of one `rocprofv3 --pmc FETCH_SIZE` run and one
`rocprofv3 --pmc WRITE_SIZE` run of the same command, groups dispatches by kernel class
(the classes the solver's live profiler uses), scales each counter by the factor a
streaming copy of known size reads for it (FETCH_SIZE and WRITE_SIZE are KiB per
dispatch), and stores them in asam_amd/traffic.json under the workload's name (entries of
other workloads are kept), which bench.py reports as roofline.traffic.

    python -m asam_amd.traffic FETCH_DIR WRITE_DIR CALIB_FETCH_DIR CALIB_WRITE_DIR \
        --workload "..." --command "..."

The calibration runs must contain launches of the copy kernel (k_copy16) of a known size
(bench.py --kernel-probe copies 1 GiB).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "traffic.json")


def kernel_class(name: str):
    """Live-profiler class of a kernel name (None: not a solver kernel)."""
    if "k_copy16" in name:
        return "copy"
    m = re.search(r"k_(apply\w*|line)<[^>]*?(double|float)", name)
    if m:
        kind = "apply" if m.group(1).startswith("apply") else "line"
        # the first floating type in the template list is the coefficient type
        return f"{kind}_{'fp64' if m.group(2) == 'double' else 'fp32'}"
    m = re.search(r"k_(bicg_update|mgs|gmres_update|cgs_dots|cgs_update)<(double|float)[,>]",
                  name)
    if m:
        return f"blas_{'fp64' if m.group(2) == 'double' else 'fp32'}"
    if "k_fill<float>" in name:
        # the FP32 correction's zero fill at each refinement step, which the live profiler
        # counts with the FP32 BLAS class
        return "blas_fp32"
    if "k_finalize" in name:
        return "reduce"
    return None


def read_counter(dirname: str):
    files = glob.glob(os.path.join(dirname, "*counter_collection.csv"))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {dirname}")
    per = collections.defaultdict(list)
    for row in csv.DictReader(open(files[0])):
        cls = kernel_class(row["Kernel_Name"])
        if cls:
            per[cls].append(float(row["Counter_Value"]))
    return per


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("calib_fetch")
    ap.add_argument("calib_write")
    ap.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    ap.add_argument("--workload", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--out", default=OUT, help="the JSON file to update (default: traffic.json "
                                               "beside this module)")
    a = ap.parse_args(argv)
    cf, cw = read_counter(a.calib_fetch), read_counter(a.calib_write)
    kib = a.copy_bytes / 1024.0
    fetch_scale = kib / (sum(cf["copy"]) / len(cf["copy"]))
    write_scale = kib / (sum(cw["copy"]) / len(cw["copy"]))
    f, w = read_counter(a.fetch), read_counter(a.write)
    out = {"_about": {"command": a.command,
                      "fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
                      "calibration": f"copy of {a.copy_bytes:.0f} bytes read and written"}}
    for cls in sorted(set(f) & set(w)):
        if cls == "copy":
            continue
        fk = sum(f[cls]) / len(f[cls])
        wk = sum(w[cls]) / len(w[cls])
        out[cls] = {"bytes_per_launch": int(round((fk * fetch_scale + wk * write_scale) * 1024)),
                    "read_bytes": int(round(fk * fetch_scale * 1024)),
                    "write_bytes": int(round(wk * write_scale * 1024)),
                    "launches": len(f[cls])}
    # one entry per workload; the other workloads' entries are kept
    try:
        with open(a.out) as fh:
            allw = json.load(fh)
    except (OSError, ValueError):
        allw = {}
    allw = {"workloads": dict(allw.get("workloads", {}))}
    allw["workloads"][a.workload] = out
    with open(a.out, "w") as fh:
        json.dump(allw, fh, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
