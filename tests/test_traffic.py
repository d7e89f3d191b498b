"""The HBM-traffic post-processing of rocprofv3 counter runs (asam_amd/traffic.py): kernel
names map to the live profiler's classes, each counter is scaled by what it reads for a copy
of known size, per-class averages are per dispatch, and other workloads' entries survive."""
import csv
import json
import os

from asam_amd import traffic

KiB = 1024.0


def write_counters(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for k, (name, value) in enumerate(rows):
            w.writerow({"Dispatch_Id": k, "Kernel_Name": name, "Counter_Name": "X",
                        "Counter_Value": value})


def test_kernel_classes():
    c = traffic.kernel_class
    assert c("void asam::k_line<2, 0, 1, 1, 8, 1, float, float>(asam::KLine<float, float>)") == "line_fp32"
    assert c("void asam::k_line<1, 1, 0, 0, 4, 0, double, double>(...)") == "line_fp64"
    assert c("void asam::k_apply_rows<2, 0, 24, 2, float, float, float, float, 2>(...)") == "apply_fp32"
    # the first floating type is the coefficient type: FP64 coefficients, FP32 output
    assert c("void asam::k_apply_rows<2, 0, 15, 2, double, double, float, double, 0>(...)") == "apply_fp64"
    assert c("void asam::k_apply_march<2, 1, 43, double, double, double, double>(...)") == "apply_fp64"
    assert c("void asam::k_bicg_update<float>(long, float*, ...)") == "blas_fp32"
    assert c("void asam::k_bicg_update<double>(long, double*, ...)") == "blas_fp64"
    assert c("void asam::k_fill<float>(long, float*, double)") == "blas_fp32"
    assert c("void asam::k_fill<double>(long, double*, double)") is None
    assert c("asam::k_finalize(int, int, ...)") == "reduce"
    assert c("asam::k_copy16(long, float4 const*, float4*)") == "copy"
    assert c("asam::k_rhs<2>(...)") is None


def test_scaled_bytes_per_launch(tmp_path):
    copy = float(1 << 20)
    # calibration: FETCH_SIZE reads half the copied bytes, WRITE_SIZE all of them
    write_counters(tmp_path / "cf", [("asam::k_copy16(...)", copy / 2 / KiB)] * 2)
    write_counters(tmp_path / "cw", [("asam::k_copy16(...)", copy / KiB)] * 2)
    line = "void asam::k_line<2, 0, 1, 1, 8, 1, float, float>(...)"
    upd = "void asam::k_bicg_update<float>(...)"
    fill = "void asam::k_fill<float>(...)"
    write_counters(tmp_path / "f", [(line, 100.0), (line, 120.0), (upd, 30.0), (fill, 0.0),
                                    ("asam::k_rhs<2>(...)", 999.0)])
    write_counters(tmp_path / "w", [(line, 50.0), (line, 50.0), (upd, 20.0), (fill, 10.0),
                                    ("asam::k_rhs<2>(...)", 999.0)])
    out = tmp_path / "traffic.json"
    out.write_text(json.dumps({"workloads": {"other": {"line_fp32": {"bytes_per_launch": 7}}}}))
    traffic.main([str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"),
                  str(tmp_path / "cw"), "--copy-bytes", str(copy), "--workload", "wl",
                  "--command", "cmd", "--out", str(out)])
    t = json.loads(out.read_text())
    assert t["workloads"]["other"] == {"line_fp32": {"bytes_per_launch": 7}}
    e = t["workloads"]["wl"]
    assert e["_about"]["fetch_scale"] == 2.0 and e["_about"]["write_scale"] == 1.0
    assert e["line_fp32"] == {"bytes_per_launch": int((110 * 2 + 50) * KiB),
                              "read_bytes": int(110 * 2 * KiB), "write_bytes": int(50 * KiB),
                              "launches": 2}
    # the refinement's zero fill of the FP32 correction averages with the update
    assert e["blas_fp32"]["launches"] == 2
    assert e["blas_fp32"]["read_bytes"] == int(15 * 2 * KiB)
    assert e["blas_fp32"]["write_bytes"] == int(15 * KiB)
    assert set(e) == {"_about", "line_fp32", "blas_fp32"}
