"""bench.py's bookkeeping, on the CPU: the roofline entry of the dominant kernel class, the
per-class table, the counter-run lookup, the tuning file reader and the one-GPU guard."""
import json
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def prof():
    # ms and bytes summed over the timed solves, as the library's profiler reports them
    return {"apply_fp32": {"ms": 0.8, "bytes": 4.0e9, "launches": 20},
            "line_fp32": {"ms": 1.1, "bytes": 6.5e9, "launches": 20},
            "reduce": {"ms": 0.2, "bytes": 1.0e6, "launches": 33},
            "line_fp64": {"ms": 0.0, "bytes": 0.0, "launches": 0}}


def test_roofline_is_the_dominant_class():
    roof, name = bench.roofline_from_profile(prof(), 2)
    assert name == "line_fp32"
    assert roof["kernel"] == "line_fp32" and roof["bound"] == "hbm"
    assert roof["avg_launch_us"] == pytest.approx(55.0)
    assert roof["bytes_per_launch"] == int(6.5e9 / 20)
    achieved = (6.5e9 / 20) / (1.1e-3 / 20) / 1e9  # GB/s
    assert roof["achieved"] == pytest.approx(achieved, abs=0.1)
    assert roof["frac"] == pytest.approx(achieved / 8000.0, abs=1e-4)
    assert roof["peak"] == 8000.0 and roof["unit"] == "GB/s"


def test_roofline_without_launches():
    roof, name = bench.roofline_from_profile(
        {"line_fp32": {"ms": 0.0, "bytes": 0.0, "launches": 0}}, 2)
    assert roof is None and name == "line_fp32"


def test_class_table_per_solve():
    t = bench.class_table(prof(), steps=2)
    assert set(t) == {"apply_fp32", "line_fp32", "reduce"}  # classes with launches only
    a = t["apply_fp32"]
    assert a["ms_per_solve"] == pytest.approx(0.4)
    assert a["launches_per_solve"] == 10
    assert a["avg_launch_us"] == pytest.approx(40.0)
    assert a["GBs"] == pytest.approx(4.0e9 / 0.8e-3 / 1e9, abs=0.1)


def test_read_traffic_by_workload(tmp_path, monkeypatch):
    os.makedirs(tmp_path / "asam_amd")
    wl = bench.traffic_workload(2, 1e-8)
    wl3 = bench.traffic_workload(2, 1e-8, 3)
    assert wl != wl3 and "256x256x64" in wl and "512x512x96" in wl3
    (tmp_path / "asam_amd" / "traffic.json").write_text(json.dumps(
        {"workloads": {wl: {"line_fp32": {"bytes_per_launch": 123}}}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.read_traffic("line_fp32", wl) == 123
    assert bench.read_traffic("apply_fp32", wl) is None
    assert bench.read_traffic("line_fp32", wl3) is None
    monkeypatch.setattr(bench, "ROOT", str(tmp_path / "nowhere"))
    assert bench.read_traffic("line_fp32", wl) is None


def test_load_tuning(tmp_path):
    f = tmp_path / "t.json"
    f.write_text(json.dumps({"fp64": {"apply_bx": 256}, "mixed": {"line_lds": 76}}))
    assert bench.load_tuning(str(f)) == ({"apply_bx": 256}, {"line_lds": 76})


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "single GPU" in r.stderr and r.stdout == ""
