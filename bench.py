#!/usr/bin/env python3
"""Benchmark of the sound stage solve on one MI355X.

A step is one SolveSound call (x = 0 initial guess, as ProjectVelFace does,
reference SRC/ADVECT/IntSub_Mod.f90:46-56) on the 256x256x64 cut-cell system
(Witch-of-Agnesi ridge, h=1 km, a=5 km; dx=dy=500 m, dz=200 m; theta=300 K
hydrostatic + 2 K bubble; b ~ N(0,1) seeded), buoyancy-coupled MultiTriTB
operator (2 unknowns per cell), BiCGStab with the vertical-line preconditioner,
to 1e-8 relative residual.  The headline solve is the FP32-inner / FP64-refined
one; the plain FP64 GPU solve is timed beside it on the same resident inputs.

Prints one JSON line.  Live kernel timing: every launch of the timed solves is
bracketed by HIP events on the solver's stream (libasam_sound profiling), and
each launch's algorithmic bytes are accounted, so roofline.achieved is measured,
not modelled.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--np1", type=int, default=2, help="1: MultiTriT, 2: MultiTriTB")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-hbm", action="store_true", help="skip the 512x512x96 kernel roofline")
    ap.add_argument("--no-transport", action="store_true",
                    help="skip the transport stage solve leg")
    ap.add_argument("--transport-only", action="store_true",
                    help="measure only the transport stage solve leg and exit")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--kernel-probe", type=int, default=0, metavar="CFG",
                    help="time the copy, apply and line kernels in isolation on config CFG "
                         "(2: 256x256x64, 3: 512x512x96) and exit (profiler runs)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--ensemble", type=int, default=0, metavar="M",
                    help="measure the batched ensemble solve of M uneven members (64: the "
                         "64-member ensemble, nx,ny~U{64..256}, nz~U{40..64}) and exit")
    ap.add_argument("--size-scale", type=float, default=1.0,
                    help="ensemble: horizontal size factor of the members (1: full size)")
    ap.add_argument("--cfg3-only", action="store_true",
                    help="measure only the 512x512x96 solves (tuned shapes, per-kernel "
                         "roofline) and exit (profiler runs)")
    ap.add_argument("--solves-only", action="store_true",
                    help="256x256x64: only the FP64 and mixed BiCGStab solves (no projection, "
                         "FGMRES, varying-b, 512x512x96, CPU or transport legs; profiler runs)")
    ap.add_argument("--tuning", default=None, metavar="JSON",
                    help="launch shapes per precision from this file (as the 'autotuned' "
                         "field prints them) instead of running the autotuner: the "
                         "256x256x64 solves, or the 512x512x96 ones with --cfg3-only "
                         "(profiler runs: the traces then hold only the tuned shapes)")
    return ap.parse_args()


def kernel_probe(cfg, np1, reps):
    """Isolated launches of the dominant kernels plus a streaming copy of known size."""
    from asam_amd import sound, workload
    p = workload.config(cfg, np1=np1)
    s = sound.SoundSolver.from_problem(p)
    N = p.grid.ncell
    out = {"config": cfg, "np1": np1, "cells": N}
    copy_bytes = 1 << 30
    ms = s.time_copy(copy_bytes, reps)
    out["copy_1GiB"] = {"avg_launch_us": round(ms * 1e3, 1),
                        "GBs": round(2 * copy_bytes / (ms * 1e-3) / 1e9, 1)}
    for which, name in ((0, "apply"), (1, "line")):
        for fp32 in (True, False):
            ms = s.time_kernel(which, fp32, reps)
            w = 4 if fp32 else 8
            words = (7 if np1 == 1 else 13) if which == 0 else (5 if np1 == 1 else 16)
            out[f"{name}_{'fp32' if fp32 else 'fp64'}"] = {
                "avg_launch_us": round(ms * 1e3, 1), "algorithmic_bytes": words * w * N,
                "GBs": round(words * w * N / (ms * 1e-3) / 1e9, 1)}
    s.close()
    print(json.dumps(out))


def timed_solves(s, xbuf, bbuf, steps, warmup, tol, profile):
    for _ in range(warmup):
        s._chk(s.L.asam_dev_memset(s.h, xbuf.ptr, 0, s.n * 8))
        s.solve_device(xbuf, bbuf, 500, tol)
    s.profile(profile)
    s.profile_read()
    infos = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s._chk(s.L.asam_dev_memset(s.h, xbuf.ptr, 0, s.n * 8))
        _, _, info = s.solve_device(xbuf, bbuf, 500, tol)
        infos.append(info.as_dict())
    t1 = time.perf_counter()  # every solve returns after a stream synchronisation
    prof = s.profile_read()
    s.profile(False)
    return (t1 - t0) / steps, infos, prof


def fgmres_leg(s, xbuf, bbuf, steps, tol):
    """FGMRES (the reference's default SolveSound method, Iter_Mod.f90:121-199) and FGMRESV
    (:201-288), m = 10, in FP64 and FP32-inner / FP64-refined, on the same resident
    system as the headline solve."""
    out = {"workload": "same system and tolerance, restart length m = 10, x = 0"}
    for method in ("FGMRES", "FGMRESV"):
        row = {}
        for prec in ("FP64", "MIXED"):
            s.set_control(method=method, step_gmres=10, precision=prec)
            t, infos, _ = timed_solves(s, xbuf, bbuf, steps, 1, tol, False)
            row[prec.lower() + "_ms"] = round(t * 1e3, 3)
            row[prec.lower() + "_steps"] = infos[-1]["iterations"]
            row[prec.lower() + "_rel_residual"] = infos[-1]["rel_residual"]
        row["speedup_mixed_vs_fp64"] = round(row["fp64_ms"] / row["mixed_ms"], 3)
        out[method] = row
    s.set_control(precision="MIXED")
    return out


def roofline_from_profile(prof, np1):
    # dominant kernel class = most device time inside the timed region
    best = max(prof.items(), key=lambda kv: kv[1]["ms"])
    name, d = best
    if d["launches"] == 0 or d["ms"] <= 0:
        return None, name
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": name, "avg_launch_us": round(avg_ms * 1e3, 2),
            "bytes_per_launch": int(bytes_per_launch)}, name


def read_traffic(kernel, workload):
    """HBM bytes per launch of a kernel class measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (each scaled by what it reads for a copy of known size), recorded in
    asam_amd/traffic.json by asam_amd/traffic.py; None unless measured on this workload."""
    path = os.path.join(ROOT, "asam_amd", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get("workloads", {}).get(workload, {}).get(kernel)
    return e["bytes_per_launch"] if e else None


def traffic_workload(np1, tol, cfg=2):
    if cfg == 3:
        return f"bench.py 512x512x96 tuned solves: cut-cell, np1={np1}, tol {tol:g}"
    return f"bench.py default solve: 256x256x64 cut-cell, np1={np1}, tol {tol:g}"


def host_cpu():
    """CPU model and core counts of the host the baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(p, tol, seconds):
    """The oracle's FP64 BiCGStab on the same system: one warm-up solve (also the reference
    answer), then the median of at least 5 solves on min(usable cores, 16) threads, and the
    median of 3 single-threaded solves."""
    from oracle import pyoracle
    cpu = host_cpu()
    threads = min(cpu["usable_cpus"] or 1, 16)
    pyoracle.set_threads(threads)
    orc = pyoracle.Oracle(p)
    x, mi, tl, st = orc.bicgstab(p.b, 500, tol)
    times = []
    while len(times) < 5 or (sum(times) < seconds and len(times) < 8):
        t0 = time.perf_counter()
        orc.bicgstab(p.b, 500, tol)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    pyoracle.set_threads(1)
    one = []
    for _ in range(3):
        t0 = time.perf_counter()
        orc.bicgstab(p.b, 500, tol)
        one.append(time.perf_counter() - t0)
    pyoracle.set_threads(threads)
    med1 = float(np.median(one))
    return {"value": round(1.0 / med, 4), "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} full FP64 BiCGStab solves of the same 256x256x64 system "
                      f"after one warm-up (C restatement of the reference algorithm, OpenMP "
                      f"over columns), median {med:.3f} s",
            "one_core": {"value": round(1.0 / med1, 4), "unit": "solves/s", "cores": 1,
                         "sample": f"3 single-threaded solves, median {med1:.3f} s"},
            "cpu": cpu}, x, mi, orc


def load_tuning(path):
    with open(path) as f:
        t = json.load(f)
    return t.get("fp64"), t.get("mixed")


def class_table(prof, steps):
    """Per kernel class of an instrumented replay: device ms and launches per solve,
    algorithmic GB/s and its fraction of the HBM peak."""
    out = {}
    for k, v in prof.items():
        if not v["launches"]:
            continue
        gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else None
        out[k] = {"ms_per_solve": round(v["ms"] / steps, 4),
                  "launches_per_solve": v["launches"] / steps,
                  "avg_launch_us": round(1e3 * v["ms"] / v["launches"], 2),
                  "GBs": round(gbs, 1) if gbs else None,
                  "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    return out


def cfg3_leg(np1, tol, steps, with_cpu, tuning=None):
    """The 512x512x96 system (25.2 M cells; one FP32 vector is 100.7 MB, so the working set
    is far beyond the 256 MB Infinity Cache and the kernels stream from HBM).  Each
    precision gets the autotuner's launch shapes (or those of `tuning`); then K timed solves
    and the same K solves replayed with HIP events around every launch on the solver's
    stream give the solve times and, per kernel class, the in-solve algorithmic GB/s.  The
    roofline entry is the dominant class of the mixed solve, with its HBM bytes per launch
    from the rocprofv3 counter runs when those were made on this workload.  Isolated
    launches of the default shapes are kept for comparison (not with `tuning`, the mode the
    profiler runs use, whose traces then hold only the solves)."""
    from asam_amd import sound, workload
    p = workload.config(3, np1=np1)
    s = sound.SoundSolver.from_problem(p)
    N = p.grid.ncell
    iso = {}
    for which, name in (((0, "apply"), (1, "line")) if tuning is None else ()):
        for fp32 in (True, False):
            ms = s.time_kernel(which, fp32, reps=20)
            w = 4 if fp32 else 8
            words = (7 if np1 == 1 else 13) if which == 0 else (5 if np1 == 1 else 16)
            gbs = words * w * N / (ms * 1e-3) / 1e9
            iso[f"{name}_{'fp32' if fp32 else 'fp64'}"] = {
                "avg_launch_us": round(ms * 1e3, 1), "achieved_GBs": round(gbs, 1),
                "frac": round(gbs / HBM_PEAK_GBS, 4)}
    xbuf, bbuf = s.vec(), s.vec(p.b)
    given = load_tuning(tuning) if tuning else (None, None)
    solves, tuned, kern = {}, {}, {}
    for prec, tg in (("FP64", given[0]), ("MIXED", given[1])):
        s.set_control(precision=prec)
        if tg:
            s.set_tuning(**tg)
            tuned[prec.lower()] = s.get_tuning()
        else:
            tuned[prec.lower()] = s.autotune(3)[0]
        t, infos, _ = timed_solves(s, xbuf, bbuf, steps, 1, tol, False)
        ti, _, prof = timed_solves(s, xbuf, bbuf, steps, 1, tol, True)
        solves[prec] = (t, infos[-1], xbuf.download(), prof, ti)
        kern[prec.lower()] = class_table(prof, steps)
    x64, xmx = solves["FP64"][2], solves["MIXED"][2]
    rbuf = s.vec()
    s.apply_device(rbuf, xbuf)
    res = float(np.linalg.norm(p.b - rbuf.download()) / np.linalg.norm(p.b))
    roof, dom = roofline_from_profile(solves["MIXED"][3], np1)
    if roof is not None:
        tr = read_traffic(dom, traffic_workload(np1, tol, 3))
        roof["traffic"] = tr
        roof["traffic_over_algorithmic"] = (round(tr / roof["bytes_per_launch"], 4)
                                            if tr else None)
    out = {"workload": f"512x512x96 cut-cell {'MultiTriTB' if np1 == 2 else 'MultiTriT'}, "
                       f"ridge 2 km, BiCGStab + vertical-line precond, tol {tol:g}, x = 0",
           "cells": N, "unknowns": p.n,
           "roofline": roof,
           "kernels_mixed": kern["mixed"], "kernels_fp64": kern["fp64"],
           "autotuned": tuned,
           "isolated_default_shapes": iso}
    out["solve"] = {
        "mixed_ms": round(solves["MIXED"][0] * 1e3, 3),
        "fp64_ms": round(solves["FP64"][0] * 1e3, 3),
        "mixed_instrumented_ms": round(solves["MIXED"][4] * 1e3, 3),
        "fp64_instrumented_ms": round(solves["FP64"][4] * 1e3, 3),
        "speedup_mixed_vs_fp64": round(solves["FP64"][0] / solves["MIXED"][0], 3),
        "mixed_inner_iterations": solves["MIXED"][1]["iterations"],
        "fp64_iterations": solves["FP64"][1]["iterations"],
        "mixed_true_rel_residual": res,
        "mixed_rel_diff_vs_fp64_gpu": float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))}
    if with_cpu:
        from oracle import pyoracle
        pyoracle.set_threads(min(host_cpu()["usable_cpus"] or 1, 16))
        orc = pyoracle.Oracle(p)
        t0 = time.perf_counter()
        xo, mo, to, so = orc.bicgstab(p.b, 500, tol)
        out["solve"]["oracle_fp64_s"] = round(time.perf_counter() - t0, 2)
        out["solve"]["oracle_maxiter_out"] = mo
        out["solve"]["mixed_err_vs_fp64_oracle"] = float(np.linalg.norm(xmx - xo) /
                                                         np.linalg.norm(xo))
        out["solve"]["fp64_err_vs_fp64_oracle"] = float(np.linalg.norm(x64 - xo) /
                                                        np.linalg.norm(xo))
        del orc
    s.close()
    return out


def varying_b_leg(s, p, steps, tol):
    """The mixed solve as a time integration calls it: every solve gets a different b
    (seeded N(0,1) on live cells), so the refinement schedule the context replays from its
    previous solve is a prediction, not a copy.  Then the operator is set up again from a
    perturbed state (theta + 0.05 K noise, the next step's system) and the varying-b solves
    are timed on it; the first solve after the new set-up is timed on its own."""
    from asam_amd import workload
    act = np.repeat(p.VolC.reshape(-1) > 0.0, p.np1)
    bs = [s.vec(np.where(act, workload.normal(p.n, 5000 + k), 0.0)) for k in range(steps)]
    xbuf = s.vec()
    s.set_control(precision="MIXED")

    def run():
        ts, its, worst = [], [], 0.0
        for bb in bs:
            s.zero(xbuf)
            t0 = time.perf_counter()
            _, _, info = s.solve_device(xbuf, bb, 500, tol)
            ts.append(time.perf_counter() - t0)
            its.append((info.restarts, info.iterations))
            worst = max(worst, info.rel_residual)
        return ts, its, worst

    s.zero(xbuf)
    s.solve_device(xbuf, bs[-1], 500, tol)  # the context's last solve: another b
    ts, its, worst = run()
    th = p.rhotheta / p.rho + 0.05 * workload.normal(p.grid.ncell, 77).reshape(p.rho.shape)
    s.setup(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho, p.rho * th, p.T)
    s.zero(xbuf)
    t0 = time.perf_counter()
    _, _, info0 = s.solve_device(xbuf, bs[0], 500, tol)
    first = time.perf_counter() - t0
    ts2, its2, worst2 = run()
    s.setup(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho, p.rhotheta, p.T)
    for b in bs:
        b.free()
    return {"workload": f"{steps} different seeded b on the same operator, then on an operator "
                        f"set up from theta + 0.05 K noise; MIXED BiCGStab, tol {tol:g}, x = 0",
            "ms_per_solve": round(1e3 * float(np.mean(ts)), 4),
            "ms_min_max": [round(1e3 * min(ts), 4), round(1e3 * max(ts), 4)],
            "refinements_inner_min_max": [list(min(its)), list(max(its))],
            "worst_rel_residual": worst,
            "perturbed_operator": {"ms_per_solve": round(1e3 * float(np.mean(ts2)), 4),
                                   "first_solve_after_setup_ms": round(1e3 * first, 4),
                                   "first_solve_refinements": info0.restarts,
                                   "worst_rel_residual": worst2}}


def ensemble_bench(a):
    """Batched ensemble: all members solved together (FP32 inner / FP64 refinement to tol,
    one HIP-graph replay per solve, autotuned tiles) vs the same members solved one after
    the other by the plain FP64 drop-in (and by the single-domain mixed solve)."""
    from asam_amd import ensemble, sound, workload
    t0 = time.perf_counter()
    ps = workload.ensemble_members(a.ensemble, seed=5, np1=a.np1, size_scale=a.size_scale)
    cells = sum(p.grid.ncell for p in ps)
    print(f"[ensemble] {len(ps)} members, {cells / 1e6:.1f} M cells, generated in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    E = ensemble.Ensemble(ps)
    tune, tms = E.autotune(reps=3)
    print(f"[ensemble] autotuned {tune}", file=sys.stderr, flush=True)
    b = E.concat([p.b for p in ps])
    bbuf, xbuf = E.buffer(b), E.buffer()
    for _ in range(a.warmup):
        xbuf.zero()
        E.solve_device(xbuf, bbuf, a.tol)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        xbuf.zero()
        infos = E.solve_device(xbuf, bbuf, a.tol)
    t_ens = (time.perf_counter() - t0) / a.steps
    st = E.stats()
    xe = E.split(xbuf.download())
    E.close()
    del bbuf, xbuf
    print(f"[ensemble] batched solve {t_ens * 1e3:.1f} ms", file=sys.stderr, flush=True)
    # one FP32 inner iteration with every member active (the tuner's objective): line
    # p-step (5 vectors) + apply v (u) + line s-step (4 vectors) + apply t (u) + update
    np1 = a.np1
    coefw = 5 if np1 == 1 else 9
    fac = (2 * np1 * np1 + (3 if np1 == 2 else 1)) * 4  # T^-1, W and the rank-one L
    it_bytes = cells * ((fac + np1 * 4 * 5) + (coefw * 4 + np1 * 12) + (fac + np1 * 4 * 4)
                        + (coefw * 4 + np1 * 12) + 8 * np1 * 4)
    gbs = it_bytes / (tms[0] * 1e-3) / 1e9
    kern = {"inner_iteration": {"ms": round(tms[0], 4), "bytes": int(it_bytes),
                                "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}}
    # sequential baselines: one drop-in context per member, solves timed on resident inputs
    t64 = tmx = 0.0
    worst_res = worst_diff = 0.0
    its64 = []
    for m, p in enumerate(ps):
        s = sound.SoundSolver.from_problem(p)
        xb, bb = s.vec(), s.vec(p.b)
        for prec in ("FP64", "MIXED"):
            s.set_control(precision=prec)
            s._chk(s.L.asam_dev_memset(s.h, xb.ptr, 0, s.n * 8))
            s.solve_device(xb, bb, 500, a.tol)  # warm-up (kernels, iteration history)
            reps = max(1, a.steps // 4)
            tt = time.perf_counter()
            for _ in range(reps):
                s._chk(s.L.asam_dev_memset(s.h, xb.ptr, 0, s.n * 8))
                _, _, inf = s.solve_device(xb, bb, 500, a.tol)
            dt = (time.perf_counter() - tt) / reps
            if prec == "FP64":
                t64 += dt
                its64.append(inf.iterations)
                x64 = xb.download()
            else:
                tmx += dt
        r = p.b - s.MatVec1(xe[m])
        worst_res = max(worst_res, float(np.linalg.norm(r) / np.linalg.norm(p.b)))
        worst_diff = max(worst_diff, float(np.linalg.norm(xe[m] - x64) / np.linalg.norm(x64)))
        s.close()
    line = {
        "metric": "ensemble members solved/s (FP32-inner + FP64-refined BiCGStab to 1e-8, "
                  "all members in one HIP-graph replay)",
        "value": round(len(ps) / t_ens, 2),
        "unit": "members/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t_ens * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 residual, dots, refinement",
        "data": "synthetic (seeded members: Agnesi ridge H~U(0,1.5) km, bubble ~U(1,3) K, "
                "b~N(0,1))",
        "config": {"workload": f"{len(ps)}-member uneven ensemble, nx,ny~U{{64..256}}x"
                               f"{a.size_scale:g}, nz~U{{40..64}}, "
                               f"{'MultiTriTB' if np1 == 2 else 'MultiTriT'}",
                   "members": len(ps), "cells": cells, "unknowns": cells * np1},
        "graph": st,
        "autotuned": tune,
        "kernels": kern,
        "sequential_fp64": {"members_per_s": round(len(ps) / t64, 2),
                            "ms_total": round(t64 * 1e3, 3),
                            "iterations_min_max": [min(its64), max(its64)]},
        "sequential_mixed_single_domain": {"members_per_s": round(len(ps) / tmx, 2),
                                           "ms_total": round(tmx * 1e3, 3)},
        "speedup_vs_sequential_fp64": round(t64 / t_ens, 3),
        "max_true_rel_residual": worst_res,
        "max_rel_diff_vs_fp64_gpu": worst_diff,
        "inner_iterations_min_max": [min(i.iterations for i in infos),
                                     max(i.iterations for i in infos)],
    }
    if not a.no_cpu:
        from oracle import pyoracle
        threads = min(os.cpu_count() or 1, 16)
        pyoracle.set_threads(threads)
        tt, done = 0.0, 0
        for p in ps:
            orc = pyoracle.Oracle(p)
            t1 = time.perf_counter()
            orc.bicgstab(p.b, 500, a.tol)
            tt += time.perf_counter() - t1
            done += 1
            if tt > a.cpu_seconds:
                break
        line["cpu_baseline"] = {"value": round(done / tt, 4), "unit": "members/s",
                                "cores": threads, "kind": "port",
                                "sample": f"first {done} members solved by the FP64 C "
                                          f"restatement (OpenMP over columns), {tt:.1f} s"}
    print(json.dumps(line))


def proj_bytes(p):
    """Algorithmic bytes of the right-hand side and of the post-solve update (both kernels)
    at problem p: every distinct array element read or written once."""
    g = p.grid
    N, np1 = g.ncell, p.np1
    nf = (g.nx + 1) * g.ny * g.nz + g.nx * (g.ny + 1) * g.nz + g.nx * g.ny * (g.nz + 1)
    rhs = 8 * (3 * nf + N * (2 + (np1 == 2)) + N * np1)
    faces = 8 * (N * (np1 + 2) + nf + 2 * nf)           # x, DUT, V; F; u read + write
    cells = 8 * (3 * nf + N * (1 + 2 + np1 + 1) + N * (2 + np1))  # u, F, theta_f; V, incr, x, DUT
    return rhs, faces + cells


def projection_leg(s, p, steps, tol):
    """One stage projection per step on resident inputs: Rhs -> SolveSound -> UpdateVelocity
    (the three calls ProjectVelFace makes, reference ADVECT/IntSub_Mod.f90:40-70)."""
    from asam_amd import workload
    U, V, W, it, ir = workload.stage_inputs(p, seed=41)
    src = [s.alloc(a.size * 8).upload(np.ascontiguousarray(a)) for a in (U, V, W, it, ir)]
    work = [s.alloc(a.size * 8) for a in (U, V, W, it, ir)]
    bbuf, xbuf = s.vec(), s.vec()
    times = []
    for step in range(steps + 1):
        for d, q in zip(work, src):
            s._chk(s.L.asam_dev_memset(s.h, d.ptr, 0, d.nbytes))
            s.axpby(1.0, q, 0.0, d, n=d.nbytes // 8)
        s.zero(xbuf)
        t0 = time.perf_counter()
        s.rhs_device(bbuf, *work)
        s.solve_device(xbuf, bbuf, 500, tol)
        s.update_device(xbuf, *work)
        if step:
            times.append(time.perf_counter() - t0)
    rb, ub = proj_bytes(p)
    ms_r = s.time_kernel(3, False, 20)
    ms_u = s.time_kernel(4, False, 20)
    return {"stage_ms": round(1e3 * float(np.median(times)), 4),
            "stages_per_s": round(1.0 / float(np.median(times)), 2),
            "rhs": {"avg_launch_us": round(ms_r * 1e3, 1), "bytes": rb,
                    "GBs": round(rb / (ms_r * 1e-3) / 1e9, 1)},
            "update": {"avg_us_two_kernels": round(ms_u * 1e3, 1), "bytes": ub,
                       "GBs": round(ub / (ms_u * 1e-3) / 1e9, 1)}}


def transport_leg(steps, with_cpu, cpu_seconds=15.0, ncomp=4, tol=1e-6, maxiter=30):
    """The transport stage solve (BiCGStabChemSp, reference MODULE/SpMatrix4CellPar_Mod.f90
    :140-243) of `ncomp` tracers on the 256x256x64 cut-cell grid: upwind advection and
    diffusion Jacobian from the stage velocities, a per-cell chemistry LU, x0 = 0 and the
    reference's BiCGStabTol / BiCGStabMaxIter defaults (Control_Mod.f90:580-581).  Solves
    on resident device vectors; the Gauss-Seidel sweep's algorithmic bytes per launch are
    ncomp * m * (7 coefficients + b + old x + new x) words; the oracle solve is timed on
    the host cores beside it."""
    from asam_amd import transport, workload
    p = workload.config(2, np1=1)
    tp = workload.transport_problem(p, ncomp=ncomp)
    s = transport.TransportSolver.from_problem(tp)
    xd, bd = s.device_vector(), s.device_vector().upload(tp.b)
    zero = np.zeros(tp.ncomp * tp.m)
    times, infos = [], []
    for step in range(steps + 1):
        xd.upload(zero)
        t0 = time.perf_counter()
        mi, tl, info = s.solve_device(xd, bd, maxiter, tol)
        if step:
            times.append(time.perf_counter() - t0)
            infos.append((mi, tl, info.as_dict()))
    x = xd.download().reshape(tp.ncomp, tp.m)
    ms_sweep = s.time_kernel(0, 10)
    ms_mv = s.time_kernel(1, 10)
    ms_ps = s.time_kernel(2, 3)
    sweep_bytes = tp.ncomp * tp.m * 10 * 8
    mv_bytes = tp.m * (7 + 3 * tp.ncomp + (tp.ncomp ** 2 if tp.lu is not None else 0)) * 8
    med = float(np.median(times))
    out = {"workload": f"256x256x64 cut-cell transport stage solve, {ncomp} tracers "
                       f"(advection + diffusion, per-cell chemistry LU), BiCGStabChem + "
                       f"{1 + tp.iterg} forward Gauss-Seidel sweeps, tol {tol:g}",
           "solves_per_s": round(1.0 / med, 3), "ms_per_solve": round(1e3 * med, 3),
           "maxiter_out": infos[-1][0], "tol_out": infos[-1][1], "info": infos[-1][2],
           "tuning": s.get_tuning().as_dict(),
           "gs_sweep": {"avg_launch_us": round(ms_sweep * 1e3, 1), "bytes": sweep_bytes,
                        "GBs": round(sweep_bytes / (ms_sweep * 1e-3) / 1e9, 1),
                        "frac": round(sweep_bytes / (ms_sweep * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "bound": "latency: one lexicographic sweep is a chain of "
                                 f"{(tp.nx + 2) + (tp.ny + 2) + (tp.nz + 2) - 2} hyperplanes"},
           "matvec": {"avg_launch_us": round(ms_mv * 1e3, 1), "bytes": mv_bytes,
                      "GBs": round(mv_bytes / (ms_mv * 1e-3) / 1e9, 1)},
           "psolve_ms": round(ms_ps, 3)}
    for v in (xd, bd):
        v.free()
    del s
    if with_cpu:
        from oracle import pyoracle, pytransport
        threads = min(host_cpu()["usable_cpus"] or 1, 16)
        pyoracle.set_threads(threads)
        o = pytransport.TransportOracle(tp)
        ct = []
        xo = None
        while len(ct) < 5:
            t0 = time.perf_counter()
            xo, mio, tlo, sto = o.bicgstab(tp.b, maxiter, tol)
            ct.append(time.perf_counter() - t0)
        cmed = float(np.median(ct))
        out["cpu_baseline"] = {"value": round(1.0 / cmed, 4), "unit": "solves/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{len(ct)} full FP64 transport solves of the same "
                                         f"system (C restatement; Gauss-Seidel single-threaded "
                                         f"as in the reference, products and BLAS on {threads} "
                                         f"threads), median {cmed:.2f} s",
                               "cpu": host_cpu()}
        out["oracle_maxiter_out"] = mio
        inner = tp.interior(x).ravel()
        innero = tp.interior(xo).ravel()
        out["err_vs_fp64_oracle"] = float(np.linalg.norm(inner - innero) /
                                          np.linalg.norm(innero))
    return out


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the sound stage solve runs on a single GPU and is "
              "not sharded (a sharded run would give each GPU horizontal tiles of full "
              "columns); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if a.kernel_probe:
        kernel_probe(a.kernel_probe, a.np1, a.reps)
        return
    if a.ensemble:
        ensemble_bench(a)
        return
    if a.transport_only:
        print(json.dumps({"transport": transport_leg(max(2, a.steps // 4), not a.no_cpu,
                                                     a.cpu_seconds)}))
        return
    if a.cfg3_only:
        print(json.dumps({"roofline_cfg3": cfg3_leg(a.np1, a.tol, a.steps, not a.no_cpu,
                                                    a.tuning)}))
        return
    from asam_amd import sound, workload
    p = workload.config(2, np1=a.np1)
    s = sound.SoundSolver.from_problem(p)
    xbuf = s.vec()
    bbuf = s.vec(p.b)

    # Timed region: K solves, uninstrumented (value, ms_per_step); then the same K solves
    # with every kernel launch bracketed by HIP events (per-kernel durations and bytes for
    # the roofline).  Both runs issue identical launch sequences.
    given = load_tuning(a.tuning) if a.tuning else (None, None)

    def tune(tg):
        if tg:
            s.set_tuning(**tg)
            return s.get_tuning()
        return s.autotune(5)[0]

    s.set_control(precision="FP64")
    tune64 = tune(given[0])
    t64, infos64, _ = timed_solves(s, xbuf, bbuf, a.steps, a.warmup, a.tol, False)
    t64_i, _, prof64 = timed_solves(s, xbuf, bbuf, a.steps, 1, a.tol, True)
    x64 = xbuf.download()

    s.set_control(precision="MIXED")
    tunemx = tune(given[1])
    tmx, infosmx, _ = timed_solves(s, xbuf, bbuf, a.steps, a.warmup, a.tol, False)
    tmx_i, _, profmx = timed_solves(s, xbuf, bbuf, a.steps, 1, a.tol, True)
    xmx = xbuf.download()

    # residual of the mixed solution with the FP64 GPU operator (size-independent check)
    rbuf = s.vec()
    s.apply_device(rbuf, xbuf)
    r = p.b - rbuf.download()
    true_rel = float(np.linalg.norm(r) / np.linalg.norm(p.b))
    diff_vs_fp64_gpu = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))

    line_proj = fgm = vary = None
    if not a.solves_only:
        line_proj = projection_leg(s, p, max(2, a.steps // 4), a.tol)
        fgm = fgmres_leg(s, xbuf, bbuf, max(2, a.steps // 4), a.tol)
        vary = varying_b_leg(s, p, max(4, a.steps // 2), a.tol)

    roof, dom = roofline_from_profile(profmx, a.np1)
    if roof is not None:
        roof["traffic"] = read_traffic(dom, traffic_workload(a.np1, a.tol))
    line = {
        "metric": "Krylov solves/s (FP32-inner + FP64-refined BiCGStab to 1e-8), 256x256x64 cut-cell",
        "value": round(1.0 / tmx, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(tmx * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 residual, dots, refinement",
        "data": "synthetic (seeded Agnesi ridge, hydrostatic bubble state, b~N(0,1))",
        "config": {"workload": f"256x256x64 cut-cell {'MultiTriTB (buoyancy, 2 unknowns/cell)' if a.np1 == 2 else 'MultiTriT'} "
                               f"stage solve, BiCGStab + vertical-line precond, tol {a.tol:g}",
                   "cells": p.grid.ncell, "unknowns": p.n, "parallelism": "1 GPU"},
        "roofline": roof,
        "fp64": {"solves_per_s": round(1.0 / t64, 3), "ms_per_step": round(t64 * 1e3, 4),
                 "iterations": infos64[-1]["iterations"],
                 "instrumented_ms_per_step": round(t64_i * 1e3, 4)},
        "mixed": {"inner_iterations": infosmx[-1]["iterations"],
                  "refinement_steps": infosmx[-1]["restarts"],
                  "true_rel_residual": true_rel,
                  "rel_diff_vs_fp64_gpu": diff_vs_fp64_gpu,
                  "instrumented_ms_per_step": round(tmx_i * 1e3, 4)},
        "speedup_mixed_vs_fp64": round(t64 / tmx, 3),
        "projection": line_proj,
        "fgmres": fgm,
        "mixed_varying_b": vary,
        "autotuned": {"fp64": tune64, "mixed": tunemx},
        "kernel_timing": "HIP events around every launch of the instrumented replay of the timed solves",
        "kernels_mixed": class_table(profmx, a.steps),
        "kernels_fp64": class_table(prof64, a.steps),
    }
    s.close()
    if a.solves_only:
        a.no_hbm = a.no_cpu = a.no_transport = True
    if not a.no_hbm:
        line["roofline_cfg3"] = cfg3_leg(a.np1, a.tol, max(4, a.steps // 4), not a.no_cpu)
    if not a.no_cpu:
        cb, xo, mo, orc = cpu_baseline(p, a.tol, a.cpu_seconds)
        line["cpu_baseline"] = cb
        line["mixed"]["err_vs_fp64_oracle"] = float(np.linalg.norm(xmx - xo) / np.linalg.norm(xo))
        line["fp64"]["err_vs_fp64_oracle"] = float(np.linalg.norm(x64 - xo) / np.linalg.norm(xo))
        line["fp64"]["oracle_maxiter_out"] = mo
    else:
        line["cpu_baseline"] = None
    if not a.no_transport:
        line["transport"] = transport_leg(max(2, a.steps // 4), not a.no_cpu, a.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
