"""The Fortran host path: ISO_C_BINDING module + SolveSound drop-in + driver built with
AMD flang; BIND(C) type sizes equal the C header's; on the GPU the driver's solve
matches the oracle (FP64 within 1e-10, iterations +-1)."""
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from asam_amd import workload as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "asam_amd", "fortran")
DRV = os.path.join(FDIR, "_build", "drv_solve_sound")
FLANG = "/opt/rocm/bin/amdflang"
LIBDIR = os.path.join(ROOT, "asam_amd", "lib")

needs_flang = pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")


def ensure_built():
    if not (os.path.exists(DRV) and os.path.exists(os.path.join(FDIR, "_build", "drv_transport"))):
        subprocess.run(["make", "-C", FDIR], check=True, capture_output=True)
    return DRV


def write_input(path, p, method=0, mixed=0, maxiter=200, tol=1e-8):
    g = p.grid
    with open(path, "wb") as f:
        f.write(struct.pack("<8i", g.nx, g.ny, g.nz, p.np1, p.grad_full, method, mixed, maxiter))
        f.write(struct.pack("<3d", p.beta0_dt, p.grav, tol))
        for a in (p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho, p.rhotheta, p.T, p.b):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def read_output(path, n):
    with open(path, "rb") as f:
        maxiter, = struct.unpack("<i", f.read(4))
        tol, rel = struct.unpack("<2d", f.read(16))
        its, = struct.unpack("<i", f.read(4))
        x = np.frombuffer(f.read(8 * n), dtype="<f8").copy()
    return maxiter, tol, rel, its, x


@needs_flang
def test_fortran_driver_builds_and_links(sound_mod):
    drv = ensure_built()
    out = subprocess.run(["ldd", drv], capture_output=True, text=True).stdout
    assert "libasam_sound.so" in out and "not found" not in out


@needs_flang
def test_bind_c_types_match_header(sound_mod):
    import ctypes as C
    src = """program sizes
  use iso_c_binding
  use asam_sound_iface
  type(asam_sound_grid) :: g
  type(asam_sound_control) :: c
  type(asam_sound_info) :: i
  print '(3i6)', c_sizeof(g), c_sizeof(c), c_sizeof(i)
end program
"""
    ensure_built()
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "s.f90")
        open(f, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run([FLANG, "-I", os.path.join(FDIR, "_build"), f,
                        os.path.join(FDIR, "_build", "asam_sound_iface.o"),
                        "-L" + LIBDIR, "-lasam_sound", "-Wl,-rpath," + LIBDIR, "-o", exe],
                       check=True, capture_output=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    S = sound_mod
    assert vals == [C.sizeof(S.Grid), C.sizeof(S.Control), C.sizeof(S.Info)]


@needs_flang
@pytest.mark.gpu
@pytest.mark.parametrize("np1,method", [(1, 0), (2, 0), (2, 1), (2, 3)])
def test_fortran_solve_sound_matches_oracle(np1, method, oracle_mod, tmp_path):
    drv = ensure_built()
    p = W.config(4, np1=np1, scale=0.5)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, p, method=method, maxiter=200 if method == 0 else 20)
    r = subprocess.run([drv, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    maxiter, tol, rel, its, x = read_output(fout, p.n)
    orc = oracle_mod.Oracle(p)
    if method == 0:
        xo, mo, to, so = orc.bicgstab(p.b, 200, 1e-8)
        assert abs(abs(maxiter) - abs(mo)) <= 1
        assert abs(tol - to) <= 1e-6 * to
    else:
        f = orc.fgmresv if method == 3 else orc.fgmres
        xo, ito, rr = f(p.b, 20, 1e-8, 10)
        assert maxiter == 20 and abs(its - ito) <= 1
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"Fortran SolveSound np1={np1} method={method}: err vs oracle {err:.3e}")
    assert err < 1e-10 and rel <= 1.0001e-8


@needs_flang
@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1])
def test_fortran_mixed_solve_reaches_1e8(method, oracle_mod, tmp_path):
    """AsamSoundInit(..., Mixed=.TRUE.) with BiCGStab or the reference's default FGMRES:
    the FP32-inner / FP64-refined solve reaches the 1e-8 relative residual of the FP64
    system, recomputed by the oracle's operator."""
    drv = ensure_built()
    p = W.config(4, np1=2, scale=0.5)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, p, method=method, mixed=1, maxiter=200 if method == 0 else 20)
    r = subprocess.run([drv, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    maxiter, tol, rel, its, x = read_output(fout, p.n)
    orc = oracle_mod.Oracle(p)
    rr = np.linalg.norm(p.b - orc.matvec1(x)) / np.linalg.norm(p.b)
    xo, _, _, _ = orc.bicgstab(p.b, 300, 1e-12)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"Fortran mixed method={method}: inner its {its}, rel residual {rr:.2e}, "
          f"err vs FP64 oracle {err:.2e}")
    assert rr <= 1.0001e-8 and err < 1e-6


@needs_flang
@pytest.mark.gpu
@pytest.mark.parametrize("np1", [1, 2])
def test_fortran_project_vel_face_matches_oracle(np1, oracle_mod, tmp_path):
    """Rhs -> SolveSound -> UpdateVelocity from Fortran, against the oracle's sequence."""
    ensure_built()
    drv = os.path.join(FDIR, "_build", "drv_project_vel_face")
    p = W.config(4, np1=np1, scale=0.5)
    U, V, Wv, it, ir = W.stage_inputs(p, seed=51)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, p, method=0, maxiter=300, tol=1e-12)
    with open(fin, "ab") as f:
        for a in (U, V, Wv, it, ir):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
    r = subprocess.run([drv, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    orc = oracle_mod.Oracle(p)
    b = orc.rhs(U, V, Wv, it, ir)
    xo, mo, to, so = orc.bicgstab(b, 300, 1e-12)
    ref = orc.update(xo, U, V, Wv, it, ir)
    with open(fout, "rb") as f:
        maxiter, = struct.unpack("<i", f.read(4))
        tol, = struct.unpack("<d", f.read(8))
        got = []
        for a in ref:
            got.append(np.frombuffer(f.read(a.nbytes), dtype="<f8").reshape(a.shape))
    assert abs(abs(maxiter) - abs(mo)) <= 1
    for nm, g_, o_ in zip(("x", "U", "V", "W", "incr_rt", "incr_r"), got, ref):
        d = np.abs(g_ - o_).max() / max(np.abs(o_).max(), 1e-300)
        print(f"np1={np1} {nm}: rel diff {d:.2e}")
        assert d < 1e-10, nm


@needs_flang
def test_bind_c_transport_types_match_header(sound_mod):
    import ctypes as C
    from asam_amd import transport as T
    src = """program sizes
  use iso_c_binding
  use asam_transport_iface
  type(asam_tr_grid) :: g
  type(asam_tr_info) :: i
  print '(2i6)', c_sizeof(g), c_sizeof(i)
end program
"""
    ensure_built()
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "s.f90")
        open(f, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run([FLANG, "-I", os.path.join(FDIR, "_build"), f,
                        os.path.join(FDIR, "_build", "asam_transport_iface.o"),
                        "-L" + LIBDIR, "-lasam_sound", "-Wl,-rpath," + LIBDIR, "-o", exe],
                       check=True, capture_output=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                               check=True).stdout.split()]
    assert vals == [C.sizeof(T.TrGrid), C.sizeof(T.TrInfo)]


@needs_flang
@pytest.mark.gpu
@pytest.mark.parametrize("met", [0, 1])
def test_fortran_transport_matches_oracle(met, oracle_mod, tmp_path):
    """BiCGStabChemSp / BiCGStabMetSp called from Fortran against the oracle's solve."""
    from oracle import pytransport
    ensure_built()
    drv = os.path.join(FDIR, "_build", "drv_transport")
    p = W.make_problem(30, 22, 14, 500.0, 500.0, 250.0, H=700.0, a=1200.0,
                       ridge_xc=15 * 500.0, np1=1, seed=5)
    tp = W.transport_problem(p, ncomp=3, fall_comp=1, mom_comps=(2,), met=bool(met),
                             wind=25.0, noise=4.0)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack("<9i", tp.nx, tp.ny, tp.nz, tp.ncomp, tp.iterg, met, 1, 1, 30))
        f.write(struct.pack("<d", 1e-10))
        for a in (tp.tmom, tp.tpot, tp.fall, tp.lu):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(tp.op_ps, dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(tp.alpha, dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(tp.b, dtype="<f8").tobytes())
    r = subprocess.run([drv, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    with open(fout, "rb") as f:
        maxiter, = struct.unpack("<i", f.read(4))
        tol, = struct.unpack("<d", f.read(8))
        its, = struct.unpack("<i", f.read(4))
        x = np.frombuffer(f.read(8 * tp.ncomp * tp.m), dtype="<f8").reshape(tp.ncomp, tp.m)
    xo, mo, to, so = pytransport.TransportOracle(tp).bicgstab(tp.b, 30, 1e-10)
    err = np.linalg.norm(tp.interior(x) - tp.interior(xo)) / np.linalg.norm(tp.interior(xo))
    print(f"Fortran BiCGStab{'Met' if met else 'Chem'}Sp: MaxIter {maxiter} (oracle {mo}), "
          f"Tol {tol:.3e} (oracle {to:.3e}), err vs oracle {err:.2e}")
    assert abs(abs(maxiter) - abs(mo)) <= 1 and (maxiter < 0) == (mo < 0)
    assert err <= 1e-10
