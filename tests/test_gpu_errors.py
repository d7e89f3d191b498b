"""Error paths of the sound C-ABI on the GPU: wrong call order, invalid arguments and
non-finite data return their ASAM_E_* codes without aborting, and the context then still
solves correctly.  The caller's MaxIter keeps capping the mixed solve."""
import ctypes as C

import numpy as np
import pytest

from asam_amd import workload as W

pytestmark = pytest.mark.gpu

E_INVALID, E_STATE, E_NAN = -1, -3, -5


def small(np1=2):
    return W.config(4, np1=np1, scale=0.25)


def raw_solve(s, x, b, maxiter, tol):
    """asam_sound_solve through the C-ABI, returning the code instead of raising."""
    from asam_amd.sound import Info, _p
    mi, tl, info = C.c_int(maxiter), C.c_double(tol), Info()
    rc = s.L.asam_sound_solve(s.h, _p(x), _p(b), C.byref(mi), C.byref(tl), C.byref(info))
    return rc, mi.value, tl.value, info


def check_solves(s, p, orc):
    xo, mo, _, _ = orc.bicgstab(p.b, 300, 1e-8)
    x = np.zeros(p.n)
    rc, mi, tl, info = raw_solve(s, x, p.b, 300, 1e-8)
    assert rc == 0 and info.status == 0
    assert abs(abs(mi) - abs(mo)) <= 1
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


def test_solve_before_setup_then_recovers(oracle_mod, sound_mod):
    p = small()
    s = sound_mod.SoundSolver.from_problem(p, setup=False)
    x = np.zeros(p.n)
    rc, *_ = raw_solve(s, x, p.b, 300, 1e-8)
    msg = s.L.asam_sound_last_error(s.h).decode()
    print(f"solve before set-up: rc {rc} ({msg})")
    assert rc == E_STATE
    y = np.empty(p.n)
    from asam_amd.sound import _p
    assert s.L.asam_sound_apply(s.h, _p(y), _p(x)) == E_STATE
    g = p.grid
    s.setup(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho, p.rhotheta, p.T)
    check_solves(s, p, oracle_mod.Oracle(p))


def test_invalid_arguments(sound_mod):
    from asam_amd.sound import Control, Grid, _vp
    L = sound_mod.load()
    h = _vp()
    for bad in (Grid(0, 4, 4, 1, 1, 1.0, 0.0, 1.0), Grid(4, 4, 4, 3, 1, 1.0, 0.0, 1.0),
                Grid(4, -1, 4, 2, 1, 1.0, 0.0, 1.0)):
        rc = L.asam_sound_create(C.byref(h), C.byref(bad))
        print(f"create nx={bad.nx} ny={bad.ny} np1={bad.np1}: rc {rc}")
        assert rc == E_INVALID
        if h:
            L.asam_sound_destroy(h)
            h = _vp()
    p = small()
    s = sound_mod.SoundSolver.from_problem(p)
    for ctl in (Control(7, 10, 0, 1e-5, 200, 12), Control(1, 0, 0, 1e-5, 200, 12),
                Control(1, 33, 0, 1e-5, 200, 12), Control(0, 10, 2, 1e-5, 200, 12)):
        assert L.asam_sound_set_control(s.h, C.byref(ctl)) == E_INVALID
    x = np.zeros(p.n)
    rc, *_ = raw_solve(s, x, p.b, 300, -1.0)
    assert rc == E_INVALID


@pytest.mark.parametrize("method,precision", [("BICGSTAB", "FP64"), ("FGMRES", "FP64"),
                                              ("GMRES", "FP64"), ("BICGSTAB", "MIXED"),
                                              ("FGMRES", "MIXED")])
def test_nonfinite_rhs_returns_nan_code_then_recovers(method, precision, oracle_mod, sound_mod):
    """A NaN (or Inf) in b stops the solve at its first residual with ASAM_E_NAN; the same
    context then solves a finite b to the oracle's answer."""
    p = small()
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, precision=precision, step_gmres=10)
    maxiter = 300 if method == "BICGSTAB" else 20
    for bad in (np.nan, np.inf):
        b = p.b.copy()
        b[p.n // 3] = bad
        x = np.zeros(p.n)
        rc, mi, tl, info = raw_solve(s, x, b, maxiter, 1e-8)
        print(f"{method} {precision} b with {bad}: rc {rc}, iterations {info.iterations}, "
              f"{info.ms:.2f} ms, |r| {info.abs_residual}")
        assert rc == E_NAN
        assert not np.isfinite(info.abs_residual)
        assert info.iterations <= 2  # stopped at once, not run to the cap
    s.set_control(method="BICGSTAB", precision="FP64")
    check_solves(s, p, orc)


def test_nonfinite_initial_guess(oracle_mod, sound_mod):
    p = small()
    s = sound_mod.SoundSolver.from_problem(p)
    x = np.zeros(p.n)
    x[5] = np.nan
    rc, mi, tl, info = raw_solve(s, x, p.b, 300, 1e-8)
    assert rc == E_NAN
    check_solves(s, p, oracle_mod.Oracle(p))


@pytest.mark.parametrize("method", ["BICGSTAB", "FGMRES", "FGMRESV"])
def test_mixed_solve_honours_maxiter(method, oracle_mod, sound_mod):
    """MIXED keeps MaxIter as a cap: BiCGStab's inner solves take at most MaxIter
    iterations each, (F)GMRES takes at most MaxIter refinement steps (one restart cycle
    each); MaxIter and Tol come back unchanged."""
    p = small()
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, precision="MIXED", step_gmres=3)
    cap = 2
    x = np.zeros(p.n)
    rc, mi, tl, info = raw_solve(s, x, p.b, cap, 1e-12)
    print(f"{method} mixed, MaxIter {cap}: rc {rc} status {info.status}, inner "
          f"{info.iterations}, refinements {info.restarts}")
    assert rc == 0 and mi == cap and tl == 1e-12
    if method == "BICGSTAB":  # each refinement step's inner solve stops at the cap
        assert info.restarts >= 1 and info.iterations <= cap * info.restarts
    else:                     # at most cap refinement steps of at most m Arnoldi steps
        assert 1 <= info.restarts <= cap and info.iterations <= 3 * info.restarts
        assert info.status == 1
    # a generous cap converges as before
    x = np.zeros(p.n)
    rc, mi, tl, info = raw_solve(s, x, p.b, 300 if method == "BICGSTAB" else 40, 1e-8)
    rr = np.linalg.norm(p.b - orc.matvec1(x)) / np.linalg.norm(p.b)
    assert rc == 0 and info.status == 0 and rr <= 1.0001e-8
