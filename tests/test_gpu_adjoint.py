"""Adjoint pair on the GPU: dot-product tests of the transposed operator, preconditioner
and solve; finite-difference checks of the theta sensitivity; and the reverse-mode
gradient of a 50-step Ros3P run with checkpointed stage states (the config-4
experiment: cfg-1 grid with a 1 km ridge, checkpoints every 10 steps,
J = sum over a 4x4x4 km box of w*u, w ~ U(0,1) seeded)."""
import numpy as np
import pytest

from asam_amd import rosenbrock as R
from asam_amd import workload as W

pytestmark = pytest.mark.gpu


def live(p, v):
    m = np.repeat((p.VolC.reshape(-1) > 0.0), p.np1)
    return np.where(m, v, 0.0)


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("np1", [1, 2])
def test_dot_product_operator_and_preconditioner(np1, sound_mod):
    p = W.config(4, np1=np1, scale=0.5)
    s = sound_mod.SoundSolver.from_problem(p)
    x = live(p, W.normal(p.n, 21))
    y = live(p, W.normal(p.n, 22))
    a, b = np.dot(s.MatVec1(x), y), np.dot(x, s.MatVec1_T(y))
    print(f"np1={np1} <Op x,y>={a:.16e} <x,Op^T y>={b:.16e} rel {rel(a, b):.2e}")
    assert rel(a, b) < 1e-13
    a, b = np.dot(s.PSolve1(x), y), np.dot(x, s.PSolve1_T(y))
    print(f"np1={np1} <P x,y> vs <x,P^T y> rel {rel(a, b):.2e}")
    assert rel(a, b) < 1e-13


@pytest.mark.parametrize("np1", [1, 2])
def test_dot_product_solve(np1, sound_mod):
    """<S b, g> = <b, S^T g> for the solve S = (-A)^-1 (forward) and its transposed
    solve (backward); to 1e-12 with both solved to 1e-14."""
    p = W.config(4, np1=np1, scale=0.5)
    s = sound_mod.SoundSolver.from_problem(p)
    b = live(p, W.normal(p.n, 31))
    g = live(p, W.normal(p.n, 32))
    x, *_ = s.SolveSound(np.zeros(p.n), b, 500, 1e-14)
    lam, *_ = s.SolveSound_T(np.zeros(p.n), g, 500, 1e-14)
    a, c = np.dot(x, g), np.dot(b, lam)
    print(f"np1={np1} <S b,g> vs <b,S^T g> rel {rel(a, c):.2e}")
    assert rel(a, c) < 1e-12


@pytest.mark.parametrize("np1", [1, 2])
def test_theta_vjp_matches_finite_difference(np1, sound_mod):
    p = W.config(4, np1=np1, scale=0.5)
    g = p.grid
    theta = p.rhotheta / p.rho
    s0 = sound_mod.SoundSolver.from_problem(p)
    x = live(p, W.normal(p.n, 41))
    mu = live(p, W.normal(p.n, 42))
    xb, mub = s0.vec(x), s0.vec(mu)
    gb = s0.alloc(g.ncell * 8)
    s0._chk(s0.L.asam_dev_memset(s0.h, gb.ptr, 0, g.ncell * 8))
    s0.theta_vjp(gb, mub, xb)
    grad = gb.download()
    dth = W.normal(g.ncell, 43).reshape(theta.shape)
    eps = 1e-6 * 300.0

    def S(th):
        s = sound_mod.SoundSolver(g.nx, g.ny, g.nz, np1, p.grad_full, p.beta0_dt, p.grav)
        s.setup_weights(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, *R.face_weights(p, th))
        return -np.dot(mu, s.MatVec1(x))  # mu^T A x
    fd = (S(theta + eps * dth) - S(theta - eps * dth)) / (2 * eps)
    an = np.dot(grad, dth.reshape(-1))
    print(f"np1={np1} theta VJP {an:.10e} FD {fd:.10e} rel {rel(an, fd):.2e}")
    assert rel(an, fd) < 1e-6


def config4(np1=2):
    dt = 2.0
    p = W.config(4, np1=np1)
    p.beta0_dt = R.ROS3P_BETA0 * dt
    g = p.grid
    w = np.zeros((g.nz, g.ny, g.nx, np1))
    # 4x4x4 km box: x, y in [30, 34) km (dx = dy = 1 km), z in [0, 4) km (dz = 250 m)
    box = (slice(0, 16), slice(30, 34), slice(30, 34))
    w[box] = W.uniform01(16 * 4 * 4 * np1, 4).reshape(16, 4, 4, np1)
    src = live(p, W.normal(p.n, 44)) * 1e-3
    # cells below the terrain (V = 0) carry no state: their weights are zero
    return p, dt, live(p, w.reshape(-1)), src


def run_J(sound_mod, p, dt, w, src, nsteps, theta=None, weights=None):
    g = p.grid
    if weights is None:
        s = sound_mod.SoundSolver.from_problem(p)
    else:
        s = sound_mod.SoundSolver(g.nx, g.ny, g.nz, p.np1, p.grad_full, p.beta0_dt, p.grav)
        s.setup_weights(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, *weights)
    rb = R.LinearRosenbrock(s, dt)
    u0 = s.vec()
    uN, _ = rb.forward(u0, s.vec(src), nsteps, 10)
    return float(np.dot(w, uN.download())), s, rb


def test_rosenbrock_adjoint_config4(sound_mod):
    p, dt, w, src = config4()
    nsteps = 50
    theta = p.rhotheta / p.rho
    s = sound_mod.SoundSolver(p.grid.nx, p.grid.ny, p.grid.nz, 2, p.grad_full, p.beta0_dt,
                              p.grav)
    s.setup_weights(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, *R.face_weights(p, theta))
    rb = R.LinearRosenbrock(s, dt)
    uN, u0bar, sbar_b, thbar_b = rb.gradient(s.vec(), s.vec(src), nsteps, 10, s.vec(w))
    J0 = float(np.dot(w, uN.download()))
    sbar = sbar_b.download()
    thbar = thbar_b.download()
    print(f"J = {J0:.12e}; {rb.solves} stage solves, {rb.iterations} BiCGStab iterations")
    assert np.isfinite(J0) and J0 != 0.0
    # dot-product test of the whole 50-step linear map s -> u_N (u0 = 0): <w, Phi ds> = <Phi^T w, ds>
    ds = live(p, W.normal(p.n, 45))
    Jd, _, _ = run_J(sound_mod, p, dt, w, ds, nsteps, weights=R.face_weights(p, theta))
    an = float(np.dot(sbar, ds))
    print(f"dot-product test: <w,Phi ds> {Jd:.15e} <Phi^T w, ds> {an:.15e} rel {rel(Jd, an):.2e}")
    assert rel(Jd, an) < 1e-12  # measured 5e-14
    # finite differences w.r.t. b (the stage forcing), eps = 1e-6
    eps = 1e-6
    Jp, _, _ = run_J(sound_mod, p, dt, w, src + eps * ds, nsteps, weights=R.face_weights(p, theta))
    Jm, _, _ = run_J(sound_mod, p, dt, w, src - eps * ds, nsteps, weights=R.face_weights(p, theta))
    fd = (Jp - Jm) / (2 * eps)
    print(f"FD d/db: {fd:.12e} adjoint {an:.12e} rel {rel(fd, an):.2e}")
    assert rel(fd, an) < 1e-9  # J is linear in b: only round-off, measured 6e-14
    # finite differences w.r.t. the background theta field (eps = 1e-6 relative)
    dth = W.normal(p.grid.ncell, 46).reshape(theta.shape)
    e = 1e-6 * 300.0
    Jp, _, _ = run_J(sound_mod, p, dt, w, src, nsteps, weights=R.face_weights(p, theta + e * dth))
    Jm, _, _ = run_J(sound_mod, p, dt, w, src, nsteps, weights=R.face_weights(p, theta - e * dth))
    fd = (Jp - Jm) / (2 * e)
    an = float(np.dot(thbar, dth.reshape(-1)))
    print(f"FD d/dtheta: {fd:.12e} adjoint {an:.12e} rel {rel(fd, an):.2e}")
    assert rel(fd, an) < 1e-7  # central-difference truncation, measured 1e-9
