"""The Ros3P stepping and its reverse mode (asam_amd/rosenbrock.py) driven on the CPU
through a dense stand-in for the solver context: the step has order 3 on the scalar test
equation (the coefficient table reproduces exp(z) to O(z^4)), and the windowed reverse
sweep is the exact transpose of the forward map (dot-product tests for the initial state
and the source)."""
import math
import types

import numpy as np

from asam_amd import rosenbrock as R


class Buf:
    def __init__(self, a):
        self.a = np.array(a, dtype=np.float64)
        self.nbytes = self.a.nbytes

    def download(self):
        return self.a.copy()


class DenseStage:
    """Stage system A = V (I - beta0 dt J) with a dense A: the operations LinearRosenbrock
    asks of a solver context, in numpy."""

    def __init__(self, A, V):
        self.A, self.V = A, V
        self.n = len(V)
        self.grid = types.SimpleNamespace(nx=self.n, ny=1, nz=1)

    def vec(self, a=None):
        return Buf(np.zeros(self.n) if a is None else a)

    def alloc(self, nbytes):
        return Buf(np.zeros(nbytes // 8))

    def zero(self, b):
        b.a[:] = 0.0

    def axpby(self, a, x, b, y, n=None):
        n = n or self.n
        y.a[:n] = a * x.a[:n] + b * y.a[:n]

    def mass(self, y, x):
        y.a[:] = self.V * x.a

    def apply_device(self, y, x, transpose=False):
        y.a[:] = -((self.A.T if transpose else self.A) @ x.a)

    def solve_device(self, x, b, maxiter, tol, transpose=False):
        M = -(self.A.T if transpose else self.A)
        x.a[:] = np.linalg.solve(M, b.a)
        return 1, 0.0, types.SimpleNamespace(iterations=1, status=0)

    def theta_vjp(self, g, mu, x):
        g.a[:] = 0.0


def step_scalar(z, beta0=R.ROS3P_BETA0):
    # u' = lambda u with V = 1, dt = 1: A = 1 - beta0 z
    s = DenseStage(np.array([[1.0 - beta0 * z]]), np.array([1.0]))
    rb = R.LinearRosenbrock(s, 1.0)
    out = s.vec()
    rb.step(out, s.vec([1.0]), s.vec())
    return out.a[0]


def test_ros3p_order_three_on_test_equation():
    errs = []
    for z in (-0.2, -0.1, -0.05):
        errs.append(abs(step_scalar(z) - math.exp(z)))
    # local error O(z^4): halving z divides the error by ~16
    assert 12.0 < errs[0] / errs[1] < 20.0 and 12.0 < errs[1] / errs[2] < 20.0
    # and the method is stable far out on the negative axis
    assert abs(step_scalar(-1e4)) < 1.0


def test_reverse_sweep_is_transpose_of_forward():
    rng = np.random.default_rng(3)
    n = 7
    V = rng.uniform(0.5, 2.0, n)
    L = rng.normal(size=(n, n))
    J = -(L @ L.T) / n - np.eye(n)            # a stable linear model u' = J u + s
    dt = 0.3
    A = np.diag(V) @ (np.eye(n) - R.ROS3P_BETA0 * dt * J)
    s = DenseStage(A, V)
    rb = R.LinearRosenbrock(s, dt)
    u0, src, w = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    uN, u0bar, sbar, _ = rb.gradient(s.vec(u0), s.vec(src), 9, 4, s.vec(w))
    du, ds = rng.normal(size=n), rng.normal(size=n)

    def J_of(u, q):
        out, _ = rb.forward(s.vec(u), s.vec(q), 9, 4)
        return float(w @ out.a)
    # the step is linear in (u0, s): directional derivatives are differences of J
    base = J_of(u0, src)
    assert np.isclose(J_of(u0 + du, src) - base, u0bar.a @ du, rtol=1e-10)
    assert np.isclose(J_of(u0, src + ds) - base, sbar.a @ ds, rtol=1e-10)
    assert np.isclose(float(w @ uN.a), base, rtol=1e-14)
