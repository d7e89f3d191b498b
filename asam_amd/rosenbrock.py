"""Linear Rosenbrock stepping of the stage system and its adjoint, on the GPU C-ABI.

Stage form and coefficients are the reference's Ros3P (SRC/ADVECT/Int_Mod.f90:149-163,
StageRos :263-367): with beta0 the method's gamma,

    (I - beta0 dt J) k_i = dt f(U_i) + sum_{j<i} beta(i,j) k_j,
    U_1 = u_n,  U_{i+1} = u_n + sum_{j<=i} alpha(i,j) k_j,  u_{n+1} = U_{nStage+1}.

The linear model is f(u) = J u + s with V (I - beta0 dt J) = A, the sound operator
assembled with beta0*dtP = beta0*dt (V = cell volumes, the mass of the stage system),
i.e. V dt J = -(A - V)/beta0.  Each stage is then one SolveSound call:

    A k_i = -(A - V) U_i / beta0 + dt V s + V sum_j beta(i,j) k_j.

Reverse mode: the adjoint of one step runs the stages backwards with the transposed
solve (asam_sound_solve_T) and the transposed operator; the sensitivity to the cell
theta that enters A through the face theta uses asam_sound_theta_vjp_device.  Forward
states are checkpointed every `window` steps; the backward sweep recomputes one window
at a time from its checkpoint (stage vectors of that window only are held).
"""
from __future__ import annotations

import numpy as np

from . import workload as W

# Ros3P, reference SRC/ADVECT/Int_Mod.f90:149-163 (alpha, beta already times beta0 there)
ROS3P_BETA0 = 7.886751345948129e-1
_A = np.zeros((3, 3))
_A[0, 0] = 1.267949192431123
_A[1, 0] = 1.267949192431123
_A[2, 0] = 2.000000000000000
_A[2, 1] = 5.773502691896258e-1
_A[2, 2] = 4.226497308103742e-1
_B = np.zeros((3, 3))
_B[1, 0] = -1.607695154586736
_B[2, 0] = -3.464101615137755
_B[2, 1] = -1.732050807568877
ROS3P_ALPHA = _A * ROS3P_BETA0
ROS3P_BETA = _B * ROS3P_BETA0


def face_weights(prob, theta):
    """Face theta (volume-weighted mean of the two cells, wall faces copy the adjacent
    cell) and DUT of the dry state -- the LaplTCompute inputs, on the host."""
    V = prob.VolC
    eps = 1e-30
    g = prob.grid
    DTU = np.empty((g.nz, g.ny, g.nx + 1))
    DTU[:, :, 0] = theta[:, :, 0]
    DTU[:, :, -1] = theta[:, :, -1]
    DTU[:, :, 1:-1] = (theta[:, :, 1:] * V[:, :, 1:] + theta[:, :, :-1] * V[:, :, :-1]) / (
        V[:, :, 1:] + V[:, :, :-1] + eps)
    DTV = np.empty((g.nz, g.ny + 1, g.nx))
    DTV[:, 0, :] = theta[:, 0, :]
    DTV[:, -1, :] = theta[:, -1, :]
    DTV[:, 1:-1, :] = (theta[:, 1:, :] * V[:, 1:, :] + theta[:, :-1, :] * V[:, :-1, :]) / (
        V[:, 1:, :] + V[:, :-1, :] + eps)
    DTW = np.empty((g.nz + 1, g.ny, g.nx))
    DTW[0] = theta[0]
    DTW[-1] = theta[-1]
    DTW[1:-1] = (theta[1:] * V[1:] + theta[:-1] * V[:-1]) / (V[1:] + V[:-1] + eps)
    kappa = W.RD / W.CPD
    rhoD = prob.rho + eps
    DUT = prob.T / (prob.rhotheta + eps) * (rhoD * W.RD) / (1.0 - kappa)
    return DTU, DTV, DTW, DUT


class LinearRosenbrock:
    """Ros3P steps u -> u' of the linear stage model on one solver context.

    The solver supplies device vectors and the operations used here: vec, alloc, zero,
    axpby, mass (V x), apply_device (MatVec1 or its transpose), solve_device (SolveSound
    or the transposed solve) and theta_vjp; SoundSolver is the GPU implementation."""

    def __init__(self, solver, dt, tol=1e-13, maxiter=500):
        self.s = solver
        self.dt = dt
        self.tol = tol
        self.maxiter = maxiter
        self.n = solver.n
        self.N = solver.grid.nx * solver.grid.ny * solver.grid.nz
        self._tmp = [solver.vec() for _ in range(4)]
        self.solves = 0
        self.iterations = 0

    # -- device helpers ------------------------------------------------------------
    def _zero(self, b):
        self.s.zero(b)

    def _copy(self, dst, src):
        self._zero(dst)
        self.s.axpby(1.0, src, 0.0, dst)

    def _K(self, y, x, transpose=False):
        """y = (A - V) x  (or (A - V)^T x); MatVec1 gives -A x."""
        t = self._tmp[0]
        self.s.apply_device(t, x, transpose)   # t = -A x
        self.s.mass(y, x)                      # y = V x
        self.s.axpby(-1.0, t, -1.0, y)         # y = A x - V x

    def _solve(self, k, R, transpose=False):
        """A k = R  (SolveSound solves -A x = b, so b = -R)."""
        b = self._tmp[1]
        self._zero(b)
        self.s.axpby(-1.0, R, 0.0, b)
        self._zero(k)
        mi, tl, info = self.s.solve_device(k, b, self.maxiter, self.tol, transpose)
        self.solves += 1
        self.iterations += info.iterations
        if info.status != 0:
            raise RuntimeError(f"stage solve did not converge: {info.as_dict()}")

    # -- forward -------------------------------------------------------------------
    def step(self, u_out, u, src, keep=None):
        """u_out = one Ros3P step from u with source src; keep: list that receives the
        stage states U_i and increments k_i (device buffers) for the adjoint."""
        s = self.s
        ks, Us = [], []
        R = self._tmp[2]
        for i in range(3):
            U = s.vec()
            self._copy(U, u)
            for j in range(i):
                s.axpby(ROS3P_ALPHA[i - 1, j], ks[j], 1.0, U)
            self._K(R, U)
            s.axpby(0.0, R, -1.0 / ROS3P_BETA0, R)           # R = -(A - V) U / beta0
            vs = self._tmp[3]
            s.mass(vs, src)
            s.axpby(self.dt, vs, 1.0, R)                      # + dt V s
            for j in range(i):
                s.mass(vs, ks[j])
                s.axpby(ROS3P_BETA[i, j], vs, 1.0, R)         # + V beta(i,j) k_j
            k = s.vec()
            self._solve(k, R)
            ks.append(k)
            Us.append(U)
        self._copy(u_out, u)
        for j in range(3):
            s.axpby(ROS3P_ALPHA[2, j], ks[j], 1.0, u_out)
        if keep is not None:
            keep.append((Us, ks))

    def forward(self, u0, src, nsteps, window):
        """Returns u_N and the checkpoints {step: state} taken every `window` steps."""
        u = self.s.vec()
        self._copy(u, u0)
        ckpt = {}
        nxt = self.s.vec()
        for n in range(nsteps):
            if n % window == 0:
                c = self.s.vec()
                self._copy(c, u)
                ckpt[n] = c
            self.step(nxt, u, src)
            u, nxt = nxt, u
        return u, ckpt

    # -- reverse -------------------------------------------------------------------
    def step_adjoint(self, ubar, sbar, thbar, Us, ks):
        """Back-propagate one step: ubar (in: dJ/du', out: dJ/du), accumulate dJ/ds in
        sbar and dJ/dtheta (cells) in thbar."""
        s = self.s
        kbar = [s.vec() for _ in range(3)]
        for j in range(3):
            s.axpby(ROS3P_ALPHA[2, j], ubar, 0.0, kbar[j])
        uacc = s.vec()
        self._copy(uacc, ubar)
        Rbar = s.vec()
        Ubar = s.vec()
        vs = self._tmp[3]
        for i in (2, 1, 0):
            self._solve(Rbar, kbar[i], transpose=True)       # A^T Rbar = kbar_i
            # k_i = A^-1 R_i: dtheta += -Rbar^T (dA/dtheta) k_i
            # R_i = -(A - V) U_i / beta0 + ...: dtheta += -Rbar^T (dA/dtheta) U_i / beta0
            self._theta_acc(thbar, Rbar, ks[i], -1.0)
            self._theta_acc(thbar, Rbar, Us[i], -1.0 / ROS3P_BETA0)
            s.mass(vs, Rbar)
            s.axpby(self.dt, vs, 1.0, sbar)                   # R_i has + dt V s
            for j in range(i):
                s.axpby(ROS3P_BETA[i, j], vs, 1.0, kbar[j])   # + V beta(i,j) k_j
            self._K(Ubar, Rbar, transpose=True)
            s.axpby(0.0, Ubar, -1.0 / ROS3P_BETA0, Ubar)      # Ubar = -(A - V)^T Rbar / beta0
            s.axpby(1.0, Ubar, 1.0, uacc)                     # U_i = u + sum alpha(i-1,j) k_j
            for j in range(i):
                s.axpby(ROS3P_ALPHA[i - 1, j], Ubar, 1.0, kbar[j])
        self._copy(ubar, uacc)

    def _theta_acc(self, thbar, mu, x, scale):
        g = self._thtmp
        self._zero(g)
        self.s.theta_vjp(g, mu, x)
        self.s.axpby(scale, g, 1.0, thbar, n=self.N)

    def gradient(self, u0, src, nsteps, window, ubarN):
        """dJ/du0, dJ/ds, dJ/dtheta for J with dJ/du_N = ubarN, recomputing each window of
        steps forward from its checkpoint."""
        s = self.s
        self._thtmp = s.alloc(self.N * 8)
        uN, ckpt = self.forward(u0, src, nsteps, window)
        ubar = s.vec()
        self._copy(ubar, ubarN)
        sbar = s.vec()
        thbar = s.alloc(self.N * 8)
        self._zero(thbar)
        starts = sorted(ckpt)
        for w0 in reversed(starts):
            w1 = min(w0 + window, nsteps)
            keep = []
            u = s.vec()
            self._copy(u, ckpt[w0])
            nxt = s.vec()
            for n in range(w0, w1):
                self.step(nxt, u, src, keep)
                u, nxt = nxt, u
            for n in reversed(range(w0, w1)):
                Us, ks = keep[n - w0]
                self.step_adjoint(ubar, sbar, thbar, Us, ks)
        return uN, ubar, sbar, thbar
