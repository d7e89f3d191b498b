"""Fused and folded scalar tails (tuning fused_tail = 1 and 2; the default is 0).  Fused: the
BiCGStab and refinement tails (alpha; half-step test and omega; exit tests and beta; the
refinement residual) run in the last workgroup of the kernel that produced their sums,
instead of in a finalize launch after it.  Folded: the three BiCGStab tails run in every
workgroup of the kernel that needs their scalars (the s-step line solve, the update, the
next p-step line solve).  Both add the partial sums in the finalize kernel's order, so all
three forms must give the same bits: solution, MaxIter/Tol returns, iteration and
refinement counts, residuals.  Three contexts run the same solve sequence, one per form,
so the first solve of a context (no refinement schedule yet) and the replayed one are both
compared."""
import numpy as np
import pytest

from asam_amd import workload as W

pytestmark = pytest.mark.gpu


def ridge_small(np1):
    return W.make_problem(70, 5, 7, 1000.0, 1000.0, 250.0, H=600.0, a=2500.0,
                          ridge_xc=35000.0, np1=np1, grad_full=1, seed=3)


CASES = {
    "64x64x40-T": lambda: W.config(1, np1=1),
    "64x64x40-TB": lambda: W.config(1, np1=2),
    "ragged-70x5x7-TB": lambda: ridge_small(2),
    "uneven-TB": lambda: W.config(4, np1=2, scale=0.25),
}


def solve_seq(s, p, transpose, calls):
    f = s.SolveSound_T if transpose else s.SolveSound
    out = []
    for b, maxiter in calls:
        x, mi, tl, info = f(np.zeros(p.n), b, maxiter, 1e-8)
        out.append((x, mi, tl, info.iterations, info.restarts, info.status,
                    info.abs_residual))
    return out


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
@pytest.mark.parametrize("transpose", [False, True], ids=["forward", "transposed"])
@pytest.mark.parametrize("case", list(CASES))
def test_fused_tails_give_the_same_bits(case, transpose, precision, sound_mod):
    p = CASES[case]()
    b2 = p.b * (1.0 + 0.25 * np.sin(np.arange(p.n)))
    calls = [(p.b, 300), (p.b, 300), (b2, 300)]
    if precision == "FP64":
        calls += [(p.b, 3), (np.zeros(p.n), 300)]  # MaxIter exhausted; rho = 0 at once
    runs = {}
    for ft in (0, 1, 2):
        s = sound_mod.SoundSolver.from_problem(p)
        assert s.get_tuning()["fused_tail"] == 0  # the default
        s.set_tuning(fused_tail=ft)
        s.set_control(method="BICGSTAB", precision=precision)
        runs[ft] = solve_seq(s, p, transpose, calls)
    for ft in (1, 2):
        for q, (a, b) in enumerate(zip(runs[0], runs[ft])):
            print(f"{case} {precision} {'T' if transpose else 'F'} fused_tail {ft} call {q}: "
                  f"MaxIter {a[1]} / {b[1]}, inner {a[3]} / {b[3]}, refinements {a[4]} / "
                  f"{b[4]}, |r| {a[6]:.6e} / {b[6]:.6e}, max|dx| {np.abs(a[0] - b[0]).max():.1e}")
            assert np.array_equal(a[0], b[0])
            assert a[1:] == b[1:]


def test_fused_tails_remove_the_finalize_launches(sound_mod):
    """Profiled FP64 BiCGStab: no finalize launch with fused tails; three per iteration
    plus the initial one without; with folded tails the initial one and the exit test of
    the last iteration of the batch (a replayed solve issues its iterations in one batch)."""
    p = W.config(1, np1=2)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method="BICGSTAB", precision="FP64")
    counts = {}
    for ft in (1, 0, 2):
        s.set_tuning(fused_tail=ft)
        s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)  # the iteration count for the batch
        s.profile(True)
        x, mi, tl, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
        fins = s.profile_read()["reduce"]["launches"]
        s.profile(False)
        counts[ft] = (fins, info.iterations)
        print(f"fused_tail {ft}: {info.iterations} iterations, {fins} finalize launches")
    assert counts[1][0] == 0
    assert counts[0][0] >= 3 * counts[0][1]
    assert counts[2][0] == 2 and counts[2][1] == counts[0][1]
