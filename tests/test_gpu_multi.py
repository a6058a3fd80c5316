"""Several right-hand sides of one system solved together (MultiSolver,
iccg_msolver_*).

Each right-hand side keeps its own CG recurrence, so the bar is the
single-system solver on that right-hand side alone, bit for bit: iterations,
restarts, recurrence residual, flags and x (and through it the oracle, which
the single-system path matches bit for bit).  The cases cover the shapes the
shared kernels take -- 3D boxes (three in-block slots), 2D boxes (two),
breadth-first blocks (six outside slots), sets of two and four, a set of four
with an idle member, a member left over for the single-system solver --
and the per-member control flow: a zero right-hand side that stops before
the first iteration, members stopped by max_iter, members restarting from
the true residual a different number of times.  (IC(0) of an SPD matrix
gives p.Ap > 0, so the breakdown branch, kept member by member as in
k_fin_alpha, has no case here.)
"""
import numpy as np
import pytest

import iccg_amd as I
from oracle import oracle as O
from tests import problems as P

pytestmark = pytest.mark.gpu


def _system(kind):
    if kind == "box3d":
        A = P.grid(16, 12, 8, coef=(-2.0, 2.0), shift=1e-6)
        o = P.abmc_order(A, grid_dims=(16, 12, 8), box=(4, 4, 4))
    elif kind == "box2d":
        A = P.grid(40, 36, coef=(-2.0, 2.0), shift=1e-6)
        o = P.abmc_order(A, grid_dims=(40, 36, 1), box=(8, 8, 1))
    elif kind == "bfs3d":
        A = P.grid(20, 18, 16, coef=(-2.0, 2.0), shift=1e-6)
        o = P.abmc_order(A, block=64)
    else:
        raise ValueError(kind)
    return o


def _rhs(n, nrhs, seed):
    return np.stack([I.gen_normal(n, seed + k) for k in range(nrhs)])


def _alone(A, M, B, X0, max_iter, tol):
    """every right-hand side through the single-system solver"""
    S = I.Solver(A, M)
    out, xs = [], []
    for k in range(B.shape[0]):
        S.set(B[k], None if X0 is None else X0[k])
        out.append(S.run(max_iter, tol))
        xs.append(S.x())
    return out, np.stack(xs)


def _same(ra, rb):
    return (ra.iters, ra.rel_resid, ra.converged, ra.breakdown, ra.restarts) == \
        (rb.iters, rb.rel_resid, rb.converged, rb.breakdown, rb.restarts)


@pytest.mark.parametrize("kind", ["box3d", "box2d", "bfs3d"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nrhs,layout", [(2, (1, 0)), (3, (1, 0)), (4, (1, 0)), (5, (1, 1)),
                                         (8, (2, 0))])
def test_multi_equals_single_bitwise(kind, precision, nrhs, layout):
    o = _system(kind)
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    assert M.plan()[0] == "abmc-sectioned-fused-r"
    B = _rhs(o.A.n, nrhs, 31)
    MS = I.MultiSolver(o.A, M, nrhs)
    assert MS.layout() == layout
    MS.set(B)
    res = MS.run(10000, 1e-8)
    X = MS.x()
    ref, Xr = _alone(o.A, M, B, None, 10000, 1e-8)
    for k in range(nrhs):
        assert _same(res[k], ref[k]), (k, res[k], ref[k])
        assert res[k].converged and res[k].true_rel_resid < 1e-8
        assert res[k].true_rel_resid == pytest.approx(ref[k].true_rel_resid, rel=1e-12)
    assert np.array_equal(X, Xr)


def test_multi_fp64_matches_oracle():
    """FP64 members against the double-double oracle's CG on each
    right-hand side: iterations, residual and x bit for bit."""
    o = _system("box3d")
    M = I.IC0(o.A, 0.0, sched=o.sched)
    B = _rhs(o.A.n, 4, 41)
    X = np.zeros_like(B)
    res = I.conjugate_gradient_multi(o.A, B, X, M, 10000, 1e-8)
    Ao = P.orc(o.A)
    f = O.ic0_factor(Ao, 0.0)
    for k in range(4):
        dd = O.cg(Ao, B[k], "ic0", ic=f, sched=P.orc_sched(o.sched), tol=1e-8, dot_mode=1)
        assert (res[k].iters, res[k].rel_resid) == (dd.iters, dd.rel_resid)
        assert np.array_equal(X[k], dd.x)


def test_multi_mixed_restarts_per_member():
    """Mixed precision at a tolerance the recurrence reaches before the true
    residual does: members restart from the FP64 true residual, each its own
    number of times (the zero member none), and each equals the oracle's
    mixed CG on its right-hand side bit for bit."""
    A = P.grid(24, 20, 16, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(24, 20, 16), box=(4, 4, 4))
    M = I.IC0(o.A, 0.0, sched=o.sched, precision="mixed")
    B = _rhs(o.A.n, 4, 51)
    B[2] = 0.0
    B[3] *= 1e3
    MS = I.MultiSolver(o.A, M, 4)
    MS.set(B)
    res = MS.run(3000, 1e-14)
    X = MS.x()
    Ao = P.orc(o.A)
    f = O.ic0_factor(Ao)
    restarts = []
    for k in range(4):
        m = O.cg_mixed(Ao, B[k], f, tol=1e-14, max_iter=3000)
        assert (res[k].iters, res[k].restarts, res[k].converged) == \
            (m.iters, m.restarts, m.converged), k
        assert res[k].rel_resid == m.rel_resid
        assert np.array_equal(X[k], m.x)
        restarts.append(res[k].restarts)
    assert restarts[2] == 0 and res[2].iters == 0
    assert max(restarts) >= 1


def test_multi_x0_max_iter_and_repeat():
    """Per-member starting vectors, members stopped by max_iter before they
    converge, a second run on the same solver (captured graphs reused) with
    new right-hand sides, and a third launched kernel by kernel with event
    timing: each equals the single-system solver."""
    o = _system("box2d")
    M = I.IC0(o.A, 0.0, sched=o.sched, precision="mixed")
    B = _rhs(o.A.n, 3, 61)
    X0 = _rhs(o.A.n, 3, 71) * 0.1
    MS = I.MultiSolver(o.A, M, 3)
    MS.set(B, X0)
    res = MS.run(25, 1e-8)
    ref, Xr = _alone(o.A, M, B, X0, 25, 1e-8)
    for k in range(3):
        assert _same(res[k], ref[k]) and res[k].iters == 25 and not res[k].converged
    assert np.array_equal(MS.x(), Xr)
    B2 = _rhs(o.A.n, 3, 81)
    MS.set(B2)
    res2 = MS.run(10000, 1e-8)
    ref2, Xr2 = _alone(o.A, M, B2, None, 10000, 1e-8)
    for k in range(3):
        assert _same(res2[k], ref2[k])
    assert np.array_equal(MS.x(), Xr2)
    # kernel-by-kernel launches with event timing: the same results, and
    # one launch group per class and set iteration in which a member ran
    MS.set(B2)
    MS.set_timing(True)
    res3 = MS.run(10000, 1e-8)
    MS.set_timing(False)
    for k in range(3):
        assert _same(res3[k], ref2[k])
    assert np.array_equal(MS.x(), Xr2)
    most = max(r.iters for r in res3)
    assert MS.timing(0)[0] == most and MS.timing(1)[0] == most
    assert all(MS.timing(c)[1] > 0.0 for c in range(4))


def test_multi_other_plans_run_alone():
    """A multicolor schedule is not the shared kernels' plan: every
    right-hand side runs on the single-system solver, with the same
    results."""
    A = P.grid(12, 12, 12, coef=(-2.0, 2.0), shift=1e-6)
    o = P.mc_order(A)
    M = I.IC0(o.A, 0.0, sched=o.sched)
    B = _rhs(o.A.n, 3, 101)
    MS = I.MultiSolver(o.A, M, 3)
    assert MS.layout() == (0, 3)
    MS.set(B)
    res = MS.run(10000, 1e-8)
    ref, Xr = _alone(o.A, M, B, None, 10000, 1e-8)
    for k in range(3):
        assert _same(res[k], ref[k])
    assert np.array_equal(MS.x(), Xr)
