"""Device memory reuse.  The library keeps the device memory of freed
objects in a pool of its own and hands it to later objects with the earlier
objects' data still in it, and keeps pinned host staging the same way; a
buffer that must start at zero is cleared explicitly.  So a system solved
again with new objects, after other systems (single-system, several
right-hand sides, batched) were solved and freed, gives the same iterations,
residuals and x bit for bit, and so it does after iccg_device_trim has
handed the idle memory back."""
import gc

import numpy as np
import pytest

import iccg_amd as I
from tests import problems as P

pytestmark = pytest.mark.gpu


def _system():
    A = P.grid(16, 12, 8, coef=(-2.0, 2.0), shift=1e-6)
    return P.abmc_order(A, grid_dims=(16, 12, 8), box=(4, 4, 4))


def _solve(o, b, precision):
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    S = I.Solver(o.A, M)
    S.set(b)
    r = S.run(10000, 1e-8)
    return (r.iters, r.rel_resid, r.converged, r.true_rel_resid), S.x()


def _churn():
    """other systems, larger and smaller, through every allocating path"""
    A = P.grid(40, 36, 20, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(40, 36, 20), box=(4, 4, 4))
    B = np.stack([I.gen_normal(o.A.n, 900 + k) for k in range(4)])
    for precision in ("fp64", "mixed"):
        M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
        S = I.Solver(o.A, M)
        S.set(B[0])
        assert S.run(10000, 1e-8).converged
        MS = I.MultiSolver(o.A, M, 4)
        MS.set(B)
        assert all(r.converged for r in MS.run(10000, 1e-8))
        del S, MS, M
    small = P.abmc_order(P.grid(9, 7, 5, coef=(-2.0, 2.0), shift=1e-6),
                         grid_dims=(9, 7, 5), box=(4, 4, 4))
    Bt = I.Batch(8)
    for k in range(3):
        Bt.add(small.A, I.IC0(small.A, 0.0, sched=small.sched, precision="mixed"),
               I.gen_normal(small.A.n, 950 + k))
    assert Bt.run() == 0
    del Bt
    gc.collect()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reused_memory_gives_the_same_solve(precision):
    o = _system()
    b = I.gen_normal(o.A.n, 77)
    ref, xref = _solve(o, b, precision)
    assert ref[2]
    _churn()
    again, x = _solve(o, b, precision)
    assert again == ref and np.array_equal(x, xref)
    gc.collect()
    I.device_trim()
    after, x2 = _solve(o, b, precision)
    assert after == ref and np.array_equal(x2, xref)
