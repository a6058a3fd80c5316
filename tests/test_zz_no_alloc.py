"""iccg_cg_solve_noalloc, the CG solve that allocates nothing (its device
buffers live in the library's per-device scratch array), against
iccg_cg_solve on the same inputs, called through the C ABI: iterations,
residuals, restarts, flags and x bit for bit, on two calls in a row, the
second with a smaller system; systems over the scratch limit are refused
and leave the entry point working."""
import ctypes as C

import numpy as np
import pytest

import iccg_amd as I
from iccg_amd import _capi
from tests import problems as P

pytestmark = pytest.mark.gpu


def _noalloc(dA, M, b, x0, max_iter=10000, tol=1e-8, true_resid=False):
    x = np.array(x0, dtype=np.float64, copy=True)
    res = _capi.CGResultC()
    flags = _capi.SOLVE_TRUE_RESID if true_resid else 0
    _capi.check(_capi.lib().iccg_cg_solve_noalloc(
        dA.handle, M.handle, b.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p),
        max_iter, tol, flags, C.byref(res)))
    return x, res


def _same(a, b):
    return ((a.iters, a.rel_resid, a.converged, a.breakdown, a.restarts, a.true_rel_resid) ==
            (b.iters, b.rel_resid, bool(b.converged), bool(b.breakdown), b.restarts,
             b.true_rel_resid))


def _case(kind):
    if kind == "abmc_mixed":  # 40^3 variable coefficients, 4x4x4 boxes, mixed
        A = P.grid(40, 40, 40, coef=(-2.0, 2.0), shift=1e-6)
        o = P.abmc_order(A, grid_dims=(40, 40, 40), box=(4, 4, 4))
        return o.A, I.IC0(o.A, 0.0, sched=o.sched, precision="mixed"), o.sched, True
    if kind == "mc_fp64":  # 2D 50x50, red-black, FP64
        o = P.mc_order(P.grid(50, 50, coef=(-1.0, 1.0)))
        return o.A, I.IC0(o.A, 0.0, sched=o.sched), o.sched, False
    A = P.grid(30, 20, coef=(-1.0, 1.0))
    return A, (I.Jacobi(A) if kind == "jacobi" else I.Identity(A)), None, False


@pytest.mark.parametrize("first,second", [("abmc_mixed", "mc_fp64"), ("mc_fp64", "jacobi"),
                                          ("abmc_mixed", "identity")])
def test_noalloc_equals_cg_solve_twice(first, second):
    """Two calls in a row, the second on a smaller system (the scratch array
    keeps the first call's contents): both equal iccg_cg_solve bit for bit."""
    for k, kind in enumerate((first, second)):
        A, M, sched, mixed = _case(kind)
        dA = I.DeviceCSR(A)
        b = I.gen_normal(A.n, 40 + k)
        x0 = np.zeros(A.n) if k == 0 else 1e-3 * I.gen_normal(A.n, 50)
        x_ref = x0.copy()
        r_ref = I.conjugate_gradient(dA, b, x_ref, M, 10000, 1e-8, sched)
        x, r = _noalloc(dA, M, b, x0, true_resid=mixed)
        assert r_ref.converged and _same(r_ref, r)
        assert np.array_equal(x, x_ref)


def test_noalloc_refuses_oversize_and_keeps_working():
    """A system needing more than the 64 MiB scratch (1200^2 rows with FP64
    IC(0): about 56 bytes per row) is refused with ICCG_ERR_ARG; the next
    call solves as before."""
    big = P.grid(1200, 1200)
    dbig = I.DeviceCSR(big)
    Mbig = I.IC0(dbig, 0.0)
    with pytest.raises(I.IccgError, match="scratch") as e:
        _noalloc(dbig, Mbig, np.ones(big.n), np.zeros(big.n))
    assert e.value.status == _capi.ICCG_ERR_ARG
    del Mbig, dbig
    A, M, sched, mixed = _case("mc_fp64")
    dA = I.DeviceCSR(A)
    b = I.gen_normal(A.n, 60)
    x_ref = np.zeros(A.n)
    r_ref = I.conjugate_gradient(dA, b, x_ref, M, 10000, 1e-8, sched)
    x, r = _noalloc(dA, M, b, np.zeros(A.n))
    assert _same(r_ref, r) and np.array_equal(x, x_ref)
