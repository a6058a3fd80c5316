"""The two GPU configurations at their full sizes, through properties that
hold at any size (the oracle's CG would take minutes there):

- the GPU IC(0) factor satisfies (L L^T)_ij = a_ij on the stored pattern, to
  rounding (the defining identity of zero-fill IC);
- a multicolor / ABMC apply equals the natural-order (level-scheduled)
  apply on the same permuted factor, bit for bit (a valid coloring only
  reorders independent rows);
- the solves converge, the FP64 true residual (oracle SpMV) meets the
  tolerance, and mixed precision agrees with FP64.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import iccg_amd as I
from oracle import oracle as O
from tests import problems as P

pytestmark = pytest.mark.gpu


def _permuted_rhs(o, seed):
    b = I.gen_normal(o.A.n, seed)
    bp = np.empty_like(b)
    bp[o.new_of_old] = b  # b'[new_of_old[i]] = b[i]
    return bp


def _ic0_identity_error(A, M):
    Lrp, Lci, Lv = M.factor()[:3]
    n = A.n
    L = sp.csr_matrix((Lv, Lci, Lrp), shape=(n, n))
    lower = sp.tril(A.to_scipy()).tocsr()
    mask = lower.copy()
    mask.data[:] = 1.0
    diff = (L @ L.T).multiply(mask) - lower
    return abs(diff).max() / abs(lower).max()


def test_config2_128cube_red_black_fp64():
    """3D 7-point Laplacian 128^3 (n = 2,097,152), red-black IC(0), FP64,
    b ~ N(0,1), tol 1e-8."""
    N = 128
    o = P.mc_order(P.grid(N, N, N))
    assert o.sched.nc == 2 and o.A.nnz == 14_581_760
    M = I.IC0(o.A, 0.0, sched=o.sched)
    assert _ic0_identity_error(o.A, M) < 1e-14
    r = I.gen_normal(o.A.n, 77)
    z_mc = M.apply(r, o.sched)
    z_nat = M.apply(r, None)
    assert np.array_equal(z_mc, z_nat)
    bp = _permuted_rhs(o, P.SEED + 2)
    x = np.zeros(o.A.n)
    res = I.conjugate_gradient(o.A, bp, x, M, 10000, 1e-8, o.sched)
    assert res.converged and res.rel_resid < 1e-8 and not res.breakdown
    assert O.true_resid(P.orc(o.A), x, bp) < 1.5e-8


def test_config3_192cube_abmc_mixed_and_fp64():
    """3D variable-coefficient 192^3 (n = 7,077,888), ABMC 4x4x4 boxes, FP64
    and mixed precision, b ~ N(0,1), tol 1e-8 (true residual for mixed)."""
    N = 192
    A = P.grid(N, N, N, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(N, N, N), box=(4, 4, 4))
    del A
    dA = I.DeviceCSR(o.A)
    M64 = I.IC0(dA, 0.0, sched=o.sched)
    r = I.gen_normal(o.A.n, 78)
    assert np.array_equal(M64.apply(r, o.sched), M64.apply(r, None))
    bp = _permuted_rhs(o, P.SEED + 1)
    x64 = np.zeros(o.A.n)
    r64 = I.conjugate_gradient(dA, bp, x64, M64, 20000, 1e-8, o.sched)
    del M64
    M32 = I.IC0(dA, 0.0, sched=o.sched, precision="mixed")
    x32 = np.zeros(o.A.n)
    r32 = I.conjugate_gradient(dA, bp, x32, M32, 20000, 1e-8, o.sched)
    assert r64.converged and r32.converged and r32.true_rel_resid < 1e-8
    Ao = P.orc(o.A)
    assert O.true_resid(Ao, x32, bp) < 1e-8
    assert O.true_resid(Ao, x64, bp) < 1.5e-8
    assert abs(r32.iters - r64.iters) <= max(3, r64.iters // 50)
    assert np.linalg.norm(x32 - x64) / np.linalg.norm(x64) < 1e-6
