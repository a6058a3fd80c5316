"""Study (test infrastructure, CPU only): does an FP32-valued operator inside
the CG recurrence reach ||b - A x|| / ||b|| < tol under true-residual
refinement, and at what iteration cost?

On the 3D 7-point variable-coefficient grid N^3 (face coefficients
10^U[-2,2], diagonal = sum + 1e-6), 4x4x4-box ABMC ordering, b ~ N(0,1):

  * the mixed solve as built (FP64 A in the recurrence, FP32 IC(0) apply,
    FP64 true-residual restarts): orc_cg_mixed;
  * refinement with A32 = A's values rounded to float in the recurrence:
    each cycle solves A32 d = b - A x (FP32 IC(0) apply, FP64 vectors,
    double-double dots) until its recurrence reaches tol ||b||, then
    x += d and the FP64 true residual is taken.

Usage: python oracle/fp32_operator_study.py N   (N = 192: about 10 minutes)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import iccg_amd as I  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests import problems as P  # noqa: E402


def main(N, tol=1e-8, cycles=10):
    A = P.grid(N, N, N, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(N, N, N), box=(4, 4, 4))
    del A
    b = I.gen_normal(o.A.n, P.SEED + 1)
    bp = np.empty_like(b)
    bp[o.new_of_old] = b
    Ao = P.orc(o.A)
    A32 = O.Csr(Ao.n, Ao.rowptr, Ao.colind, Ao.val.astype(np.float32).astype(np.float64))
    print(f"max relative change of the values in float: "
          f"{np.max(np.abs(A32.val - Ao.val) / np.abs(Ao.val)):.3e}", flush=True)
    f = O.ic0_factor(Ao)
    t = time.time()
    ref = O.cg_mixed(Ao, bp, f, tol=tol, max_iter=20000)
    print(f"[{N}^3] FP64 A in the recurrence: {ref.iters} iterations, restarts {ref.restarts}, "
          f"true {ref.true_rel_resid:.3e} ({time.time() - t:.0f}s)", flush=True)
    nb = np.linalg.norm(bp)
    x = np.zeros(Ao.n)
    total = 0
    for cyc in range(cycles):
        rt = bp - O.spmv(Ao, x)
        tr = O.true_resid_dd(Ao, x, bp)
        print(f"[{N}^3] A32 cycle {cyc}: true {tr:.3e} after {total} iterations", flush=True)
        if tr < tol:
            break
        t = time.time()
        c = O.cg(A32, rt, "ic0_f32", ic=f, tol=tol * nb / np.linalg.norm(rt), max_iter=20000,
                 dot_mode=1)
        x = x + c.x
        total += c.iters
        print(f"[{N}^3]   inner solve: {c.iters} iterations, converged {c.converged}, "
              f"rel {c.rel_resid:.3e} ({time.time() - t:.0f}s)", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 64)
