"""Writes the golden fixtures of tests/golden/ from the reference itself.

Run in a container where the reference sources exist and oracle/_ref is
built (make -C oracle).  The reference is compiled from its own sources by
oracle/Makefile; this script only calls it and stores inputs/outputs:

  ref_saad.npz      Saad.mm CRS, and for Identity / Jacobi / IC0 at tol 1e-8
                    and 1e-14: iters, rel_resid, x; the IC(0) factor L and
                    L^T (CSC) values.
  ref_grids.npz     for small synthetic grids (2D 5-point 20x20 variable
                    coefficient, 3D 7-point 8x7x6 variable coefficient) in
                    natural, red-black/greedy MC and ABMC orderings with
                    b ~ N(0,1): permutation, schedule arrays, iters,
                    rel_resid, x, and the factor values.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref as R  # noqa: E402
from oracle.oracle import Csr, Schedule  # noqa: E402
from tests import problems as P  # noqa: E402


def main():
    if not R.available():
        raise SystemExit("oracle/_ref/libref.so is not built")
    out = {}
    A = R.read_matrix_market(os.path.join(HERE, "Saad.mm"))
    out.update(rowptr=A.rowptr, colind=A.colind, val=A.val)
    for kind in ("none", "jacobi", "ic0"):
        for tol in (1e-8, 1e-14):
            it, rel, conv, x = R.cg(A, np.ones(A.n), kind, tol=tol)
            key = f"{kind}_{tol:g}"
            out[key + "_iters"] = it
            out[key + "_rel"] = rel
            out[key + "_x"] = x
    f = R.ic0_factor(A, 0.0)
    out.update(L_val=f.Lv, LT_val=f.LTv, L_colind=f.Lci, L_rowptr=f.Lrp)
    np.savez(os.path.join(HERE, "ref_saad.npz"), **out)

    g = {}
    cases = {"v2d": P.grid(20, 20, coef=(-1.0, 1.0)),
             "v3d": P.grid(8, 7, 6, coef=(-2.0, 2.0), shift=1e-6)}
    for name, A in cases.items():
        b = P.I.gen_normal(A.n, P.SEED + 1)
        for order in ("natural", "mc", "abmc"):
            if order == "natural":
                Ar, s, perm = A, None, np.arange(A.n, dtype=np.int32)
            elif order == "mc":
                o = P.mc_order(A)
                Ar, s, perm = o.A, o.sched, o.new_of_old
            else:
                o = P.abmc_order(A, block=8)
                Ar, s, perm = o.A, o.sched, o.new_of_old
            Ao = Csr(Ar.n, Ar.rowptr, Ar.colind, Ar.val)
            so = P.orc_sched(s)
            it, rel, conv, x = R.cg(Ao, b, "ic0", sched=so, tol=1e-8)
            f = R.ic0_factor(Ao, 0.0)
            k = f"{name}_{order}"
            g[k + "_perm"] = perm
            g[k + "_rowptr"], g[k + "_colind"], g[k + "_val"] = Ar.rowptr, Ar.colind, Ar.val
            if s is not None:
                for fld in ("color_ptr", "rows_of_color", "color_ptr_blk", "blocks_of_color",
                            "block_ptr", "rows_of_block"):
                    a = getattr(s, fld)
                    if a is not None:
                        g[k + "_" + fld] = a
                g[k + "_nc"], g[k + "_nb"], g[k + "_mode"] = s.nc, s.nb, s.mode
            g[k + "_b"] = b
            g[k + "_iters"], g[k + "_rel"], g[k + "_x"] = it, rel, x
            g[k + "_Lv"], g[k + "_LTv"] = f.Lv, f.LTv
    np.savez(os.path.join(HERE, "ref_grids.npz"), **g)
    print("wrote ref_saad.npz, ref_grids.npz")


if __name__ == "__main__":
    main()
