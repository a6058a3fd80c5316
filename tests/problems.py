"""Problem builders shared by the tests: the reference's driver pipeline
(cg_crs.cpp:376-484) run through the product's host functions, plus the
oracle's view of the same arrays."""
from __future__ import annotations

import os

import numpy as np

import iccg_amd as I
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SAAD = os.path.join(GOLDEN, "Saad.mm")
SEED = 0x1CC6


def orc(A) -> O.Csr:
    return O.Csr(A.n, np.asarray(A.rowptr, np.int32), np.asarray(A.colind, np.int32),
                 np.asarray(A.val, np.float64))


def grid(nx, ny=1, nz=1, coef=(0.0, 0.0), shift=0.0, seed=SEED):
    return I.gen_grid(nx, ny, nz, coef_exp=coef, diag_shift=shift, seed=seed)


class Ordered:
    """A permuted system A_run = P^T A P and its schedule, built as main()
    does (cg_crs.cpp:413-484)."""

    def __init__(self, A, sched, new_of_old, color, extra=None):
        self.A = A
        self.sched = sched
        self.new_of_old = new_of_old
        self.color = color
        self.extra = extra or {}


def mc_order(A, color=None):
    if color is None:
        color, nc = I.greedy_color(A)
    else:
        color = np.asarray(color, np.int32)
        nc = int(color.max()) + 1
    p = I.BuildPermutationByColor(color)
    Ap = I.expand_lower_to_full(I.Permute_PtAP_LowerCRS_to_LowerCRS(A, p))
    cp, rc = I.BuildColorPtrRows_FromPermutationAndColor(color, nc, p)
    s = I.ColorSchedule.mc(nc, cp, rc)
    return Ordered(Ap, s, p, color, {"nc": nc})


def abmc_order(A, block=None, grid_dims=None, box=(4, 4, 4)):
    if grid_dims is not None:
        bo, nb, bc, nc = I.abmc_blocking_grid(*grid_dims, *box)
    else:
        bo, nb, bc, nc = I.abmc_blocking(A, block)
    color = bc[bo]
    p = I.BuildPermutationByBlockColor(bo, bc)
    Ap = I.expand_lower_to_full(I.Permute_PtAP_LowerCRS_to_LowerCRS(A, p))
    cpb, boc, bp, rob = I.BuildABMCSchedule(p, bo, bc, nb, nc)
    s = I.ColorSchedule.abmc(nc, nb, cpb, boc, bp, rob)
    return Ordered(Ap, s, p, color, {"nb": nb, "nc": nc, "block_of_old": bo, "block_color": bc})


def orc_sched(s: I.ColorSchedule) -> O.Schedule:
    if s is None:
        return None
    return O.Schedule(nc=s.nc, color_ptr=s.color_ptr, rows_of_color=s.rows_of_color, mode=s.mode,
                      nb=s.nb, color_ptr_blk=s.color_ptr_blk, blocks_of_color=s.blocks_of_color,
                      block_ptr=s.block_ptr, rows_of_block=s.rows_of_block)


def rhs_ones_times_A(A):
    """b = A * 1 (the 100x100 Laplacian case's right-hand side) computed by the oracle's SpMV."""
    return O.spmv(orc(A), np.ones(A.n))
