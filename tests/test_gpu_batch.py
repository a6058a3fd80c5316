"""Multi-system solves (iccg_batch_*): each system gives in a batch the
iterates it gives alone -- whatever else is in the batch, whenever it was
admitted, and whatever work decomposition the batch runs."""
import numpy as np
import pytest

import iccg_amd as I
from oracle import oracle as O
from tests import problems as P

pytestmark = pytest.mark.gpu


def _f32_grid(nx, ny, nz, seed, coef=(-1.0, 1.0)):
    return I.gen_grid(nx, ny, nz, coef_exp=coef, seed=seed, fp32_values=True)


def _system(nx, ny, nz, seed, kind="mc", f32=True):
    A = _f32_grid(nx, ny, nz, seed) if f32 else I.gen_grid(nx, ny, nz, coef_exp=(-1.0, 1.0),
                                                           seed=seed)
    o = P.mc_order(A) if kind == "mc" else P.abmc_order(A, grid_dims=(nx, ny, nz),
                                                        box=(4, 4, 4) if nz > 1 else (8, 8, 1))
    b = I.gen_normal(A.n, seed + 7)
    return o, b


def _alone(o, b, precision):
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    x = np.zeros(o.A.n)
    r = I.conjugate_gradient(o.A, b, x, M, 10000, 1e-8, o.sched)
    return r, x


def _check_same(B, slots, alone):
    for s, (r, x) in zip(slots, alone):
        fin, rb = B.result(s)
        assert fin and rb.converged
        assert (rb.iters, rb.rel_resid, rb.restarts) == (r.iters, r.rel_resid, r.restarts)
        assert rb.true_rel_resid == r.true_rel_resid
        assert np.array_equal(B.x(s), x)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_equals_alone(precision):
    systems = [_system(n, m, 1, 100 + k) for k, (n, m) in
               enumerate([(32, 32), (48, 40), (17, 90), (64, 64), (20, 20), (33, 47)])]
    alone = [_alone(o, b, precision) for o, b in systems]
    B = I.Batch(16)
    slots = [B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched, precision=precision), b)
             for o, b in systems]
    assert B.run() == 0
    _check_same(B, slots, alone)
    for s, (o, b) in zip(slots, systems):
        assert O.true_resid(P.orc(o.A), B.x(s), b) < (1e-8 if precision == "mixed" else 1e-7)


def test_batch_mixes_orderings_and_precisions():
    """MC and ABMC systems, FP64 and mixed preconditioners, FP32-exact and
    FP64 matrix values, 2D and 3D, all in one batch."""
    specs = [((40, 40, 1), "mc", "mixed", True), ((16, 16, 16), "abmc", "mixed", False),
             ((24, 24, 1), "abmc", "fp64", True), ((12, 10, 9), "mc", "fp64", False),
             ((20, 20, 20), "abmc", "fp64", False), ((64, 16, 1), "mc", "mixed", False)]
    systems, alone, Ms = [], [], []
    for k, (dims, kind, prec, f32) in enumerate(specs):
        o, b = _system(*dims, 300 + k, kind=kind, f32=f32)
        systems.append((o, b))
        alone.append(_alone(o, b, prec))
        Ms.append(I.IC0(o.A, 0.0, sched=o.sched, precision=prec))
    B = I.Batch(len(specs))
    slots = [B.add(o.A, M, b) for (o, b), M in zip(systems, Ms)]
    assert B.run() == 0
    _check_same(B, slots, alone)


@pytest.mark.parametrize("tuning", [(4, 1), (32, 4), (1, 2)])
def test_batch_work_decomposition_does_not_change_results(tuning):
    systems = [_system(n, n, 1, 400 + n, kind=k) for n, k in ((48, "mc"), (40, "abmc"),
                                                              (72, "mc"))]
    alone = [_alone(o, b, "mixed") for o, b in systems]
    B = I.Batch(4)
    B.tuning = tuning
    assert B.tuning == tuning
    slots = [B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched, precision="mixed"), b)
             for o, b in systems]
    assert B.run() == 0
    _check_same(B, slots, alone)


def test_batch_autotune_keeps_results():
    """Enough rows (>= 65,536) for the autotuner to run; results unchanged."""
    systems = [_system(n, n, 1, 500 + n) for n in (160, 128, 200)]
    alone = [_alone(o, b, "mixed") for o, b in systems]
    B = I.Batch(4)
    B.set_autotune(True)
    slots = [B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched, precision="mixed"), b)
             for o, b in systems]
    assert B.run() == 0
    ch, cw = B.tuning
    assert ch in (4, 8, 16, 32) and cw in (1, 2, 4)
    _check_same(B, slots, alone)


def test_batch_fp64_matches_oracle_and_continuous_admission():
    systems = [_system(n, n, 1, 200 + n, f32=False) for n in (24, 40, 56, 30)]
    B = I.Batch(4)
    first = []
    for o, b in systems[:2]:
        first.append(B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), b))
    B.run(max_steps=5)  # the first two are mid-solve when the others arrive
    later = []
    for o, b in systems[2:]:
        later.append(B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), b))
    while B.run(max_steps=7):
        pass
    for s, (o, b) in zip(first + later, systems):
        fin, r = B.result(s)
        Ao = P.orc(o.A)
        dd = O.cg(Ao, b, "ic0", ic=O.ic0_factor(Ao), sched=P.orc_sched(o.sched), tol=1e-8,
                  dot_mode=1)
        assert fin and (r.iters, r.rel_resid) == (dd.iters, dd.rel_resid)
        assert np.array_equal(B.x(s), dd.x)
    # a removed slot takes a new system (its buffers are reused)
    B.remove(first[0])
    o, b = systems[1]
    s = B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), b)
    assert B.run() == 0
    r1, x1 = _alone(o, b, "fp64")
    _check_same(B, [s], [(r1, x1)])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_bfs_blocks_with_boxes(precision):
    """Breadth-first ABMC blocks (up to six outside entries per row in the
    sectioned layout) beside 4x4x4-box systems (three) in one batch: every
    system gives its single-system iterates, and in FP64 the double-double
    oracle's."""
    systems = []
    for k, dims in enumerate([(20, 18, 16), (16, 16, 16), (60, 50, 1), (24, 20, 12)]):
        A = I.gen_grid(*dims, coef_exp=(-1.0, 1.0), seed=700 + k)
        o = (P.abmc_order(A, block=64) if k % 2 == 0 else
             P.abmc_order(A, grid_dims=dims, box=(4, 4, 4)))
        systems.append((o, I.gen_normal(A.n, 710 + k)))
    Ms = [I.IC0(o.A, 0.0, sched=o.sched, precision=precision) for o, _ in systems]
    assert [M.plan() for M in Ms] == [("abmc-sectioned-fused-r", w) for w in (3, 3, 2, 3)]
    alone = [_alone(o, b, precision) for o, b in systems]
    B = I.Batch(len(systems))
    slots = [B.add(o.A, M, b) for (o, b), M in zip(systems, Ms)]
    assert B.run() == 0
    _check_same(B, slots, alone)
    if precision == "fp64":
        for s, (o, b) in zip(slots, systems):
            Ao = P.orc(o.A)
            dd = O.cg(Ao, b, "ic0", ic=O.ic0_factor(Ao), sched=P.orc_sched(o.sched), tol=1e-8,
                      dot_mode=1)
            assert B.result(s)[1].iters == dd.iters
            assert np.array_equal(B.x(s), dd.x)


def test_batch_remove_while_running():
    systems = [_system(n, n, 1, 600 + n) for n in (64, 48, 56)]
    B = I.Batch(3)
    slots = [B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched, precision="mixed"), b)
             for o, b in systems]
    B.run(max_steps=4)
    B.remove(slots[1])  # dropped mid-solve; the others are unaffected
    assert B.run() == 0
    alone = [_alone(o, b, "mixed") for o, b in systems]
    _check_same(B, [slots[0], slots[2]], [alone[0], alone[2]])


def test_batch_rejects_unsupported():
    A = P.grid(10, 10)
    B = I.Batch(1)
    with pytest.raises(I.IccgError, match="multicolor"):
        B.add(A, I.IC0(A, 0.0), np.ones(A.n))  # natural order: no sliced plan
    o = P.mc_order(A)
    B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), np.ones(A.n))
    with pytest.raises(I.IccgError, match="full"):
        B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), np.ones(A.n))


@pytest.mark.parametrize("kind,precision", [("mc", "mixed"), ("abmc", "mixed"), ("abmc", "fp64")])
def test_batch_shared_matrix_groups(kind, precision):
    """Right-hand sides of one matrix run as groups (up to 8 per group, here
    11 right-hand sides = 8 + 3) that read the matrix once; every one gives
    its solo iterates and the oracle's (FP64: the double-double CG; mixed:
    orc_cg_mixed) bit for bit.  A second system with 2 right-hand sides
    joins mid-run."""
    dims = (40, 36, 1) if kind == "mc" else (12, 12, 12)
    o, _ = _system(*dims, 700, kind=kind)
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    rhs = [I.gen_normal(o.A.n, 800 + k) for k in range(11)]
    o2, _ = _system(30, 30, 1, 701, kind="abmc")
    M2 = I.IC0(o2.A, 0.0, sched=o2.sched, precision=precision)
    rhs2 = [I.gen_normal(o2.A.n, 900 + k) for k in range(2)]
    B = I.Batch(16)
    B.set_max_group(8)
    slots = [B.add(o.A, M, b) for b in rhs]
    B.run(max_steps=6)
    slots2 = [B.add(o2.A, M2, b) for b in rhs2]
    while B.run(max_steps=9):
        pass
    _check_same(B, slots, [_alone(o, b, precision) for b in rhs])
    _check_same(B, slots2, [_alone(o2, b, precision) for b in rhs2])
    # and against the oracle: the double-double FP64 CG, or the mixed CG
    # with its restarts
    for sl, (oo, bs) in ((slots, (o, rhs)), (slots2, (o2, rhs2))):
        Ao = P.orc(oo.A)
        f = O.ic0_factor(Ao)
        for s, b in zip(sl, bs):
            fin, rb = B.result(s)
            if precision == "fp64":
                d = O.cg(Ao, b, "ic0", ic=f, tol=1e-8, dot_mode=1)
                assert (rb.iters, rb.rel_resid) == (d.iters, d.rel_resid)
                assert np.array_equal(B.x(s), d.x)
            else:
                m = O.cg_mixed(Ao, b, f, tol=1e-8)
                assert (rb.iters, rb.restarts, rb.rel_resid) == (m.iters, m.restarts,
                                                                 m.rel_resid)
                assert np.array_equal(B.x(s), m.x)


def test_batch_timing_mode_counts_and_keeps_results():
    systems = [_system(n, n, 1, 950 + n) for n in (48, 64)]
    alone = [_alone(o, b, "mixed") for o, b in systems]
    B = I.Batch(4)
    B.set_timing(True)
    slots = [B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched, precision="mixed"), b)
             for o, b in systems]
    assert B.run() == 0
    _check_same(B, slots, alone)
    iters = max(B.result(s)[1].iters for s in slots)
    n_spmv, ms_spmv = B.timing(0)
    assert n_spmv >= iters and ms_spmv > 0
    assert all(B.timing(c)[0] > 0 for c in range(4))
    B.reset_timing()
    assert B.timing(0) == (0, 0.0)


def test_batch_retired_system_may_be_freed_while_others_run():
    """A finished and removed system's matrix and preconditioner can be
    destroyed at once while the others keep running: retired items left in
    the running list must not touch them.  The retired system is small, so
    the running list is not rebuilt when it leaves."""
    big = [_system(n, n, 1, 1000 + n) for n in (96, 100, 104)]
    small_o, small_b = _system(12, 12, 1, 1099)
    Ms = [I.IC0(o.A, 0.0, sched=o.sched, precision="mixed") for o, _ in big]
    dA_small = I.DeviceCSR(small_o.A)
    M_small = I.IC0(dA_small, 0.0, sched=small_o.sched, precision="mixed")
    B = I.Batch(4)
    slots = [B.add(o.A, M, b) for (o, b), M in zip(big, Ms)]
    s_small = B.add(dA_small, M_small, small_b)
    while not B.result(s_small)[0]:
        B.run(max_steps=2)
    B.remove(s_small)
    del M_small, dA_small  # destroyed now (the batch keeps no reference)
    assert B.run() == 0
    _check_same(B, slots, [_alone(o, b, "mixed") for o, b in big])
