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
