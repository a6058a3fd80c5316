"""Multi-system solves (iccg_batch_*): each system gives in a batch the
iterates it gives alone, including systems admitted while others run."""
import numpy as np
import pytest

import iccg_amd as I
from oracle import oracle as O
from tests import problems as P

pytestmark = pytest.mark.gpu


def _system(nx, ny, nz, seed, f32=True):
    A = I.gen_grid(nx, ny, nz, coef_exp=(-1.0, 1.0), seed=seed)
    if f32:  # exactly representable values (the batched config's FP32 CRS)
        A = I.CRS(A.n, A.rowptr, A.colind, A.val.astype(np.float32).astype(np.float64))
    o = P.mc_order(A)
    b = I.gen_normal(A.n, seed + 7)
    return o, b


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_batch_equals_alone(precision):
    systems = [_system(n, m, 1, 100 + k) for k, (n, m) in
               enumerate([(32, 32), (48, 40), (17, 90), (64, 64), (20, 20), (33, 47)])]
    alone = []
    for o, b in systems:
        M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
        x = np.zeros(o.A.n)
        r = I.conjugate_gradient(o.A, b, x, M, 10000, 1e-8, o.sched)
        alone.append((r, x))
    B = I.Batch(16)
    slots = []
    for o, b in systems:
        M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
        slots.append(B.add(o.A, M, b, max_iter=10000, tol=1e-8))
    assert B.run() == 0
    for s, (r, x), (o, b) in zip(slots, alone, systems):
        fin, rb = B.result(s)
        assert fin and rb.converged
        assert (rb.iters, rb.rel_resid) == (r.iters, r.rel_resid)
        xb = B.x(s)
        assert np.array_equal(xb, x)
        assert O.true_resid(P.orc(o.A), xb, b) < (1e-8 if precision == "mixed" else 1e-7)


def test_batch_fp64_matches_oracle_and_continuous_admission():
    systems = [_system(n, n, 1, 200 + n, f32=False) for n in (24, 40, 56, 30)]
    B = I.Batch(4)
    first = []
    for o, b in systems[:2]:
        M = I.IC0(o.A, 0.0, sched=o.sched)
        first.append(B.add(o.A, M, b))
    B.run(max_steps=5)  # the first two are mid-solve when the others arrive
    later = []
    for o, b in systems[2:]:
        M = I.IC0(o.A, 0.0, sched=o.sched)
        later.append(B.add(o.A, M, b))
    while B.run(max_steps=7):
        pass
    for s, (o, b) in zip(first + later, systems):
        fin, r = B.result(s)
        Ao = P.orc(o.A)
        dd = O.cg(Ao, b, "ic0", ic=O.ic0_factor(Ao), sched=P.orc_sched(o.sched), tol=1e-8,
                  dot_mode=1)
        assert fin and (r.iters, r.rel_resid) == (dd.iters, dd.rel_resid)
        assert np.array_equal(B.x(s), dd.x)
    # slots are reusable after removal
    B.remove(first[0])
    o, b = systems[0]
    s = B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), b)
    assert B.run() == 0 and B.result(s)[1].converged


def test_batch_rejects_unsupported():
    A = P.grid(10, 10)
    B = I.Batch(2)
    with pytest.raises(I.IccgError, match="multicolor"):
        B.add(A, I.IC0(A, 0.0), np.ones(A.n))  # natural order: no sliced MC plan
    o = P.mc_order(A)
    B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched, precision="mixed"), np.ones(A.n))
    with pytest.raises(I.IccgError, match="precision"):
        B.add(o.A, I.IC0(o.A, 0.0, sched=o.sched), np.ones(A.n))
