"""The batched and served configurations at their stated workloads:

- the 512-system batch (2D 5-point variable-coefficient systems, grid side
  uniform in [32, 256], FP32 values, mixed precision, tol 1e-8) in one
  Batch: every system converges with the oracle's true residual below tol,
  equal to the solver's own within 2 ulps; every x equals the system solved
  alone bit for bit; eight systems across the size range equal the oracle's
  mixed-precision CG bit for bit, and FP64 copies of them, run in the same
  batch, equal the oracle's double-double FP64 CG bit for bit;
- requests of the serving stream (2D and 3D variable-coefficient
  Laplacians, coefficients 10^U[-2,2], 1-8 right-hand sides), among them a
  1.75M-row 2D and a 1.8M-row 3D request, served batched alone and in a
  mix: the same x bit for bit, every right-hand side meeting tol on the
  oracle's true residual, and close to the plain FP64 one-at-a-time path.
"""
import time

import numpy as np
import pytest

import bench
import iccg_amd as I
from oracle import oracle as O
from tests import problems as P
from tests.test_gpu_server import _settled

pytestmark = pytest.mark.gpu


def _ulps2(got, want):
    return abs(got - want) <= 4.5e-16 * want


def test_batch_of_512_2d_systems_sides_32_to_256():
    sides, systems = bench.batch_problem(512)
    assert len(systems) == 512 and 1024 <= min(A.n for A, _, _ in systems)
    assert max(A.n for A, _, _ in systems) <= 65536
    dA = [I.DeviceCSR(A) for A, _, _ in systems]
    M32 = [I.IC0(d, 0.0, sched=s, precision="mixed") for d, (_, s, _) in zip(dA, systems)]
    order = np.argsort([A.n for A, _, _ in systems], kind="stable")
    pick = [int(k) for k in order[np.linspace(0, 511, 8).astype(int)]]
    M64 = {k: I.IC0(dA[k], 0.0, sched=systems[k][1]) for k in pick}
    B = I.Batch(512 + len(pick))
    slots = [B.add(dA[k], M32[k], systems[k][2]) for k in range(512)]
    slots64 = {k: B.add(dA[k], M64[k], systems[k][2]) for k in pick}
    while B.run():
        pass
    for k in range(512):
        A, s, b = systems[k]
        fin, r = B.result(slots[k])
        assert fin and r.converged and r.true_rel_resid < 1e-8
        xb = B.x(slots[k])
        assert _ulps2(r.true_rel_resid, O.true_resid_dd(P.orc(A), xb, b))
        # alone, through the single-system solver: bit for bit
        x = np.zeros(A.n)
        ra = I.conjugate_gradient(dA[k], b, x, M32[k], 10000, 1e-8, s)
        assert (ra.iters, ra.rel_resid, ra.restarts) == (r.iters, r.rel_resid, r.restarts)
        assert np.array_equal(x, xb)
    for k in pick:
        A, s, b = systems[k]
        Ao = P.orc(A)
        f = O.ic0_factor(Ao)
        m = O.cg_mixed(Ao, b, f, tol=1e-8)
        fin, r = B.result(slots[k])
        assert (r.iters, r.restarts, r.rel_resid) == (m.iters, m.restarts, m.rel_resid)
        assert np.array_equal(B.x(slots[k]), m.x)
        d = O.cg(Ao, b, "ic0", ic=f, tol=1e-8, dot_mode=1)
        fin, r = B.result(slots64[k])
        assert fin and (r.iters, r.rel_resid) == (d.iters, d.rel_resid)
        assert np.array_equal(B.x(slots64[k]), d.x)


# the first requests of the serving stream: 0 (2D 276^2, 8 right-hand
# sides), 3 (2D 1321^2 = 1,745,041 rows, 2), 4 (3D 122^3 = 1,815,848 rows,
# 6), 6 (3D 22^3 = 10,648 rows, 2)
PICK = (0, 3, 4, 6)


def _collect(S, ids):
    out = {}
    for rid in ids:
        assert S.wait(rid, timeout_ms=100000) == rid
        out[rid] = S.result(rid)
        S.release(rid)
    return out


def test_serving_stream_requests_up_to_1_8m_rows():
    stream = bench.serve_requests(max(PICK) + 1, bench.SEED)
    reqs = [stream[k] for k in PICK]
    assert [A.n for A, _, _ in reqs] == [76176, 1745041, 1815848, 10648]
    assert [b.shape[0] for _, b, _ in reqs] == [8, 2, 6, 2]
    # alone: each request on an idle server
    alone = []
    for A, b, dims in reqs:
        S = I.Server("batched", precision="mixed", max_slots=16, workers=4)
        rid = S.submit(A, b, grid=dims)
        alone.append(_collect(S, [rid])[rid])
        S.close()
    # in the mix: all of them on one server, arriving while others run
    S = I.Server("batched", precision="mixed", max_slots=16, workers=4)
    ids = []
    for A, b, dims in reqs:
        ids.append(S.submit(A, b, grid=dims))
        time.sleep(0.02)
    mix = _collect(S, ids)
    st = _settled(S)
    assert st["failed"] == 0 and st["hbm_in_use"] == 0
    S.close()
    for (A, b, dims), (xa, ra, _, _), rid in zip(reqs, alone, ids):
        xm, rm, lat, prep = mix[rid]
        assert np.array_equal(xm, xa)
        assert [(r.iters, r.rel_resid, r.restarts) for r in rm] == \
            [(r.iters, r.rel_resid, r.restarts) for r in ra]
        # the server's ordering (4x4x4 / 8x8 boxes), for the oracle's true
        # residual on the system it solved and the FP64 path on the same one
        box = (4, 4, 4) if dims[2] > 1 else (8, 8, 1)
        o = P.abmc_order(A, grid_dims=dims, box=box)
        Ao = P.orc(o.A)
        M64 = I.IC0(o.A, 0.0, sched=o.sched)
        for k, r in enumerate(rm):
            assert r.converged and r.true_rel_resid < 1e-8
            bp = np.empty_like(b[k])
            bp[o.new_of_old] = b[k]
            xp = np.empty_like(xm[k])
            xp[o.new_of_old] = xm[k]
            assert _ulps2(r.true_rel_resid, O.true_resid_dd(Ao, xp, bp))
            x64 = np.zeros(A.n)
            r64 = I.conjugate_gradient(o.A, bp, x64, M64, 10000, 1e-8, o.sched)
            assert r64.converged
            e = np.linalg.norm(xp - x64) / np.linalg.norm(x64)
            print(f"[request n={A.n} rhs {k}] mixed iters {r.iters} true "
                  f"{r.true_rel_resid:.3e}; fp64 iters {r64.iters}; ||x-x64||/||x64|| = {e:.3e}")
            assert abs(r.iters - r64.iters) <= max(3, r64.iters // 50) and e < 1e-4
