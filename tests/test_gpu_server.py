"""Request loop (iccg_server_*): every right-hand side gives the same x, bit
for bit, alone and in the mix, batched and one at a time, and equal to the
single-system solver on the same ABMC-ordered system."""
import time

import numpy as np
import pytest

import iccg_amd as I
from tests import problems as P

pytestmark = pytest.mark.gpu


def _request(dims, seed, nrhs):
    A = I.gen_grid(*dims, coef_exp=(-2.0, 2.0), seed=seed)
    b = np.stack([I.gen_normal(A.n, seed * 16 + k) for k in range(nrhs)])
    return A, b


def _direct(A, b, precision, dims=None, block=64):
    """The request through the single-system API, ordered as the server
    orders it (box blocks for a grid, breadth-first blocks otherwise)."""
    if dims is not None:
        box = (4, 4, 4) if dims[2] > 1 else (8, 8, 1)
        o = P.abmc_order(A, grid_dims=dims, box=box)
    else:
        o = P.abmc_order(A, block=block)
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    xs, rs = [], []
    for bk in b:
        bp = np.empty_like(bk)
        bp[o.new_of_old] = bk
        x = np.zeros(A.n)
        rs.append(I.conjugate_gradient(o.A, bp, x, M, 10000, 1e-8, o.sched))
        xs.append(x[o.new_of_old])
    return np.array(xs), rs


REQS = [((24, 20, 1), 1, 2), ((8, 8, 8), 2, 1), ((40, 40, 1), 3, 3), ((12, 10, 14), 4, 2),
        ((64, 16, 1), 5, 1), ((16, 16, 16), 6, 2)]


def _settled(S, timeout_s=60.0):
    """Stats once the reaper thread has freed every finished request's
    device memory and returned its HBM budget (it runs after completion)."""
    t0 = time.time()
    while True:
        st = S.stats()
        if st["hbm_in_use"] == 0 or time.time() - t0 > timeout_s:
            return st
        time.sleep(0.01)


def _collect(S, ids):
    out = {}
    for rid in ids:
        assert S.wait(rid, timeout_ms=60000) == rid
        out[rid] = S.result(rid)
    return out


@pytest.mark.parametrize("multi_rows", [0, 500])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_server_batched_equals_direct(precision, multi_rows):
    """multi_rows=500: the requests of 512 to 4,096 rows run on the runner
    threads beside the batch (those of two or three right-hand sides on the
    multi-right-hand-side solver, the single ones on the single-system
    solver), the 480-row one in the batch; every result as the
    single-system solver's."""
    reqs = [_request(d, s, k) for d, s, k in REQS]
    S = I.Server("batched", precision=precision, max_slots=8, workers=4,
                 multi_rows=multi_rows)
    ids = [S.submit(A, b, grid=d) for (A, b), (d, _, _) in zip(reqs, REQS)]
    got = _collect(S, ids)
    for rid, (A, b), (d, _, _) in zip(ids, reqs, REQS):
        x, res, lat, prep = got[rid]
        xd, rd = _direct(A, b, precision, dims=d)
        assert np.array_equal(x, xd)
        for r, q in zip(res, rd):
            assert r.converged and (r.iters, r.rel_resid, r.restarts) == (q.iters, q.rel_resid,
                                                                         q.restarts)
        assert lat >= prep >= 0.0
    st = _settled(S)
    assert st["finished"] == len(REQS) and st["failed"] == 0 and st["hbm_in_use"] == 0


@pytest.mark.parametrize("multi_rows", [0, 1000])
def test_server_alone_equals_in_the_mix(multi_rows):
    """One request served alone, then again while others arrive before,
    during and after it: identical results (multi_rows=1000: it and the
    other requests of 1,000 rows or more on the runner threads)."""
    A0, b0 = _request((48, 40, 1), 11, 3)
    S1 = I.Server("batched", max_slots=8, workers=2, multi_rows=multi_rows)
    r0 = S1.submit(A0, b0, grid=(48, 40, 1))
    x_alone, res_alone, _, _ = _collect(S1, [r0])[r0]
    del S1
    others = [_request(d, 20 + s, k) for d, s, k in REQS]
    S = I.Server("batched", max_slots=6, workers=3, multi_rows=multi_rows)
    ids = [S.submit(A, b, grid=d) for (A, b), (d, _, _) in zip(others[:3], REQS[:3])]
    time.sleep(0.05)
    rid = S.submit(A0, b0, grid=(48, 40, 1))
    ids += [S.submit(A, b, grid=d) for (A, b), (d, _, _) in zip(others[3:], REQS[3:])]
    got = _collect(S, ids + [rid])
    x_mix, res_mix, _, _ = got[rid]
    assert np.array_equal(x_mix, x_alone)
    assert [(r.iters, r.rel_resid) for r in res_mix] == [(r.iters, r.rel_resid)
                                                         for r in res_alone]


def test_server_sequential_equals_batched():
    reqs = [_request(d, 40 + s, k) for d, s, k in REQS[:4]]
    outs = []
    for mode in ("sequential", "batched"):
        S = I.Server(mode, precision="fp64", max_slots=4)
        ids = [S.submit(A, b, grid=d) for (A, b), (d, _, _) in zip(reqs, REQS)]
        got = _collect(S, ids)
        outs.append([got[i][0] for i in ids])
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


def test_server_hbm_budget_serializes_without_changing_results():
    reqs = [_request(d, 60 + s, k) for d, s, k in REQS[:4]]
    S = I.Server("batched", max_slots=8, hbm_budget=1)  # every request alone
    ids = [S.submit(A, b, grid=d) for (A, b), (d, _, _) in zip(reqs, REQS)]
    st = S.stats()
    assert st["waiting"] + st["preparing"] + st["ready"] + st["running"] + st["finished"] == 4
    got = _collect(S, ids)
    for rid, (A, b), (d, _, _) in zip(ids, reqs, REQS):
        xd, _ = _direct(A, b, "mixed", dims=d)
        assert np.array_equal(got[rid][0], xd)


def test_server_general_matrix_and_failures():
    A, b = _request((30, 30, 1), 80, 2)  # served without its grid shape: BFS blocks
    bad = I.gen_grid(10, 10, 1)
    bad = I.CRS(bad.n, bad.rowptr, bad.colind, -bad.val)  # negative definite: IC(0) fails
    S = I.Server("batched", max_slots=4)
    r_bad = S.submit(bad, np.ones((1, bad.n)), grid=(10, 10, 1))
    r_ok = S.submit(A, b)
    got = _collect(S, [r_ok])
    assert S.wait(r_bad, timeout_ms=60000) == r_bad
    with pytest.raises(I.IccgError):
        S.result(r_bad)
    xd, _ = _direct(A, b, "mixed", dims=None, block=64)
    assert np.array_equal(got[r_ok][0], xd)
    assert S.stats()["failed"] == 1
    with pytest.raises(I.IccgError, match="grid"):
        S.submit(A, b, grid=(31, 30, 1))


def test_server_wide_rows_fail_alone():
    """A request whose rows exceed the batched kernels' width fails with its
    own error; a request served beside it is unaffected."""
    n, bw = 300, 10
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(max(0, i - bw), min(n, i + bw + 1)):
            rows.append(i); cols.append(j); vals.append(4.0 * bw if i == j else -1.0)
    wide = I.coo_to_crs(n, np.array(rows), np.array(cols), np.array(vals))
    A, b = _request((24, 24, 1), 90, 1)
    S = I.Server("batched", max_slots=4)
    r_wide = S.submit(wide, np.ones((1, n)))
    r_ok = S.submit(A, b, grid=(24, 24, 1))
    got = _collect(S, [r_ok])
    assert S.wait(r_wide, timeout_ms=60000) == r_wide
    with pytest.raises(I.IccgError, match="16 entries"):
        S.result(r_wide)
    xd, _ = _direct(A, b, "mixed", dims=(24, 24, 1))
    assert np.array_equal(got[r_ok][0], xd)
