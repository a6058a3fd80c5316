"""GPU parity of the HIP solve path against the CPU oracle (same inputs).

Bar: bit-exact for SpMV, the Jacobi scale, the IC(0) factor and the
triangular solves (every row sum is one lane's ascending fma chain, as the
reference's loops compile); CG matches the oracle run with double-double
dots bit for bit, and the oracle with the reference's 80-bit dots to within
one iteration, with the final residual within 1e-12 when the iteration
counts agree.
"""
import numpy as np
import pytest

import iccg_amd as I
from oracle import oracle as O
from tests import problems as P

pytestmark = pytest.mark.gpu


def _cases():
    return {
        "saad": I.read_matrix_market_crs(P.SAAD),
        "lap2d_30": P.grid(30, 30),
        "lap3d_10": P.grid(10, 10, 10),
        "var3d_9": P.grid(9, 8, 7, coef=(-2.0, 2.0), shift=1e-6),
        "var2d_17": P.grid(17, 13, coef=(-1.0, 1.0)),
    }


@pytest.mark.parametrize("name", ["saad", "lap2d_30", "lap3d_10", "var3d_9"])
def test_spmv_bitwise(name):
    A = _cases()[name]
    x = I.gen_normal(A.n, 11)
    y = I.DeviceCSR(A).spmv(x)
    assert np.array_equal(y, O.spmv(P.orc(A), x))


@pytest.mark.parametrize("nblk,per_row,most", [(12, 3, 12), (48, 8, 48), (100, 8, 100)])
def test_spmv_16bit_columns_and_fallback(nblk, per_row, most):
    """The SpMV reads 16-bit columns with a per-slice table of the 64-row
    column blocks its entries reach: 16 entries when every slice reaches at
    most 16 blocks, 64 when every slice reaches at most 64; beyond that the
    columns stay 32-bit.  Rows here reach per_row random blocks of the first
    nblk, so every slice reaches at most (and here exactly) `most` of them.
    Every case bitwise."""
    rng = np.random.default_rng(nblk)
    n = 64 * max(nblk, 48)
    rows, cols = [], []
    for i in range(n):
        blocks = rng.choice(nblk, size=per_row, replace=False)
        js = np.unique(blocks * 64 + rng.integers(0, 64, size=len(blocks)))
        rows += [i] * len(js)
        cols += list(js)
    rows, cols = np.array(rows), np.array(cols)
    reach = [len(np.unique(cols[rows // 64 == s] // 64)) for s in range(n // 64)]
    assert max(reach) == most
    v = rng.normal(size=len(rows))
    A = I.coo_to_crs(n, rows, cols, v)
    x = I.gen_normal(A.n, 12)
    assert np.array_equal(I.DeviceCSR(A).spmv(x), O.spmv(P.orc(A), x))


def test_spmv_empty_rows():
    # a row with no entries gives y_i = 0 (sum starts at 0.0)
    A = I.CRS(3, np.array([0, 1, 1, 3], np.int32), np.array([0, 0, 2], np.int32),
              np.array([2.0, 1.0, 3.0]))
    y = I.DeviceCSR(A).spmv(np.array([1.0, 5.0, 2.0]))
    assert np.array_equal(y, [2.0, 0.0, 7.0])


def test_spmv_empty_slices():
    """Whole 64-row slices without entries, the last one included: their
    rows give 0 and no entry of the matrix is read for them."""
    n = 130
    rp = np.zeros(n + 1, np.int32)
    rp[1:65] = np.arange(1, 65)
    rp[65:] = 64
    A = I.CRS(n, rp, np.arange(64, dtype=np.int32), np.arange(1.0, 65.0))
    x = I.gen_normal(n, 4)
    y = I.DeviceCSR(A).spmv(x)
    assert np.array_equal(y, O.spmv(P.orc(A), x))
    assert np.array_equal(y[64:], np.zeros(n - 64))


@pytest.mark.parametrize("name", ["saad", "var3d_9"])
def test_jacobi_bitwise(name):
    A = _cases()[name]
    r = I.gen_normal(A.n, 3)
    z = I.Jacobi(A).apply(r)
    assert np.array_equal(z, O.jacobi_inv_diag(P.orc(A)) * r)


def _factor_equal(M, f):
    Lrp, Lci, Lv, LTcp, LTri, LTv = M.factor()
    assert np.array_equal(Lrp, f.Lrp) and np.array_equal(Lci, f.Lci)
    assert np.array_equal(LTcp, f.LTcp) and np.array_equal(LTri, f.LTri)
    assert np.array_equal(Lv, f.Lv), np.max(np.abs(Lv - f.Lv))
    assert np.array_equal(LTv, f.LTv)


@pytest.mark.parametrize("name", ["saad", "lap2d_30", "var3d_9", "var2d_17"])
@pytest.mark.parametrize("order", ["natural", "mc", "abmc"])
def test_ic0_factor_and_apply_bitwise(name, order):
    A = _cases()[name]
    if order == "natural":
        Ar, s = A, None
    elif order == "mc":
        o = P.mc_order(A)
        Ar, s = o.A, o.sched
    else:
        o = P.abmc_order(A, block=8)
        Ar, s = o.A, o.sched
    f = O.ic0_factor(P.orc(Ar), 0.0)
    M = I.IC0(Ar, 0.0, sched=s)
    _factor_equal(M, f)
    r = I.gen_normal(Ar.n, 5)
    z = M.apply(r, s)
    assert np.array_equal(z, O.ic0_apply(f, r, P.orc_sched(s)))
    # the MC / ABMC result equals the sequential one bit for bit (valid coloring)
    assert np.array_equal(z, O.ic0_apply(f, r, None))


@pytest.mark.parametrize("B", [4, 16, 64])
def test_abmc_block_sizes(B):
    A = P.grid(12, 11, 10, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, block=B)
    f = O.ic0_factor(P.orc(o.A), 0.0)
    M = I.IC0(o.A, 0.0, sched=o.sched)
    r = I.gen_normal(A.n, 8)
    assert np.array_equal(M.apply(r, o.sched), O.ic0_apply(f, r, None))


@pytest.mark.parametrize("B", [128, 512, 2047])
@pytest.mark.parametrize("dims", [(24, 22, 20), (100, 100, 1)])
def test_abmc_large_blocks(B, dims):
    """Breadth-first blocks of more than 64 rows, the reference's protocol
    sizes (run_abmc.sh:15: B from 32 to 2047), one workgroup per block (two
    rows per thread past 1024): the factor, the FP64 and FP32 applies, and
    FP64 and mixed CG, each bitwise equal to the oracle."""
    A = P.grid(*dims, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, block=B)
    assert int(np.diff(o.sched.block_ptr).max()) > 64
    Ao = P.orc(o.A)
    f = O.ic0_factor(Ao, 0.0)
    M = I.IC0(o.A, 0.0, sched=o.sched)
    assert M.plan()[0] == "abmc-blocks"
    _factor_equal(M, f)
    r = I.gen_normal(A.n, 8)
    assert np.array_equal(M.apply(r, o.sched), O.ic0_apply(f, r, None))
    b = I.gen_normal(A.n, 9)
    x = np.zeros(A.n)
    res = I.conjugate_gradient(o.A, b, x, M, 10000, 1e-8, o.sched)
    dd = O.cg(Ao, b, "ic0", ic=f, sched=P.orc_sched(o.sched), tol=1e-8, dot_mode=1)
    assert res.converged and (res.iters, res.rel_resid) == (dd.iters, dd.rel_resid)
    assert np.array_equal(x, dd.x)
    _assert_true_resid(o.A, x, b, res)
    M32 = I.IC0(o.A, 0.0, sched=o.sched, precision="mixed")
    assert np.array_equal(M32.apply(r, o.sched), O.ic0_apply_f32(f, r))
    x32 = np.zeros(A.n)
    r32 = I.conjugate_gradient(o.A, b, x32, M32, 10000, 1e-8, o.sched)
    m = O.cg_mixed(Ao, b, f, tol=1e-8)
    assert (r32.iters, r32.restarts, r32.rel_resid) == (m.iters, m.restarts, m.rel_resid)
    assert np.array_equal(x32, m.x)


@pytest.mark.parametrize("dims", [(24, 22, 20), (100, 100, 1)])
def test_abmc_bfs_blocks_sectioned(dims):
    """Breadth-first blocks of at most 64 rows with single-row fragments
    among them: rows of later colors have up to six (3D) or four (2D)
    earlier-color neighbours, so the sectioned layout keeps three (3D) or
    two (2D) in-block and six outside slots per row.  The factor, the FP64
    and FP32 applies, and FP64 and mixed CG are each bitwise equal to the
    oracle."""
    A = P.grid(*dims, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, block=64)
    sizes = np.diff(o.sched.block_ptr)
    assert sizes.max() == 64 and sizes.min() == 1
    Ao = P.orc(o.A)
    f = O.ic0_factor(Ao, 0.0)
    M = I.IC0(o.A, 0.0, sched=o.sched)
    plan = ("abmc-sectioned-fused-r", 3 if dims[2] > 1 else 2)
    assert M.plan() == plan
    _factor_equal(M, f)
    r = I.gen_normal(A.n, 8)
    assert np.array_equal(M.apply(r, o.sched), O.ic0_apply(f, r, None))
    b = I.gen_normal(A.n, 9)
    x = np.zeros(A.n)
    res = I.conjugate_gradient(o.A, b, x, M, 10000, 1e-8, o.sched)
    dd = O.cg(Ao, b, "ic0", ic=f, sched=P.orc_sched(o.sched), tol=1e-8, dot_mode=1)
    assert res.converged and (res.iters, res.rel_resid) == (dd.iters, dd.rel_resid)
    assert np.array_equal(x, dd.x)
    _assert_true_resid(o.A, x, b, res)
    M32 = I.IC0(o.A, 0.0, sched=o.sched, precision="mixed")
    assert M32.plan() == plan
    assert np.array_equal(M32.apply(r, o.sched), O.ic0_apply_f32(f, r))
    x32 = np.zeros(A.n)
    r32 = I.conjugate_gradient(o.A, b, x32, M32, 10000, 1e-8, o.sched)
    m = O.cg_mixed(Ao, b, f, tol=1e-8)
    assert (r32.iters, r32.restarts, r32.rel_resid) == (m.iters, m.restarts, m.rel_resid)
    assert np.array_equal(x32, m.x)


def test_abmc_grid_boxes():
    A = P.grid(16, 12, 8, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(16, 12, 8), box=(4, 4, 4))
    assert o.extra["nc"] == 2
    f = O.ic0_factor(P.orc(o.A), 0.0)
    M = I.IC0(o.A, 0.0, sched=o.sched)
    assert M.plan() == ("abmc-sectioned-fused-r", 3)
    _factor_equal(M, f)
    r = I.gen_normal(A.n, 9)
    assert np.array_equal(M.apply(r, o.sched), O.ic0_apply(f, r, None))


@pytest.mark.parametrize("name", ["lap2d_30", "var3d_9", "var2d_17"])
@pytest.mark.parametrize("order", ["natural", "mc", "abmc", "boxes"])
def test_ic0_apply_mixed_bitwise(name, order):
    """FP32 factor and sweeps: the GPU apply equals the oracle's FP32 apply
    bit for bit in every order (boxes: the sectioned ABMC layout)."""
    A = _cases()[name]
    if order == "natural":
        Ar, s = A, None
    elif order == "mc":
        o = P.mc_order(A)
        Ar, s = o.A, o.sched
    elif order == "abmc":
        o = P.abmc_order(A, block=8)
        Ar, s = o.A, o.sched
    else:
        dims = {"lap2d_30": (30, 30, 1), "var3d_9": (9, 8, 7), "var2d_17": (17, 13, 1)}[name]
        box = (4, 4, 4) if dims[2] > 1 else (8, 8, 1)
        o = P.abmc_order(A, grid_dims=dims, box=box)
        Ar, s = o.A, o.sched
    f = O.ic0_factor(P.orc(Ar), 0.0)
    M = I.IC0(Ar, 0.0, sched=s, precision="mixed")
    r = I.gen_normal(Ar.n, 6)
    assert np.array_equal(M.apply(r, s), O.ic0_apply_f32(f, r))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sectioned_sweeps_out_of_range_numerators(precision):
    """Right-hand sides spread over many decades, with +0 and -0 entries:
    numerators outside the range where the hoisted division reproduces
    a / d send their blocks through the exact-division pass; the result
    still equals the sequential oracle bit for bit."""
    A = P.grid(16, 12, 8, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(16, 12, 8), box=(4, 4, 4))
    f = O.ic0_factor(P.orc(o.A), 0.0)
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    rng = np.random.default_rng(11)
    span = 120 if precision == "fp64" else 18
    ref = O.ic0_apply if precision == "fp64" else O.ic0_apply_f32
    r = rng.standard_normal(o.A.n) * 10.0 ** rng.uniform(-span, span, o.A.n)
    r[rng.choice(o.A.n, 40, replace=False)] = 0.0
    r[rng.choice(o.A.n, 40, replace=False)] = -0.0
    z = M.apply(r, o.sched)
    assert np.array_equal(z, ref(f, r))
    assert np.array_equal(np.signbit(z), np.signbit(ref(f, r)))
    # one out-of-range row among ordinary ones
    r2 = rng.standard_normal(o.A.n)
    r2[77] = 1e-200 if precision == "fp64" else 1e-30
    assert np.array_equal(M.apply(r2, o.sched), ref(f, r2))
    # an all-zero right-hand side gives +0 everywhere
    z0 = M.apply(np.zeros(o.A.n), o.sched)
    assert np.array_equal(z0, ref(f, np.zeros(o.A.n))) and not np.signbit(z0).any()


def _slab_schedule(n, nblocks):
    """ABMC schedule of a natural-order matrix: nblocks contiguous slabs of
    rows, slab k in color k (every dependency lies in an earlier slab or in
    its own)."""
    bp = np.linspace(0, n, nblocks + 1).astype(np.int32)
    ar = np.arange(nblocks + 1, dtype=np.int32)
    return I.ColorSchedule.abmc(nblocks, nblocks, ar, ar[:-1].copy(), bp,
                                np.arange(n, dtype=np.int32))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_schedule_switch_lays_out_factor_again(precision):
    """The sliced factor of an ABMC schedule is laid out by its blocks
    (sectioned layout, or the plain one): switching the schedule of the same
    preconditioner lays it out again, and every apply stays bitwise equal to
    the oracle."""
    A = P.grid(4, 4, 8, coef=(-2.0, 2.0), shift=1e-6)
    f = O.ic0_factor(P.orc(A), 0.0)
    ref = O.ic0_apply if precision == "fp64" else O.ic0_apply_f32
    two, four = _slab_schedule(A.n, 2), _slab_schedule(A.n, 4)
    one = _slab_schedule(A.n, 1)  # one block of 128 rows: a workgroup per block
    M = I.IC0(A, 0.0, sched=two, precision=precision)
    r = I.gen_normal(A.n, 12)
    expect = ref(f, r)
    for sched, plan in [(two, ("abmc-sectioned-fused-r", 3)), (four, ("abmc-sliced", 0)),
                        (one, ("abmc-blocks", 0)), (two, ("abmc-sectioned-fused-r", 3)),
                        (one, ("abmc-blocks", 0)), (None, ("levels", 0)),
                        (four, ("abmc-sliced", 0))]:
        z = M.apply(r, sched)
        assert M.plan() == plan
        assert np.array_equal(z, expect)


def _solve(A, kind, sched=None, b=None, tol=1e-8, max_iter=10000, precision="fp64"):
    if b is None:
        b = np.ones(A.n)
    x = np.zeros(A.n)
    if kind == "none":
        M = I.Identity(A)
    elif kind == "jacobi":
        M = I.Jacobi(A)
    else:
        M = I.IC0(A, 0.0, sched=sched, precision=precision)
    res = I.conjugate_gradient(A, b, x, M, max_iter, tol, sched)
    return res, x


def _assert_true_resid(A, x, b, res):
    """The solver's true_rel_resid (k_resid + k_fin_resid) against the
    oracle's ||A x - b|| / ||b||: within 2 ulps of the double-double form
    (the GPU's SpMV is the oracle's bit for bit, and both sums are rounded
    once), and within the recursive-summation bound 2 n u of the reference's
    sequential form (cg_crs.cpp:517-528)."""
    Ao = P.orc(A)
    want = O.true_resid_dd(Ao, x, b)
    assert abs(res.true_rel_resid - want) <= 4.5e-16 * want
    seq = O.true_resid(Ao, x, b)
    assert abs(res.true_rel_resid - seq) <= 2 * A.n * 1.12e-16 * seq


def _oracle(A, kind, sched=None, b=None, tol=1e-8, max_iter=10000, dot_mode=0):
    if b is None:
        b = np.ones(A.n)
    Ao = P.orc(A)
    f = O.ic0_factor(Ao, 0.0) if kind == "ic0" else None
    return O.cg(Ao, b, kind, ic=f, sched=P.orc_sched(sched), tol=tol, max_iter=max_iter,
                dot_mode=dot_mode)


@pytest.mark.parametrize("kind", ["none", "jacobi", "ic0"])
@pytest.mark.parametrize("name", ["saad", "lap2d_30", "var3d_9"])
def test_cg_matches_oracle(kind, name):
    A = _cases()[name]
    res, x = _solve(A, kind)
    dd = _oracle(A, kind, dot_mode=1)
    ld = _oracle(A, kind, dot_mode=0)
    # same arithmetic as the double-double oracle: identical run
    assert res.iters == dd.iters and res.converged == dd.converged
    assert res.rel_resid == dd.rel_resid
    assert np.array_equal(x, dd.x)
    _assert_true_resid(A, x, np.ones(A.n), res)
    # against the reference's 80-bit dots: +-1 iteration
    assert abs(res.iters - ld.iters) <= 1
    if res.iters == ld.iters:
        assert abs(res.rel_resid - ld.rel_resid) <= 1e-12
    err = np.linalg.norm(x - ld.x) / np.linalg.norm(ld.x)
    assert err < 1e-6


def test_cg_100x100_laplacian_b_is_a_times_ones():
    """2D 100x100 Laplacian (the reference's CPU-sized case), b = A*1, x0 = 0, tol 1e-8."""
    A = P.grid(100, 100)
    b = P.rhs_ones_times_A(A)
    for kind in ("none", "ic0"):
        res, x = _solve(A, kind, b=b)
        ld = _oracle(A, kind, b=b)
        assert res.converged and abs(res.iters - ld.iters) <= 1
        if res.iters == ld.iters:
            assert abs(res.rel_resid - ld.rel_resid) <= 1e-12
        assert np.linalg.norm(x - ld.x) / np.linalg.norm(ld.x) < 1e-6


@pytest.mark.parametrize("order", ["mc", "abmc8", "abmcbox"])
def test_cg_colored_matches_oracle(order):
    A = P.grid(20, 18, 16, coef=(-2.0, 2.0), shift=1e-6)
    if order == "mc":
        o = P.mc_order(A)
    elif order == "abmc8":
        o = P.abmc_order(A, block=8)
    else:
        o = P.abmc_order(A, grid_dims=(20, 18, 16), box=(4, 4, 4))
    b = I.gen_normal(A.n, P.SEED + 1)
    if order != "abmc8":  # the multicolor and box plans take the r update into the sweeps
        plan = I.IC0(o.A, 0.0, sched=o.sched).plan()[0]
        assert plan == ("mc-sliced-r" if order == "mc" else "abmc-sectioned-fused-r")
    res, x = _solve(o.A, "ic0", o.sched, b=b)
    dd = _oracle(o.A, "ic0", o.sched, b=b, dot_mode=1)
    ld = _oracle(o.A, "ic0", o.sched, b=b, dot_mode=0)
    assert res.iters == dd.iters and np.array_equal(x, dd.x)
    assert abs(res.iters - ld.iters) <= 1
    assert res.converged
    _assert_true_resid(o.A, x, b, res)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_sectioned_three_colors(precision):
    """Three colors on the sectioned plan (three 64-row slabs, one color
    each), so the sweeps' r update and r.z leave partials from several color
    launches: FP64 equals the oracle's double-double solve bitwise; mixed
    equals the batch engine's solve of the same system (which forms r and
    r.z in passes of their own) bitwise and meets the tolerance on the true
    residual."""
    A = P.grid(4, 4, 12, coef=(-2.0, 2.0), shift=1e-6)
    sched = _slab_schedule(A.n, 3)
    M = I.IC0(A, 0.0, sched=sched, precision=precision)
    assert M.plan() == ("abmc-sectioned-fused-r", 3)
    b = I.gen_normal(A.n, 21)
    x = np.zeros(A.n)
    res = I.conjugate_gradient(A, b, x, M, 10000, 1e-8, sched)
    assert res.converged
    if precision == "fp64":
        dd = _oracle(A, "ic0", sched, b=b, dot_mode=1)
        assert res.iters == dd.iters and np.array_equal(x, dd.x)
        _assert_true_resid(A, x, b, res)
    else:
        B = I.Batch(2)
        slot = B.add(A, M, b)
        assert B.run() == 0
        fin, rb = B.result(slot)
        assert fin and rb.converged
        assert (rb.iters, rb.restarts) == (res.iters, res.restarts)
        assert np.array_equal(B.x(slot), x)
        _assert_true_resid(A, x, b, res)
        assert res.true_rel_resid < 1e-8


def test_mixed_precision_true_residual():
    A = P.grid(24, 20, 16, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(24, 20, 16))
    b = I.gen_normal(A.n, 4)
    res, x = _solve(o.A, "ic0", o.sched, b=b, precision="mixed")
    assert res.converged
    assert res.true_rel_resid < 1e-8
    _assert_true_resid(o.A, x, b, res)
    ld = _oracle(o.A, "ic0", o.sched, b=b)
    assert np.linalg.norm(x - ld.x) / np.linalg.norm(ld.x) < 1e-5


@pytest.mark.parametrize("order,dims,c,tol,restarts", [
    ("abmcbox", (24, 20, 16), 2.0, 1e-8, 0), ("abmcbox", (24, 20, 16), 2.0, 1e-14, 1),
    ("mc", (12, 12, 12), 3.0, 5e-15, 3), ("abmc8", (12, 12, 12), 3.0, 5e-15, 1),
    ("natural", (40, 40, 1), 2.0, 3e-14, 1), ("abmcbox", (40, 40, 1), 2.0, 3e-14, 1)])
def test_mixed_cg_matches_oracle(order, dims, c, tol, restarts):
    """Mixed precision (FP32 factor and sweeps, FP64 CG, stop on the FP64 true
    residual with restarts from the current x) against the oracle's
    orc_cg_mixed bit for bit: iterations, restarts, recurrence residual and
    x; the true residual within 2 ulps.  The tight tolerances make the
    recurrence stop while the true residual still misses tol (1 to 3
    restarts), on every plan: sectioned ABMC, plain ABMC, multicolor,
    natural-order levels."""
    A = P.grid(*dims, coef=(-c, c), shift=1e-6)
    if order == "mc":
        o = P.mc_order(A)
    elif order == "abmc8":
        o = P.abmc_order(A, block=8)
    elif order == "abmcbox":
        o = P.abmc_order(A, grid_dims=dims, box=(4, 4, 4) if dims[2] > 1 else (8, 8, 1))
    else:
        o = None
    As, sched = (o.A, o.sched) if o else (A, None)
    b = I.gen_normal(As.n, 4)
    res, x = _solve(As, "ic0", sched, b=b, tol=tol, max_iter=3000, precision="mixed")
    Ao = P.orc(As)
    m = O.cg_mixed(Ao, b, O.ic0_factor(Ao), tol=tol, max_iter=3000)
    assert m.restarts == restarts
    assert (res.iters, res.restarts, res.converged) == (m.iters, m.restarts, m.converged)
    assert res.rel_resid == m.rel_resid
    assert np.array_equal(x, m.x)
    _assert_true_resid(As, x, b, res)


def test_solver_resident_repeat():
    A = P.grid(32, 32)
    M = I.IC0(A, 0.0)
    S = I.Solver(A, M)
    b = I.gen_normal(A.n, 2)
    S.set(b)
    r1 = S.run(10000, 1e-8)
    x1 = S.x()
    S.set(b)
    r2 = S.run(10000, 1e-8)
    assert r1.iters == r2.iters and np.array_equal(x1, S.x())


def test_max_iter_and_zero_rhs():
    A = P.grid(20, 20)
    res, _ = _solve(A, "none", max_iter=5)
    ld = _oracle(A, "none", max_iter=5)
    assert res.iters == 5 == ld.iters and not res.converged
    assert abs(res.rel_resid - ld.rel_resid) <= 1e-12
    # b = 0: normb -> 1, r0 = 0 -> converged at iteration 0 (cg_crs.cpp:109-123)
    res, x = _solve(A, "ic0", b=np.zeros(A.n))
    assert res.iters == 0 and res.converged and not x.any()


def test_breakdown_indefinite():
    # indefinite diagonal matrix with Identity: p.Ap <= eps |p||Ap| -> breakdown
    n = 4
    A = I.CRS(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32),
              np.array([1.0, -1.0, 1.0, -1.0]))
    res, _ = _solve(A, "none")
    ld = _oracle(A, "none")
    assert res.iters == ld.iters and res.converged == ld.converged == False  # noqa: E712
    assert res.breakdown


def test_pivot_errors_match_reference_messages():
    A = P.grid(6, 6)
    v = A.val.copy()
    d = np.where(A.colind == np.repeat(np.arange(A.n), np.diff(A.rowptr)))[0]
    v[d[7]] = -1.0
    B = I.CRS(A.n, A.rowptr, A.colind, v)
    with pytest.raises(I.IccgError, match="IC\\(0\\): non-positive pivot at row 7"):
        I.IC0(B, 0.0)
    with pytest.raises(O.PivotError, match="non-positive pivot at row 7"):
        O.ic0_factor(P.orc(B), 0.0)


def test_invalid_schedule_rejected():
    A = P.grid(8, 8)
    o = P.mc_order(A)
    bad = np.zeros(A.n, np.int32)  # one color: every dependency is a race
    s = I.ColorSchedule.mc(1, np.array([0, A.n], np.int32), np.arange(A.n, dtype=np.int32))
    with pytest.raises(I.IccgError, match="not a valid MC schedule"):
        I.IC0(o.A, 0.0, sched=s)
    del bad


def test_graph_replay_matches_eager():
    """Iterations replayed from captured hipGraphs give the eager launches'
    result bit for bit, on every replay, with eager work in between, and after
    the preconditioner is re-planned for another schedule."""
    A = P.grid(24, 20, 16, coef=(-2.0, 2.0), shift=1e-6)
    o = P.abmc_order(A, grid_dims=(24, 20, 16))
    b = I.gen_normal(A.n, 12)
    M = I.IC0(o.A, 0.0, sched=o.sched)
    S = I.Solver(o.A, M)
    S.set(b)
    S.set_graphs(False)
    ref = S.run(10000, 1e-8, zero_x0=True)
    x_ref = S.x()
    S.set_graphs(True)
    for _ in range(3):
        r = S.run(10000, 1e-8, zero_x0=True)
        assert (r.iters, r.rel_resid) == (ref.iters, ref.rel_resid)
        assert np.array_equal(S.x(), x_ref)
        S.set_graphs(False)
        assert S.run(10000, 1e-8, zero_x0=True).iters == ref.iters  # eager in between
        S.set_graphs(True)
    # re-plan M for the natural order: same factor, same bits (valid schedules)
    M.use_schedule(None)
    r = S.run(10000, 1e-8, zero_x0=True)
    assert (r.iters, r.rel_resid) == (ref.iters, ref.rel_resid)
    assert np.array_equal(S.x(), x_ref)
    # max_iter not a multiple of the chunk size
    r = S.run(7, 1e-8, zero_x0=True)
    assert r.iters == 7 and not r.converged


def _random_lower(n, per_row, band, seed):
    """Lower-triangular CRS (diagonal last) of a symmetric matrix with
    per_row random neighbours j in [i - band, i) per row, entries
    -10^U[-1,1], diagonal 1 + the row's absolute sum over both triangles."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        lo = max(0, i - band)
        cols = np.unique(rng.integers(lo, i, size=min(per_row, i))) if i > 0 else np.zeros(0, int)
        rows.append(cols)
    off = [(-(10.0 ** rng.uniform(-1, 1, len(c)))) for c in rows]
    absrow = np.zeros(n)
    for i, (c, v) in enumerate(zip(rows, off)):
        absrow[i] += np.abs(v).sum()
        np.add.at(absrow, c, np.abs(v))
    rp = np.zeros(n + 1, np.int32)
    ci, va = [], []
    for i, (c, v) in enumerate(zip(rows, off)):
        ci.extend(c.tolist() + [i])
        va.extend(v.tolist() + [1.0 + absrow[i]])
        rp[i + 1] = len(ci)
    return I.CRS(n, rp, np.asarray(ci, np.int32), np.asarray(va, np.float64))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cg_multicolor_many_colors(precision):
    """A multicolor plan with several colors (greedy coloring of a random
    banded pattern), so the forward sweep's r update and the backward
    sweep's r.z leave partials from several color launches: FP64 equals the
    oracle's double-double solve bitwise; mixed equals the batch engine's
    solve (which forms r and r.z in passes of their own) bitwise."""
    o = P.mc_order(_random_lower(3000, 3, 48, 5))
    assert o.extra["nc"] >= 3
    M = I.IC0(o.A, 0.0, sched=o.sched, precision=precision)
    assert M.plan()[0] == "mc-sliced-r"
    b = I.gen_normal(o.A.n, 22)
    x = np.zeros(o.A.n)
    res = I.conjugate_gradient(o.A, b, x, M, 10000, 1e-8, o.sched)
    assert res.converged
    if precision == "fp64":
        dd = _oracle(o.A, "ic0", o.sched, b=b, dot_mode=1)
        assert res.iters == dd.iters and np.array_equal(x, dd.x)
    else:
        B = I.Batch(2)
        slot = B.add(o.A, M, b)
        assert B.run() == 0
        fin, rb = B.result(slot)
        assert fin and rb.converged
        assert (rb.iters, rb.restarts) == (res.iters, res.restarts)
        assert np.array_equal(B.x(slot), x)
        assert O.true_resid(P.orc(o.A), x, b) < 1e-8
