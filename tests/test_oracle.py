"""Pins the CPU oracle (oracle/iccg_oracle.c) before the GPU path is
compared against it.

Sources of truth, strongest first:
  * anchors recorded from the reference's own binaries (iteration counts and
    printed residuals of cg/iccg/piccg/abmc on Saad.mm and the 100x100 grid);
  * tests/golden/*.npz written by the reference (make_golden.py);
  * oracle/_ref (the reference compiled from its sources) when it is built;
  * closed forms: scipy/numpy permutations, tril, triangular solves,
    Cholesky (IC(0) has no fill on tree patterns), exact rational dots,
    numpy.linalg.solve.
"""
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.io
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from oracle import ref as R
from tests import problems as P

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built")
GS = np.load(os.path.join(P.GOLDEN, "ref_saad.npz"))
GG = np.load(os.path.join(P.GOLDEN, "ref_grids.npz"))


def saad():
    return O.read_matrix_market(P.SAAD)


# ------------------------------------------------------------------ anchors
@pytest.mark.parametrize("kind,tol,iters,rel", [
    ("none", 1e-8, 10, 4.01652e-09), ("jacobi", 1e-8, 10, 4.01652e-09),
    ("ic0", 1e-8, 5, 1.87367e-10), ("none", 1e-14, 14, 3.68685e-15),
    ("ic0", 1e-14, 7, 4.40064e-27)])
def test_anchor_saad(kind, tol, iters, rel):
    A = saad()
    f = O.ic0_factor(A) if kind == "ic0" else None
    o = O.cg(A, np.ones(A.n), kind, ic=f, tol=tol)
    assert o.iters == iters and o.converged
    assert float(f"{o.rel_resid:.5e}") == rel


def test_anchor_grid100():
    """Reference binaries on the 2D 100x100 Laplacian with b = 1: cg 187
    (8.59665e-09), iccg 79 (8.09918e-09), red-black piccg 94 (8.14029e-09),
    8x8-tile 2-color ABMC 86 (8.74622e-09)."""
    A = P.grid(100, 100)
    Ao = P.orc(A)
    b = np.ones(A.n)
    o = O.cg(Ao, b, "none", tol=1e-8)
    assert (o.iters, f"{o.rel_resid:.5e}") == (187, "8.59665e-09")
    o = O.cg(Ao, b, "ic0", ic=O.ic0_factor(Ao), tol=1e-8)
    assert (o.iters, f"{o.rel_resid:.5e}") == (79, "8.09918e-09")
    col = np.array([(i % 100 + i // 100) % 2 for i in range(A.n)], np.int32)
    m = P.mc_order(A, col)
    Am = P.orc(m.A)
    o = O.cg(Am, b, "ic0", ic=O.ic0_factor(Am), sched=P.orc_sched(m.sched), tol=1e-8)
    assert (o.iters, f"{o.rel_resid:.5e}") == (94, "8.14029e-09")
    a = P.abmc_order(A, grid_dims=(100, 100, 1), box=(8, 8, 1))
    Aa = P.orc(a.A)
    o = O.cg(Aa, b, "ic0", ic=O.ic0_factor(Aa), sched=P.orc_sched(a.sched), tol=1e-8)
    assert (o.iters, f"{o.rel_resid:.5e}") == (86, "8.74622e-09")


# ------------------------------------------------------------------ golden fixtures
def test_golden_saad_reader_and_cg():
    A = saad()
    assert np.array_equal(A.rowptr, GS["rowptr"]) and np.array_equal(A.colind, GS["colind"])
    assert np.array_equal(A.val, GS["val"])
    for kind in ("none", "jacobi", "ic0"):
        for tol in (1e-8, 1e-14):
            f = O.ic0_factor(A) if kind == "ic0" else None
            o = O.cg(A, np.ones(A.n), kind, ic=f, tol=tol)
            k = f"{kind}_{tol:g}"
            assert o.iters == GS[k + "_iters"] and o.rel_resid == GS[k + "_rel"]
            assert np.array_equal(o.x, GS[k + "_x"])
    f = O.ic0_factor(A)
    assert np.array_equal(f.Lv, GS["L_val"]) and np.array_equal(f.LTv, GS["LT_val"])


@pytest.mark.parametrize("name", ["v2d", "v3d"])
@pytest.mark.parametrize("order", ["natural", "mc", "abmc"])
def test_golden_grids(name, order):
    k = f"{name}_{order}"
    A = O.Csr(len(GG[k + "_rowptr"]) - 1, GG[k + "_rowptr"], GG[k + "_colind"], GG[k + "_val"])
    s = None
    if order != "natural":
        s = O.Schedule(nc=int(GG[k + "_nc"]), mode=int(GG[k + "_mode"]), nb=int(GG[k + "_nb"]))
        for fld in ("color_ptr", "rows_of_color", "color_ptr_blk", "blocks_of_color",
                    "block_ptr", "rows_of_block"):
            if k + "_" + fld in GG:
                setattr(s, fld, np.ascontiguousarray(GG[k + "_" + fld], np.int32))
    f = O.ic0_factor(A)
    assert np.array_equal(f.Lv, GG[k + "_Lv"]) and np.array_equal(f.LTv, GG[k + "_LTv"])
    o = O.cg(A, GG[k + "_b"], "ic0", ic=f, sched=s, tol=1e-8)
    assert o.iters == GG[k + "_iters"] and o.rel_resid == GG[k + "_rel"]
    assert np.array_equal(o.x, GG[k + "_x"])


# ------------------------------------------------------------------ against oracle/_ref
def _grids():
    return [P.grid(17, 13, coef=(-1.0, 1.0)), P.grid(9, 8, 7, coef=(-2.0, 2.0), shift=1e-6),
            P.grid(30, 30)]


@needs_ref
def test_ref_reader_expand_ptap():
    A = R.read_matrix_market(P.SAAD)
    B = saad()
    assert all(np.array_equal(getattr(A, f), getattr(B, f)) for f in ("rowptr", "colind", "val"))
    for G in _grids():
        Go = P.orc(G)
        color, nc = P.I.greedy_color(G)
        p = O.perm_by_color(color)
        assert np.array_equal(p, R.perm_by_color(color))
        L1, L2 = O.ptap_lower(Go, p), R.ptap_lower(Go, p)
        assert all(np.array_equal(getattr(L1, f), getattr(L2, f)) for f in ("rowptr", "colind", "val"))
        F1, F2 = O.expand_lower_to_full(L1), R.expand_lower_to_full(L2)
        assert all(np.array_equal(getattr(F1, f), getattr(F2, f)) for f in ("rowptr", "colind", "val"))
        cp1, rc1 = O.color_ptr_rows(color, nc, p)
        cp2, rc2 = R.color_ptr_rows(color, nc, p)
        assert np.array_equal(cp1, cp2) and np.array_equal(rc1, rc2)
        bo, nb, bc, nc2 = P.I.abmc_blocking(G, 6)
        q = O.perm_by_block_color(bo, bc)
        assert np.array_equal(q, R.perm_by_block_color(bo, bc))
        s1 = O.abmc_schedule(q, bo, bc, nb, nc2)
        s2 = R.abmc_schedule(q, bo, bc, nb, nc2)
        assert all(np.array_equal(a, b) for a, b in zip(s1, s2))


@needs_ref
def test_ref_precond_and_cg_bitwise():
    for G in _grids() + [saad()]:
        Go = P.orc(G) if not isinstance(G, O.Csr) else G
        assert np.array_equal(O.jacobi_inv_diag(Go), R.jacobi_inv_diag(Go))
        x = P.I.gen_normal(Go.n, 3)
        assert np.array_equal(O.spmv(Go, x), R.spmv(Go, x))
        assert O.dot_ld(x, x[::-1].copy()) == R.dot(x, x[::-1].copy())
        f1, f2 = O.ic0_factor(Go), R.ic0_factor(Go)
        for fld in ("Lrp", "Lci", "Lv", "diag_pos", "LTcp", "LTri", "LTv", "LT_diag_pos"):
            assert np.array_equal(getattr(f1, fld), getattr(f2, fld)), fld
        assert np.array_equal(O.ic0_apply(f1, x), R.ic0_apply(Go.n, x))
        b = P.I.gen_normal(Go.n, 4)
        for kind in ("none", "jacobi", "ic0"):
            o = O.cg(Go, b, kind, ic=f1 if kind == "ic0" else None, tol=1e-10)
            it, rel, conv, xr = R.cg(Go, b, kind, tol=1e-10)
            assert (o.iters, o.rel_resid, o.converged) == (it, rel, conv)
            assert np.array_equal(o.x, xr)


@needs_ref
@pytest.mark.parametrize("order", ["mc", "abmc"])
def test_ref_colored_cg_bitwise(order):
    G = P.grid(12, 10, 9, coef=(-2.0, 2.0), shift=1e-6)
    o = P.mc_order(G) if order == "mc" else P.abmc_order(G, block=16)
    Ao = P.orc(o.A)
    s = P.orc_sched(o.sched)
    f = O.ic0_factor(Ao)
    R.ic0_factor(Ao)
    r = P.I.gen_normal(Ao.n, 5)
    assert np.array_equal(O.ic0_apply(f, r, s), R.ic0_apply(Ao.n, r, s))
    b = P.I.gen_normal(Ao.n, 6)
    oo = O.cg(Ao, b, "ic0", ic=f, sched=s, tol=1e-8)
    it, rel, conv, xr = R.cg(Ao, b, "ic0", sched=s, tol=1e-8)
    assert (oo.iters, oo.rel_resid) == (it, rel) and np.array_equal(oo.x, xr)


@needs_ref
def test_ref_validators_agree():
    G = P.grid(12, 12)
    Go = P.orc(G)
    bo, nb, bc, nc = P.I.abmc_blocking(G, 8)
    tot, per = O.check_interblock(Go, bo, bc, nb, nc)
    rep = R.check_interblock_report(Go, bo, bc, nb, nc)
    assert f"TOTAL conflicts={tot} " in rep and tot == 0
    bad = np.zeros(nb, np.int32)  # one color for every block
    tot, per = O.check_interblock(Go, bo, bad, nb, 1)
    rep = R.check_interblock_report(Go, bo, bad, nb, 1)
    assert f"TOTAL conflicts={tot} " in rep and tot > 0
    o = P.abmc_order(G, block=8)
    color = bc[bo]
    th, hz, sb, cb = O.check_hazards(P.orc(o.A), o.new_of_old, color, bo)
    rep = R.check_hazards_report(P.orc(o.A), o.new_of_old, color, bo, bc)
    assert f"TOTAL hazards={th}\n" in rep
    for c in range(len(hz)):
        assert f"color {c} : hazards={hz[c]} (same-block={sb[c]}, cross-block={cb[c]})" in rep
    assert cb.sum() == 0


# ------------------------------------------------------------------ closed forms
def test_reader_vs_scipy():
    A = saad()
    S = scipy.io.mmread(P.SAAD).tocsr()
    S.sort_indices()
    assert np.array_equal(A.rowptr, S.indptr) and np.array_equal(A.colind, S.indices)
    assert np.array_equal(A.val, S.data)


def test_coo_to_crs_sums_duplicates():
    rng = np.random.default_rng(0)
    I = rng.integers(0, 30, 400)
    J = rng.integers(0, 30, 400)
    V = rng.integers(-5, 6, 400).astype(float)  # integer values: any order sums exactly
    A = O.coo_to_crs(30, I, J, V)
    S = sp.coo_matrix((V, (I, J)), shape=(30, 30)).tocsr()
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(A.rowptr, S.indptr) and np.array_equal(A.colind, S.indices)
    assert np.array_equal(A.val, S.data)  # explicit zeros kept, like the reference


def test_permutation_and_ptap_closed_form():
    G = P.grid(11, 9, coef=(-1.0, 1.0))
    Go = P.orc(G)
    rng = np.random.default_rng(1)
    color = rng.integers(0, 4, G.n).astype(np.int32)
    p = O.perm_by_color(color)
    assert np.array_equal(np.argsort(p), np.lexsort((np.arange(G.n), color)))
    B = O.ptap_lower(Go, p)
    S = Go.to_scipy()
    Pm = sp.csr_matrix((np.ones(G.n), (p, np.arange(G.n))), shape=(G.n, G.n))
    T = sp.tril(Pm @ S @ Pm.T).tocsr()
    T.eliminate_zeros()
    T.sort_indices()
    assert np.array_equal(B.rowptr, T.indptr) and np.array_equal(B.colind, T.indices)
    assert np.allclose(B.val, T.data, rtol=0, atol=0)
    F = O.expand_lower_to_full(B)
    Ts = (T + sp.tril(T, -1).T).tocsr()
    Ts.sort_indices()
    assert np.array_equal(F.colind, Ts.indices) and np.array_equal(F.val, Ts.data)
    cp, rc = O.color_ptr_rows(color, 4, p)
    assert np.array_equal(rc, np.arange(G.n))  # new ids are contiguous per color
    assert np.array_equal(np.diff(cp), np.bincount(color, minlength=4))


def test_ptap_drops_zero_sums():
    # a stored exact zero off-diagonal disappears from P^T A P (color.cpp:123)
    A = O.Csr(3, np.array([0, 1, 3, 5], np.int32), np.array([0, 0, 1, 1, 2], np.int32),
              np.array([2.0, 0.0, 3.0, -1.0, 4.0]))
    B = O.ptap_lower(A, np.array([0, 1, 2], np.int32))
    assert list(B.colind) == [0, 1, 1, 2] and list(B.rowptr) == [0, 1, 2, 4]


def test_abmc_schedule_invariants():
    G = P.grid(13, 11, 5, coef=(-1.0, 1.0))
    bo, nb, bc, nc = P.I.abmc_blocking(G, 10)
    q = O.perm_by_block_color(bo, bc)
    order = np.lexsort((np.arange(G.n), bo, bc[bo]))
    assert np.array_equal(np.argsort(q), order)
    cpb, boc, bp, rob = O.abmc_schedule(q, bo, bc, nb, nc)
    assert np.array_equal(np.diff(cpb), np.bincount(bc, minlength=nc))
    for b in range(nb):  # each block is a contiguous ascending run of new ids
        rows = rob[bp[b]:bp[b + 1]]
        assert np.array_equal(rows, np.arange(rows[0], rows[0] + len(rows)))
        assert len(rows) == np.sum(bo == b)


def test_ic0_equals_cholesky_on_tree_patterns():
    # tridiagonal (1D Laplacian) and an arrow (star) pattern: no fill, so
    # IC(0) is the exact Cholesky factor
    n = 40
    A1 = P.grid(n, 1, 1, coef=(-1.0, 1.0))
    star = np.zeros((n, n))
    star[-1, :] = star[:, -1] = 0.1
    np.fill_diagonal(star, np.linspace(2, 5, n))
    S2 = sp.csr_matrix(star)
    S2.sort_indices()
    A2 = O.Csr(n, S2.indptr.astype(np.int32), S2.indices.astype(np.int32), S2.data)
    for A in (P.orc(A1), A2):
        f = O.ic0_factor(A)
        Lf = sp.csr_matrix((f.Lv, f.Lci, f.Lrp), shape=(A.n, A.n)).toarray()
        C = np.linalg.cholesky(A.to_scipy().toarray())
        assert np.allclose(Lf, C, rtol=1e-13, atol=1e-15)


def test_ic0_bruteforce_masked():
    # IC(0) by the textbook definition on a dense copy (drop fill outside
    # the pattern), compared on a small 2D variable-coefficient grid
    G = P.grid(6, 5, coef=(-1.0, 1.0))
    A = P.orc(G).to_scipy().toarray()
    n = A.shape[0]
    mask = np.tril(A != 0)
    L = np.zeros_like(A)
    for i in range(n):
        for j in range(i):
            if mask[i, j]:
                L[i, j] = (A[i, j] - L[i, :j] @ L[j, :j]) / L[j, j]
        L[i, i] = np.sqrt(A[i, i] - L[i, :i] @ L[i, :i])
    f = O.ic0_factor(P.orc(G))
    Lf = sp.csr_matrix((f.Lv, f.Lci, f.Lrp), shape=(n, n)).toarray()
    assert np.allclose(Lf, L, rtol=1e-13, atol=1e-15)


def test_ic0_apply_vs_triangular_solves():
    G = P.grid(9, 8, 7, coef=(-2.0, 2.0), shift=1e-6)
    f = O.ic0_factor(P.orc(G))
    L = sp.csr_matrix((f.Lv, f.Lci, f.Lrp), shape=(G.n, G.n))
    r = P.I.gen_normal(G.n, 9)
    y = spla.spsolve_triangular(L, r, lower=True)
    z = spla.spsolve_triangular(L.T.tocsr(), y, lower=False)
    assert np.allclose(O.ic0_apply(f, r), z, rtol=1e-10, atol=0)
    o = P.mc_order(G)
    fo = O.ic0_factor(P.orc(o.A))
    assert np.array_equal(O.ic0_apply(fo, r, P.orc_sched(o.sched)), O.ic0_apply(fo, r))


def _round_f32(x):
    """Nearest float32 to the rational x, ties to even (normal range)."""
    if x == 0:
        return Fraction(0)
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while a >= Fraction(2) ** (e + 1):
        e += 1
    while a < Fraction(2) ** e:
        e -= 1
    assert -126 <= e <= 127
    m = a / Fraction(2) ** (e - 23)  # in [2^23, 2^24)
    q, rem = divmod(m.numerator, m.denominator)
    half = Fraction(rem, m.denominator) - Fraction(1, 2)
    if half > 0 or (half == 0 and q % 2 == 1):
        q += 1
    return (1 if x > 0 else -1) * q * Fraction(2) ** (e - 23)


def _apply_f32_exact(f, r):
    """The FP32 sweeps in exact rationals, every operation rounded once to
    float32: fmaf(a, b, c) = round(a b + c), round(a / d)."""
    rd = _round_f32
    y = [Fraction(0)] * f.n
    for i in range(f.n):
        acc, di = rd(Fraction(float(r[i]))), Fraction(1)
        for k in range(f.Lrp[i], f.Lrp[i + 1]):
            j, v = int(f.Lci[k]), rd(Fraction(float(f.Lv[k])))
            if j < i:
                acc = rd(-v * y[j] + acc)
            elif j == i:
                di = v
        y[i] = rd(acc / di)
    z = [Fraction(0)] * f.n
    for i in reversed(range(f.n)):
        acc, di = y[i], Fraction(1)
        for k in range(f.LTcp[i], f.LTcp[i + 1]):
            j, v = int(f.LTri[k]), rd(Fraction(float(f.LTv[k])))
            if j > i:
                acc = rd(-v * z[j] + acc)
            elif j == i:
                di = v
        z[i] = rd(acc / di)
    return np.array([float(v) for v in z])


def test_ic0_apply_f32_exact_rational():
    """The mixed-precision apply equals float32 arithmetic done exactly and
    rounded once per operation; FP64 arithmetic does not."""
    G = P.grid(6, 5, 4, coef=(-2.0, 2.0), shift=1e-6)
    f = O.ic0_factor(P.orc(G))
    r = P.I.gen_normal(G.n, 4)
    z = O.ic0_apply_f32(f, r)
    assert np.array_equal(z, _apply_f32_exact(f, r))
    assert np.all(z.astype(np.float32).astype(np.float64) == z)
    assert not np.array_equal(z, O.ic0_apply(f, r))


def test_pivot_errors():
    A = O.Csr(2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
              np.array([1.0, 2.0, 2.0, 1.0]))
    with pytest.raises(O.PivotError, match="non-positive pivot at row 1"):
        O.ic0_factor(A)
    B = O.Csr(2, np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32),
              np.array([1.0, 1.0]))
    with pytest.raises(O.PivotError, match="diagonal missing at row 1"):
        O.ic0_factor(B)


def test_dots_exact_rational():
    rng = np.random.default_rng(3)
    for n in (1, 7, 1000):
        a = rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)
        b = rng.standard_normal(n)
        exact = sum(Fraction(x) * Fraction(y) for x, y in zip(a, b))
        assert O.dot_dd(a, b) == float(exact)  # correctly rounded
        assert abs(O.dot_ld(a, b) - float(exact)) <= 1e-12 * float(
            sum(abs(Fraction(x) * Fraction(y)) for x, y in zip(a, b)))


def test_spmv_and_cg_textbook():
    G = P.grid(7, 6, 5, coef=(-1.0, 1.0))
    Go = P.orc(G)
    x = P.I.gen_normal(G.n, 2)
    assert np.allclose(O.spmv(Go, x), Go.to_scipy() @ x, rtol=1e-14, atol=1e-14)
    b = P.I.gen_normal(G.n, 3)
    xs = np.linalg.solve(Go.to_scipy().toarray(), b)
    for kind in ("none", "jacobi", "ic0"):
        o = O.cg(Go, b, kind, ic=O.ic0_factor(Go) if kind == "ic0" else None, tol=1e-12)
        assert o.converged and np.allclose(o.x, xs, rtol=1e-9, atol=1e-11)
        assert O.true_resid(Go, o.x, b) < 1e-11


# ------------------------------------------------------------------ true residual
def _exact_true_resid(Ao, x, b):
    """sqrt(sum d_i^2) / sqrt(sum b_i^2) for d = fl(A x - b), both sums exact
    (rational), each rounded once, then the IEEE square roots and quotient."""
    import math
    d = O.spmv(Ao, x) - b
    nr = sum(Fraction(float(v)) ** 2 for v in d)
    nb = sum(Fraction(float(v)) ** 2 for v in b)
    return math.sqrt(float(nr)) / math.sqrt(float(nb))


@pytest.mark.parametrize("scale,noise", [(1e3, 1e-6), (1e-4, 1e-2), (1.0, 0.5)])
def test_true_resid_exact_rational(scale, noise):
    """||A x - b|| / ||b|| (cg_crs.cpp:517-528) against exact rational sums of
    the same d_i, with ||b|| far from 1 and the residual far from 0, so a
    missing division by ||b||, a missing square root or a wrong sign fails."""
    G = P.grid(6, 5, 4, coef=(-1.0, 1.0))
    Go = P.orc(G)
    b = scale * P.I.gen_normal(G.n, 21)
    xs = np.linalg.solve(Go.to_scipy().toarray(), b)
    x = xs * (1.0 + noise * P.I.gen_normal(G.n, 22))
    want = _exact_true_resid(Go, x, b)
    assert 1e-9 < want < 10.0 and abs(np.linalg.norm(b) - 1.0) > 0.5
    assert abs(O.true_resid_dd(Go, x, b) - want) <= 4.5e-16 * want
    assert abs(O.true_resid(Go, x, b) - want) <= 1e-13 * want
    # the same quotient from numpy norms of the same d, a formula of its own
    d = O.spmv(Go, x) - b
    assert abs(np.linalg.norm(d) / np.linalg.norm(b) - want) <= 1e-13 * want


def test_true_resid_zero_rhs():
    """||b|| = 0: the double-double form divides by 1, as the solver's normb
    does (cg_crs.cpp:109-110)."""
    G = P.grid(5, 4)
    Go = P.orc(G)
    x = P.I.gen_normal(G.n, 23)
    assert O.true_resid_dd(Go, x, np.zeros(G.n)) == pytest.approx(
        np.linalg.norm(O.spmv(Go, x)), rel=1e-15)


@needs_ref
@pytest.mark.parametrize("prog,tol,maxit", [("iccg_crs", "1e-14", "60"),
                                            ("cg_crs", "1e-14", "200"),
                                            ("iccg_crs", "1e-3", "10000")])
def test_true_resid_matches_reference_binary(tmp_path, prog, tol, maxit):
    """The reference binary's printed ||Ax-b||/||b|| (cg_crs.cpp:517-528) on
    a 30x30 Laplacian with b = 1, x0 = 0: at tol 1e-14 it is the rounding
    floor of A x - b (5.6e-14 against a recurrence residual of 4e-15), so it
    pins the exact bits of x and of the SpMV as well as the quotient."""
    import re
    import subprocess
    path = tmp_path / "g30.mtx"
    scipy.io.mmwrite(str(path), P.grid(30, 30).to_scipy())
    out = subprocess.run([R.bin_path(prog), str(path), tol, maxit], capture_output=True,
                         text=True, check=True, env=dict(os.environ, OMP_NUM_THREADS="1")).stdout
    m = re.search(r"iters=(\d+), rel_resid=(\S+), \|\|Ax-b\|\|/\|\|b\|\| = (\S+)", out)
    A = O.read_matrix_market(str(path))
    kind = "ic0" if prog == "iccg_crs" else "none"
    o = O.cg(A, np.ones(A.n), kind, ic=O.ic0_factor(A) if kind == "ic0" else None,
             tol=float(tol), max_iter=int(maxit))
    assert o.iters == int(m.group(1)) and f"{o.rel_resid:g}" == m.group(2)
    assert f"{O.true_resid(A, o.x, np.ones(A.n)):g}" == m.group(3)


# ------------------------------------------------------------------ mixed precision
@pytest.mark.parametrize("dims,c,tol,restarts", [((24, 20, 16), 2.0, 1e-8, 0),
                                                 ((24, 20, 16), 2.0, 1e-14, 1),
                                                 ((12, 12, 12), 3.0, 5e-15, 2),
                                                 ((40, 40, 1), 2.0, 1e-14, 100)])
def test_cg_mixed_restart_invariants(dims, c, tol, restarts):
    """orc_cg_mixed: FP64 CG with the FP32 IC(0) apply, restarted from the
    current x while the true residual misses tol.  Without a restart it is
    orc_cg with the FP32 apply; each restart is orc_cg started again from the
    x the previous pass left (an independent run, segment by segment); a
    solve under the 100-restart cap meets tol on the true residual."""
    G = P.grid(*dims, coef=(-c, c), shift=1e-6)
    Go = P.orc(G)
    f = O.ic0_factor(Go)
    b = P.I.gen_normal(G.n, 4)
    m = O.cg_mixed(Go, b, f, tol=tol, max_iter=3000)
    assert m.converged and m.restarts == restarts
    assert m.true_rel_resid == O.true_resid_dd(Go, m.x, b)
    if restarts < 100:
        assert m.true_rel_resid < tol
    # segment by segment
    x, total = None, 0
    for k in range(restarts + 1):
        o = O.cg(Go, b, "ic0_f32", ic=f, x0=x, tol=tol, max_iter=3000 - total, dot_mode=1)
        total += o.iters
        x = o.x
        tr = O.true_resid_dd(Go, x, b)
        assert (tr >= tol) == (k < restarts) or restarts == 100
    assert total == m.iters and np.array_equal(x, m.x) and o.rel_resid == m.rel_resid
    # the FP32 apply with FP64 dots and recurrence differs from FP64 IC(0)-CG
    d = O.cg(Go, b, "ic0", ic=f, tol=tol, max_iter=3000, dot_mode=1)
    assert not np.array_equal(d.x, m.x)
