"""ctypes bindings for the CPU oracle (liboracle.so) and the reference harness
(_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, for comparison.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class Csr:
    n: int
    rowptr: np.ndarray
    colind: np.ndarray
    val: np.ndarray

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.colind, self.rowptr), shape=(self.n, self.n))


@dataclass
class IC0Factor:
    n: int
    Lrp: np.ndarray
    Lci: np.ndarray
    Lv: np.ndarray
    diag_pos: np.ndarray
    LTcp: np.ndarray
    LTri: np.ndarray
    LTv: np.ndarray
    LT_diag_pos: np.ndarray


class _IC0F(C.Structure):
    _fields_ = [("n", C.c_int), ("Lrp", _vp), ("Lci", _vp), ("Lv", _vp),
                ("diag_pos", _vp), ("LTcp", _vp), ("LTri", _vp), ("LTv", _vp),
                ("LT_diag_pos", _vp)]


class _Sched(C.Structure):
    _fields_ = [("nc", C.c_int), ("color_ptr", _vp), ("rows_of_color", _vp),
                ("mode", C.c_int), ("nb", C.c_int), ("color_ptr_blk", _vp),
                ("blocks_of_color", _vp), ("block_ptr", _vp), ("rows_of_block", _vp)]


@dataclass
class Schedule:
    """Mirror of ColorSchedule (color.hpp:9-23); arrays owned here."""
    nc: int = 0
    color_ptr: np.ndarray | None = None
    rows_of_color: np.ndarray | None = None
    mode: int = 0  # 0 MC_NODE, 1 ABMC_BLOCK
    nb: int = 0
    color_ptr_blk: np.ndarray | None = None
    blocks_of_color: np.ndarray | None = None
    block_ptr: np.ndarray | None = None
    rows_of_block: np.ndarray | None = None

    def c_struct(self):
        return _Sched(self.nc, _ptr(self.color_ptr), _ptr(self.rows_of_color), self.mode,
                      self.nb, _ptr(self.color_ptr_blk), _ptr(self.blocks_of_color),
                      _ptr(self.block_ptr), _ptr(self.rows_of_block))


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _lib = C.CDLL(path)
        L = _lib
        L.orc_coo_to_crs.restype = C.c_long
        L.orc_expand_lower_to_full.restype = C.c_long
        L.orc_ptap_lower.restype = C.c_long
        L.orc_ic0_factor.restype = C.c_long
        L.orc_dot_ld.restype = C.c_double
        L.orc_dot_dd.restype = C.c_double
        L.orc_true_resid.restype = C.c_double
        L.orc_true_resid_dd.restype = C.c_double
        L.orc_check_interblock.restype = C.c_longlong
        L.orc_check_hazards.restype = C.c_longlong
        L.orc_ic0_factor.argtypes = [C.c_int, _vp, _vp, _vp, C.c_double] + [_vp] * 9
        L.orc_true_resid.argtypes = [C.c_int, _vp, _vp, _vp, _vp, _vp]
        L.orc_true_resid_dd.argtypes = [C.c_int, _vp, _vp, _vp, _vp, _vp]
        L.orc_dot_ld.argtypes = [C.c_long, _vp, _vp]
        L.orc_dot_dd.argtypes = [C.c_long, _vp, _vp]
        L.orc_cg_mixed.argtypes = [C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double,
                                   _vp, _vp, _vp, _vp]
        L.orc_cg.argtypes = [C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, _vp, _vp,
                             C.c_int, C.c_double, C.c_int, _vp, _vp, _vp]
    return _lib


# ---------------------------------------------------------------- L0 CRS
def coo_to_crs(nrows, I, J, V):
    I, J, V = i32(I), i32(J), f64(V)
    rp = np.zeros(nrows + 1, np.int32)
    ci = np.zeros(max(len(I), 1), np.int32)
    v = np.zeros(max(len(I), 1), np.float64)
    nnz = lib().orc_coo_to_crs(C.c_int(nrows), C.c_long(len(I)), _ptr(I), _ptr(J), _ptr(V),
                               _ptr(rp), _ptr(ci), _ptr(v))
    return Csr(nrows, rp, ci[:nnz].copy(), v[:nnz].copy())


def read_matrix_market(path):
    """crs_io.cpp:73-159 parse (banner/format checks, comment skip, 1->0 based,
    symmetric mirror, pattern -> 1.0) followed by the oracle's coo_to_crs."""
    with open(path) as f:
        lines = f.read().split("\n")
    head = lines[0].split()
    if len(head) < 5:
        raise ValueError("Invalid MatrixMarket header")
    if head[0] != "%%MatrixMarket":
        raise ValueError("Not a MatrixMarket file")
    mtx, storage, field, sym = (h.lower() for h in head[1:5])
    if mtx != "matrix" or storage != "coordinate":
        raise ValueError("Only 'matrix coordinate' is supported")
    if field not in ("real", "integer", "pattern"):
        raise ValueError("Unsupported field: " + field)
    if sym not in ("general", "symmetric"):
        raise ValueError("Unsupported symmetry: " + sym)
    k = 1
    while k < len(lines) and lines[k][:1] == "%":
        k += 1
    nr, nc_, nnz = (int(t) for t in lines[k].split()[:3])
    k += 1
    I, J, V = [], [], []
    cnt = 0
    while cnt < nnz:
        if k >= len(lines):
            raise ValueError("Unexpected EOF")
        ln = lines[k]
        k += 1
        if ln == "" or ln[0] == "%":
            continue
        t = ln.split()
        i, j = int(t[0]) - 1, int(t[1]) - 1
        v = float(t[2]) if field != "pattern" else 1.0
        if not (0 <= i < nr and 0 <= j < nc_):
            raise ValueError("Index out of range")
        I.append(i); J.append(j); V.append(v)
        if sym == "symmetric" and i != j:
            I.append(j); J.append(i); V.append(v)
        cnt += 1
    return coo_to_crs(nr, I, J, V)


def expand_lower_to_full(A: Csr) -> Csr:
    cap = 2 * A.nnz + 1
    rp = np.zeros(A.n + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    v = np.zeros(cap, np.float64)
    nnz = lib().orc_expand_lower_to_full(C.c_int(A.n), _ptr(A.rowptr), _ptr(A.colind),
                                         _ptr(A.val), _ptr(rp), _ptr(ci), _ptr(v))
    return Csr(A.n, rp, ci[:nnz].copy(), v[:nnz].copy())


# ---------------------------------------------------------------- L1 orderings
def perm_by_color(color):
    color = i32(color)
    out = np.zeros(len(color), np.int32)
    lib().orc_perm_by_color(C.c_int(len(color)), _ptr(color), _ptr(out))
    return out


def perm_by_block_color(block_of_old, block_color):
    bo, bc = i32(block_of_old), i32(block_color)
    out = np.zeros(len(bo), np.int32)
    lib().orc_perm_by_block_color(C.c_int(len(bo)), _ptr(bo), _ptr(bc), _ptr(out))
    return out


def ptap_lower(A: Csr, new_of_old) -> Csr:
    p = i32(new_of_old)
    cap = A.nnz + 1
    rp = np.zeros(A.n + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    v = np.zeros(cap, np.float64)
    nnz = lib().orc_ptap_lower(C.c_int(A.n), _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val),
                               _ptr(p), _ptr(rp), _ptr(ci), _ptr(v))
    if nnz < 0:
        raise ValueError("Permute_PtAP: size mismatch")
    return Csr(A.n, rp, ci[:nnz].copy(), v[:nnz].copy())


def color_ptr_rows(color, nc, new_of_old):
    color, p = i32(color), i32(new_of_old)
    cp = np.zeros(nc + 1, np.int32)
    rc = np.zeros(len(color), np.int32)
    if lib().orc_color_ptr_rows(C.c_int(len(color)), _ptr(color), C.c_int(nc), _ptr(p),
                                _ptr(cp), _ptr(rc)) != 0:
        raise ValueError("color out of range")
    return cp, rc


def abmc_schedule(new_of_old, block_of_old, block_color, nb, nc):
    p, bo, bc = i32(new_of_old), i32(block_of_old), i32(block_color)
    n = len(p)
    cpb = np.zeros(nc + 1, np.int32)
    boc = np.zeros(nb, np.int32)
    bp = np.zeros(nb + 1, np.int32)
    rob = np.zeros(n, np.int32)
    lib().orc_abmc_schedule(C.c_int(n), _ptr(p), _ptr(bo), _ptr(bc), C.c_int(nb), C.c_int(nc),
                            _ptr(cpb), _ptr(boc), _ptr(bp), _ptr(rob))
    return cpb, boc, bp, rob


# ---------------------------------------------------------------- L2 precond
def jacobi_inv_diag(A: Csr):
    out = np.zeros(A.n, np.float64)
    lib().orc_jacobi_inv_diag(C.c_int(A.n), _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(out))
    return out


class PivotError(RuntimeError):
    pass


def ic0_factor(A: Csr, shift=0.0) -> IC0Factor:
    n = A.n
    cap = A.nnz + 1
    Lrp = np.zeros(n + 1, np.int32); Lci = np.zeros(cap, np.int32); Lv = np.zeros(cap)
    dp = np.zeros(n, np.int32); LTcp = np.zeros(n + 1, np.int32)
    LTri = np.zeros(cap, np.int32); LTv = np.zeros(cap); LTdp = np.zeros(n, np.int32)
    err = C.c_int(-1)
    nnz = lib().orc_ic0_factor(n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), shift,
                               _ptr(Lrp), _ptr(Lci), _ptr(Lv), _ptr(dp), _ptr(LTcp), _ptr(LTri),
                               _ptr(LTv), _ptr(LTdp), C.byref(err))
    if nnz == -1:
        raise PivotError(f"IC(0): diagonal missing at row {err.value}")
    if nnz == -2:
        raise PivotError(f"IC(0): zero/near-zero pivot at row {err.value}")
    if nnz == -3:
        raise PivotError(f"IC(0): non-positive pivot at row {err.value}")
    return IC0Factor(n, Lrp, Lci[:nnz].copy(), Lv[:nnz].copy(), dp, LTcp, LTri[:nnz].copy(),
                     LTv[:nnz].copy(), LTdp)


def _ic0f_struct(f: IC0Factor):
    return _IC0F(f.n, _ptr(f.Lrp), _ptr(f.Lci), _ptr(f.Lv), _ptr(f.diag_pos), _ptr(f.LTcp),
                 _ptr(f.LTri), _ptr(f.LTv), _ptr(f.LT_diag_pos))


def ic0_apply(f: IC0Factor, r, sched: Schedule | None = None):
    r = f64(r)
    z = np.zeros(f.n)
    y = np.zeros(f.n)
    fs = _ic0f_struct(f)
    ss = sched.c_struct() if sched is not None else None
    lib().orc_ic0_apply(C.byref(fs), C.byref(ss) if ss is not None else None, _ptr(r), _ptr(z), _ptr(y))
    return z


def ic0_apply_f32(f: IC0Factor, r):
    """The mixed-precision apply (FP32 sweeps, sequential order)."""
    r = f64(r)
    z = np.zeros(f.n)
    fs = _ic0f_struct(f)
    lib().orc_ic0_apply_f32(C.byref(fs), _ptr(r), _ptr(z))
    return z


# ---------------------------------------------------------------- L3 Krylov
def dot_ld(a, b):
    a, b = f64(a), f64(b)
    return lib().orc_dot_ld(len(a), _ptr(a), _ptr(b))


def dot_dd(a, b):
    a, b = f64(a), f64(b)
    return lib().orc_dot_dd(len(a), _ptr(a), _ptr(b))


def spmv(A: Csr, x):
    x = f64(x)
    y = np.zeros(A.n)
    lib().orc_spmv(C.c_int(A.n), _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(x), _ptr(y))
    return y


def true_resid(A: Csr, x, b):
    """||A x - b|| / ||b|| summed as main() does (cg_crs.cpp:517-528)."""
    x, b = f64(x), f64(b)
    return lib().orc_true_resid(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(x), _ptr(b))


def true_resid_dd(A: Csr, x, b):
    """The same quotient with double-double sums (||b|| = 0 taken as 1)."""
    x, b = f64(x), f64(b)
    return lib().orc_true_resid_dd(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(x),
                                   _ptr(b))


@dataclass
class CGOut:
    iters: int
    rel_resid: float
    converged: bool
    x: np.ndarray
    hist: np.ndarray


def cg(A: Csr, b, kind, *, ic: IC0Factor | None = None, sched: Schedule | None = None,
       x0=None, max_iter=10000, tol=1e-8, dot_mode=0):
    """conjugate_gradient<Precond> (cg_crs.cpp:89-162).  kind: 'none', 'jacobi', 'ic0',
    'ic0_f32' (the IC(0) apply in FP32, orc_ic0_apply_f32)."""
    k = {"none": 0, "jacobi": 1, "ic0": 2, "ic0_f32": 3}[kind]
    b = f64(b)
    x = np.zeros(A.n) if x0 is None else f64(x0).copy()
    inv = jacobi_inv_diag(A) if k == 1 else None
    fs = _ic0f_struct(ic) if k >= 2 else None
    ss = sched.c_struct() if (sched is not None and k == 2) else None
    rel = C.c_double(0.0)
    conv = C.c_int(0)
    hist = np.full(max_iter + 1, np.nan)
    it = lib().orc_cg(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(b), _ptr(x), k,
                      _ptr(inv), C.byref(fs) if fs is not None else None,
                      C.byref(ss) if ss is not None else None, max_iter, tol, dot_mode,
                      C.byref(rel), C.byref(conv), _ptr(hist))
    return CGOut(it, rel.value, bool(conv.value), x, hist[: it + 1].copy())


@dataclass
class CGMixedOut:
    iters: int
    rel_resid: float
    converged: bool
    restarts: int
    true_rel_resid: float
    x: np.ndarray


def cg_mixed(A: Csr, b, ic: IC0Factor, *, x0=None, max_iter=10000, tol=1e-8):
    """FP64 CG with the FP32 IC(0) apply and double-double dots, stopped on
    the FP64 true residual with restarts from the current x."""
    b = f64(b)
    x = np.zeros(A.n) if x0 is None else f64(x0).copy()
    fs = _ic0f_struct(ic)
    rel, tr = C.c_double(0.0), C.c_double(0.0)
    conv, rs = C.c_int(0), C.c_int(0)
    it = lib().orc_cg_mixed(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(b), _ptr(x),
                            C.byref(fs), max_iter, tol, C.byref(rel), C.byref(conv),
                            C.byref(rs), C.byref(tr))
    return CGMixedOut(it, rel.value, bool(conv.value), rs.value, tr.value, x)


# ---------------------------------------------------------------- validators
def check_interblock(A: Csr, block_of_old, block_color, nb, nc):
    bo, bc = i32(block_of_old), i32(block_color)
    per = np.zeros(nc, np.int64)
    tot = lib().orc_check_interblock(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(bo), _ptr(bc),
                                     nb, nc, _ptr(per))
    if tot < 0:
        raise ValueError("id out of range while scanning edges")
    return int(tot), per


def check_hazards(Ap_full: Csr, new_of_old, color_old, block_of_old=None, nc_cap=4096):
    p, c = i32(new_of_old), i32(color_old)
    bo = i32(block_of_old) if block_of_old is not None else None
    h = np.zeros(nc_cap, np.int64); s = np.zeros(nc_cap, np.int64); x = np.zeros(nc_cap, np.int64)
    nc = C.c_int(0)
    tot = lib().orc_check_hazards(Ap_full.n, _ptr(Ap_full.rowptr), _ptr(Ap_full.colind), _ptr(p),
                                  _ptr(c), _ptr(bo), C.byref(nc), _ptr(h), _ptr(s), _ptr(x), nc_cap)
    if tot < 0:
        raise ValueError("check_hazards: bad input")
    k = nc.value
    return int(tot), h[:k].copy(), s[:k].copy(), x[:k].copy()
