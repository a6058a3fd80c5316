"""ctypes bindings for oracle/_ref/libref.so: the reference's own sources
compiled in place (oracle/Makefile) behind oracle/ref_harness.cpp.

TEST INFRASTRUCTURE ONLY (comparison / CPU baseline).  available() is False when
the library was not built (no reference sources at build time).
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

from .oracle import Csr, IC0Factor, Schedule, _ptr, f64, i32

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libref.so")
_vp = C.c_void_p
_lib = None


def available():
    return os.path.exists(_PATH)


def bin_path(name):
    return os.path.join(_HERE, "_ref", name)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError("oracle/_ref/libref.so not built")
        _lib = C.CDLL(_PATH)
        L = _lib
        L.ref_last_error.restype = C.c_char_p
        L.ref_read_mm.argtypes = [C.c_char_p]
        L.ref_ic0_nnz.restype = C.c_long
        L.ref_dot.restype = C.c_double
        L.ref_dot.argtypes = [C.c_long, _vp, _vp]
        L.ref_ic0_create.argtypes = [C.c_int, _vp, _vp, _vp, C.c_double]
        L.ref_cg.argtypes = [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_double,
                             C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp,
                             _vp, _vp]
        L.ref_ic0_apply.argtypes = [C.c_int, _vp, _vp, C.c_int, C.c_int, _vp, _vp, C.c_int,
                                    C.c_int, _vp, _vp, _vp, _vp]
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_set_threads(1)  # sequential reductions unless set_threads() says otherwise
    return _lib


def set_threads(n):
    """OpenMP threads of the reference's loops (1 = sequential reductions)."""
    lib().ref_set_threads(int(n))


def _err():
    return lib().ref_last_error().decode()


def _take_crs():
    n = C.c_int(0)
    nnz = C.c_long(0)
    lib().ref_crs_dims(C.byref(n), C.byref(nnz))
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1))
    lib().ref_crs_copy(_ptr(rp), _ptr(ci), _ptr(v))
    return Csr(n.value, rp, ci[: nnz.value].copy(), v[: nnz.value].copy())


def read_matrix_market(path):
    if lib().ref_read_mm(str(path).encode()) != 0:
        raise RuntimeError(_err())
    return _take_crs()


def expand_lower_to_full(A: Csr):
    lib().ref_expand_lower_to_full(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val))
    return _take_crs()


def ptap_lower(A: Csr, new_of_old):
    p = i32(new_of_old)
    if lib().ref_ptap(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(p)) != 0:
        raise RuntimeError(_err())
    return _take_crs()


def perm_by_color(color):
    c = i32(color)
    out = np.zeros(len(c), np.int32)
    lib().ref_perm_by_color(len(c), _ptr(c), _ptr(out))
    return out


def perm_by_block_color(block_of_old, block_color):
    bo, bc = i32(block_of_old), i32(block_color)
    out = np.zeros(len(bo), np.int32)
    lib().ref_perm_by_block_color(len(bo), _ptr(bo), len(bc), _ptr(bc), _ptr(out))
    return out


def color_ptr_rows(color, nc, new_of_old):
    c, p = i32(color), i32(new_of_old)
    cp = np.zeros(nc + 1, np.int32)
    rc = np.zeros(len(c), np.int32)
    if lib().ref_color_ptr_rows(len(c), _ptr(c), nc, _ptr(p), _ptr(cp), _ptr(rc)) != 0:
        raise RuntimeError(_err())
    return cp, rc


def abmc_schedule(new_of_old, block_of_old, block_color, nb, nc):
    p, bo, bc = i32(new_of_old), i32(block_of_old), i32(block_color)
    n = len(p)
    cpb = np.zeros(nc + 1, np.int32); boc = np.zeros(nb, np.int32)
    bp = np.zeros(nb + 1, np.int32); rob = np.zeros(n, np.int32)
    lib().ref_abmc_schedule(n, _ptr(p), _ptr(bo), _ptr(bc), nb, nc, _ptr(cpb), _ptr(boc),
                            _ptr(bp), _ptr(rob))
    return cpb, boc, bp, rob


def _read_ids(fn, path, n):
    out = np.zeros(n, np.int32)
    cnt = C.c_int(0)
    if fn(str(path).encode(), n, _ptr(out), C.byref(cnt)) != 0:
        raise RuntimeError(_err())
    return out, cnt.value


def read_col(path, n):
    return _read_ids(lib().ref_read_col, path, n)


def read_blk(path, n):
    return _read_ids(lib().ref_read_blk, path, n)


def read_bcol(path, nb):
    return _read_ids(lib().ref_read_bcol, path, nb)


def jacobi_inv_diag(A: Csr):
    out = np.zeros(A.n)
    lib().ref_jacobi(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(out))
    return out


def ic0_factor(A: Csr, shift=0.0) -> IC0Factor:
    if lib().ref_ic0_create(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), shift) != 0:
        raise RuntimeError(_err())
    nnz = lib().ref_ic0_nnz()
    n = A.n
    f = IC0Factor(n, np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz),
                  np.zeros(n, np.int32), np.zeros(n + 1, np.int32), np.zeros(nnz, np.int32),
                  np.zeros(nnz), np.zeros(n, np.int32))
    lib().ref_ic0_copy(_ptr(f.Lrp), _ptr(f.Lci), _ptr(f.Lv), _ptr(f.diag_pos), _ptr(f.LTcp),
                       _ptr(f.LTri), _ptr(f.LTv), _ptr(f.LT_diag_pos))
    return f


def _sched_args(s: Schedule | None):
    if s is None:
        return [0, 0, None, None, 0, 0, None, None, None, None]
    return [1, s.nc, _ptr(s.color_ptr), _ptr(s.rows_of_color), s.mode, s.nb,
            _ptr(s.color_ptr_blk), _ptr(s.blocks_of_color), _ptr(s.block_ptr),
            _ptr(s.rows_of_block)]


def ic0_apply(n, r, sched: Schedule | None = None):
    """Applies the factor built by the last ic0_factor call."""
    r = f64(r)
    z = np.zeros(n)
    lib().ref_ic0_apply(n, _ptr(r), _ptr(z), *_sched_args(sched))
    return z


def cg(A: Csr, b, kind, *, sched: Schedule | None = None, x0=None, max_iter=10000, tol=1e-8,
       shift=0.0):
    """conjugate_gradient<Precond> of the reference (cg_crs.cpp:89-162)."""
    k = {"none": 0, "jacobi": 1, "ic0": 2}[kind]
    if k == 2:
        ic0_factor(A, shift)
    b = f64(b)
    x = np.zeros(A.n) if x0 is None else f64(x0).copy()
    rel = C.c_double(0.0)
    conv = C.c_int(0)
    it = lib().ref_cg(k, A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(b), _ptr(x),
                      max_iter, tol, *_sched_args(sched if k == 2 else None),
                      C.byref(rel), C.byref(conv))
    return it, rel.value, bool(conv.value), x


def cg_timed_ic0(A: Csr, b, sched: Schedule | None, max_iter, tol=1e-8, shift=0.0,
                 build_factor=True):
    """IC(0)-CG of the reference, timing only the CG phase as its main()
    does.  Returns (iters, rel_resid, converged, x, seconds)."""
    if build_factor:
        ic0_factor_only(A, shift)
    b = f64(b)
    x = np.zeros(A.n)
    rel = C.c_double(0.0)
    conv = C.c_int(0)
    sec = C.c_double(0.0)
    f = lib().ref_cg_timed_ic0
    f.argtypes = [C.c_int] + [_vp] * 5 + [C.c_int, C.c_double, C.c_int, C.c_int, _vp, _vp,
                                          C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    it = f(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(b), _ptr(x), max_iter, tol,
           *_sched_args(sched), C.byref(rel), C.byref(conv), C.byref(sec))
    return it, rel.value, bool(conv.value), x, sec.value


def ic0_factor_only(A: Csr, shift=0.0):
    if lib().ref_ic0_create(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), shift) != 0:
        raise RuntimeError(_err())


def spmv(A: Csr, x):
    x = f64(x)
    y = np.zeros(A.n)
    lib().ref_spmv(A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(A.val), _ptr(x), _ptr(y))
    return y


def dot(a, b):
    a, b = f64(a), f64(b)
    return lib().ref_dot(len(a), _ptr(a), _ptr(b))


def _capture_stderr(fn):
    """Runs fn() with fd 2 redirected to a temp file; returns the text."""
    import sys
    sys.stderr.flush()
    saved = os.dup(2)
    with tempfile.TemporaryFile(mode="w+b") as tf:
        os.dup2(tf.fileno(), 2)
        try:
            fn()
        finally:
            os.dup2(saved, 2)
            os.close(saved)
        tf.seek(0)
        return tf.read().decode(errors="replace")


def check_interblock_report(A: Csr, block_of_old, block_color, nb, nc):
    bo, bc = i32(block_of_old), i32(block_color)
    return _capture_stderr(lambda: lib().ref_check_interblock(
        A.n, _ptr(A.rowptr), _ptr(A.colind), _ptr(bo), nb, _ptr(bc), nc))


def check_hazards_report(Ap_full: Csr, new_of_old, color_old, block_of_old=None,
                         block_color=None):
    p, c = i32(new_of_old), i32(color_old)
    bo = i32(block_of_old) if block_of_old is not None else None
    bc = i32(block_color) if block_color is not None else None
    nb = len(bc) if bc is not None else 0
    return _capture_stderr(lambda: lib().ref_check_hazards(
        Ap_full.n, _ptr(Ap_full.rowptr), _ptr(Ap_full.colind), _ptr(Ap_full.val), _ptr(p),
        _ptr(c), _ptr(bo), nb, _ptr(bc)))
