"""Python mirror of the reference's solve-path interface over the C ABI.

Names, argument meaning and error behaviour follow stomo0511/ICCG:
  CRS, read_matrix_market_crs, expand_lower_to_full        (crs_io.hpp)
  ReadColorFile_1Based, BuildPermutationByColor,
  Permute_PtAP_LowerCRS_to_LowerCRS,
  BuildColorPtrRows_FromPermutationAndColor, ColorSchedule  (color.hpp)
  ReadBlockFile_1Based, ReadBlockColorFile_1Based,
  BuildPermutationByBlockColor, BuildABMCSchedule           (block.hpp)
  Identity, Jacobi                                          (precond.hpp)
  IC0                                                       (ic0.hpp)
  CGResult, conjugate_gradient                              (cg_crs.cpp:82-162)
Errors the reference raises as std::runtime_error raise RuntimeError
(IccgError) with the same message.  Every numeric routine runs in
libiccg_hip.so; nothing here computes on the CPU except array bookkeeping.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import IccgError, check, lib

_vp = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ----------------------------------------------------------------------------- L0
@dataclass
class CRS:
    """struct CRS (crs_io.hpp:10-15): 0-based CSR, int32 indices, FP64 values."""
    n: int
    rowptr: np.ndarray
    colind: np.ndarray
    val: np.ndarray

    @property
    def nnz(self):
        return int(self.rowptr[-1]) if len(self.rowptr) else 0

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.colind, self.rowptr), shape=(self.n, self.n))


def read_matrix_market_crs(path) -> CRS:
    """read_matrix_market_crs (crs_io.hpp:52)."""
    n = C.c_int(0)
    nnz = C.c_long(0)
    rp = C.POINTER(C.c_int)()
    ci = C.POINTER(C.c_int)()
    v = C.POINTER(C.c_double)()
    check(lib().iccg_mm_read(str(path).encode(), C.byref(n), C.byref(nnz), C.byref(rp),
                             C.byref(ci), C.byref(v)))
    try:
        m = nnz.value
        rowptr = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
        colind = np.ctypeslib.as_array(ci, shape=(max(m, 1),))[:m].copy()
        val = np.ctypeslib.as_array(v, shape=(max(m, 1),))[:m].copy()
    finally:
        for q in (rp, ci, v):
            lib().iccg_free_host(C.cast(q, _vp))
    return CRS(n.value, rowptr, colind, val)


def coo_to_crs(nrows, I, J, V) -> CRS:
    I, J, V = _i32(I), _i32(J), _f64(V)
    rp = np.zeros(nrows + 1, np.int32)
    ci = np.zeros(max(len(I), 1), np.int32)
    v = np.zeros(max(len(I), 1))
    nnz = C.c_long(0)
    check(lib().iccg_coo_to_crs(nrows, len(I), _p(I), _p(J), _p(V), _p(rp), _p(ci), _p(v),
                                C.byref(nnz)))
    return CRS(nrows, rp, ci[: nnz.value].copy(), v[: nnz.value].copy())


def expand_lower_to_full(L: CRS) -> CRS:
    """expand_lower_to_full (crs_io.hpp:56)."""
    cap = max(2 * L.nnz, 1)
    rp = np.zeros(L.n + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    v = np.zeros(cap)
    nnz = C.c_long(0)
    check(lib().iccg_expand_lower_to_full(L.n, _p(_i32(L.rowptr)), _p(_i32(L.colind)),
                                          _p(_f64(L.val)), _p(rp), _p(ci), _p(v), C.byref(nnz)))
    return CRS(L.n, rp, ci[: nnz.value].copy(), v[: nnz.value].copy())


# ----------------------------------------------------------------------------- L1
def _read_ids(fn, path, n):
    out = np.zeros(max(n, 1), np.int32)
    cnt = C.c_int(0)
    check(fn(str(path).encode(), n, _p(out), C.byref(cnt)))
    return out[:n].copy(), cnt.value


def ReadColorFile_1Based(path, n_expected):
    """color.hpp:4 -> (color[n] 0-based, nc)."""
    return _read_ids(lib().iccg_read_color_file, path, n_expected)


def ReadBlockFile_1Based(path, n_expected):
    """block.hpp:10 -> (block_of_old[n], nb)."""
    return _read_ids(lib().iccg_read_block_file, path, n_expected)


def ReadBlockColorFile_1Based(path, nb_expected):
    """block.hpp:18 -> (block_color[nb], nc)."""
    return _read_ids(lib().iccg_read_block_color_file, path, nb_expected)


def write_color_file(path, color, nc):
    c = _i32(color)
    check(lib().iccg_write_color_file(str(path).encode(), len(c), _p(c), nc))


def write_block_file(path, block_of_old, nb):
    c = _i32(block_of_old)
    check(lib().iccg_write_block_file(str(path).encode(), len(c), _p(c), nb))


def write_block_color_file(path, block_color, nc):
    c = _i32(block_color)
    check(lib().iccg_write_block_color_file(str(path).encode(), len(c), _p(c), nc))


def BuildPermutationByColor(color):
    c = _i32(color)
    out = np.zeros(len(c), np.int32)
    check(lib().iccg_perm_by_color(len(c), _p(c), _p(out)))
    return out


def BuildPermutationByBlockColor(block_of_old, block_color):
    bo, bc = _i32(block_of_old), _i32(block_color)
    out = np.zeros(len(bo), np.int32)
    check(lib().iccg_perm_by_block_color(len(bo), _p(bo), len(bc), _p(bc), _p(out)))
    return out


def Permute_PtAP_LowerCRS_to_LowerCRS(A: CRS, new_of_old) -> CRS:
    p = _i32(new_of_old)
    if len(p) != A.n:
        raise IccgError(_capi.ICCG_ERR_ARG, "Permute_PtAP: new_of_old size mismatch")
    cap = max(A.nnz, 1)
    rp = np.zeros(A.n + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    v = np.zeros(cap)
    nnz = C.c_long(0)
    check(lib().iccg_ptap_lower(A.n, _p(_i32(A.rowptr)), _p(_i32(A.colind)), _p(_f64(A.val)),
                                _p(p), _p(rp), _p(ci), _p(v), C.byref(nnz)))
    return CRS(A.n, rp, ci[: nnz.value].copy(), v[: nnz.value].copy())


def BuildColorPtrRows_FromPermutationAndColor(color, nc, new_of_old):
    c, p = _i32(color), _i32(new_of_old)
    if len(p) != len(c):
        raise IccgError(_capi.ICCG_ERR_ARG, "size mismatch")
    cp = np.zeros(nc + 1, np.int32)
    rc = np.zeros(max(len(c), 1), np.int32)
    check(lib().iccg_color_ptr_rows(len(c), _p(c), nc, _p(p), _p(cp), _p(rc)))
    return cp, rc[: len(c)].copy()


def BuildABMCSchedule(new_of_old, block_of_old, block_color, nb, nc):
    p, bo, bc = _i32(new_of_old), _i32(block_of_old), _i32(block_color)
    n = len(p)
    cpb = np.zeros(nc + 1, np.int32)
    boc = np.zeros(max(nb, 1), np.int32)
    bp = np.zeros(nb + 1, np.int32)
    rob = np.zeros(max(n, 1), np.int32)
    check(lib().iccg_abmc_schedule(n, _p(p), _p(bo), _p(bc), nb, nc, _p(cpb), _p(boc), _p(bp),
                                   _p(rob)))
    return cpb, boc[:nb].copy(), bp, rob[:n].copy()


@dataclass
class ColorSchedule:
    """struct ColorSchedule (color.hpp:9-23)."""
    MC_NODE = 0
    ABMC_BLOCK = 1
    nc: int = 0
    color_ptr: np.ndarray | None = None
    rows_of_color: np.ndarray | None = None
    mode: int = 0
    nb: int = 0
    color_ptr_blk: np.ndarray | None = None
    blocks_of_color: np.ndarray | None = None
    block_ptr: np.ndarray | None = None
    rows_of_block: np.ndarray | None = None
    _handle: object = field(default=None, repr=False, compare=False)

    def handle(self, n):
        """Device schedule object for an n-row system (created once)."""
        if self._handle is None:
            h = _vp()
            if self.mode == self.ABMC_BLOCK and all(
                    a is not None for a in (self.color_ptr_blk, self.blocks_of_color,
                                            self.block_ptr, self.rows_of_block)):
                self._arrs = [_i32(self.color_ptr_blk), _i32(self.blocks_of_color),
                              _i32(self.block_ptr), _i32(self.rows_of_block)]
                check(lib().iccg_schedule_create_abmc(n, self.nc, self.nb,
                                                      *[_p(a) for a in self._arrs], C.byref(h)))
            elif self.color_ptr is not None and self.rows_of_color is not None:
                self._arrs = [_i32(self.color_ptr), _i32(self.rows_of_color)]
                check(lib().iccg_schedule_create_mc(n, self.nc, *[_p(a) for a in self._arrs],
                                                    C.byref(h)))
            else:  # IC0::apply falls back to the sequential solve (ic0.cpp:389)
                check(lib().iccg_schedule_create_mc(n, 0, None, None, C.byref(h)))
            self._handle = _Handle(h, lib().iccg_schedule_destroy)
        return self._handle.h

    @classmethod
    def mc(cls, nc, color_ptr, rows_of_color):
        return cls(nc=nc, color_ptr=_i32(color_ptr), rows_of_color=_i32(rows_of_color))

    @classmethod
    def abmc(cls, nc, nb, color_ptr_blk, blocks_of_color, block_ptr, rows_of_block):
        return cls(nc=nc, mode=cls.ABMC_BLOCK, nb=nb, color_ptr_blk=_i32(color_ptr_blk),
                   blocks_of_color=_i32(blocks_of_color), block_ptr=_i32(block_ptr),
                   rows_of_block=_i32(rows_of_block))


class _Handle:
    def __init__(self, h, destroy):
        self.h = h
        self._destroy = destroy

    def __del__(self):
        try:
            if self.h:
                self._destroy(self.h)
        except Exception:
            pass


# ----------------------------------------------------------------------------- generators
def greedy_color(A: CRS):
    c = np.zeros(max(A.n, 1), np.int32)
    nc = C.c_int(0)
    check(lib().iccg_greedy_color(A.n, _p(_i32(A.rowptr)), _p(_i32(A.colind)), _p(c),
                                  C.byref(nc)))
    return c[: A.n].copy(), nc.value


def abmc_blocking(A: CRS, block_size):
    bo = np.zeros(max(A.n, 1), np.int32)
    bc = np.zeros(max(A.n, 1), np.int32)
    nb = C.c_int(0)
    nc = C.c_int(0)
    check(lib().iccg_abmc_blocking(A.n, _p(_i32(A.rowptr)), _p(_i32(A.colind)), block_size,
                                   _p(bo), C.byref(nb), _p(bc), C.byref(nc)))
    return bo[: A.n].copy(), nb.value, bc[: nb.value].copy(), nc.value


def abmc_blocking_grid(nx, ny, nz, bx, by, bz):
    n = nx * ny * nz
    bo = np.zeros(n, np.int32)
    bc = np.zeros(n, np.int32)
    nb = C.c_int(0)
    nc = C.c_int(0)
    check(lib().iccg_abmc_blocking_grid(nx, ny, nz, bx, by, bz, _p(bo), C.byref(nb), _p(bc),
                                        C.byref(nc)))
    return bo, nb.value, bc[: nb.value].copy(), nc.value


def gen_grid(nx, ny=1, nz=1, coef_exp=(0.0, 0.0), diag_shift=0.0, seed=0x1CC6,
             fp32_values=False) -> CRS:
    """2D 5-point / 3D 7-point finite-volume matrix (see iccg_gen_grid).

    fp32_values: round every value to the nearest float, the diagonal
    upwards (so it still dominates), for systems stored with FP32 values."""
    n = C.c_int(0)
    nnz = C.c_long(0)
    lo, hi = coef_exp
    check(lib().iccg_gen_grid(nx, ny, nz, lo, hi, diag_shift, seed, C.byref(n), C.byref(nnz),
                              None, None, None))
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    v = np.zeros(nnz.value)
    check(lib().iccg_gen_grid(nx, ny, nz, lo, hi, diag_shift, seed, C.byref(n), C.byref(nnz),
                              _p(rp), _p(ci), _p(v)))
    if fp32_values:
        v32 = v.astype(np.float32)
        diag = ci == np.repeat(np.arange(n.value, dtype=np.int32), np.diff(rp))
        low = diag & (v32.astype(np.float64) < v)
        v32[low] = np.nextafter(v32[low], np.float32(np.inf))
        v = v32.astype(np.float64)
    return CRS(n.value, rp, ci, v)


def gen_normal(n, seed):
    out = np.zeros(max(n, 1))
    check(lib().iccg_gen_normal(n, seed, _p(out)))
    return out[:n].copy()


def check_interblock(A: CRS, block_of_old, block_color, nb, nc):
    per = np.zeros(max(nc, 1), np.int64)
    tot = C.c_longlong(0)
    check(lib().iccg_check_interblock(A.n, _p(_i32(A.rowptr)), _p(_i32(A.colind)),
                                      _p(_i32(block_of_old)), nb, _p(_i32(block_color)), nc,
                                      _p(per), C.byref(tot)))
    return tot.value, per[:nc].copy()


def check_hazards(Ap_full: CRS, new_of_old, color_old, block_of_old=None, nc_cap=4096):
    h = np.zeros(nc_cap, np.int64)
    s = np.zeros(nc_cap, np.int64)
    x = np.zeros(nc_cap, np.int64)
    nc = C.c_int(0)
    tot = C.c_longlong(0)
    bo = _i32(block_of_old) if block_of_old is not None else None
    check(lib().iccg_check_hazards(Ap_full.n, _p(_i32(Ap_full.rowptr)), _p(_i32(Ap_full.colind)),
                                   _p(_i32(new_of_old)), _p(_i32(color_old)), _p(bo), nc_cap,
                                   C.byref(nc), _p(h), _p(s), _p(x), C.byref(tot)))
    k = nc.value
    return tot.value, h[:k].copy(), s[:k].copy(), x[:k].copy()


# ----------------------------------------------------------------------------- device
class DeviceCSR:
    """A CRS uploaded to HBM (iccg_csr_create)."""

    def __init__(self, A: CRS):
        self.n = A.n
        self.nnz = A.nnz
        h = _vp()
        check(lib().iccg_csr_create(A.n, A.nnz, _p(_i32(A.rowptr)), _p(_i32(A.colind)),
                                    _p(_f64(A.val)), C.byref(h)))
        self._h = _Handle(h, lib().iccg_csr_destroy)
        self.crs = A

    @property
    def handle(self):
        return self._h.h

    def spmv(self, x):
        x = _f64(x)
        y = np.zeros(self.n)
        check(lib().iccg_spmv(self.handle, _p(x), _p(y)))
        return y


def _device_csr(A):
    if isinstance(A, DeviceCSR):
        return A
    dev = getattr(A, "_device", None)
    if dev is None:
        dev = DeviceCSR(A)
        try:
            object.__setattr__(A, "_device", dev)
        except Exception:
            pass
    return dev


class _Precond:
    kind = None

    def __init__(self, A, shift=0.0, sched=None, precision="fp64"):
        self.A = _device_csr(A)
        self.n = self.A.n
        self.precision = {"fp64": _capi.FP64, "mixed": _capi.MIXED}[precision]
        self._sched = sched
        h = _vp()
        sh = sched.handle(self.n) if sched is not None else None
        check(lib().iccg_precond_create(self.A.handle, self.kind, shift, sh, self.precision,
                                        C.byref(h)))
        self._h = _Handle(h, lib().iccg_precond_destroy)

    @property
    def handle(self):
        return self._h.h

    def use_schedule(self, sched):
        if sched is not self._sched:
            sh = sched.handle(self.n) if sched is not None else None
            check(lib().iccg_precond_set_schedule(self.handle, sh))
            self._sched = sched

    def apply(self, r, sched=None):
        """z = M^-1 r (precond.hpp:12,38; ic0.hpp:34); returns z."""
        if sched is not None or self.kind == _capi.PRECOND_IC0:
            self.use_schedule(sched)
        r = _f64(r)
        z = np.zeros(self.n)
        check(lib().iccg_precond_apply(self.handle, _p(r), _p(z)))
        return z


class Identity(_Precond):
    """precond.hpp:8-17."""
    kind = _capi.PRECOND_IDENTITY

    def __init__(self, A):
        super().__init__(A)

    def name(self):
        return "No preconditioner"


class Jacobi(_Precond):
    """precond.hpp:20-45."""
    kind = _capi.PRECOND_JACOBI

    def __init__(self, A):
        super().__init__(A)

    def name(self):
        return "Jacobi preconditioner"


class IC0(_Precond):
    """ic0.hpp:11-76.  `sched` (optional) plans the factorization and the
    triangular solves; apply()/conjugate_gradient() re-plan when called with
    another schedule, as the reference passes it per call."""
    kind = _capi.PRECOND_IC0

    def __init__(self, A, shift=1e-12, sched=None, precision="fp64"):
        super().__init__(A, shift, sched, precision)

    def name(self):
        return ("IC preconditioner with multi-coloring" if self._sched is not None
                else "IC preconditioner")

    def plan(self):
        """(plan name, section width) of the device triangular solves."""
        p, w = C.c_int(0), C.c_int(0)
        check(lib().iccg_precond_get_plan(self.handle, C.byref(p), C.byref(w)))
        return _capi.PLAN_NAMES[p.value], w.value

    def factor(self):
        """(Lrp, Lci, Lv, LTcp, LTri, LTv): IC0's members (ic0.hpp:38-49)."""
        nnz = C.c_long(0)
        check(lib().iccg_precond_factor_nnz(self.handle, C.byref(nnz)))
        m = nnz.value
        out = [np.zeros(self.n + 1, np.int32), np.zeros(m, np.int32), np.zeros(m),
               np.zeros(self.n + 1, np.int32), np.zeros(m, np.int32), np.zeros(m)]
        check(lib().iccg_precond_get_factor(self.handle, *[_p(a) for a in out]))
        return tuple(out)


@dataclass
class CGResult:
    """CGResult (cg_crs.cpp:82-86) + the GPU path's extra report fields."""
    iters: int = 0
    rel_resid: float = float("nan")
    converged: bool = False
    breakdown: bool = False
    true_rel_resid: float = -1.0
    restarts: int = 0
    solve_ms: float = 0.0


def _result(c: _capi.CGResultC) -> CGResult:
    return CGResult(c.iters, c.rel_resid, bool(c.converged), bool(c.breakdown),
                    c.true_rel_resid, c.restarts, c.solve_ms)


def conjugate_gradient(A, b, x, M, max_iter, tol, sched=None, *, true_residual=None):
    """conjugate_gradient<Precond> (cg_crs.cpp:89-162).  x (numpy float64) is
    updated in place, as the reference's std::vector& x."""
    if sched is not None or isinstance(M, IC0):
        M.use_schedule(sched)
    dA = _device_csr(A)
    if not isinstance(x, np.ndarray) or x.dtype != np.float64 or not x.flags.c_contiguous:
        raise TypeError("x must be a contiguous float64 numpy array (updated in place)")
    b = _f64(b)
    flags = 0
    if true_residual is None:
        true_residual = M.precision == _capi.MIXED
    if true_residual:
        flags |= _capi.SOLVE_TRUE_RESID
    res = _capi.CGResultC()
    check(lib().iccg_cg_solve(dA.handle, M.handle, _p(b), _p(x), max_iter, tol, flags,
                              C.byref(res)))
    return _result(res)


class Solver:
    """Device-resident solver: b, x and the Krylov vectors stay in HBM."""

    def __init__(self, A, M):
        self.A = _device_csr(A)
        self.M = M
        h = _vp()
        check(lib().iccg_solver_create(self.A.handle, M.handle, C.byref(h)))
        self._h = _Handle(h, lib().iccg_solver_destroy)
        self.n = self.A.n

    def set(self, b, x0=None):
        b = _f64(b)
        x0 = None if x0 is None else _f64(x0)
        check(lib().iccg_solver_set(self._h.h, _p(b), _p(x0)))

    def run(self, max_iter, tol, true_residual=None, zero_x0=False) -> CGResult:
        if true_residual is None:
            true_residual = self.M.precision == _capi.MIXED
        flags = (_capi.SOLVE_TRUE_RESID if true_residual else 0) | \
            (_capi.SOLVE_ZERO_X0 if zero_x0 else 0)
        res = _capi.CGResultC()
        check(lib().iccg_solver_run(self._h.h, max_iter, tol, flags, C.byref(res)))
        return _result(res)

    def x(self):
        out = np.zeros(self.n)
        check(lib().iccg_solver_get_x(self._h.h, _p(out)))
        return out

    def set_timing(self, on=True):
        check(lib().iccg_solver_set_timing(self._h.h, int(on)))

    def set_graphs(self, on=True):
        check(lib().iccg_solver_set_graphs(self._h.h, int(on)))

    def reset_timing(self):
        check(lib().iccg_solver_reset_timing(self._h.h))

    def timing(self, cls):
        n = C.c_long(0)
        ms = C.c_double(0)
        check(lib().iccg_solver_get_timing(self._h.h, cls, C.byref(n), C.byref(ms)))
        return n.value, ms.value


class MultiSolver:
    """Several right-hand sides of one system solved together
    (iccg_msolver_*): each keeps its own CG recurrence and gets the iterates
    Solver gives it alone, bit for bit, while the kernels read each matrix
    and factor entry once for up to four of them.  b, x0 and x() are
    (nrhs, n) arrays."""

    def __init__(self, A, M, nrhs):
        self.A = _device_csr(A)
        self.M = M
        self.nrhs = int(nrhs)
        h = _vp()
        check(lib().iccg_msolver_create(self.A.handle, M.handle, self.nrhs, C.byref(h)))
        self._h = _Handle(h, lib().iccg_msolver_destroy)
        self.n = self.A.n

    def layout(self):
        """(shared-kernel sets, right-hand sides on the single-system solver)"""
        s, l = C.c_int(0), C.c_int(0)
        check(lib().iccg_msolver_native(self._h.h, C.byref(s), C.byref(l)))
        return s.value, l.value

    def _rhs(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.nrhs, self.n):
            raise ValueError(f"expected shape {(self.nrhs, self.n)}, got {a.shape}")
        return a

    def set(self, b, x0=None):
        b = self._rhs(b)
        x0 = None if x0 is None else self._rhs(x0)
        check(lib().iccg_msolver_set(self._h.h, _p(b), _p(x0)))

    def run(self, max_iter, tol, true_residual=None, zero_x0=False):
        """Returns one CGResult per right-hand side."""
        if true_residual is None:
            true_residual = self.M.precision == _capi.MIXED
        flags = (_capi.SOLVE_TRUE_RESID if true_residual else 0) | \
            (_capi.SOLVE_ZERO_X0 if zero_x0 else 0)
        res = (_capi.CGResultC * self.nrhs)()
        check(lib().iccg_msolver_run(self._h.h, max_iter, tol, flags, res))
        return [_result(r) for r in res]

    def x(self):
        out = np.zeros((self.nrhs, self.n))
        check(lib().iccg_msolver_get_x(self._h.h, _p(out)))
        return out

    def set_timing(self, on=True):
        check(lib().iccg_msolver_set_timing(self._h.h, int(on)))

    def reset_timing(self):
        check(lib().iccg_msolver_reset_timing(self._h.h))

    def timing(self, cls):
        n = C.c_long(0)
        ms = C.c_double(0)
        check(lib().iccg_msolver_get_timing(self._h.h, cls, C.byref(n), C.byref(ms)))
        return n.value, ms.value


def conjugate_gradient_multi(A, B, X, M, max_iter, tol, *, true_residual=None):
    """conjugate_gradient for the rows of B (nrhs, n) at once
    (iccg_cg_solve_multi); X (nrhs, n) float64 is updated in place."""
    dA = _device_csr(A)
    B = np.ascontiguousarray(B, dtype=np.float64)
    if (not isinstance(X, np.ndarray) or X.dtype != np.float64 or not X.flags.c_contiguous
            or X.shape != B.shape or B.ndim != 2):
        raise TypeError("X must be a contiguous float64 (nrhs, n) array shaped like B")
    if true_residual is None:
        true_residual = M.precision == _capi.MIXED
    flags = _capi.SOLVE_TRUE_RESID if true_residual else 0
    res = (_capi.CGResultC * B.shape[0])()
    check(lib().iccg_cg_solve_multi(dA.handle, M.handle, B.shape[0], _p(B), _p(X), max_iter,
                                    tol, flags, res))
    return [_result(r) for r in res]


def device_count():
    c = C.c_int(0)
    check(lib().iccg_device_count(C.byref(c)))
    return c.value


def device_trim():
    """Hands the device memory and pinned host staging that the library
    keeps idle for reuse (from objects already freed) back to the driver
    (iccg_device_trim); live objects keep theirs."""
    check(lib().iccg_device_trim())


def bench_copy(nbytes, reps):
    ms = C.c_double(0)
    check(lib().iccg_bench_copy(nbytes, reps, C.byref(ms)))
    return ms.value


class Batch:
    """Many independent systems advanced together (iccg_batch_*): every CG
    phase is one launch over all running systems; systems may be added while
    others run.  Each system needs an IC0 preconditioner built with a
    contiguous multicolor schedule (e.g. from greedy_color) or an ABMC
    schedule of contiguous blocks of at most 64 rows."""

    def __init__(self, capacity):
        h = _vp()
        check(lib().iccg_batch_create(int(capacity), C.byref(h)))
        self._h = _Handle(h, lib().iccg_batch_destroy)
        self._keep = {}

    def add(self, A, M, b, x0=None, max_iter=10000, tol=1e-8, true_residual=None):
        dA = _device_csr(A)
        if true_residual is None:
            true_residual = M.precision == _capi.MIXED
        b = _f64(b)
        x0 = None if x0 is None else _f64(x0)
        slot = C.c_int(-1)
        check(lib().iccg_batch_add(self._h.h, dA.handle, M.handle, _p(b), _p(x0), max_iter, tol,
                                   _capi.SOLVE_TRUE_RESID if true_residual else 0,
                                   C.byref(slot)))
        self._keep[slot.value] = (dA, M)
        return slot.value

    def run(self, max_steps=1 << 30):
        n = C.c_int(0)
        check(lib().iccg_batch_run(self._h.h, int(max_steps), C.byref(n)))
        return n.value

    def result(self, slot):
        fin = C.c_int(0)
        res = _capi.CGResultC()
        check(lib().iccg_batch_result(self._h.h, slot, C.byref(fin), C.byref(res)))
        return bool(fin.value), _result(res)

    def x(self, slot):
        n = self._keep[slot][0].n
        out = np.zeros(n)
        check(lib().iccg_batch_get_x(self._h.h, slot, _p(out)))
        return out

    def remove(self, slot):
        check(lib().iccg_batch_remove(self._h.h, slot))
        self._keep.pop(slot, None)

    def set_graphs(self, on=True):
        check(lib().iccg_batch_set_graphs(self._h.h, int(on)))

    def set_autotune(self, on=True):
        check(lib().iccg_batch_set_autotune(self._h.h, int(on)))

    def set_max_group(self, k):
        check(lib().iccg_batch_set_max_group(self._h.h, int(k)))

    def set_timing(self, on=True):
        check(lib().iccg_batch_set_timing(self._h.h, int(on)))

    def reset_timing(self):
        check(lib().iccg_batch_reset_timing(self._h.h))

    def timing(self, cls):
        """(launch groups, total ms) of kernel class cls: 0 SpMV, 1 forward
        sweeps, 2 backward sweeps, 3 vector updates and finish kernels."""
        n, ms = C.c_long(0), C.c_double(0)
        check(lib().iccg_batch_get_timing(self._h.h, cls, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    @property
    def tuning(self):
        """(64-row slices per workgroup of the SpMV/vector kernels, slices or
        ABMC blocks per wavefront of the sweeps)"""
        a, b = C.c_int(0), C.c_int(0)
        check(lib().iccg_batch_get_tuning(self._h.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    @tuning.setter
    def tuning(self, v):
        check(lib().iccg_batch_set_tuning(self._h.h, int(v[0]), int(v[1])))


class Server:
    """Request loop over one GPU (iccg_server_*).

    mode "batched": requests are admitted by HBM budget, prepared (ABMC
    ordering, P^T A P, IC(0)) on `workers` host threads and solved with
    continuous batching, every right-hand side a slot of one Batch.
    Requests of at least `multi_rows` rows (multi_rows=0: none) are solved
    instead by `multi_runners` threads with the multi-right-hand-side solver
    (one right-hand side: the single-system solver), beside the batch.
    mode "sequential": the plain one-at-a-time path (prepare, then solve each
    right-hand side alone) for comparison.  submit() returns at once; wait()
    and result() collect finished requests."""

    def __init__(self, mode="batched", precision="mixed", max_slots=256, workers=8,
                 block_size=64, true_residual=None, hbm_budget=64 << 30,
                 multi_rows=50000, multi_runners=6):
        prec = {"fp64": _capi.FP64, "mixed": _capi.MIXED}[precision]
        if true_residual is None:
            true_residual = prec == _capi.MIXED
        cfg = _capi.ServerConfigC(
            {"batched": _capi.SERVE_BATCHED, "sequential": _capi.SERVE_SEQUENTIAL}[mode], prec,
            int(max_slots), int(workers), int(block_size),
            _capi.SOLVE_TRUE_RESID if true_residual else 0, int(hbm_budget), int(multi_rows),
            int(multi_runners))
        h = _vp()
        check(lib().iccg_server_create(C.byref(cfg), C.byref(h)))
        self._h = _Handle(h, lib().iccg_server_destroy)
        self._shape = {}

    def submit(self, A: CRS, b, grid=None, max_iter=10000, tol=1e-8):
        b = np.ascontiguousarray(np.atleast_2d(np.asarray(b, np.float64)))
        if b.shape[1] != A.n:
            raise ValueError("b must be (n,) or (nrhs, n)")
        g = None if grid is None else np.asarray(list(grid) + [1] * (3 - len(grid)), np.int32)
        rid = C.c_longlong(0)
        rp, ci, v = _i32(A.rowptr), _i32(A.colind), _f64(A.val)
        check(lib().iccg_server_submit(self._h.h, A.n, len(ci), _p(rp), _p(ci), _p(v), _p(g),
                                       b.shape[0], _p(b), int(max_iter), float(tol),
                                       C.byref(rid)))
        self._shape[rid.value] = b.shape
        return rid.value

    def wait(self, rid=-1, timeout_ms=-1):
        """Request id once finished (rid < 0: any finished request not yet
        returned), or None on timeout."""
        out = C.c_longlong(-1)
        check(lib().iccg_server_wait(self._h.h, int(rid), int(timeout_ms), C.byref(out)))
        return None if out.value < 0 else out.value

    def result(self, rid):
        """(x of shape (nrhs, n), [CGResult], latency_ms, prepare_ms)"""
        shape = self._shape[rid]
        x = np.zeros(shape)
        res = (_capi.CGResultC * shape[0])()
        lat, prep = C.c_double(0), C.c_double(0)
        check(lib().iccg_server_result(self._h.h, rid, _p(x), C.cast(res, _vp), C.byref(lat),
                                       C.byref(prep)))
        return x, [_result(r) for r in res], lat.value, prep.value

    def release(self, rid):
        check(lib().iccg_server_release(self._h.h, rid))
        self._shape.pop(rid, None)

    def stats(self):
        st = _capi.ServerStatsC()
        check(lib().iccg_server_get_stats(self._h.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        self._h = None
