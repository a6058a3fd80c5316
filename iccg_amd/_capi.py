"""ctypes declarations of libiccg_hip.so (include/iccg_hip.h).

The library is loaded from iccg_amd/lib/ (built in-tree by
`make -C iccg_amd/csrc`).  A missing library is an error: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libiccg_hip.so")

ICCG_OK = 0
ICCG_ERR_ARG, ICCG_ERR_IO, ICCG_ERR_PIVOT, ICCG_ERR_SCHEDULE = 1, 2, 3, 4
ICCG_ERR_DEVICE, ICCG_ERR_NOMEM, ICCG_ERR_STATE = 5, 6, 7
PRECOND_IDENTITY, PRECOND_JACOBI, PRECOND_IC0 = 0, 1, 2
PLAN_NAMES = {0: "none", 1: "levels", 2: "csr", 3: "mc-sliced", 4: "abmc-sliced",
              5: "abmc-sectioned", 6: "abmc-sectioned-fused", 7: "abmc-sectioned-fused-r",
              8: "mc-sliced-r", 9: "abmc-blocks"}
FP64, MIXED = 0, 1
SCHED_MC, SCHED_ABMC = 0, 1
SOLVE_TRUE_RESID = 1
SOLVE_ZERO_X0 = 2

_vp = C.c_void_p
_i = C.c_int
_l = C.c_long
_d = C.c_double
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_dp = C.POINTER(C.c_double)


class CGResultC(C.Structure):
    _fields_ = [("iters", C.c_int), ("rel_resid", C.c_double), ("converged", C.c_int),
                ("breakdown", C.c_int), ("true_rel_resid", C.c_double),
                ("restarts", C.c_int), ("solve_ms", C.c_double)]


SERVE_BATCHED, SERVE_SEQUENTIAL = 0, 1


class ServerConfigC(C.Structure):
    _fields_ = [("mode", C.c_int), ("precision", C.c_int), ("max_slots", C.c_int),
                ("workers", C.c_int), ("block_size", C.c_int), ("flags", C.c_int),
                ("hbm_budget", C.c_longlong), ("multi_rows", C.c_int),
                ("multi_runners", C.c_int)]


class ServerStatsC(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("finished", C.c_longlong),
                ("failed", C.c_longlong), ("hbm_in_use", C.c_longlong), ("waiting", C.c_int),
                ("preparing", C.c_int), ("ready", C.c_int), ("running", C.c_int),
                ("engine_run_ms", C.c_double), ("engine_admit_ms", C.c_double),
                ("engine_harvest_ms", C.c_double)]


_llp = C.POINTER(C.c_longlong)

# name -> (restype, argtypes)
_SIGS = {
    "iccg_last_error": (C.c_char_p, []),
    "iccg_status_string": (C.c_char_p, [_i]),
    "iccg_device_count": (_i, [_ip]),
    "iccg_device_set": (_i, [_i]),
    "iccg_device_trim": (_i, []),
    "iccg_mm_read": (_i, [C.c_char_p, _ip, _lp, C.POINTER(_ip), C.POINTER(_ip), C.POINTER(_dp)]),
    "iccg_free_host": (None, [_vp]),
    "iccg_coo_to_crs": (_i, [_i, _l, _vp, _vp, _vp, _vp, _vp, _vp, _lp]),
    "iccg_expand_lower_to_full": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _lp]),
    "iccg_read_color_file": (_i, [C.c_char_p, _i, _vp, _ip]),
    "iccg_read_block_file": (_i, [C.c_char_p, _i, _vp, _ip]),
    "iccg_read_block_color_file": (_i, [C.c_char_p, _i, _vp, _ip]),
    "iccg_write_color_file": (_i, [C.c_char_p, _i, _vp, _i]),
    "iccg_write_block_file": (_i, [C.c_char_p, _i, _vp, _i]),
    "iccg_write_block_color_file": (_i, [C.c_char_p, _i, _vp, _i]),
    "iccg_perm_by_color": (_i, [_i, _vp, _vp]),
    "iccg_perm_by_block_color": (_i, [_i, _vp, _i, _vp, _vp]),
    "iccg_ptap_lower": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _lp]),
    "iccg_color_ptr_rows": (_i, [_i, _vp, _i, _vp, _vp, _vp]),
    "iccg_abmc_schedule": (_i, [_i, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp]),
    "iccg_check_interblock": (_i, [_i, _vp, _vp, _vp, _i, _vp, _i, _vp, C.POINTER(C.c_longlong)]),
    "iccg_check_hazards": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _i, _ip, _vp, _vp, _vp,
                                C.POINTER(C.c_longlong)]),
    "iccg_greedy_color": (_i, [_i, _vp, _vp, _vp, _ip]),
    "iccg_abmc_blocking": (_i, [_i, _vp, _vp, _i, _vp, _ip, _vp, _ip]),
    "iccg_abmc_blocking_grid": (_i, [_i, _i, _i, _i, _i, _i, _vp, _ip, _vp, _ip]),
    "iccg_gen_grid": (_i, [_i, _i, _i, _d, _d, _d, C.c_ulonglong, _ip, _lp, _vp, _vp, _vp]),
    "iccg_gen_normal": (_i, [_l, C.c_ulonglong, _vp]),
    "iccg_csr_create": (_i, [_i, _l, _vp, _vp, _vp, C.POINTER(_vp)]),
    "iccg_csr_destroy": (_i, [_vp]),
    "iccg_spmv": (_i, [_vp, _vp, _vp]),
    "iccg_schedule_create_mc": (_i, [_i, _i, _vp, _vp, C.POINTER(_vp)]),
    "iccg_schedule_create_abmc": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "iccg_schedule_destroy": (_i, [_vp]),
    "iccg_precond_create": (_i, [_vp, _i, _d, _vp, _i, C.POINTER(_vp)]),
    "iccg_precond_destroy": (_i, [_vp]),
    "iccg_precond_set_schedule": (_i, [_vp, _vp]),
    "iccg_precond_apply": (_i, [_vp, _vp, _vp]),
    "iccg_precond_factor_nnz": (_i, [_vp, _lp]),
    "iccg_precond_get_factor": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "iccg_precond_get_plan": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "iccg_cg_solve": (_i, [_vp, _vp, _vp, _vp, _i, _d, _i, C.POINTER(CGResultC)]),
    "iccg_cg_solve_noalloc": (_i, [_vp, _vp, _vp, _vp, _i, _d, _i, C.POINTER(CGResultC)]),
    "iccg_solver_create": (_i, [_vp, _vp, C.POINTER(_vp)]),
    "iccg_solver_destroy": (_i, [_vp]),
    "iccg_solver_set": (_i, [_vp, _vp, _vp]),
    "iccg_solver_run": (_i, [_vp, _i, _d, _i, C.POINTER(CGResultC)]),
    "iccg_solver_get_x": (_i, [_vp, _vp]),
    "iccg_solver_set_timing": (_i, [_vp, _i]),
    "iccg_solver_get_timing": (_i, [_vp, _i, _lp, _dp]),
    "iccg_solver_set_graphs": (_i, [_vp, _i]),
    "iccg_solver_set_blocking": (_i, [_vp, _i]),
    "iccg_batch_set_blocking": (_i, [_vp, _i]),
    "iccg_solver_reset_timing": (_i, [_vp]),
    "iccg_msolver_create": (_i, [_vp, _vp, _i, C.POINTER(_vp)]),
    "iccg_msolver_destroy": (_i, [_vp]),
    "iccg_msolver_native": (_i, [_vp, _ip, _ip]),
    "iccg_msolver_set": (_i, [_vp, _vp, _vp]),
    "iccg_msolver_run": (_i, [_vp, _i, _d, _i, _vp]),
    "iccg_msolver_get_x": (_i, [_vp, _vp]),
    "iccg_msolver_set_timing": (_i, [_vp, _i]),
    "iccg_msolver_get_timing": (_i, [_vp, _i, _lp, _dp]),
    "iccg_msolver_reset_timing": (_i, [_vp]),
    "iccg_cg_solve_multi": (_i, [_vp, _vp, _i, _vp, _vp, _i, _d, _i, _vp]),
    "iccg_bench_copy": (_i, [_l, _i, _dp]),
    "iccg_batch_create": (_i, [_i, C.POINTER(_vp)]),
    "iccg_batch_destroy": (_i, [_vp]),
    "iccg_batch_add": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _d, _i, _ip]),
    "iccg_batch_run": (_i, [_vp, _i, _ip]),
    "iccg_batch_result": (_i, [_vp, _i, _ip, C.POINTER(CGResultC)]),
    "iccg_batch_get_x": (_i, [_vp, _i, _vp]),
    "iccg_batch_remove": (_i, [_vp, _i]),
    "iccg_batch_set_graphs": (_i, [_vp, _i]),
    "iccg_batch_set_autotune": (_i, [_vp, _i]),
    "iccg_batch_get_tuning": (_i, [_vp, _ip, _ip]),
    "iccg_batch_set_tuning": (_i, [_vp, _i, _i]),
    "iccg_batch_trim": (_i, [_vp]),
    "iccg_batch_set_max_group": (_i, [_vp, _i]),
    "iccg_batch_set_timing": (_i, [_vp, _i]),
    "iccg_batch_get_timing": (_i, [_vp, _i, _lp, _dp]),
    "iccg_batch_reset_timing": (_i, [_vp]),
    "iccg_device_get": (_i, [_ip]),
    "iccg_server_create": (_i, [C.POINTER(ServerConfigC), C.POINTER(_vp)]),
    "iccg_server_destroy": (_i, [_vp]),
    "iccg_server_submit": (_i, [_vp, _i, _l, _vp, _vp, _vp, _vp, _i, _vp, _i, _d, _llp]),
    "iccg_server_wait": (_i, [_vp, C.c_longlong, _i, _llp]),
    "iccg_server_result": (_i, [_vp, C.c_longlong, _vp, _vp, _dp, _dp]),
    "iccg_server_release": (_i, [_vp, C.c_longlong]),
    "iccg_server_get_stats": (_i, [_vp, C.POINTER(ServerStatsC)]),
}

_lib = None


class IccgError(RuntimeError):
    """Error raised for a non-zero status; .status holds the ICCG_* code."""

    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `make -C iccg_amd/csrc` (or "
                "__graft_entry__.build()); the HIP library is required")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(status):
    if status != ICCG_OK:
        msg = lib().iccg_last_error().decode(errors="replace")
        raise IccgError(status, msg)
