"""
ctypes binding of libflowperm.so (include/flowperm.h).

The library is built in-tree (``make -C flow-anizotropy_amd/csrc``) into
``flowperm/_lib/libflowperm.so``.  There is no fallback: if the library is
missing, or no HIP device is present, every solve raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libflowperm.so"

FP_OK = 0
FP_PENDING = 1
FP_NOT_CONVERGED = 3
FP_ERR_BAD_DIMS = 10
FP_ERR_NON_BINARY = 11
FP_ERR_NO_PORE = 12
FP_ERR_NONFINITE = 13
FP_ERR_OOM = 14
FP_ERR_DEVICE = 15
FP_ERR_ARG = 16
FP_ERR_NOT_FOUND = 17

FP_F64, FP_F32, FP_F32_REF2D = 0, 1, 2
PRECISIONS = {"f64": FP_F64, "fp64": FP_F64, "f32": FP_F32, "fp32": FP_F32,
              "f32ref": FP_F32_REF2D, "ref2d": FP_F32_REF2D}


class Config3D(C.Structure):
    _fields_ = [("ff", C.c_double), ("eps", C.c_double),
                ("lag", C.c_int64), ("min_steps", C.c_int64),
                ("max_steps", C.c_int64), ("avg_steps", C.c_int64),
                ("hold", C.c_int32), ("validate", C.c_int32),
                ("precision", C.c_int32), ("concurrent", C.c_int32)]


class SolveOut(C.Structure):
    _fields_ = [("q", C.c_double * 3), ("se", C.c_double * 3), ("steps", C.c_int64),
                ("converged", C.c_int32), ("status", C.c_int32), ("fieldres", C.c_double),
                ("check_step", C.c_int64), ("checks", C.c_int64), ("check_q", C.c_double * 3),
                ("check_ratio", C.c_double * 2), ("check_hold", C.c_int32), ("pad", C.c_int32)]


# void (*fp_progress_fn)(void *user, int run, const fp_solve_out *out)
PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(SolveOut))


class Result3D(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad0", C.c_int32),
                ("n_pore", C.c_int64),
                ("porosity", C.c_double), ("force", C.c_double), ("mu", C.c_double),
                ("run", SolveOut * 3),
                ("K", C.c_double * 9),
                ("k_xy", C.c_double), ("k_xz", C.c_double), ("k_yz", C.c_double),
                ("recip_xy", C.c_double), ("recip_xz", C.c_double), ("recip_yz", C.c_double),
                ("recip_resid", C.c_double),
                ("k", C.c_double * 3), ("k_mean", C.c_double), ("fa", C.c_double),
                ("v", C.c_double * 9),
                ("validated", C.c_int32), ("pad1", C.c_int32),
                ("val", SolveOut), ("val_pred", C.c_double * 3), ("val_err", C.c_double),
                ("seconds", C.c_double), ("fluid_updates", C.c_int64)]


class Config2D(C.Structure):
    _fields_ = [("ff", C.c_double), ("eps", C.c_double), ("qfloor", C.c_double),
                ("ksigma", C.c_double), ("L0", C.c_double), ("validate_deg", C.c_double),
                ("lag", C.c_int64), ("min_steps", C.c_int64), ("max_steps", C.c_int64),
                ("avg_steps", C.c_int64), ("avg_sample", C.c_int64),
                ("hold", C.c_int32), ("precision", C.c_int32), ("concurrent", C.c_int32),
                ("pad", C.c_int32)]


class Result2D(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32),
                ("porosity", C.c_double), ("force", C.c_double), ("mu", C.c_double),
                ("run", SolveOut * 2),
                ("k_xx", C.c_double), ("k_yx", C.c_double), ("k_xy", C.c_double),
                ("k_yy", C.c_double), ("k_off", C.c_double), ("recip_resid", C.c_double),
                ("k1", C.c_double), ("k2", C.c_double), ("k_mean", C.c_double),
                ("anisotropy", C.c_double), ("theta_deg", C.c_double),
                ("alpha_x_deg", C.c_double), ("alpha_y_deg", C.c_double),
                ("validated", C.c_int32), ("pad1", C.c_int32),
                ("val", SolveOut), ("val_pred", C.c_double * 2), ("val_err", C.c_double),
                ("seconds", C.c_double), ("fluid_updates", C.c_int64)]


class BenchOut(C.Structure):
    _fields_ = [("n_fluid", C.c_int64), ("n_cells", C.c_int64), ("steps", C.c_int64),
                ("total_ms", C.c_double), ("local_ms", C.c_double), ("nbr_ms", C.c_double),
                ("monitor_ms", C.c_double), ("bytes_per_update", C.c_double),
                ("eager_total_ms", C.c_double), ("deferred", C.c_int32), ("pad", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("local_cpt", C.c_int32), ("nbr_waves", C.c_int32), ("chunk_steps", C.c_int32),
                ("nontemporal", C.c_int32), ("local_bpc", C.c_int32), ("nbr_bpc", C.c_int32),
                ("nbr_mode", C.c_int32), ("nbr_segs", C.c_int32), ("local_wgs", C.c_int32),
                ("nbr_wgs", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("lattice", C.c_int32), ("precision", C.c_int32), ("budget_bytes", C.c_int64),
                ("chunk_steps", C.c_int32), ("max_bypass", C.c_int32)]


class RequestInfo(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("device_bytes", C.c_int64), ("fluid_updates", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in
                ("submitted", "admitted", "completed", "rejected", "queued", "resident",
                 "live_instances", "chunks", "steps", "fluid_updates", "bytes_in_use",
                 "bytes_peak", "budget_bytes")] + [("busy_seconds", C.c_double)]


# every symbol include/flowperm.h declares, with its ctypes signature
_u8p = C.POINTER(C.c_uint8)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
SIGNATURES = {
    "fp_exit_code": (C.c_int, [C.c_int]),
    "fp_abi_version": (C.c_int, []),
    "fp_set_progress": (C.c_int, [C.c_void_p, PROGRESS_FN, C.c_void_p]),
    "fp_create": (C.c_void_p, [C.c_int, _ip]),
    "fp_destroy": (None, [C.c_void_p]),
    "fp_last_error": (C.c_char_p, [C.c_void_p]),
    "fp_config3d_default": (None, [C.POINTER(Config3D)]),
    "fp_check_mask3d": (C.c_int, [_u8p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    "fp_solve3d": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int,
                             C.POINTER(Config3D), C.POINTER(Result3D), _fp]),
    "fp_solve_dir3d": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(Config3D), _dp, C.POINTER(SolveOut), _fp]),
    "fp_jacobi3": (None, [_dp, _dp, _dp]),
    "fp_verify_links": (C.c_int, [_u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    "fp_config2d_default": (None, [C.POINTER(Config2D)]),
    "fp_solve2d": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int,
                             C.POINTER(Config2D), C.POINTER(Result2D), _fp]),
    "fp_solve_dir2d": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int,
                                 C.POINTER(Config2D), _dp, C.POINTER(SolveOut), _fp]),
    "fp_batch_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, _ip]),
    "fp_batch_destroy": (None, [C.c_void_p]),
    "fp_batch_add3d": (C.c_int64, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(Config3D), _dp]),
    "fp_batch_add2d": (C.c_int64, [C.c_void_p, _u8p, C.c_int, C.c_int,
                                   C.POINTER(Config2D), _dp]),
    "fp_batch_run": (C.c_int, [C.c_void_p]),
    "fp_batch_result": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(SolveOut)]),
    "fp_batch_stats": (None, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                              C.POINTER(C.c_double)]),
    "fp_batch_set_chunk": (None, [C.c_void_p, C.c_int]),
    "fp_bench_steps": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 _dp, C.c_int, C.c_int, C.POINTER(Tuning), C.c_int,
                                 C.POINTER(BenchOut)]),
    "fp_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "fp_autotune": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_int, C.POINTER(Tuning), _dp, _dp]),
    "fp_tune_shape": (None, [C.c_int, C.c_int, _ip, _ip, _ip]),
    "fp_get_tuning": (None, [C.c_void_p, C.POINTER(Tuning)]),
    "fp_bench_copy": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, _dp]),
    "fp_now": (C.c_double, []),
    "fp_server_create": (C.c_void_p, [C.c_void_p, C.POINTER(ServerConfig), _ip]),
    "fp_server_destroy": (None, [C.c_void_p]),
    "fp_server_submit3d": (C.c_int64, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(Config3D)]),
    "fp_server_submit2d": (C.c_int64, [C.c_void_p, _u8p, C.c_int, C.c_int, C.POINTER(Config2D)]),
    "fp_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "fp_server_poll3d": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Result3D),
                                   C.POINTER(RequestInfo)]),
    "fp_server_poll2d": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Result2D),
                                   C.POINTER(RequestInfo)]),
    "fp_server_get_stats": (None, [C.c_void_p, C.POINTER(ServerStats)]),
    "fp_replay_convergence3d": (C.c_int, [C.POINTER(Config3D), _dp, C.c_int64, C.POINTER(SolveOut)]),
    "fp_replay_convergence2d": (C.c_int, [C.POINTER(Config2D), _dp, C.c_int64, C.POINTER(SolveOut)]),
    "fp_field_residual2d": (C.c_double, [C.c_int, C.c_int, _fp, _fp]),
    "fp_test_fail_rebuilds": (None, [C.c_int]),
    "fp_solve_dir3d_noalloc": (C.c_int, [C.c_void_p, _u8p, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(Config3D), _dp, C.POINTER(SolveOut)]),
    "fp_verify_batch_tables": (C.c_int, [C.POINTER(_u8p), C.POINTER(C.c_int32), C.c_int, C.c_int,
                                         C.c_int, C.POINTER(Tuning), C.POINTER(C.c_int64)]),
    "fp_verify_finish_plan": (C.c_int, [C.POINTER(Config3D), C.POINTER(Config2D), C.c_int, _dp,
                                        C.c_int64, C.POINTER(C.c_int64), C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
}

_lib = None


class FlowpermError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"{message} (status {status})" if message else f"status {status}")


def load(path=None):
    """Load libflowperm.so (once) and attach the declared signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FlowpermError(FP_ERR_DEVICE,
                            f"{p} is not built: run `make -C flow-anizotropy_amd/csrc` "
                            f"(or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def u8ptr(arr):
    return arr.ctypes.data_as(_u8p)


def f32ptr(arr):
    return None if arr is None else arr.ctypes.data_as(_fp)


def f64ptr(arr):
    return arr.ctypes.data_as(_dp)


def exit_code(status):
    return 0 if status == FP_OK else (3 if status == FP_NOT_CONVERGED else 2)


def device_index():
    return int(os.environ.get("FLOWPERM_DEVICE", "0"))
