"""ctypes binding of libuegccd.so (the C ABI in include/uegccd.h).

The library is the only compute path: if it is missing or no HIP device is
usable, every call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libuegccd.so")

OK = 0
ERR_INVALID, ERR_ALLOC, ERR_HIP, ERR_UNSUPPORTED = -1, -2, -3, -4
ERR_NOT_CONVERGED, ERR_NONFINITE, ERR_BUSY, ERR_NOT_FOUND = -5, -6, -7, -8
ERR_SCRATCH = -9
NOALLOC_SCRATCH_MB = 64
MODE_REFERENCE, MODE_DIIS = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

# Every symbol include/uegccd.h declares.
EXPORTS = (
    "uegccd_default_options", "uegccd_ctx_create", "uegccd_ctx_destroy", "uegccd_ctx_info",
    "uegccd_residual", "uegccd_energy", "uegccd_solve", "uegccd_trace", "uegccd_mp2",
    "uegccd_heg_build", "uegccd_strerror", "uegccd_solve_batch", "uegccd_solve_resident",
    "uegccd_get_amplitudes", "uegccd_kernel_time", "uegccd_footprint",
    "uegccd_server_create", "uegccd_server_destroy", "uegccd_server_submit",
    "uegccd_server_step", "uegccd_server_drain", "uegccd_server_poll",
    "uegccd_server_clock_ms", "uegccd_time_kernel", "uegccd_kernel_stats",
    "uegccd_heg_build_device", "uegccd_connectivity", "uegccd_drpa",
    "uegccd_structure_factor", "uegccd_solve_noalloc", "uegccd_mp2_structure_factor",
    "uegccd_exchange_structure_factor", "uegccd_set_tiles", "uegccd_get_tiles",
)
REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2


class System(ctypes.Structure):
    _fields_ = [("nocc", ctypes.c_int32), ("nao", ctypes.c_int32),
                ("nvec", ctypes.POINTER(ctypes.c_int32)), ("eigen", ctypes.POINTER(ctypes.c_double)),
                ("L", ctypes.c_double), ("madelung", ctypes.c_double), ("rs", ctypes.c_double),
                ("gapii", ctypes.c_double), ("coren", ctypes.c_int32), ("safe_eri", ctypes.c_int32),
                ("irange", ctypes.c_int32 * 15), ("do_ring", ctypes.c_int32),
                ("do_xring", ctypes.c_int32), ("do_ladder", ctypes.c_int32),
                ("do_mosaic", ctypes.c_int32), ("screening", ctypes.c_double)]


class Options(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("ndiis", ctypes.c_int32),
                ("diis_start", ctypes.c_int32), ("max_iter", ctypes.c_int32),
                ("tol_dt", ctypes.c_double), ("tol_e", ctypes.c_double), ("tol_r", ctypes.c_double),
                ("damping", ctypes.c_double), ("precision", ctypes.c_int32),
                ("mixed_switch", ctypes.c_double), ("device", ctypes.c_int32),
                ("iters_per_sync", ctypes.c_int32), ("graph", ctypes.c_int32),
                ("persistent", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("ecorr", ctypes.c_double), ("ehf_plus_ec", ctypes.c_double),
                ("dt", ctypes.c_double), ("maxres", ctypes.c_double), ("niter", ctypes.c_int32),
                ("failed", ctypes.c_int32), ("iters_fp32", ctypes.c_int32),
                ("status", ctypes.c_int32), ("solve_ms", ctypes.c_double)]


class Tiles(ctypes.Structure):
    _fields_ = [("ladder_group", ctypes.c_int32), ("quad_steps", ctypes.c_int32),
                ("hh_steps", ctypes.c_int32)]


class HegParams(ctypes.Structure):
    _fields_ = [("nel", ctypes.c_int32), ("maxkpoint", ctypes.c_int32), ("rs", ctypes.c_double),
                ("twist", ctypes.c_double * 3), ("fcut", ctypes.c_double),
                ("corefcut", ctypes.c_double), ("madelungfactor", ctypes.c_double),
                ("gap", ctypes.c_double), ("exchangefactor", ctypes.c_double),
                ("cutoff_twisted", ctypes.c_int32)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("hbm_budget_bytes", ctypes.c_int64),
                ("max_active", ctypes.c_int32), ("iters_per_step", ctypes.c_int32),
                ("planner_threads", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("heg", HegParams), ("gapii", ctypes.c_double), ("coren", ctypes.c_int32),
                ("safe_eri", ctypes.c_int32), ("irange", ctypes.c_int32 * 15),
                ("do_ring", ctypes.c_int32), ("do_xring", ctypes.c_int32),
                ("do_ladder", ctypes.c_int32), ("do_mosaic", ctypes.c_int32),
                ("opt", Options)]


class RequestResult(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32),
                ("nao", ctypes.c_int32), ("nocc", ctypes.c_int32),
                ("niter", ctypes.c_int32), ("iters_fp32", ctypes.c_int32),
                ("ehf", ctypes.c_double), ("e2", ctypes.c_double), ("ecorr", ctypes.c_double),
                ("maxres", ctypes.c_double), ("t_submit_ms", ctypes.c_double),
                ("t_start_ms", ctypes.c_double), ("t_done_ms", ctypes.c_double),
                ("footprint_bytes", ctypes.c_int64), ("plan_ms", ctypes.c_double)]


class UegccdError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = strerror(status) if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


_LIB = None


def load():
    """Load libuegccd.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libuegccd.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    P, D = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    I32 = ctypes.POINTER(ctypes.c_int32)
    lib.uegccd_default_options.argtypes = [ctypes.POINTER(Options)]
    lib.uegccd_default_options.restype = None
    lib.uegccd_ctx_create.argtypes = [ctypes.POINTER(System), ctypes.POINTER(Options),
                                      ctypes.POINTER(P)]
    lib.uegccd_ctx_destroy.argtypes = [P]
    lib.uegccd_ctx_destroy.restype = None
    lib.uegccd_ctx_info.argtypes = [P, I32, I32, ctypes.POINTER(ctypes.c_int64), I32]
    lib.uegccd_residual.argtypes = [P, D, D, D, D, D, D]
    lib.uegccd_energy.argtypes = [P, D, D, D]
    lib.uegccd_solve.argtypes = [P, D, D, D, ctypes.POINTER(Result)]
    lib.uegccd_solve_noalloc.argtypes = [ctypes.POINTER(System), ctypes.POINTER(Options), D, D, D,
                                         ctypes.POINTER(Result)]
    lib.uegccd_trace.argtypes = [P, D, D, ctypes.c_int32]
    lib.uegccd_mp2.argtypes = [P, D, D, D, D]
    lib.uegccd_drpa.argtypes = [P, D, ctypes.POINTER(Result), D]
    I32 = ctypes.POINTER(ctypes.c_int32)
    lib.uegccd_structure_factor.argtypes = [P, D, D, I32, I32, D, D]
    lib.uegccd_mp2_structure_factor.argtypes = [P, I32, I32, D, D]
    lib.uegccd_exchange_structure_factor.argtypes = [ctypes.POINTER(HegParams), I32, I32, I32, D]
    lib.uegccd_heg_build.argtypes = [ctypes.POINTER(HegParams), I32, I32, D, D]
    lib.uegccd_heg_build_device.argtypes = [ctypes.POINTER(HegParams), ctypes.c_int32, I32, I32, D, D]
    lib.uegccd_strerror.argtypes = [ctypes.c_int]
    lib.uegccd_strerror.restype = ctypes.c_char_p
    lib.uegccd_solve_batch.argtypes = [ctypes.POINTER(P), ctypes.c_int32, ctypes.POINTER(D),
                                       ctypes.POINTER(D), ctypes.POINTER(Result)]
    lib.uegccd_solve_resident.argtypes = [P, ctypes.c_int32, ctypes.POINTER(Result)]
    lib.uegccd_get_amplitudes.argtypes = [P, D, D, D]
    lib.uegccd_kernel_time.argtypes = [P, ctypes.c_int32, ctypes.c_int32, D]
    lib.uegccd_set_tiles.argtypes = [P, ctypes.POINTER(Tiles)]
    lib.uegccd_get_tiles.argtypes = [P, ctypes.POINTER(Tiles), I32, I32]
    I64 = ctypes.POINTER(ctypes.c_int64)
    lib.uegccd_footprint.argtypes = [ctypes.POINTER(System), ctypes.POINTER(Options), I64]
    lib.uegccd_server_create.argtypes = [ctypes.POINTER(ServerOptions), ctypes.POINTER(P)]
    lib.uegccd_server_destroy.argtypes = [P]
    lib.uegccd_server_destroy.restype = None
    lib.uegccd_server_submit.argtypes = [P, ctypes.POINTER(Request), I64]
    lib.uegccd_server_step.argtypes = [P, I32, I32]
    lib.uegccd_server_drain.argtypes = [P]
    lib.uegccd_server_poll.argtypes = [P, ctypes.c_int64, ctypes.POINTER(RequestResult)]
    lib.uegccd_time_kernel.argtypes = [P, ctypes.c_int32]
    lib.uegccd_connectivity.argtypes = [P, I32, ctypes.c_int32]
    lib.uegccd_kernel_stats.argtypes = [P, D, I64]
    lib.uegccd_server_clock_ms.argtypes = [P]
    lib.uegccd_server_clock_ms.restype = ctypes.c_double
    for name in ("uegccd_ctx_create", "uegccd_ctx_info", "uegccd_residual", "uegccd_energy",
                 "uegccd_solve", "uegccd_trace", "uegccd_mp2", "uegccd_heg_build",
                 "uegccd_solve_batch", "uegccd_solve_resident", "uegccd_get_amplitudes",
                 "uegccd_kernel_time", "uegccd_footprint", "uegccd_server_create",
                 "uegccd_server_submit", "uegccd_server_step", "uegccd_server_drain",
                 "uegccd_server_poll", "uegccd_time_kernel", "uegccd_kernel_stats",
                 "uegccd_heg_build_device", "uegccd_connectivity", "uegccd_drpa",
                 "uegccd_structure_factor"):
        getattr(lib, name).restype = ctypes.c_int
    _LIB = lib
    return lib


def strerror(status: int) -> str:
    return load().uegccd_strerror(int(status)).decode()


def dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def check(status: int, what: str = ""):
    if status != OK:
        raise UegccdError(status, what)
