"""ctypes binding of include/gexs_hip.h (libgexs_hip.so, built in-tree).

There is no fallback: if the HIP library is missing or does not load, every
entry point raises.  The library is located next to this file.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgexs_hip.so")

# profile ids (common/Input.h:10-17)
GAUSSIAN, DOPPLER, VOIGT, VOIGT_QUAD = 0, 1, 2, 4
PROFILE_IDS = {"gaussian": GAUSSIAN, "doppler": DOPPLER, "voigt": VOIGT, "voigt_quad": VOIGT_QUAD}
FP32C, FP64 = 0, 1

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESTATE", -5: "EBUSY", -6: "ENOTFOUND"}

# exported symbols, in the order include/gexs_hip.h declares them
EXPORTS = [
    "gexs_create", "gexs_destroy", "gexs_init_vectors", "gexs_init_constants",
    "gexs_transfer_freq", "gexs_transfer_lines", "gexs_transfer_lines_device", "gexs_execute",
    "gexs_transfer_results", "gexs_ready", "gexs_sync", "gexs_last_error", "gexs_last_stats",
    "gexs_timing", "gexs_line_windows", "gexs_pass_lines", "gexs_pass_items", "gexs_set_tuning", "gexs_load_list", "gexs_line_count",
    "gexs_partition", "gexs_run_batch", "gexs_batch_device_out", "gexs_stream_xsec",
    "gexs_set_launch", "gexs_autotune", "gexs_run_fixed",
    "gexs_serve_start", "gexs_submit", "gexs_poll", "gexs_wait", "gexs_release",
    "gexs_serve_stop",
]


class GexsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Constants(C.Structure):
    _fields_ = [("half_width", C.c_double), ("temperature", C.c_double),
                ("partition", C.c_double), ("dfreq", C.c_double), ("mean_mass", C.c_double),
                ("pressure", C.c_double), ("ref_temp", C.c_double), ("ref_press", C.c_double),
                ("cutoff", C.c_double), ("gauss_hwhm", C.c_double),
                ("width_conv", C.c_int32), ("pressure_mode", C.c_int32),
                ("precision", C.c_int32), ("reserved", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("n_eval", C.c_int64), ("n_lines", C.c_int64), ("n_tiles", C.c_int64),
                ("kernel_ms", C.c_double), ("total_ms", C.c_double)]


class States(C.Structure):
    _fields_ = [("n", C.c_int64), ("E", C.c_void_p), ("g", C.c_void_p)]


class Transitions(C.Structure):
    _fields_ = [("n", C.c_int64), ("upper", C.c_void_p), ("lower", C.c_void_p),
                ("A", C.c_void_p), ("gamma", C.c_void_p), ("nexp", C.c_void_p),
                ("A64", C.c_void_p), ("gamma64", C.c_void_p), ("nexp64", C.c_void_p)]


class Grid(C.Structure):
    _fields_ = [("start", C.c_double), ("end", C.c_double), ("npoints", C.c_int32),
                ("reserved", C.c_int32)]


class TP(C.Structure):
    _fields_ = [("T", C.c_double), ("P", C.c_double), ("Q", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("points_per_lane", C.c_int32), ("order", C.c_int32),
                ("persistent_blocks", C.c_int32), ("waves_per_tile", C.c_int32),
                ("kernel_ms", C.c_double)]


class StreamStats(C.Structure):
    _fields_ = [("n_transitions", C.c_int64), ("n_lines", C.c_int64), ("n_eval", C.c_int64),
                ("chunk", C.c_int64), ("device_bytes", C.c_int64), ("n_chunks", C.c_int32),
                ("reserved", C.c_int32), ("wall_ms", C.c_double), ("device_ms", C.c_double),
                ("stage_ms", C.c_double), ("copy_ms", C.c_double), ("sort_ms", C.c_double),
                ("pass_ms", C.c_double)]


class LinesDev(C.Structure):
    _fields_ = [("n", C.c_int64), ("nu", C.c_void_p), ("E", C.c_void_p), ("A", C.c_void_p),
                ("g", C.c_void_p), ("gamma", C.c_void_p), ("nexp", C.c_void_p)]


FIXED_MAX_LINES = 1 << 20
FIXED_MAX_TILES = 1 << 17


class Request(C.Structure):
    _fields_ = [("T", C.c_double), ("P", C.c_double), ("start", C.c_double),
                ("end", C.c_double), ("dfreq", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("npoints", C.c_int32), ("sigma", C.POINTER(C.c_double)),
                ("latency_ms", C.c_double), ("device_ms", C.c_double),
                ("tile_points", C.c_int32), ("waves_per_tile", C.c_int32),
                ("line_groups", C.c_int32), ("reserved", C.c_int32), ("n_eval", C.c_int64)]


_lib = None


def load():
    """The loaded library; raises when it is absent (never falls back)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m gexs_amd.build` "
                          "(the HIP path has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i32, i64, dbl = C.c_int, C.c_int64, C.c_double
    sig = {
        "gexs_create": (i32, [i32, i32, C.POINTER(P)]),
        "gexs_destroy": (i32, [P]),
        "gexs_init_vectors": (i32, [P, i32, i32, i64, C.POINTER(i64)]),
        "gexs_init_constants": (i32, [P, C.POINTER(Constants)]),
        "gexs_transfer_freq": (i32, [P, P, i32]),
        "gexs_transfer_lines": (i32, [P, i64, P, P, P, P, P, P]),
        "gexs_transfer_lines_device": (i32, [P, i64, P, P, P, P, P, P]),
        "gexs_execute": (i32, [P, i32, i64, i32, dbl, dbl, dbl]),
        "gexs_transfer_results": (i32, [P, P, i32, i32, i32]),
        "gexs_ready": (i32, [P]),
        "gexs_sync": (i32, [P]),
        "gexs_last_error": (C.c_char_p, [P]),
        "gexs_last_stats": (i32, [P, C.POINTER(Stats)]),
        "gexs_timing": (i32, [P, i32, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(i32)]),
        "gexs_line_windows": (i32, [P, i64, P, P, P]),
        "gexs_pass_lines": (i32, [P, C.POINTER(i64), C.POINTER(i64)]),
        "gexs_pass_items": (i32, [P, i64, P, C.POINTER(i64)]),
        "gexs_set_tuning": (i32, [P, i32, i32, i32]),
        "gexs_load_list": (i32, [P, C.POINTER(States), C.POINTER(Transitions)]),
        "gexs_line_count": (i32, [P, C.POINTER(i64)]),
        "gexs_partition": (i32, [P, P, i32, P]),
        "gexs_run_batch": (i32, [P, C.POINTER(TP), i32, C.POINTER(Grid), P]),
        "gexs_batch_device_out": (i32, [P, C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
        "gexs_stream_xsec": (i32, [P, C.POINTER(States), C.POINTER(Transitions), C.POINTER(TP),
                                   i32, C.POINTER(Grid), i64, P, C.POINTER(StreamStats)]),
        "gexs_set_launch": (i32, [P, i32, i32]),
        "gexs_autotune": (i32, [P, C.POINTER(TP), i32, C.POINTER(Grid), C.POINTER(Tuning),
                                C.POINTER(Tuning), i32, C.POINTER(i32)]),
        "gexs_run_fixed": (i32, [P, C.POINTER(LinesDev), C.POINTER(TP), C.POINTER(Grid), P,
                                 C.POINTER(i64)]),
        "gexs_serve_start": (i32, [P, i32, i64]),
        "gexs_submit": (i32, [P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
        "gexs_poll": (i32, [P, C.c_uint64, C.POINTER(Result)]),
        "gexs_wait": (i32, [P, C.c_uint64, C.POINTER(Result)]),
        "gexs_release": (i32, [P, C.c_uint64]),
        "gexs_serve_stop": (i32, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
