"""ctypes binding of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker.  ``ref()`` returns the binding of
``oracle/_ref/libgexs_ref.so`` (the reference's own CPU path compiled from its
sources by ``make -C oracle ref``) or None where that library was not built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libgexs_ref.so")

VOIGT, DOPPLER, GAUSSIAN, VOIGT_QUAD = 0, 1, 2, 3
PROFILES = {"voigt": VOIGT, "doppler": DOPPLER, "gaussian": GAUSSIAN, "voigt_quad": VOIGT_QUAD}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("profile", C.c_int32), ("width_conv", C.c_int32),
                ("pressure_mode", C.c_int32), ("reserved", C.c_int32),
                ("T", C.c_double), ("P", C.c_double), ("Q", C.c_double),
                ("mass", C.c_double), ("t_ref", C.c_double), ("p_scale", C.c_double),
                ("cutoff", C.c_double), ("hwhm", C.c_double)]


def params(profile="voigt", T=296.0, P=1.0, Q=1.0, mass=18.0153, t_ref=296.0,
           p_scale=1.0 / 1.01325, cutoff=25.0, hwhm=0.1, width_conv=0, pressure_mode=1):
    prof = PROFILES[profile] if isinstance(profile, str) else int(profile)
    return Params(prof, int(width_conv), int(pressure_mode), 0, float(T), float(P), float(Q),
                  float(mass), float(t_ref), float(p_scale), float(cutoff), float(hwhm))


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    L = C.CDLL(LIB)
    P = C.POINTER(Params)
    L.gxo_partition.restype = C.c_double
    L.gxo_partition.argtypes = [_dp, _ip, C.c_int64, C.c_double]
    L.gxo_select.restype = C.c_int
    L.gxo_select.argtypes = [C.c_double] * 4
    L.gxo_windows.restype = None
    L.gxo_windows.argtypes = [C.c_int64, _dp, C.c_double, C.c_double, C.c_double,
                              C.c_int32, C.c_int32, _ip, _ip]
    for fn in ("gxo_humlic", "gxo_voigt916", "gxo_gh11", "gxo_voigt_quad"):
        getattr(L, fn).restype = C.c_double
        getattr(L, fn).argtypes = [C.c_double, C.c_double]
    L.gxo_erfcx.restype = C.c_double
    L.gxo_erfcx.argtypes = [C.c_double]
    L.gxo_line_params.restype = None
    L.gxo_line_params.argtypes = [P, C.c_int64, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp]
    L.gxo_cross_section.restype = C.c_int64
    L.gxo_cross_section.argtypes = [P, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int64,
                                    _dp, _dp, _dp, _ip, _dp, _dp, _dp, C.c_int32]
    L.gxo_cross_section_at.restype = None
    L.gxo_cross_section_at.argtypes = [P, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                       C.c_int64, _dp, _dp, _dp, _ip, _dp, _dp, C.c_int64,
                                       _ip, _dp, C.c_int32]
    L.gxo_neval.restype = C.c_int64
    L.gxo_neval.argtypes = [P, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int64,
                            _dp, _dp, _dp, _ip]
    L.gxo_total_intensity.restype = C.c_double
    L.gxo_total_intensity.argtypes = [C.c_double, C.c_double, C.c_int64, _dp, _dp, _dp, _ip]
    _lib = L
    return L


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def partition(E, g, T):
    E, g = _f(E), _i(g)
    return lib().gxo_partition(E, g, E.shape[0], float(T))


def windows(nu, cut, start, dnu, lo, hi):
    nu = _f(nu)
    ib = np.empty(nu.shape[0], np.int32)
    ie = np.empty(nu.shape[0], np.int32)
    lib().gxo_windows(nu.shape[0], nu, float(cut), float(start), float(dnu), int(lo), int(hi), ib, ie)
    return ib, ie


def humlic(x, y):
    return lib().gxo_humlic(float(x), float(y))


def voigt916(x, y):
    return lib().gxo_voigt916(float(x), float(y))


def gh11(x, y):
    return lib().gxo_gh11(float(x), float(y))


def voigt_quad(x, y):
    return lib().gxo_voigt_quad(float(x), float(y))


def line_params(p, nu, E, A, g, gamma, nexp):
    nu, E, A, g, gamma, nexp = _f(nu), _f(E), _f(A), _i(g), _f(gamma), _f(nexp)
    n = nu.shape[0]
    S, gL, iw = np.empty(n), np.empty(n), np.empty(n)
    lib().gxo_line_params(C.byref(p), n, nu, E, A, g, gamma, nexp, S, gL, iw)
    return S, gL, iw


def cross_section(p, start, dnu, npoints, nu, E, A, g, gamma=None, nexp=None, nthreads=1,
                  start_idx=0, nwin=None, out=None):
    nu, E, A, g = _f(nu), _f(E), _f(A), _i(g)
    n = nu.shape[0]
    gamma = _f(gamma) if gamma is not None else np.zeros(n)
    nexp = _f(nexp) if nexp is not None else np.zeros(n)
    if out is None:
        out = np.zeros(npoints)
    if nwin is None:
        nwin = npoints - start_idx
    neval = lib().gxo_cross_section(C.byref(p), float(start), float(dnu), int(start_idx),
                                    int(nwin), n, nu, E, A, g, gamma, nexp, out, int(nthreads))
    return out, neval


def cross_section_at(p, start, dnu, npoints, nu, E, A, g, gamma, nexp, pts, nthreads=1):
    """Values at grid points ``pts``; nu must be ascending."""
    nu, E, A, g = _f(nu), _f(E), _f(A), _i(g)
    gamma, nexp, pts = _f(gamma), _f(nexp), _i(pts)
    out = np.empty(pts.shape[0])
    lib().gxo_cross_section_at(C.byref(p), float(start), float(dnu), 0, int(npoints),
                               nu.shape[0], nu, E, A, g, gamma, nexp, pts.shape[0], pts, out,
                               int(nthreads))
    return out


def neval(p, start, dnu, npoints, nu, E, A, g):
    nu, E, A, g = _f(nu), _f(E), _f(A), _i(g)
    return lib().gxo_neval(C.byref(p), float(start), float(dnu), 0, int(npoints),
                           nu.shape[0], nu, E, A, g)


def total_intensity(T, Q, nu, E, A, g):
    nu, E, A, g = _f(nu), _f(E), _f(A), _i(g)
    return lib().gxo_total_intensity(float(T), float(Q), nu.shape[0], nu, E, A, g)


# ------------------------------------------------------------------ reference
_ref = None


def ref():
    """Binding of oracle/_ref/libgexs_ref.so, or None when it is not built."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_LIB):
        return None
    L = C.CDLL(REF_LIB)
    L.gxr_humlic.restype = C.c_double
    L.gxr_humlic.argtypes = [C.c_double, C.c_double]
    L.gxr_partition.restype = C.c_double
    L.gxr_partition.argtypes = [_dp, _ip, C.c_long, C.c_double]
    L.gxr_voigt_tp.restype = None
    L.gxr_voigt_tp.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_double, C.c_double,
                               C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_long, _dp, _dp, _dp, _ip, _dp, _dp, C.c_int, C.c_int, _dp]
    L.gxr_doppler.restype = None
    L.gxr_doppler.argtypes = [C.c_int, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_long, _dp, _dp, _dp, _ip, C.c_int, C.c_int, _dp]
    _ref = L
    return L


def ref_voigt_tp(start, dnu, npoints, T, P, Q, mass, t_ref, p_ref, nu, E, A, g, gamma, nexp,
                 nthreads=1, parallel=False):
    """OpenMPManager::ComputeVoigtTP over the full grid.  Its pressure-free
    gamma_L and min(25P, 100) cutoff are the reference CPU conventions."""
    L = ref()
    freq = start + np.arange(npoints, dtype=np.float64) * dnu
    out = np.zeros(npoints)
    nu, E, A, g, gamma, nexp = _f(nu), _f(E), _f(A), _i(g), _f(gamma), _f(nexp)
    L.gxr_voigt_tp(int(nthreads), int(parallel), freq, int(npoints), float(dnu), float(T),
                   float(P), float(Q), float(mass), float(t_ref), float(p_ref), nu.shape[0],
                   E, nu, A, g, gamma, nexp, 0, int(npoints), out)
    return out
