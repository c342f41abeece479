"""HipManager: the BaseManager plugin API over the gexs HIP C ABI.

Method names, arguments and call order follow Managers/BaseManager.h:22-35 of
the reference (InitializeVectors, InitializeConstants, TransferFreq,
TransferVectors, ExecuteCrossSection, TransferResults, ReadyForWork,
GetNtrans, Cleanup), so a driver written against BaseManager (as
Profiles/BaseProfile.cpp:60-112,281 is) runs unchanged.  Arrays are numpy;
every call goes through libgexs_hip.so.  Errors raise GexsError (the reference
calls exit(-1) instead, compute/CUDA/cuda_utils.cu:8-17).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Constants, GexsError, Stats


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class HipManager:
    def __init__(self, profile="voigt", device=0, precision="fp32c"):
        L = _lib.load()
        self._L = L
        self.profile = _lib.PROFILE_IDS[profile] if isinstance(profile, str) else int(profile)
        h = C.c_void_p()
        rc = L.gexs_create(int(device), self.profile, C.byref(h))
        if rc != 0:
            raise GexsError(rc, "gexs_create failed (is a GPU visible?)")
        self._h = h
        self._precision = _lib.FP64 if precision == "fp64" else _lib.FP32C
        self.N_trans = 0
        self.npoints = 0
        self._keep = []

    # ------------------------------------------------------------ helpers
    def _chk(self, rc):
        if rc != 0:
            msg = self._L.gexs_last_error(self._h).decode()
            raise GexsError(rc, msg)
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._L.gexs_destroy(self._h)
            self._h = None

    Cleanup = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------- BaseManager API
    def InitializeVectors(self, Npoints, num_cs=1, mem_budget=0):
        cap = C.c_int64()
        self._chk(self._L.gexs_init_vectors(self._h, int(Npoints), int(num_cs), int(mem_budget),
                                            C.byref(cap)))
        self.N_trans = cap.value
        self.npoints = int(Npoints)

    def GetNtrans(self):
        return self.N_trans

    def InitializeConstants(self, half_width, temperature, partition, dfreq, meanmass,
                            pressure=10.0, ref_temp=296.0, ref_press=1.0, cutoff=0.0,
                            width_conv=0, pressure_mode=1, gauss_hwhm=None):
        k = Constants(float(half_width), float(temperature), float(partition), float(dfreq),
                      float(meanmass), float(pressure), float(ref_temp), float(ref_press),
                      float(cutoff),
                      float(gauss_hwhm if gauss_hwhm is not None else half_width),
                      int(width_conv), int(pressure_mode), int(self._precision), 0)
        self._chk(self._L.gexs_init_constants(self._h, C.byref(k)))

    def TransferFreq(self, freq, intens=None, N=None):
        freq = np.ascontiguousarray(freq, dtype=np.float64)
        n = int(N if N is not None else freq.shape[0])
        self._chk(self._L.gexs_transfer_freq(self._h, _ptr(freq), n))

    def TransferVectors(self, Nener, energies, nu, aif, gns, gamma=None, nexp=None):
        E = np.ascontiguousarray(energies, dtype=np.float64)
        v = np.ascontiguousarray(nu, dtype=np.float64)
        A = np.ascontiguousarray(aif, dtype=np.float64)
        g = np.ascontiguousarray(gns, dtype=np.int32)
        ga = np.ascontiguousarray(gamma, dtype=np.float64) if gamma is not None else None
        nx = np.ascontiguousarray(nexp, dtype=np.float64) if nexp is not None else None
        self._chk(self._L.gexs_transfer_lines(self._h, int(Nener), _ptr(E), _ptr(v), _ptr(A),
                                              _ptr(g), _ptr(ga), _ptr(nx)))
        self._n_lines = int(Nener)

    def TransferVectorsDevice(self, Nener, E_ptr, nu_ptr, A_ptr, g_ptr, gamma_ptr=None,
                              nexp_ptr=None):
        """TransferVectors from device pointers (ints) already in HBM."""
        vp = lambda p: C.c_void_p(p) if p else None  # noqa: E731
        self._chk(self._L.gexs_transfer_lines_device(self._h, int(Nener), vp(E_ptr), vp(nu_ptr),
                                                     vp(A_ptr), vp(g_ptr), vp(gamma_ptr),
                                                     vp(nexp_ptr)))
        self._n_lines = int(Nener)

    def ExecuteCrossSection(self, N, N_ener, start_idx, temperature, pressure, partition):
        self._chk(self._L.gexs_execute(self._h, int(N), int(N_ener), int(start_idx),
                                       float(temperature), float(pressure), float(partition)))

    def TransferResults(self, freq, intens, N, I=0, accumulate=True):
        out = intens
        assert out.dtype == np.float64 and out.flags["C_CONTIGUOUS"]
        self._chk(self._L.gexs_transfer_results(self._h, _ptr(out), int(N), int(I),
                                                1 if accumulate else 0))
        return out

    def ReadyForWork(self):
        return self._L.gexs_ready(self._h) == 1

    # ------------------------------------------------------------ extras
    def sync(self):
        self._chk(self._L.gexs_sync(self._h))

    def stats(self):
        s = Stats()
        self._chk(self._L.gexs_last_stats(self._h, C.byref(s)))
        return dict(n_eval=s.n_eval, n_lines=s.n_lines, n_tiles=s.n_tiles,
                    kernel_ms=s.kernel_ms, total_ms=s.total_ms)

    def timing(self, reset=False):
        """(sum of kernel ms, sum of total ms, count) since the last reset: per
        execute its profile kernel and whole pass, per run_batch the batch's
        device work and the same from before its table upload."""
        k, t, n = C.c_double(), C.c_double(), C.c_int()
        self._chk(self._L.gexs_timing(self._h, 1 if reset else 0, C.byref(k), C.byref(t),
                                      C.byref(n)))
        return k.value, t.value, n.value

    def pass_lines(self):
        """(first, count) of the nu-sorted lines the last pass used (its selection)."""
        a, b = C.c_int64(), C.c_int64()
        self._chk(self._L.gexs_pass_lines(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def line_windows(self, n=None):
        """(sorted nu, ib, ie) of the last pass's lines as the device holds them."""
        if n is None:
            n = self.pass_lines()[1]
        nu = np.empty(n, np.float64)
        ib = np.empty(n, np.int32)
        ie = np.empty(n, np.int32)
        self._chk(self._L.gexs_line_windows(self._h, int(n), _ptr(nu), _ptr(ib), _ptr(ie)))
        return nu, ib, ie

    def pass_items(self):
        """Work items of the last FP32C pass as an int32 array (n, 14): lo, hi,
        w0, w1, s1, s2, s3, s4, job, group, f2, f3, e2, e3 (line range, grid
        points, segments)."""
        n = C.c_int64()
        self._chk(self._L.gexs_pass_items(self._h, 0, None, C.byref(n)))
        out = np.empty((n.value, 14), np.int32)
        if n.value:
            self._chk(self._L.gexs_pass_items(self._h, n.value, _ptr(out), C.byref(n)))
        return out

    def set_tuning(self, points_per_lane=4, order=1, waves_per_tile=1):
        self._chk(self._L.gexs_set_tuning(self._h, int(points_per_lane), int(order),
                                          int(waves_per_tile)))

    # ------------------------------------------------- resident list / batch
    def load_list(self, lst):
        """Upload a synth.LineList (states + transitions) and sort it on the device.
        A, gamma and n go up in the list's own precision (FP32 or FP64); the
        device holds them in FP64 either way."""
        st, tr, keep = self._list_structs(lst)
        self._chk(self._L.gexs_load_list(self._h, C.byref(st), C.byref(tr)))
        del keep
        n = C.c_int64()
        self._chk(self._L.gexs_line_count(self._h, C.byref(n)))
        self._n_lines = n.value
        return n.value

    @staticmethod
    def _list_structs(lst):
        """(States, Transitions, arrays kept alive) of a synth.LineList; A,
        gamma and n go in the list's own precision."""
        E = np.ascontiguousarray(lst.E, dtype=np.float64)
        g = np.ascontiguousarray(lst.g, dtype=np.int32)
        up = np.ascontiguousarray(lst.upper, dtype=np.uint32)
        lo = np.ascontiguousarray(lst.lower, dtype=np.uint32)
        f64 = lst.A.dtype == np.float64
        dt = np.float64 if f64 else np.float32
        cols = [np.ascontiguousarray(a, dtype=dt) if a is not None else None
                for a in (lst.A, lst.gamma, lst.nexp)]
        ptrs = [a.ctypes.data if a is not None else None for a in cols]
        none = [None, None, None]
        st = _lib.States(E.shape[0], E.ctypes.data, g.ctypes.data)
        tr = _lib.Transitions(up.shape[0], up.ctypes.data, lo.ctypes.data,
                              *(none + ptrs if f64 else ptrs + none))
        return st, tr, (E, g, up, lo, cols)

    def stream_xsec(self, lst, tps, start, end, npoints, mem_budget, out=None):
        """gexs_stream_xsec: cross-sections at [(T, P[, Q]), ...] of a list that
        stays in host memory, streamed through the device in chunks that fit
        mem_budget bytes.  Returns (grids, stats): grids a (len(tps), npoints)
        float64 array (``out`` if given; out="device" leaves them in the
        library's device buffer and returns None)."""
        st, tr, keep = self._list_structs(lst)
        arr = (_lib.TP * len(tps))()
        for i, t in enumerate(tps):
            arr[i] = _lib.TP(float(t[0]), float(t[1]), float(t[2]) if len(t) > 2 else 0.0)
        grid = _lib.Grid(float(start), float(end), int(npoints), 0)
        stats = _lib.StreamStats()
        dev = isinstance(out, str) and out == "device"
        if not dev and out is None:
            out = np.empty((len(tps), int(npoints)), dtype=np.float64)
        self._chk(self._L.gexs_stream_xsec(self._h, C.byref(st), C.byref(tr), arr, len(tps),
                                           C.byref(grid), int(mem_budget),
                                           None if dev else _ptr(out), C.byref(stats)))
        del keep
        info = {k: getattr(stats, k) for k, _ in _lib.StreamStats._fields_ if k != "reserved"}
        return (None if dev else out), info

    def set_launch(self, persistent_blocks=0, use_graph=False):
        self._chk(self._L.gexs_set_launch(self._h, int(persistent_blocks), 1 if use_graph else 0))

    def autotune(self, tps, start, end, npoints):
        """Pick the fastest profile-kernel launch shape for this batch; returns
        (best, [all candidates]) as dicts."""
        arr = (_lib.TP * len(tps))()
        for i, t in enumerate(tps):
            arr[i] = _lib.TP(float(t[0]), float(t[1]), float(t[2]) if len(t) > 2 else 0.0)
        grid = _lib.Grid(float(start), float(end), int(npoints), 0)
        best = _lib.Tuning()
        cap = 128
        rep = (_lib.Tuning * cap)()
        nrep = C.c_int()
        self._chk(self._L.gexs_autotune(self._h, arr, len(tps), C.byref(grid), C.byref(best), rep,
                                        cap, C.byref(nrep)))
        conv = lambda t: dict(points_per_lane=t.points_per_lane,  # noqa: E731
                              waves_per_tile=t.waves_per_tile, order=t.order,
                              persistent_blocks=t.persistent_blocks, kernel_ms=t.kernel_ms)
        return conv(best), [conv(rep[i]) for i in range(min(nrep.value, cap))]

    def batch_device_out(self):
        """(device address, capacity in doubles) of the last run_batch's results."""
        p, n = C.c_void_p(), C.c_int64()
        self._chk(self._L.gexs_batch_device_out(self._h, C.byref(p), C.byref(n)))
        return p.value, n.value

    def partition(self, temps):
        T = np.ascontiguousarray(temps, dtype=np.float64)
        Q = np.empty_like(T)
        self._chk(self._L.gexs_partition(self._h, _ptr(T), T.shape[0], _ptr(Q)))
        return Q

    def run_batch(self, tps, start, end, npoints, out=None):
        """Cross-sections at [(T, P[, Q]), ...] over grid [start, end) with npoints.
        ``out`` is a (len(tps), npoints) float64 array, an int (the device
        address of len(tps)*npoints doubles), or "device": the results stay in
        the library's device buffer (batch_device_out) and None is returned."""
        arr = (_lib.TP * len(tps))()
        for i, t in enumerate(tps):
            arr[i] = _lib.TP(float(t[0]), float(t[1]), float(t[2]) if len(t) > 2 else 0.0)
        grid = _lib.Grid(float(start), float(end), int(npoints), 0)
        if isinstance(out, str) and out == "device":
            self._chk(self._L.gexs_run_batch(self._h, arr, len(tps), C.byref(grid), None))
            return None
        if isinstance(out, int):
            self._chk(self._L.gexs_run_batch(self._h, arr, len(tps), C.byref(grid),
                                             C.c_void_p(out)))
            return out
        if out is None:
            out = np.empty((len(tps), int(npoints)), dtype=np.float64)
        self._chk(self._L.gexs_run_batch(self._h, arr, len(tps), C.byref(grid), _ptr(out)))
        return out

    def run_fixed(self, lines_dev, T, P, Q, start, end, npoints, d_out, n_eval=True):
        """gexs_run_fixed: one cross-section of lines held in device memory
        (lines_dev = (n, nu, E, A, g, gamma, nexp) device addresses, nu
        ascending) into the device grid at d_out, with no device allocation.
        Returns N_eval (waits) or None (asynchronous) when n_eval is False."""
        L = _lib.LinesDev(*[int(v) if v else None for v in lines_dev])
        tp = _lib.TP(float(T), float(P), float(Q))
        grid = _lib.Grid(float(start), float(end), int(npoints), 0)
        ne = C.c_int64()
        self._chk(self._L.gexs_run_fixed(self._h, C.byref(L), C.byref(tp), C.byref(grid),
                                         C.c_void_p(int(d_out)),
                                         C.byref(ne) if n_eval else None))
        return ne.value if n_eval else None

    # ------------------------------------------------------------- serving
    def serve_start(self, max_inflight=2, mem_budget=4 << 30):
        """Start the continuous-batching server over the resident list."""
        self._chk(self._L.gexs_serve_start(self._h, int(max_inflight), int(mem_budget)))

    def submit(self, T, P, start, end, dfreq):
        rid = C.c_uint64()
        req = _lib.Request(float(T), float(P), float(start), float(end), float(dfreq))
        self._chk(self._L.gexs_submit(self._h, C.byref(req), C.byref(rid)))
        return rid.value

    def _result(self, r):
        if r.status != 0:
            return dict(status=r.status)
        sigma = np.ctypeslib.as_array(r.sigma, shape=(r.npoints,)).copy()
        return dict(status=0, sigma=sigma, latency_ms=r.latency_ms, device_ms=r.device_ms,
                    shape=(r.tile_points, r.waves_per_tile, r.line_groups), n_eval=r.n_eval)

    def poll(self, rid):
        r = _lib.Result()
        self._chk(self._L.gexs_poll(self._h, C.c_uint64(rid), C.byref(r)))
        return self._result(r)

    def wait(self, rid, release=True):
        r = _lib.Result()
        self._chk(self._L.gexs_wait(self._h, C.c_uint64(rid), C.byref(r)))
        out = self._result(r)
        if release:
            self._chk(self._L.gexs_release(self._h, C.c_uint64(rid)))
        return out

    def release(self, rid):
        self._chk(self._L.gexs_release(self._h, C.c_uint64(rid)))

    def serve_stop(self):
        self._chk(self._L.gexs_serve_stop(self._h))
