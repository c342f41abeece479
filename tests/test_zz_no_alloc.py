"""gexs_run_fixed, the entry point that allocates nothing, against
gexs_run_batch on the same lines.

The lines go to device memory through the HIP runtime the library links
(ctypes, test plumbing only).  Two calls in a row, the second with fewer
lines and a smaller grid, each bit for bit equal to gexs_run_batch for the
same (T,P), grid and launch shape (a tile sums its lines in the same order in
both), with N_eval equal to the oracle's; the second also against the oracle
pointwise.  Sizes beyond the maxima are refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from gexs_amd import _lib, synth
from gexs_amd._lib import GexsError
from gexs_amd.manager import HipManager

pytestmark = pytest.mark.gpu

P_SCALE = synth.P_SCALE_BAR
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Dev:
    """Device buffers made with the HIP runtime (freed at close)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(int(nbytes), 8))) == 0
        self.bufs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(C.c_void_p(p), a.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(a.nbytes), H2D) == 0
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p),
                                  C.c_size_t(out.nbytes), D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def selected_lines(L, start, end, hw=100.0):
    """The lines gexs_run_batch takes for the grid: nu != 0, ascending in nu
    (stable, as the device sort), start - hw <= nu <= end + hw."""
    nu, E, A, g, gam, nex = L.lines()
    keep = nu != 0.0
    nu, E, A, g, gam, nex = nu[keep], E[keep], A[keep], g[keep], gam[keep], nex[keep]
    o = np.argsort(nu, kind="stable")
    nu, E, A, g, gam, nex = nu[o], E[o], A[o], g[o], gam[o], nex[o]
    s = (nu >= start - hw) & (nu <= end + hw)
    return nu[s], E[s], A[s], g[s], gam[s], nex[s]


@pytest.fixture(scope="module")
def setup():
    L = synth.workload_list("voigt_1e6", scale=0.05)
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 296.0, 1.0, 0.01, 18.0153, 1.0, 296.0, P_SCALE, cutoff=25.0)
    m.load_list(L)
    dev = Dev()
    yield L, m, dev
    dev.close()
    m.close()


@pytest.mark.parametrize("ppl,split,order", [(2, 4, 3), (1, 16, 1), (4, 1, 0)])
def test_run_fixed_equals_run_batch(O, setup, ppl, split, order):
    L, m, dev = setup
    m.set_tuning(ppl, order, split)
    T, P = 296.0, 1.0
    Q = float(m.partition([T])[0])
    d_out = dev.alloc(1_000_000 * 8)
    for k, (start, end, N) in enumerate([(0.0, 10000.0, 1_000_000), (2000.0, 2600.0, 60_000)]):
        lines = selected_lines(L, start, end)
        n = lines[0].shape[0]
        ptrs = [dev.put(a) for a in lines]
        print(f"call {k}: ppl={ppl} split={split} order={order} {n} lines, {N} points",
              flush=True)
        ne = m.run_fixed([n] + ptrs, T, P, Q, start, end, N, d_out)
        got = dev.get(d_out, N)
        ref = m.run_batch([(T, P, Q)], start, end, N)[0]
        assert m.stats()["n_eval"] == ne
        p = O.params("voigt", T=T, P=P, Q=Q, cutoff=25.0, p_scale=P_SCALE)
        assert ne == O.neval(p, start, (end - start) / N, N, *lines[:4])
        assert np.array_equal(got, ref), k
        if k == 1:
            want, _ = O.cross_section(p, start, (end - start) / N, N, *lines, nthreads=8)
            err = np.max(np.abs(got - want)) / want.max()
            band = abs(got.sum() - want.sum()) / want.sum()
            print(f"  vs oracle: peak-rel {err:.2e} band {band:.2e}", flush=True)
            assert err <= 1e-4 and band <= 1e-6
    # asynchronous form (no N_eval wait): the same grid again after a sync
    m.run_fixed([n] + ptrs, T, P, Q, start, end, N, d_out, n_eval=False)
    m.sync()
    assert np.array_equal(dev.get(d_out, N), ref)
    m.set_tuning(4, 1, 1)


def test_run_fixed_refuses_oversized(setup):
    L, m, dev = setup
    d = dev.alloc(64)
    m.set_tuning(2, 3, 4)
    too_many = [_lib.FIXED_MAX_LINES + 1, d, d, d, d, d, d]
    with pytest.raises(GexsError, match="maximum"):
        m.run_fixed(too_many, 296.0, 1.0, 100.0, 0.0, 10.0, 100, d)
    # 128-point tiles: one tile more than the maximum
    npts = 128 * (_lib.FIXED_MAX_TILES + 1)
    with pytest.raises(GexsError, match="maximum"):
        m.run_fixed([1, d, d, d, d, d, d], 296.0, 1.0, 100.0, 0.0, 10.0, npts, d)
    with pytest.raises(GexsError):
        m.run_fixed([1, d, d, d, d, d, d], 296.0, 1.0, 0.0, 0.0, 10.0, 100, d)  # Q = 0
    m.set_tuning(4, 1, 1)
