"""Full-size workloads against the oracle.

The 10^6-line Voigt workload is compared point by point over the whole grid
(the multi-threaded oracle finishes in seconds).  For the 10^7-line (T,P)
table and the 10^8-line list the oracle is evaluated at sampled points: 10^4
random grid points, every point within 1 cm^-1 of the 100 strongest lines
(they set the spectrum peak) and three random 10 cm^-1 windows, whose sums
stand in for the band strength.
"""
import os

import numpy as np
import pytest

from gexs_amd import synth
from gexs_amd.manager import HipManager

pytestmark = pytest.mark.gpu

P_SCALE = synth.P_SCALE_BAR
NTHREADS = min(16, os.cpu_count() or 1)


def log(msg):
    print(msg, flush=True)


def test_voigt_1e6_full_grid(O):
    spec = synth.WORKLOADS["voigt_1e6"]
    L = synth.workload_list("voigt_1e6")
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    T, P = spec["T"], spec["P"]
    m = HipManager("voigt")
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, 296.0, P_SCALE, cutoff=25.0)
    m.load_list(L)
    out = m.run_batch([(T, P)], start, end, N)[0]
    ne_gpu = m.stats()["n_eval"]
    nu, E, A, g, gam, nex = L.lines()
    Q = O.partition(L.E, L.g, T)
    p = O.params("voigt", T=T, P=P, Q=Q, mass=spec["mass"], cutoff=25.0, p_scale=P_SCALE)
    ref, ne = O.cross_section(p, start, dnu, N, nu, E, A, g, gam, nex, nthreads=NTHREADS)
    assert ne_gpu == ne
    err = np.max(np.abs(out - ref)) / ref.max()
    band = abs(out.sum() - ref.sum()) / ref.sum()
    log(f"voigt_1e6 full grid: N_eval {ne:.4e} peak-rel {err:.3e} band {band:.3e}")
    assert err <= 1e-4 and band <= 1e-6


def sample_points(rng, N, start, dnu, nu_strong, windows=3):
    pts = [rng.integers(0, N, 10_000)]
    for v in nu_strong:
        c = int(round((v - start) / dnu))
        w = int(round(1.0 / dnu))
        pts.append(np.arange(max(c - w, 0), min(c + w + 1, N)))
    win = []
    for _ in range(windows):
        a = int(rng.integers(0, N - int(10.0 / dnu)))
        win.append(np.arange(a, a + int(10.0 / dnu)))
    allp = np.unique(np.concatenate(pts + win)).astype(np.int32)
    return allp, win


def sampled_check(O, L, spec, T, P, out, seed):
    """Oracle at sampled points of the grid; returns (peak-relative error,
    worst relative error of the three 10 cm^-1 window sums)."""
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    nu, E, _, g, _, _ = L.lines()
    A, gam, nex = L.A.astype(float), L.gamma.astype(float), L.nexp.astype(float)
    o = np.argsort(nu, kind="stable")
    keep = (nu[o] >= start - 100.0) & (nu[o] <= end + 100.0) & (nu[o] != 0)
    o = o[keep]
    nu, E, A, g, gam, nex = nu[o], E[o], A[o], g[o], gam[o], nex[o]
    del o, keep
    Q = O.partition(L.E, L.g, T)
    p = O.params("voigt", T=T, P=P, Q=Q, mass=spec["mass"], cutoff=25.0, p_scale=P_SCALE)
    S, _, _ = O.line_params(p, nu, E, A, g, gam, nex)
    inside = (nu > start + 1) & (nu < end - 1)
    strong = nu[inside][np.argsort(S[inside])[-100:]]
    del S, inside
    rng = np.random.default_rng(seed)
    pts, wins = sample_points(rng, N, start, dnu, strong)
    ref = O.cross_section_at(p, start, dnu, N, nu, E, A, g, gam, nex, pts, nthreads=NTHREADS)
    got = out[pts]
    err = np.max(np.abs(got - ref)) / ref.max()
    pos = {j: i for i, j in enumerate(pts)}
    wrel = 0.0
    for w in wins:
        r = ref[[pos[j] for j in w]].sum()
        if r > 0:
            wrel = max(wrel, abs(out[w].sum() - r) / r)
    log(f"T={T:.1f} P={P:.2e}: {pts.shape[0]} points, peak-rel {err:.3e}, window sums {wrel:.2e}")
    return err, wrel


def test_stream_1e8_sampled(O):
    """The 10^8-transition hot list (2e6 states, FP32 A, T = 2000 K, P = 0.1 bar,
    grid 100-30000 cm^-1 at 0.01): the whole selected list resident in HBM,
    one pass."""
    name = "voigt_1e8_stream"
    spec = synth.WORKLOADS[name]
    L = synth.workload_list(name)
    log(f"{name}: {L.n_trans} transitions generated")
    start, end, N = spec["grid"]
    T, P = spec["T"], spec["P"]
    m = HipManager("voigt")
    m.InitializeConstants(100.0, T, 1.0, (end - start) / N, spec["mass"], P, 296.0, P_SCALE,
                          cutoff=25.0)
    m.load_list(L)
    out = m.run_batch([(T, P)], start, end, N)[0]
    log(f"N_eval {m.stats()['n_eval']:.4e}, {m.stats()['n_lines']} lines selected")
    m.close()
    err, wrel = sampled_check(O, L, spec, T, P, out, 77)
    assert err <= 1e-4 and wrel <= 1e-6


@pytest.mark.parametrize("k", [0, 58, 119])
def test_tp_table_1e7_sampled(O, k):
    spec = synth.WORKLOADS["voigt_tp_table_1e7"]
    L = synth.workload_list("voigt_tp_table_1e7")
    start, end, N = spec["grid"]
    T, P = synth.tp_table()[k]
    m = HipManager("voigt")
    m.InitializeConstants(100.0, T, 1.0, (end - start) / N, spec["mass"], P, 296.0, P_SCALE,
                          cutoff=25.0)
    m.load_list(L)
    out = m.run_batch([(T, P)], start, end, N)[0]
    m.close()
    err, wrel = sampled_check(O, L, spec, T, P, out, 1000 + k)
    assert err <= 1e-4 and wrel <= 1e-6
