"""Full-size workloads against the oracle.

The 10^6-line Voigt workload is compared point by point over the whole grid
(the multi-threaded oracle finishes in seconds).  For the 10^7-line (T,P)
table and the 10^8-line list (streamed from host memory in chunks) the
oracle is evaluated at sampled points: 10^4
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


@pytest.fixture(scope="module")
def c1_full(O):
    spec = synth.WORKLOADS["voigt_1e6"]
    L = synth.workload_list("voigt_1e6")
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    T, P = spec["T"], spec["P"]
    nu, E, A, g, gam, nex = L.lines()
    Q = O.partition(L.E, L.g, T)
    p = O.params("voigt", T=T, P=P, Q=Q, mass=spec["mass"], cutoff=25.0, p_scale=P_SCALE)
    ref, ne = O.cross_section(p, start, dnu, N, nu, E, A, g, gam, nex, nthreads=NTHREADS)
    return spec, L, Q, ref, ne


# (path, points per lane, waves per tile, tile order): the batch path at its
# default shape, and the execute path (ExecuteCrossSection over the resident
# list) at the shape the autotuner picks on this workload and the bench times
@pytest.mark.parametrize("path,ppl,split,order", [("batch", 4, 1, 1), ("execute", 4, 4, 3),
                                                  ("execute", 2, 4, 3), ("execute", 2, 1, 1),
                                                  ("execute", 8, 16, 1)])
def test_voigt_1e6_full_grid(c1_full, path, ppl, split, order):
    spec, L, Q, ref, ne = c1_full
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    T, P = spec["T"], spec["P"]
    m = HipManager("voigt")
    m.set_tuning(ppl, order, split)
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, 296.0, P_SCALE, cutoff=25.0)
    n = m.load_list(L)
    if path == "batch":
        out = m.run_batch([(T, P)], start, end, N)[0]
    else:
        m.InitializeVectors(N, 1)
        m.TransferFreq(start + np.arange(N, dtype=np.float64) * dnu)
        m.ExecuteCrossSection(N, n, 0, T, P, Q)
        out = np.zeros(N)
        m.TransferResults(None, out, N, 0)
    ne_gpu = m.stats()["n_eval"]
    m.close()
    assert ne_gpu == ne
    err = np.max(np.abs(out - ref)) / ref.max()
    band = abs(out.sum() - ref.sum()) / ref.sum()
    log(f"voigt_1e6 full grid {path} ppl={ppl} split={split} order={order}: N_eval {ne:.4e} "
        f"peak-rel {err:.3e} band {band:.3e}")
    assert err <= 1e-4 and band <= 1e-6


@pytest.mark.parametrize("ppl,split,order", [(4, 4, 1), (2, 4, 3), (1, 4, 3)])
def test_voigt_quad_1e6_full_grid(O, ppl, split, order):
    """The 10^6-line workload with the Voigt-Quad profile (the bench's second
    profile line) over the whole grid, at the launch shapes its autotuner picks."""
    spec = synth.WORKLOADS["voigt_1e6"]
    L = synth.workload_list("voigt_1e6")
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    T, P = spec["T"], spec["P"]
    nu, E, A, g, gam, nex = L.lines()
    Q = O.partition(L.E, L.g, T)
    p = O.params("voigt_quad", T=T, P=P, Q=Q, mass=spec["mass"], cutoff=25.0, p_scale=P_SCALE)
    ref, ne = O.cross_section(p, start, dnu, N, nu, E, A, g, gam, nex, nthreads=NTHREADS)
    m = HipManager("voigt_quad")
    m.set_tuning(ppl, order, split)
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, 296.0, P_SCALE, cutoff=25.0)
    n = m.load_list(L)
    m.InitializeVectors(N, 1)
    m.TransferFreq(start + np.arange(N, dtype=np.float64) * dnu)
    m.ExecuteCrossSection(N, n, 0, T, P, Q)
    out = np.zeros(N)
    m.TransferResults(None, out, N, 0)
    ne_gpu = m.stats()["n_eval"]
    m.close()
    assert ne_gpu == ne
    err = np.max(np.abs(out - ref)) / ref.max()
    band = abs(out.sum() - ref.sum()) / ref.sum()
    log(f"voigt_1e6 voigt-quad full grid ppl={ppl} split={split} order={order}: N_eval {ne:.4e} "
        f"peak-rel {err:.3e} band {band:.3e}")
    assert err <= 1e-4 and band <= 1e-6


def sample_points(rng, N, start, dnu, nu_strong, windows=3, nrand=10_000, wwidth=10.0):
    pts = [rng.integers(0, N, nrand)]
    for v in nu_strong:
        c = int(round((v - start) / dnu))
        w = int(round(1.0 / dnu))
        pts.append(np.arange(max(c - w, 0), min(c + w + 1, N)))
    win = []
    wn = max(1, min(int(round(wwidth / dnu)), N))
    for _ in range(windows):
        a = int(rng.integers(0, N - wn + 1))
        win.append(np.arange(a, a + wn))
    allp = np.unique(np.concatenate(pts + win)).astype(np.int32)
    return allp, win


class SortedList:
    """A resident list's lines (nu != 0) in nu order, as the oracle takes them."""

    def __init__(self, L):
        nu, E, _, g, _, _ = L.lines()
        A, gam, nex = L.A.astype(float), L.gamma.astype(float), L.nexp.astype(float)
        o = np.argsort(nu, kind="stable")
        o = o[nu[o] != 0]
        self.nu, self.E, self.A, self.g = nu[o], E[o], A[o], g[o]
        self.gam, self.nex = gam[o], nex[o]
        self.states = (L.E, L.g)

    def select(self, lo, hi):
        """Lines with lo <= nu <= hi (Profiles/BaseProfile.cpp:152-157 as a filter)."""
        a = np.searchsorted(self.nu, lo, side="left")
        b = np.searchsorted(self.nu, hi, side="right")
        s = slice(a, b)
        return (self.nu[s], self.E[s], self.A[s], self.g[s], self.gam[s], self.nex[s])


def sampled_check(O, SL, mass, T, P, start, end, N, out, seed, nrand=10_000, nstrong=100,
                  windows=3, wwidth=10.0, ne_gpu=None):
    """Oracle at sampled points of the grid [start, end) of N points: nrand
    random points, +-1 cm^-1 around the nstrong strongest lines inside it
    (they set the peak) and `windows` random windows of wwidth cm^-1 whose
    sums stand in for the band strength.  Returns (peak-relative error, worst
    relative error of the window sums); with ne_gpu also checks N_eval."""
    dnu = (end - start) / N
    nu, E, A, g, gam, nex = SL.select(start - 100.0, end + 100.0)
    Q = O.partition(SL.states[0], SL.states[1], T)
    p = O.params("voigt", T=T, P=P, Q=Q, mass=mass, cutoff=25.0, p_scale=P_SCALE)
    if ne_gpu is not None:
        ne = O.neval(p, start, dnu, N, nu, E, A, g)
        assert ne_gpu == ne, (ne_gpu, ne)
    S, _, _ = O.line_params(p, nu, E, A, g, gam, nex)
    inside = (nu > start + 1) & (nu < end - 1)
    strong = nu[inside][np.argsort(S[inside])[-nstrong:]]
    del S, inside
    rng = np.random.default_rng(seed)
    pts, wins = sample_points(rng, N, start, dnu, strong, windows, nrand, wwidth)
    ref = O.cross_section_at(p, start, dnu, N, nu, E, A, g, gam, nex, pts, nthreads=NTHREADS)
    got = out[pts]
    if ref.max() == 0:
        assert np.all(got == 0)
        return 0.0, 0.0
    err = np.max(np.abs(got - ref)) / ref.max()
    pos = {j: i for i, j in enumerate(pts)}
    wrel = 0.0
    for w in wins:
        r = ref[[pos[j] for j in w]].sum()
        if r > 0:
            wrel = max(wrel, abs(out[w].sum() - r) / r)
    log(f"T={T:.1f} P={P:.2e} [{start:.2f}, {end:.2f}) x {N}: {pts.shape[0]} points, "
        f"peak-rel {err:.3e}, window sums {wrel:.2e}")
    return err, wrel


def test_stream_1e8_sampled(O):
    """The 10^8-transition hot list (2e6 states, FP32 A, T = 2000 K, P = 0.1 bar,
    grid 100-30000 cm^-1 at 0.01) streamed from host memory through a 2 GiB
    device budget in list-order chunks (gexs_stream_xsec): N_eval exact and
    sampled points and window sums against the oracle; every point against
    the same list resident in HBM and run in one pass."""
    name = "voigt_1e8_stream"
    spec = synth.WORKLOADS[name]
    L = synth.workload_list(name)
    log(f"{name}: {L.n_trans} transitions generated")
    start, end, N = spec["grid"]
    T, P = spec["T"], spec["P"]
    m = HipManager("voigt")
    m.InitializeConstants(100.0, T, 1.0, (end - start) / N, spec["mass"], P, 296.0, P_SCALE,
                          cutoff=25.0)
    out, st = m.stream_xsec(L, [(T, P)], start, end, N, mem_budget=2 << 30)
    out = out[0]
    log(f"streamed: {st['n_chunks']} chunks of {st['chunk']} transitions, "
        f"{st['device_bytes'] / 2**30:.2f} GiB, {st['n_lines']} lines selected, "
        f"N_eval {st['n_eval']:.4e}, {st['wall_ms']:.1f} ms")
    assert st["n_chunks"] >= 8 and st["device_bytes"] <= 2 << 30
    m.load_list(L)
    res = m.run_batch([(T, P)], start, end, N)[0]
    assert m.stats()["n_eval"] == st["n_eval"]
    m.close()
    d = np.max(np.abs(out - res)) / res.max()
    log(f"streamed vs resident one pass: {d:.2e} of the peak")
    assert d <= 2e-6
    del res
    err, wrel = sampled_check(O, SortedList(L), spec["mass"], T, P, start, end, N, out, 77,
                              ne_gpu=st["n_eval"])
    assert err <= 1e-4 and wrel <= 1e-6


@pytest.fixture(scope="module")
def c3_list():
    L = synth.workload_list("voigt_tp_table_1e7")
    return L, SortedList(L)


@pytest.mark.parametrize("k", [0, 58, 119])
def test_tp_table_1e7_sampled(O, c3_list, k):
    L, SL = c3_list
    spec = synth.WORKLOADS["voigt_tp_table_1e7"]
    start, end, N = spec["grid"]
    T, P = synth.tp_table()[k]
    m = HipManager("voigt")
    m.InitializeConstants(100.0, T, 1.0, (end - start) / N, spec["mass"], P, 296.0, P_SCALE,
                          cutoff=25.0)
    m.load_list(L)
    out = m.run_batch([(T, P)], start, end, N)[0]
    m.close()
    err, wrel = sampled_check(O, SL, spec["mass"], T, P, start, end, N, out, 1000 + k)
    assert err <= 1e-4 and wrel <= 1e-6


def test_tp_table_1e7_graph_batch_120(O, c3_list):
    """The whole 120-(T,P) table as the bench times it: one graph-captured
    launch sequence at 2 points per lane and 4 waves per tile, Q(T) on the
    device.  The replay equals the capture run, four rows equal the same (T,P)
    run alone bit for bit (a job's tiles sum in the same order in any batch),
    and the four rows match the oracle at sampled points with N_eval exact."""
    L, SL = c3_list
    spec = synth.WORKLOADS["voigt_tp_table_1e7"]
    start, end, N = spec["grid"]
    tps = synth.tp_table()
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 1000.0, 1.0, (end - start) / N, spec["mass"], 1.0, 296.0,
                          P_SCALE, cutoff=25.0)
    m.load_list(L)
    m.set_tuning(2, 1, 4)
    m.set_launch(0, True)
    first = m.run_batch(tps, start, end, N)  # captures the graph
    ne_all = m.stats()["n_eval"]
    out = m.run_batch(tps, start, end, N)    # replays it
    assert np.array_equal(out, first)
    del first
    assert m.stats()["n_eval"] == ne_all
    log(f"120 (T,P): N_eval {ne_all:.4e}")
    m.set_launch(0, False)
    rows = [0, 37, 58, 119]
    for k in rows:
        alone = m.run_batch([tps[k]], start, end, N)[0]
        assert np.array_equal(alone, out[k]), k
        ne_k = m.stats()["n_eval"]
        T, P = tps[k]
        err, wrel = sampled_check(O, SL, spec["mass"], T, P, start, end, N, out[k], 2000 + k,
                                  ne_gpu=ne_k)
        assert err <= 1e-4 and wrel <= 1e-6, k
    m.close()


def serve_requests(rng, n):
    """Requests of the serving workload: T ~ U[200, 3000] K, P log-uniform in
    [1e-5, 100] bar, a range of width log-uniform in [100, 5000] cm^-1 placed
    uniformly inside 500-15000 cm^-1, resolution in {0.01, 0.1, 1.0}."""
    out = []
    for _ in range(n):
        T = rng.uniform(200.0, 3000.0)
        P = 10.0 ** rng.uniform(-5.0, 2.0)
        w = float(np.exp(rng.uniform(np.log(100.0), np.log(5000.0))))
        lo = rng.uniform(500.0, 15000.0 - w)
        res = float(rng.choice([0.01, 0.1, 1.0]))
        w = round(w / res) * res
        out.append((T, P, lo, lo + w, res))
    return out


def test_serve_full_list(O, c3_list):
    """Serving over the whole 10^7-line list held on the device, as the bench
    serves it (4 overlapping batch slots, 24 GiB arena): 20 requests of the
    serving distribution plus two 0.01 cm^-1 requests of >= 3e5 points (their
    tiles run 2 or 4 points per lane), all submitted at once.  Each result
    matches the oracle at sampled points with N_eval exact, and a subset
    served again one at a time gives the same bits."""
    L, SL = c3_list
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 1000.0, 1.0, 0.01, 28.0, 1.0, 296.0, P_SCALE, cutoff=25.0)
    m.load_list(L)
    m.set_tuning(4, 3, 4)
    m.serve_start(max_inflight=4, mem_budget=24 << 30)
    reqs = serve_requests(np.random.default_rng(4242), 20)
    reqs += [(2400.0, 3e-3, 2000.0, 6500.0, 0.01), (700.0, 8.0, 9000.0, 12500.0, 0.01)]
    ids = [m.submit(*r) for r in reqs]
    res = [m.wait(i) for i in ids]
    big = 0
    for k, (r, q) in enumerate(zip(reqs, res)):
        T, P, a, b, d = r
        N = int(round((b - a) / d))
        assert q["status"] == 0 and q["sigma"].shape == (N,)
        tile_points = q["shape"][0]
        big += N >= 300_000 and tile_points >= 128
        log(f"request {k}: T={T:.0f} P={P:.2e} [{a:.1f}, {b:.1f}) d={d} N={N} shape "
            f"{q['shape']} latency {q['latency_ms']:.2f} ms")
        nr = min(4000, N)
        err, wrel = sampled_check(O, SL, 28.0, T, P, a, b, N, q["sigma"], 3000 + k, nrand=nr,
                                  nstrong=20, windows=2, wwidth=min(10.0, (b - a) / 4),
                                  ne_gpu=q["n_eval"])
        assert err <= 1e-4 and wrel <= 1e-6, k
    assert big >= 2
    for k in (0, 5, 20, 21):
        again = m.wait(m.submit(*reqs[k]))
        assert np.array_equal(again["sigma"], res[k]["sigma"]), k
    m.serve_stop()
    m.close()
