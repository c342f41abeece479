"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bars (per profile and precision):
  window indices ib/ie and the nu-sorted line order      bit-exact
  N_eval                                                  exact
  FP32C  max|sigma - oracle| / max(oracle)                <= 1e-4
         |sum sigma - sum oracle| / sum oracle            <= 1e-6
  FP64   max|sigma - oracle| / max(oracle)                <= 1e-10
Sizes are the workloads' own grids with line lists the oracle finishes in
seconds; the full-size lists are covered by sampled points in
test_gpu_fullsize.py.
"""
import os

import numpy as np
import pytest

from gexs_amd import synth
from gexs_amd.manager import HipManager

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
P_SCALE = synth.P_SCALE_BAR


def log(msg):
    print(msg, flush=True)


def run_manager(profile, lines, T, P, Q, start, end, N, cutoff, precision="fp32c", mass=18.0153,
                width_conv=0, pressure_mode=1, hwhm=0.1, ppl=4, order=1, start_idx=0, nwin=None,
                split=1):
    nu, E, A, g, gam, nex = lines
    dnu = (end - start) / N
    m = HipManager(profile, precision=precision)
    m.set_tuning(ppl, order, split)
    m.InitializeVectors(N, 1)
    m.InitializeConstants(100.0, T, Q, dnu, mass, P, 296.0, P_SCALE, cutoff=cutoff,
                          width_conv=width_conv, pressure_mode=pressure_mode, gauss_hwhm=hwhm)
    freq = start + np.arange(N, dtype=np.float64) * dnu
    m.TransferFreq(freq)
    voigt = profile in ("voigt", "voigt_quad")
    m.TransferVectors(nu.shape[0], E, nu, A, g, gam if voigt else None, nex if voigt else None)
    m.ExecuteCrossSection(N if nwin is None else nwin, nu.shape[0], start_idx, T, P, Q)
    out = np.zeros(N)
    m.TransferResults(freq, out, N, 0)
    st = m.stats()
    return out, st, m


def run_oracle(O, profile, lines, T, P, Q, start, end, N, cutoff, mass=18.0153, width_conv=0,
               pressure_mode=1, hwhm=0.1, start_idx=0, nwin=None):
    nu, E, A, g, gam, nex = lines
    dnu = (end - start) / N
    p = O.params(profile, T=T, P=P, Q=Q, mass=mass, cutoff=cutoff, width_conv=width_conv,
                 pressure_mode=pressure_mode, p_scale=P_SCALE, hwhm=hwhm)
    out, ne = O.cross_section(p, start, dnu, N, nu, E, A, g, gam, nex, nthreads=NTHREADS,
                              start_idx=start_idx, nwin=nwin)
    return out, ne, p


def check(out, ref, tol, band_tol=None, what=""):
    peak = np.max(np.abs(ref))
    assert peak > 0
    err = np.max(np.abs(out - ref)) / peak
    log(f"{what}: peak-relative error {err:.3e}")
    assert err <= tol, f"{what}: {err}"
    if band_tol is not None:
        b = abs(out.sum() - ref.sum()) / ref.sum()
        log(f"{what}: band-strength error {b:.3e}")
        assert b <= band_tol, f"{what}: band {b}"


@pytest.fixture(scope="module")
def c1_small():
    L = synth.workload_list("voigt_1e6", scale=0.02)
    return L, L.lines()


@pytest.fixture(scope="module")
def c1_small_ref(O, c1_small):
    L, lines = c1_small
    T, P = 296.0, 1.0
    Q = O.partition(L.E, L.g, T)
    ref, ne, _ = run_oracle(O, "voigt", lines, T, P, Q, 0.0, 10000.0, 1_000_000, 25.0)
    return Q, ref, ne


# every (points per lane, waves per tile) the autotuner can pick, with the tile
# orders of the launch shapes it chose on the 10^6-line workload (order 3:
# dispatch order, the timed shape at 2 points per lane and 4 waves per tile)
SHAPES = [(ppl, split, order) for ppl in (1, 2, 4, 8) for split in (1, 4, 16)
          for order in ((1, 3) if (ppl, split) == (2, 4) else (1,))]


@pytest.mark.parametrize("ppl,split,order", SHAPES)
def test_voigt_fp32c_vs_oracle(O, c1_small, c1_small_ref, ppl, split, order):
    L, lines = c1_small
    T, P = 296.0, 1.0
    Q, ref, ne = c1_small_ref
    log(f"voigt fp32c ppl={ppl} split={split} order={order} lines={lines[0].shape[0]}")
    out, st, m = run_manager("voigt", lines, T, P, Q, 0.0, 10000.0, 1_000_000, 25.0, ppl=ppl,
                             split=split, order=order)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"voigt fp32c ppl={ppl} split={split} order={order}")
    check_item_ranges(m, lines[0].shape[0])
    if split > 1:
        # the waves of a tile take three-piece shares (before, inside and after
        # the near-centre lines [s2, s3)) where a tile has >= split near lines;
        # 64-point tiles of this 2e4-line list never hold 16 (that shape is
        # covered on the 0.1 cm^-1 grid of test_voigt_coarse_grids)
        three = three_piece_tiles(m, st, ppl, split)
        assert three > 0 or (ppl, split) == (1, 16)


def check_item_ranges(m, n):
    """Every work item's line range [lo, hi) and whole-tile lines [s1, s4)
    equal numpy's bisection of the pass's windows (ib and ie are
    non-decreasing in nu): lo = first ie > w0, hi = first ib >= w1, and the
    lines covering [w0, w1) run from the first ie >= w1 to the first ib > w0,
    within [lo, hi) (s1 = s4 = hi when none does)."""
    _, ib, ie = m.line_windows(n)
    it = m.pass_items()
    lo, hi, w0, w1, s1, s4 = (it[:, k].astype(np.int64) for k in (0, 1, 2, 3, 4, 7))
    elo = np.searchsorted(ie, w0, side="right")
    ehi = np.searchsorted(ib, w1, side="left")
    lf = np.clip(np.searchsorted(ie, w1, side="left"), elo, ehi)
    rf = np.clip(np.searchsorted(ib, w0, side="right"), elo, ehi)
    cover = lf < rf
    assert np.array_equal(lo, elo) and np.array_equal(hi, ehi)
    assert np.array_equal(s1, np.where(cover, lf, ehi))
    assert np.array_equal(s4, np.where(cover, rf, ehi))
    assert np.any(cover) and np.any(hi > s4)


def three_piece_tiles(m, st, ppl, split):
    """Tiles of the last pass whose waves took three-piece line shares."""
    it = m.pass_items()
    assert it.shape[0] == st["n_tiles"] and np.all(it[:, 3] - it[:, 2] <= 64 * ppl)
    three = int(np.sum((it[:, 4] < it[:, 7]) & (it[:, 6] - it[:, 5] >= split)))
    log(f"  {three} of {it.shape[0]} tiles take three-piece shares")
    return three


def test_voigt_fp64_vs_oracle(O, c1_small):
    L, lines = c1_small
    T, P = 296.0, 1.0
    Q = O.partition(L.E, L.g, T)
    out, st, _ = run_manager("voigt", lines, T, P, Q, 0.0, 10000.0, 1_000_000, 25.0,
                             precision="fp64")
    ref, ne, _ = run_oracle(O, "voigt", lines, T, P, Q, 0.0, 10000.0, 1_000_000, 25.0)
    assert st["n_eval"] == ne
    check(out, ref, 1e-10, 1e-10, "voigt fp64")


def test_windows_and_sort_bit_exact(O, c1_small):
    L, lines = c1_small
    nu = lines[0]
    T, P = 296.0, 1.0
    Q = O.partition(L.E, L.g, T)
    _, _, m = run_manager("voigt", lines, T, P, Q, 0.0, 10000.0, 1_000_000, 25.0)
    nus, ib, ie = m.line_windows(nu.shape[0])
    assert np.array_equal(nus, np.sort(nu, kind="stable"))
    rib, rie = O.windows(nus, 25.0, 0.0, 10000.0 / 1_000_000, 0, 1_000_000)
    assert np.array_equal(ib, rib) and np.array_equal(ie, rie)


def test_windows_round_half_boundaries(O):
    # positions whose ((nu - cut) - start)/dnu is exactly k + 0.5 in FP64
    start, dnu, N, cut = 0.0, 0.5, 4000, 1.0
    k = np.arange(10, 3000, 7, dtype=np.float64)
    nu = (k + 0.5) * dnu + cut  # exact binary fractions
    n = nu.shape[0]
    lines = (nu, np.zeros(n), np.ones(n), np.ones(n, np.int32), np.full(n, 0.05), np.full(n, 0.5))
    _, _, m = run_manager("voigt", lines, 296.0, 1.0, 1.0, start, start + N * dnu, N, cut)
    nus, ib, ie = m.line_windows(n)
    rib, rie = O.windows(nus, cut, start, dnu, 0, N)
    assert np.array_equal(ib, rib) and np.array_equal(ie, rie)
    assert np.all(ib == (k + 1).astype(np.int32))  # half away from zero


@pytest.fixture(scope="module")
def c0_full(O):
    L = synth.workload_list("doppler_2e5")
    lines = L.lines()
    T = 1000.0
    Q = O.partition(L.E, L.g, T)
    ref, ne, _ = run_oracle(O, "doppler", lines, T, 0.0, Q, 0.0, 10000.0, 1_000_000, 1.0)
    return lines, T, Q, ref, ne


# the bin-averaged Gaussian at every tile width: 64-point tiles (the shape the
# bench's autotuner picks) run the window-edge segments with one-sided masks
# and the covering lines unmasked; 512-point tiles are wider than the
# 200-point windows, so every line is masked at both edges
@pytest.mark.parametrize("ppl,split,order", [(1, 1, 1), (1, 4, 3), (2, 4, 3), (4, 1, 1),
                                             (8, 16, 1)])
def test_doppler_full_config_vs_oracle(c0_full, ppl, split, order):
    lines, T, Q, ref, ne = c0_full
    out, st, m = run_manager("doppler", lines, T, 0.0, Q, 0.0, 10000.0, 1_000_000, 1.0,
                             ppl=ppl, split=split, order=order)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"doppler fp32c ppl={ppl} split={split} order={order}")
    # each tile sums only the lines centred within reach of it, [f2, f3) of
    # its window's lines [lo, hi): at 1000 K a line at 1e4 cm^-1 reaches 23
    # points of 0.01 cm^-1, against the 100 of its cutoff (the wider the
    # tile, the smaller the share of lines the cut removes)
    it = m.pass_items()
    lo, hi, w0, w1, f2, f3 = (it[:, k] for k in (0, 1, 2, 3, 10, 11))
    assert np.all((lo <= f2) & (f2 <= f3) & (f3 <= np.maximum(hi, lo)))
    kept, window = int(np.sum(f3 - f2)), int(np.sum(np.maximum(hi - lo, 0)))
    log(f"  lines within reach: {kept} of {window} line-tile pairs")
    assert kept < (0.75 if ppl <= 2 else 0.9) * window


def test_doppler_fp64_vs_oracle(c0_full):
    lines, T, Q, ref, ne = c0_full
    out, st, _ = run_manager("doppler", lines, T, 0.0, Q, 0.0, 10000.0, 1_000_000, 1.0,
                             precision="fp64")
    assert st["n_eval"] == ne
    check(out, ref, 1e-10, 1e-10, "doppler fp64")


@pytest.mark.parametrize("ppl,split", [(4, 1), (1, 1), (2, 16)])
def test_gaussian_vs_oracle(O, c1_small, ppl, split):
    L, lines = c1_small
    T, hwhm = 296.0, 0.05
    Q = O.partition(L.E, L.g, T)
    ref, ne, _ = run_oracle(O, "gaussian", lines, T, 1.0, Q, 0.0, 10000.0, 200_000, 0.5,
                            hwhm=hwhm)
    precs = (("fp32c", 1e-4, 1e-6), ("fp64", 1e-10, 1e-10)) if ppl == 4 else \
        (("fp32c", 1e-4, 1e-6),)
    for prec, tol, btol in precs:
        out, st, _ = run_manager("gaussian", lines, T, 1.0, Q, 0.0, 10000.0, 200_000, 0.5,
                                 precision=prec, hwhm=hwhm, ppl=ppl, split=split)
        assert st["n_eval"] == ne
        check(out, ref, tol, btol, f"gaussian {prec} ppl={ppl} split={split}")


@pytest.fixture(scope="module")
def quad_small():
    L = synth.workload_list("voigt_1e6", scale=0.002)
    return L, L.lines(), {}


# Voigt-Quad runs the Voigt tile segments: lines beyond x^2 + y^2 >= 1600 of
# every point take the staged region-I loop, the others the Zaghloul-Ali core
# and 4-point Gauss-Hermite wings per point; every tile width and split (at one
# point per lane the inner points are evaluated compacted across the wave).
# Regimes: 1 bar at 296 K (the 10^6-line workload), low pressure (cores of
# narrow lines), 100 bar (y > 40: region I everywhere) and P = 0 (y = 0).
@pytest.mark.parametrize("ppl,split,order", [(1, 1, 1), (1, 4, 3), (2, 4, 3), (4, 16, 1),
                                             (8, 1, 1)])
@pytest.mark.parametrize("T,P", [(1000.0, 0.05), (296.0, 1.0), (2000.0, 1e-4), (3000.0, 100.0),
                                 (296.0, 0.0)])
def test_voigt_quad_vs_oracle(O, quad_small, T, P, ppl, split, order):
    L, lines, cache = quad_small
    Q = O.partition(L.E, L.g, T)
    if (T, P) not in cache:
        cache[(T, P)] = run_oracle(O, "voigt_quad", lines, T, P, Q, 0.0, 10000.0, 500_000,
                                   10.0)[:2]
    ref, ne = cache[(T, P)]
    out, st, _ = run_manager("voigt_quad", lines, T, P, Q, 0.0, 10000.0, 500_000, 10.0, ppl=ppl,
                             split=split, order=order)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"voigt-quad T={T} P={P} ppl={ppl} split={split} order={order}")


def test_voigt_quad_fp64_vs_oracle(O, quad_small):
    L, lines, cache = quad_small
    T, P = 1000.0, 0.05
    Q = O.partition(L.E, L.g, T)
    if (T, P) not in cache:
        cache[(T, P)] = run_oracle(O, "voigt_quad", lines, T, P, Q, 0.0, 10000.0, 500_000,
                                   10.0)[:2]
    ref, ne = cache[(T, P)]
    out, st, _ = run_manager("voigt_quad", lines, T, P, Q, 0.0, 10000.0, 500_000, 10.0,
                             precision="fp64")
    assert st["n_eval"] == ne
    check(out, ref, 1e-10, 1e-10, "voigt-quad fp64")


@pytest.mark.parametrize("dnu,ppl,split", [(1.0, 1, 1), (1.0, 1, 16), (0.1, 1, 4), (1.0, 4, 1)])
def test_voigt_quad_coarse_grids(O, dnu, ppl, split):
    # windows shorter than a tile: every line is an edge line, checked per point
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.001)
    lines = L.lines()
    T, P = 2500.0, 1e-3
    N = int(round(14500.0 / dnu))
    Q = O.partition(L.E, L.g, T)
    ref, ne, _ = run_oracle(O, "voigt_quad", lines, T, P, Q, 500.0, 15000.0, N, 25.0, mass=28.0)
    out, st, _ = run_manager("voigt_quad", lines, T, P, Q, 500.0, 15000.0, N, 25.0, mass=28.0,
                             ppl=ppl, split=split)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"voigt-quad dnu={dnu} ppl={ppl} split={split}")


@pytest.fixture(scope="module")
def c3_small():
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.001)
    return L, L.lines(), {}


@pytest.mark.parametrize("ppl,split,order", [(4, 1, 1), (2, 4, 3), (1, 16, 1)])
@pytest.mark.parametrize("T,P", [(2000.0, 1e-4), (200.0, 1e-5), (3000.0, 100.0), (296.0, 0.0)])
def test_voigt_pressure_regimes(O, c3_small, T, P, ppl, split, order):
    # low P: Humlicek regions III/IV at line centres; high P: region I everywhere;
    # P = 0: the Doppler limit (y = 0).
    L, lines, cache = c3_small
    Q = O.partition(L.E, L.g, T)
    out, st, m = run_manager("voigt", lines, T, P, Q, 500.0, 15000.0, 1_450_000, 25.0,
                             mass=28.0, ppl=ppl, split=split, order=order)
    if (T, P) not in cache:
        cache[(T, P)] = run_oracle(O, "voigt", lines, T, P, Q, 500.0, 15000.0, 1_450_000, 25.0,
                                   mass=28.0)[:2]
    ref, ne = cache[(T, P)]
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"voigt T={T} P={P} ppl={ppl} split={split} order={order}")


@pytest.mark.parametrize("dnu,ppl,split", [(1.0, 1, 1), (1.0, 1, 16), (0.1, 1, 4), (0.1, 1, 16),
                                           (1.0, 4, 1)])
@pytest.mark.parametrize("T,P", [(2500.0, 1e-4), (900.0, 50.0)])
def test_voigt_coarse_grids(O, T, P, dnu, ppl, split):
    # windows of +-25 points at 1 cm^-1 are shorter than a tile, so every line
    # is masked at both edges, and a line has one or two points outside
    # Humlicek region I: the compacted W4 path of one-point-per-lane tiles
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.002)
    lines = L.lines()
    Q = O.partition(L.E, L.g, T)
    N = int(round(14500.0 / dnu))
    out, st, m = run_manager("voigt", lines, T, P, Q, 500.0, 15000.0, N, 25.0, mass=28.0,
                             ppl=ppl, split=split)
    ref, ne, _ = run_oracle(O, "voigt", lines, T, P, Q, 500.0, 15000.0, N, 25.0, mass=28.0)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"voigt T={T} P={P} dnu={dnu} ppl={ppl} split={split}")
    if split > 1 and dnu == 0.1 and P < 1.0:
        # 6.4 cm^-1 tiles of the densest part of the list hold ~18 lines
        # centred near them (at 1 cm^-1 no window covers a whole tile)
        assert three_piece_tiles(m, st, ppl, split) > 0


def test_reference_cpu_conventions(O, c1_small):
    # the reference CPU path: no pressure in gammaL, cutoff min(25P, 100),
    # 1/w = 1/(nu*dpwcoef); and the textbook width convention
    L, lines = c1_small
    T, P = 500.0, 0.5
    Q = O.partition(L.E, L.g, T)
    for wc in (0, 1):
        out, st, _ = run_manager("voigt", lines, T, P, Q, 0.0, 10000.0, 500_000, 0.0,
                                 pressure_mode=0, width_conv=wc)
        ref, ne, _ = run_oracle(O, "voigt", lines, T, P, Q, 0.0, 10000.0, 500_000,
                                min(25 * P, 100.0), pressure_mode=0, width_conv=wc)
        assert st["n_eval"] == ne
        check(out, ref, 1e-4, 1e-6, f"reference conventions width_conv={wc}")


def test_window_subset_and_chunk_accumulation(O, c1_small):
    L, (nu, E, A, g, gam, nex) = c1_small
    T, P, N = 296.0, 1.0, 300_000
    Q = O.partition(L.E, L.g, T)
    dnu = 10000.0 / N
    m = HipManager("voigt")
    m.InitializeVectors(N, 2)
    m.InitializeConstants(100.0, T, Q, dnu, 18.0153, P, 296.0, P_SCALE, cutoff=25.0)
    freq = np.arange(N) * dnu
    m.TransferFreq(freq)
    half = nu.shape[0] // 2
    # two chunks into result grid 0 (a chunk restarts the execute counter), a
    # window subset into grid 1 of the second chunk
    m.TransferVectors(half, E[:half], nu[:half], A[:half], g[:half], gam[:half], nex[:half])
    m.ExecuteCrossSection(N, half, 0, T, P, Q)
    m.TransferVectors(nu.shape[0] - half, E[half:], nu[half:], A[half:], g[half:], gam[half:],
                      nex[half:])
    m.ExecuteCrossSection(N, nu.shape[0] - half, 0, T, P, Q)
    m.ExecuteCrossSection(50_000, nu.shape[0] - half, 100_000, T, P, Q)
    out0 = np.zeros(N)
    out1 = np.zeros(N)
    m.TransferResults(freq, out0, N, 0)
    m.TransferResults(freq, out1, N, 1)
    ref, _, _ = run_oracle(O, "voigt", (nu, E, A, g, gam, nex), T, P, Q, 0.0, 10000.0, N, 25.0)
    check(out0, ref, 1e-4, 1e-6, "two chunks")
    sub = (nu[half:], E[half:], A[half:], g[half:], gam[half:], nex[half:])
    ref1, _, _ = run_oracle(O, "voigt", sub, T, P, Q, 0.0, 10000.0, N, 25.0, start_idx=100_000,
                            nwin=50_000)
    assert np.all(out1[:100_000] == 0) and np.all(out1[150_000:] == 0)
    check(out1, ref1, 1e-4, 1e-6, "window subset")


def test_execute_passes_equal_single_executes(O, c1_small):
    """Back-to-back executes alternate between the two streams (odd passes on
    the second), each with its own buffer set; other calls in between join
    the streams, and a window change rebuilds the positions while a pass may
    still run.  Every pass must equal the same execute run alone, bit for
    bit: eight passes at two temperatures, an item inspection (a kernel on
    the main stream) after the third, a stats read after the fifth and a
    window subset at the seventh."""
    L, (nu, E, A, g, gam, nex) = c1_small
    P, N = 1.0, 200_000
    dnu = 10000.0 / N
    Ts = [296.0, 700.0] * 4
    Qs = {T: float(O.partition(L.E, L.g, T)) for T in set(Ts)}
    wins = [(N, 0)] * 6 + [(60_000, 70_000), (N, 0)]
    m = HipManager("voigt")
    m.InitializeVectors(N, len(Ts))
    m.InitializeConstants(100.0, Ts[0], Qs[Ts[0]], dnu, 18.0153, P, 296.0, P_SCALE, cutoff=25.0)
    freq = np.arange(N) * dnu
    m.TransferFreq(freq)
    m.TransferVectors(nu.shape[0], E, nu, A, g, gam, nex)
    for k, T in enumerate(Ts):
        nwin, s0 = wins[k]
        m.ExecuteCrossSection(nwin, nu.shape[0], s0, T, P, Qs[T])
        if k == 2:
            assert m.pass_items().shape[0] > 0  # a kernel on the main stream
        if k == 4:
            m.stats()
    outs = []
    for k in range(len(Ts)):
        o = np.zeros(N)
        m.TransferResults(freq, o, N, k)
        outs.append(o)
    for k, T in enumerate(Ts):
        nwin, s0 = wins[k]
        alone, _, _ = run_manager("voigt", (nu, E, A, g, gam, nex), T, P, Qs[T], 0.0, 10000.0, N,
                                  25.0, start_idx=s0, nwin=nwin)
        assert np.array_equal(outs[k], alone), (k, T)
    log(f"  {len(Ts)} executes equal single executes bit for bit")


@pytest.mark.parametrize("ppl,split,order", [(4, 4, 3), (1, 1, 1), (2, 16, 1)])
def test_far_wings_pointwise(O, ppl, split, order):
    """Points whose whole value comes from far wings: 20,000 lines fill
    1000-1400 cm^-1 and 1600-2000 cm^-1.  The points at 1405-1420 cm^-1 are
    within the 25 cm^-1 cutoff of the lower band's upper lines only, 500 or
    more points from any of them (windows that cover their tiles, and windows
    that end inside them); those at 1580-1595 cm^-1 likewise of the upper
    band's lower lines (windows that start inside their tiles).  At 64- and
    256-point tiles those sums come from the far-field expansion alone.
    There the GPU must equal the FP64 oracle point by point, within 2e-6 of
    each point's own value, not only of the spectrum's peak."""
    rng = np.random.default_rng(31)
    n = 20000
    nu = np.sort(np.concatenate([rng.uniform(1000.0, 1400.0, n // 2),
                                 rng.uniform(1600.0, 2000.0, n // 2)]))
    lines = (nu, rng.uniform(0.0, 5000.0, n), 10 ** rng.uniform(-3.0, 1.0, n),
             rng.integers(1, 10, n).astype(np.int32), rng.uniform(0.02, 0.10, n),
             rng.uniform(0.3, 0.8, n))
    T, P, Q, N = 296.0, 1.0, 1000.0, 100_000
    out, st, _ = run_manager("voigt", lines, T, P, Q, 1000.0, 2000.0, N, 25.0, ppl=ppl,
                             split=split, order=order)
    ref, ne, _ = run_oracle(O, "voigt", lines, T, P, Q, 1000.0, 2000.0, N, 25.0)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, f"far wings ppl={ppl} split={split} order={order}")
    for gap, what in ((slice(40_500, 42_000), "1405-1420"), (slice(58_000, 59_500), "1580-1595")):
        assert np.all(ref[gap] > 0)
        rel = np.max(np.abs(out[gap] - ref[gap]) / ref[gap])
        log(f"  far wings ppl={ppl}: pointwise relative error at {what} cm^-1 {rel:.3e}")
        assert rel < 2e-6


def test_edge_cases(O):
    T, P, N = 296.0, 1.0, 1000
    start, end = 100.0, 110.0
    nu = np.array([0.0, 99.0, 100.0, 104.999, 105.0, 105.0, 109.99, 111.0, 500.0])
    n = nu.shape[0]
    lines = (nu, np.full(n, 100.0), np.full(n, 1.0), np.full(n, 3, np.int32), np.full(n, 0.07),
             np.full(n, 0.6))
    out, st, _ = run_manager("voigt", lines, T, P, 50.0, start, end, N, 0.5, ppl=1)
    ref, ne, _ = run_oracle(O, "voigt", lines, T, P, 50.0, start, end, N, 0.5)
    assert st["n_eval"] == ne
    check(out, ref, 1e-4, 1e-6, "edge lines")
    # a grid smaller than one wave, and an empty chunk
    out, st, _ = run_manager("voigt", lines, T, P, 50.0, 104.9, 105.1, 20, 0.5, ppl=1)
    ref, ne, _ = run_oracle(O, "voigt", lines, T, P, 50.0, 104.9, 105.1, 20, 0.5)
    check(out, ref, 1e-4, 1e-6, "tiny grid")
    empty = tuple(a[:0] for a in lines)
    out, st, _ = run_manager("voigt", empty, T, P, 50.0, start, end, N, 0.5)
    assert np.all(out == 0) and st["n_eval"] == 0


def test_deterministic_repeat(c1_small, O):
    L, lines = c1_small
    Q = O.partition(L.E, L.g, 296.0)
    a, _, _ = run_manager("voigt", lines, 296.0, 1.0, Q, 0.0, 10000.0, 1_000_000, 25.0)
    b, _, _ = run_manager("voigt", lines, 296.0, 1.0, Q, 0.0, 10000.0, 1_000_000, 25.0, order=0)
    assert np.array_equal(a, b)


def test_device_partition(O):
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.01)
    m = HipManager("voigt")
    m.load_list(L)
    Ts = np.array([200.0, 296.0, 1000.0, 3000.0])
    Q = m.partition(Ts)
    for t, q in zip(Ts, Q):
        assert abs(q - O.partition(L.E, L.g, t)) / q < 1e-13


def test_batch_equals_manager(O):
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.002)
    nu, E, A, g, gam, nex = L.lines()
    A32 = L.A.astype(np.float64)  # the resident list stores FP32 A, gamma, n
    lines = (nu, E, A32, g, L.gamma.astype(np.float64), L.nexp.astype(np.float64))
    tps = [(300.0, 1.0), (1500.0, 0.01), (2500.0, 30.0)]
    N, start, end = 145_000, 500.0, 15000.0
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 296.0, 1.0, (end - start) / N, 28.0, 1.0, 296.0, P_SCALE,
                          cutoff=25.0)
    m.load_list(L)
    Q = m.partition([t for t, _ in tps])
    batch = m.run_batch([(t, p, q) for (t, p), q in zip(tps, Q)], start, end, N)
    for k, (t, p) in enumerate(tps):
        out, _, _ = run_manager("voigt", lines, t, p, Q[k], start, end, N, 25.0, mass=28.0)
        assert np.array_equal(batch[k], out), k
        ref, _, _ = run_oracle(O, "voigt", lines, t, p, Q[k], start, end, N, 25.0, mass=28.0)
        check(batch[k], ref, 1e-4, 1e-6, f"batch T={t} P={p}")


def test_batch_line_selection_bit_exact(O):
    # selection of the resident-list path = the driver's filter (BaseProfile.cpp:152-157)
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.002)
    nu = L.lines()[0]
    N, start, end, hw = 145_000, 500.0, 15000.0, 100.0
    m = HipManager("voigt")
    m.InitializeConstants(hw, 296.0, 1.0, (end - start) / N, 28.0, 1.0, 296.0, P_SCALE,
                          cutoff=25.0)
    m.load_list(L)
    m.run_batch([(1000.0, 1.0)], start, end, N)
    first, count = m.pass_lines()
    sel = np.array([O.lib().gxo_select(v, start, end, hw) for v in nu], dtype=bool)
    want = np.sort(nu[sel], kind="stable")
    nus, ib, ie = m.line_windows(count)
    assert count == sel.sum() and np.array_equal(nus, want)
    rib, rie = O.windows(want, 25.0, start, (end - start) / N, 0, N)
    assert np.array_equal(ib, rib) and np.array_equal(ie, rie)
