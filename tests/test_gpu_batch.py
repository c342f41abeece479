"""Batched (T,P) table: one launch sequence for many cross-sections of one
resident list; persistent launch; HIP-graph capture and replay; autotuner.

Graph replays are interleaved with eager work and synchronisations and each
replay is compared with the eager result bit for bit (the captured work holds
no memset nodes: buffers are cleared by fill kernels).
"""
import numpy as np
import pytest

from gexs_amd import synth
from gexs_amd.manager import HipManager

pytestmark = pytest.mark.gpu

P_SCALE = synth.P_SCALE_BAR
GRID = (500.0, 15000.0, 145_000)


@pytest.fixture(scope="module")
def batch_mgr():
    L = synth.workload_list("voigt_tp_table_1e7", scale=0.002)
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 296.0, 1.0, (GRID[1] - GRID[0]) / GRID[2], 28.0, 1.0, 296.0,
                          P_SCALE, cutoff=25.0)
    m.load_list(L)
    return L, m


def table(k=12):
    tp = synth.tp_table()
    idx = np.linspace(0, len(tp) - 1, k).astype(int)
    return [tp[i] for i in idx]


def test_table_vs_oracle(batch_mgr, O):
    L, m = batch_mgr
    tps = table(6)
    out = m.run_batch(tps, *GRID)
    nu, E, A, g, _, _ = L.lines()
    A, gam, nex = L.A.astype(float), L.gamma.astype(float), L.nexp.astype(float)
    keep = (nu >= GRID[0] - 100) & (nu <= GRID[1] + 100)
    for k, (T, P) in enumerate(tps):
        Q = O.partition(L.E, L.g, T)
        p = O.params("voigt", T=T, P=P, Q=Q, mass=28.0, cutoff=25.0, p_scale=P_SCALE)
        ref, ne = O.cross_section(p, GRID[0], (GRID[1] - GRID[0]) / GRID[2], GRID[2], nu[keep],
                                  E[keep], A[keep], g[keep], gam[keep], nex[keep], nthreads=8)
        err = np.max(np.abs(out[k] - ref)) / ref.max()
        band = abs(out[k].sum() - ref.sum()) / ref.sum()
        print(f"T={T:.1f} P={P:.2e}: err {err:.2e} band {band:.2e}", flush=True)
        assert err <= 1e-4 and band <= 1e-6


def test_persistent_equals_wave_per_tile(batch_mgr):
    L, m = batch_mgr
    tps = table(12)
    m.set_launch(0, False)
    a = m.run_batch(tps, *GRID)
    m.set_launch(1024, False)
    b = m.run_batch(tps, *GRID)
    m.set_launch(0, False)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("split", [1, 4, 16])
def test_launch_modes_agree_bitwise(batch_mgr, split):
    """For one tile width and split, every launch shape and tile order sums
    each point in the same order."""
    L, m = batch_mgr
    tps = table(6)
    outs = []
    for order, pb in [(0, 0), (1, 0), (3, 0), (0, 1024), (1, 2048), (2, 1024)]:
        m.set_tuning(4, order, split)
        m.set_launch(pb, False)
        outs.append(m.run_batch(tps, *GRID))
    m.set_tuning(4, 1, 1)
    m.set_launch(0, False)
    for k, o in enumerate(outs[1:]):
        assert np.array_equal(o, outs[0]), k


@pytest.mark.parametrize("ppl", [1, 2, 4])
@pytest.mark.parametrize("split", [4, 16])
def test_split_close_to_single_wave(batch_mgr, ppl, split):
    """Waves of a split tile take equal shares of the lines before, inside
    and after the tile's near-centre lines (three pieces each when the tile
    has at least `split` near lines, as many tiles of this table do): every
    line is summed once, so only the grouping of the FP32 partial sums
    differs from one wave per tile."""
    L, m = batch_mgr
    tps = table(6)
    m.set_tuning(ppl, 1, 1)
    a = m.run_batch(tps, *GRID)
    m.set_tuning(ppl, 1, split)
    b = m.run_batch(tps, *GRID)
    it = m.pass_items()
    three = int(np.sum((it[:, 4] < it[:, 7]) & (it[:, 6] - it[:, 5] >= split)))
    print(f"ppl={ppl} split={split}: {three} of {it.shape[0]} items take three-piece shares",
          flush=True)
    assert three > 0
    m.set_tuning(4, 1, 1)
    assert np.max(np.abs(a - b)) / np.max(a) < 1e-6
    # and the integrated band strengths within the band tolerance
    assert np.max(np.abs(a.sum(axis=1) - b.sum(axis=1)) / a.sum(axis=1)) < 1e-6


def test_graph_follows_tuning_and_buffer_changes(batch_mgr):
    """A captured batch is replayed only while everything it was captured
    with still holds: changing the split, or growing the buffers with a
    larger batch in between, makes the next graph run capture afresh."""
    L, m = batch_mgr
    tps = table(4)
    m.set_launch(1024, False)
    eager = {}
    for split in (1, 4):
        m.set_tuning(4, 1, split)
        eager[split] = m.run_batch(tps, *GRID)
    m.set_launch(1024, True)
    for split in (1, 4, 1):
        m.set_tuning(4, 1, split)
        assert np.array_equal(m.run_batch(tps, *GRID), eager[split]), split
    m.set_launch(1024, False)
    m.run_batch(table(12), *GRID)  # grows the per-pass tables
    m.set_launch(1024, True)
    assert np.array_equal(m.run_batch(tps, *GRID), eager[1])
    # one (T,P) more grows the line records' workgroup-sum scratch by a whole
    # job's share, while the 256-byte-rounded tables may keep their size
    m.set_launch(1024, False)
    m.run_batch(table(13), *GRID)
    m.set_launch(1024, True)
    assert np.array_equal(m.run_batch(tps, *GRID), eager[1])
    m.set_tuning(4, 1, 1)
    m.set_launch(0, False)


def test_device_resident_output(batch_mgr):
    """out="device" leaves the results in the library's buffer, equal to the
    host copy of the same batch."""
    import ctypes as C
    L, m = batch_mgr
    tps = table(5)
    host = m.run_batch(tps, *GRID)
    assert m.run_batch(tps, *GRID, out="device") is None
    ptr, cap = m.batch_device_out()
    assert ptr and cap >= host.size
    hip = C.CDLL("libamdhip64.so.7")  # the runtime the library is linked to
    back = np.empty_like(host)
    assert hip.hipMemcpy(back.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                         C.c_size_t(host.nbytes), 2) == 0  # hipMemcpyDeviceToHost
    assert np.array_equal(back, host)


def test_graph_replays_equal_eager(batch_mgr):
    L, m = batch_mgr
    tps = table(12)
    m.set_launch(0, False)
    eager = m.run_batch(tps, *GRID)
    other = [(t * 1.1, p) for t, p in tps]
    eager_other = m.run_batch(other, *GRID)
    m.set_launch(0, True)
    for rep in range(3):
        g = m.run_batch(tps, *GRID)
        assert np.array_equal(g, eager), rep
        # the same graph with another (T,P) table, then eager work in between
        g2 = m.run_batch(other, *GRID)
        assert np.array_equal(g2, eager_other), rep
        m.set_launch(0, False)
        m.run_batch(tps[:3], *GRID)
        m.sync()
        m.set_launch(0, True)
    m.set_launch(0, False)


def test_autotune_picks_a_candidate(batch_mgr):
    L, m = batch_mgr
    tps = table(8)
    ref = m.run_batch(tps, *GRID)
    best, cands = m.autotune(tps, *GRID)
    # ppl {1,2,4,8} x waves per tile {1,4,16} x 9 launch shapes
    assert len(cands) == 108
    assert best["kernel_ms"] == min(c["kernel_ms"] for c in cands)
    print("autotune best", best, flush=True)
    out = m.run_batch(tps, *GRID)
    # a different wave-tile width changes summation grouping only
    assert np.max(np.abs(out - ref)) / np.max(ref) < 1e-6
    m.set_tuning(4, 1)
    m.set_launch(0, False)
