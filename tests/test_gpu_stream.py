"""A list that stays in host memory, streamed through the device in chunks
(gexs_stream_xsec), against the oracle and against the same list resident in
HBM (gexs_run_batch).

The reference driver reads transitions in file order and hands a manager a
chunk whenever it holds the manager's capacity (Profiles/BaseProfile.cpp:174-181,
capacity from free device memory, Managers/Accel/GPU/GPUManager.cpp:98-101);
each chunk's cross-section adds into the grid.  The streamed result therefore
equals the one-pass result up to the grouping of the FP32 partial sums, and
its N_eval is exactly the one-pass N_eval.
"""
import os

import numpy as np
import pytest

from gexs_amd import synth
from gexs_amd._lib import GexsError
from gexs_amd.manager import HipManager

pytestmark = pytest.mark.gpu

P_SCALE = synth.P_SCALE_BAR
NTHREADS = min(16, os.cpu_count() or 1)
# the 10^8-transition hot list's distributions (2e6 states, FP32 A), at 1/500
NAME = "voigt_1e8_stream"
SPEC = synth.WORKLOADS[NAME]
T0, P0 = SPEC["T"], SPEC["P"]
GRID = (100.0, 30000.0, 2_990_000)


def log(msg):
    print(msg, flush=True)


def manager(precision="fp32c", profile="voigt"):
    start, end, N = GRID
    m = HipManager(profile, precision=precision)
    m.InitializeConstants(100.0, T0, 1.0, (end - start) / N, SPEC["mass"], P0, 296.0, P_SCALE,
                          cutoff=25.0)
    return m


@pytest.fixture(scope="module")
def small():
    return synth.workload_list(NAME, scale=0.002)  # 2e5 transitions, 4000 states


@pytest.fixture(scope="module")
def small_ref(O, small):
    L = small
    start, end, N = GRID
    nu, E, A, g, gam, nex = L.lines()
    keep = (nu >= start - 100.0) & (nu <= end + 100.0) & (nu != 0.0)
    o = np.argsort(nu[keep], kind="stable")
    cols = [a[keep][o] for a in (nu, E, A, g, gam, nex)]
    Q = O.partition(L.E, L.g, T0)
    p = O.params("voigt", T=T0, P=P0, Q=Q, mass=SPEC["mass"], cutoff=25.0, p_scale=P_SCALE)
    ref, ne = O.cross_section(p, start, (end - start) / N, N, *cols, nthreads=NTHREADS)
    return Q, ref, ne, int(keep.sum())


def check(out, ref, what):
    err = np.max(np.abs(out - ref)) / ref.max()
    band = abs(out.sum() - ref.sum()) / ref.sum()
    log(f"{what}: peak-relative {err:.3e}, band {band:.3e}")
    assert err <= 1e-4 and band <= 1e-6


# budgets: about 8 chunks, about 2 chunks, the whole list in one chunk
BUDGETS = [36 << 20, 56 << 20, 1 << 30]


def sub(L, n):
    return L.__class__(E=L.E, g=L.g, J=L.J, upper=L.upper[:n], lower=L.lower[:n], A=L.A[:n],
                       gamma=None if L.gamma is None else L.gamma[:n],
                       nexp=None if L.nexp is None else L.nexp[:n])


def budget_for(m, L, tps, grid, chunks):
    """The budget that gives `chunks` chunks: the device bytes of a one-chunk
    call are fixed + per-transition bytes x n, so two one-chunk calls of
    different n give both."""
    n1, n2 = L.n_trans, L.n_trans // 2
    _, a = m.stream_xsec(L, tps, *grid, mem_budget=1 << 40)
    _, b = m.stream_xsec(sub(L, n2), tps, *grid, mem_budget=1 << 40)
    assert a["n_chunks"] == 1 and b["n_chunks"] == 1
    per = (a["device_bytes"] - b["device_bytes"]) // (n1 - n2)
    fixed = a["device_bytes"] - per * n1
    return fixed + per * (-(-n1 // chunks))


@pytest.mark.parametrize("budget", BUDGETS)
def test_stream_vs_oracle(O, small, small_ref, budget):
    Q, ref, ne, nsel = small_ref
    m = manager()
    out, st = m.stream_xsec(small, [(T0, P0)], *GRID, mem_budget=budget)
    m.close()
    log(f"budget {budget >> 20} MiB: {st['n_chunks']} chunks of {st['chunk']}, "
        f"{st['device_bytes'] >> 20} MiB, {st['wall_ms']:.1f} ms")
    assert st["device_bytes"] <= budget
    assert st["n_transitions"] == small.n_trans and st["n_lines"] == nsel
    assert st["n_eval"] == ne
    if budget == BUDGETS[0]:
        assert st["n_chunks"] >= 6
    if budget == BUDGETS[-1]:
        assert st["n_chunks"] == 1
    check(out[0], ref, f"stream {st['n_chunks']} chunks")


@pytest.mark.parametrize("ppl,split,order", [(2, 4, 3), (8, 1, 1), (4, 16, 0)])
def test_stream_equals_resident(small, ppl, split, order):
    """Several (T,P) at once (Q from the streamed states on the device for two
    of them) at three launch shapes: each row as the resident list's batch
    gives it, to the grouping of the FP32 partial sums, with the same N_eval."""
    tps = [(T0, P0), (700.0, 3e-3), (2600.0, 20.0)]
    m = manager()
    m.set_tuning(ppl, order, split)
    out, st = m.stream_xsec(small, tps, *GRID, mem_budget=budget_for(m, small, tps, GRID, 7))
    assert st["n_chunks"] == 7
    m.load_list(small)
    res = m.run_batch(tps, *GRID)
    ne_res = m.stats()["n_eval"]
    m.close()
    assert st["n_eval"] == ne_res
    for k in range(len(tps)):
        d = np.max(np.abs(out[k] - res[k])) / res[k].max()
        log(f"{tps[k]}: streamed vs resident {d:.2e}")
        assert d <= 2e-6, k


def test_stream_keeps_the_resident_list(small):
    """A stream call leaves the resident list of load_list untouched: the
    batch after it equals the batch before it bit for bit."""
    m = manager()
    m.load_list(sub(small, 50_000))
    before = m.run_batch([(T0, P0)], *GRID)
    m.stream_xsec(small, [(T0, P0)], *GRID, mem_budget=BUDGETS[0])
    after = m.run_batch([(T0, P0)], *GRID)
    m.close()
    assert np.array_equal(before, after)


def test_stream_nu_sorted_file_and_empty_chunks(small):
    """A list stored in nu order (as ExoMol lists split into wavenumber-range
    files are) streams as nu-range chunks; over a narrow grid most chunks
    select no line.  The result equals the resident list's."""
    nu = small.E[small.upper.astype(np.int64)] - small.E[small.lower.astype(np.int64)]
    o = np.argsort(nu, kind="stable")
    L = small.__class__(E=small.E, g=small.g, J=small.J, upper=small.upper[o],
                        lower=small.lower[o], A=small.A[o], gamma=small.gamma[o],
                        nexp=small.nexp[o])
    grid = (4000.0, 4500.0, 50_000)
    m = manager()
    out, st = m.stream_xsec(L, [(T0, P0)], *grid,
                            mem_budget=budget_for(m, L, [(T0, P0)], grid, 8))
    m.load_list(L)
    res = m.run_batch([(T0, P0)], *grid)
    ne_res = m.stats()["n_eval"]
    m.close()
    log(f"nu-sorted: {st['n_chunks']} chunks, {st['n_lines']} lines selected")
    assert st["n_chunks"] == 8 and 0 < st["n_lines"] < st["n_transitions"] // 4
    assert st["n_eval"] == ne_res
    assert np.max(np.abs(out[0] - res[0])) / res[0].max() <= 2e-6


def test_stream_fp64_columns_and_precision(small):
    """FP64 A / gamma / n columns stream as they are (the chunk gather reads
    them without narrowing), in the FP32C and the plain FP64 precision."""
    L = small.__class__(E=small.E, g=small.g, J=small.J, upper=small.upper, lower=small.lower,
                        A=small.A.astype(np.float64) * (1.0 + 1e-9), gamma=small.gamma.astype(np.float64),
                        nexp=small.nexp.astype(np.float64))
    for prec in ("fp32c", "fp64"):
        m = manager(prec)
        out, st = m.stream_xsec(L, [(T0, P0)], *GRID, mem_budget=BUDGETS[0])
        m.load_list(L)
        res = m.run_batch([(T0, P0)], *GRID)
        ne_res = m.stats()["n_eval"]
        m.close()
        d = np.max(np.abs(out[0] - res[0])) / res[0].max()
        log(f"{prec} FP64 columns: {st['n_chunks']} chunks, streamed vs resident {d:.2e}")
        assert st["n_eval"] == ne_res
        assert d <= (2e-6 if prec == "fp32c" else 1e-12)


def test_stream_doppler(O):
    """A Doppler-only list (no gamma / n columns) streamed in chunks against
    the oracle (the plain Doppler workload's list and grid, cutoff 1 cm^-1)."""
    spec = synth.WORKLOADS["doppler_2e5"]
    L = synth.workload_list("doppler_2e5")
    L = L.__class__(E=L.E, g=L.g, J=L.J, upper=L.upper, lower=L.lower, A=L.A, gamma=None,
                    nexp=None)
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    T = spec["T"]
    m = HipManager("doppler")
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], 0.0, 296.0, 1.0, cutoff=spec["cutoff"])
    out, st = m.stream_xsec(L, [(T, 0.0)], start, end, N,
                            mem_budget=budget_for(m, L, [(T, 0.0)], (start, end, N), 5))
    m.close()
    assert st["n_chunks"] == 5
    nu = L.E[L.upper.astype(np.int64)] - L.E[L.lower.astype(np.int64)]
    keep = (nu >= start - 100.0) & (nu <= end + 100.0) & (nu != 0.0)
    o = np.argsort(nu[keep], kind="stable")
    E = L.E[L.lower.astype(np.int64)][keep][o]
    A = L.A.astype(float)[keep][o]
    g = L.g[L.upper.astype(np.int64)][keep][o]
    Q = O.partition(L.E, L.g, T)
    p = O.params("doppler", T=T, P=0.0, Q=Q, mass=spec["mass"], cutoff=spec["cutoff"])
    ref, ne = O.cross_section(p, start, dnu, N, nu[keep][o], E, A, g, nthreads=NTHREADS)
    assert st["n_eval"] == ne
    check(out[0], ref, f"doppler stream {st['n_chunks']} chunks")


def test_stream_errors(small):
    m = manager()
    # a budget below 4096 transitions per chunk is refused before any work
    with pytest.raises(GexsError, match="mem_budget"):
        m.stream_xsec(small, [(T0, P0)], *GRID, mem_budget=1 << 20)
    # an index outside the state table stops the call on the host, before the
    # chunk that holds it reaches the device
    up = small.upper.copy()
    up[-3] = small.n_states
    bad = small.__class__(E=small.E, g=small.g, J=small.J, upper=up, lower=small.lower,
                          A=small.A, gamma=small.gamma, nexp=small.nexp)
    with pytest.raises(GexsError, match="outside the state table"):
        m.stream_xsec(bad, [(T0, P0)], *GRID, mem_budget=BUDGETS[0])
    # an empty list gives zero grids; the ctx still works afterwards
    empty = sub(small, 0)
    out, st = m.stream_xsec(empty, [(T0, P0)], *GRID, mem_budget=BUDGETS[0])
    assert st["n_chunks"] == 0 and st["n_eval"] == 0 and not np.any(out)
    out, st = m.stream_xsec(small, [(T0, P0)], *GRID, mem_budget=BUDGETS[1])
    assert st["n_chunks"] >= 2 and out.max() > 0
    m.close()
