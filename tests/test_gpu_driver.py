"""The C++ driver end to end: keyword input -> reader -> selection -> chunks ->
HipManager -> output text, against the FP64 oracle.

ExoMol inputs are written from a seeded list (states, unsorted transitions and
a broadener file); HITRAN inputs are 160-character records built here.
"""
import os
import subprocess

import numpy as np
import pytest

from gexs_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gexs_amd", "gexs_xsec")


def run_cli(text, cwd, raw=False):
    r = subprocess.run([CLI], input=text, capture_output=True, text=True, cwd=cwd, timeout=300)
    assert r.returncode == 0, r.stderr
    if raw:
        return r.stdout, r.stderr
    blocks = {}
    band = None
    for line in r.stdout.splitlines():
        if line.startswith("["):
            k, rest = line[1:].split("]", 1)
            f, s = rest.split()
            blocks.setdefault(int(k), []).append(float(s))
        elif line.startswith("Our summed intensity"):
            band = line
    return {k: np.array(v) for k, v in blocks.items()}, band, r.stderr


def band_sums(stderr):
    """(summed line intensity, sum(sigma) * dnu) as gexs_xsec reports them at
    full precision on stderr (the stdout report line prints 7 digits)."""
    for line in stderr.splitlines():
        if line.startswith("band strength:"):
            w = line.split()
            return float(w[3]), float(w[5])
    raise AssertionError("no band strength line on stderr")


def report_numbers(line):
    """The three numbers of the report line (BaseProfile.cpp:289)."""
    import re
    m = re.match(r"Our summed intensity = +(\S+) compared to the integrated intensity = +(\S+) "
                 r"with a difference of +(\S+) %$", line)
    assert m, line
    return tuple(float(v) for v in m.groups())


def broadeners(J):
    """a0 per lower J and a1 for a few (J'', J') pairs; the reader's lookup."""
    a0 = {j: (0.03 + 0.001 * j, 0.5 + 0.004 * j) for j in range(0, 61, 1) if j % 3 != 2}
    a1 = {(j, j + 1): (0.09, 0.7) for j in range(0, 60, 5)}
    return a0, a1


def gamma_n(L, a0, a1, dg=0.07, dn=0.5):
    Jl = L.J[L.lower.astype(np.int64)]
    Ju = L.J[L.upper.astype(np.int64)]
    g = np.full(L.n_trans, dg)
    n = np.full(L.n_trans, dn)
    for i, (jl, ju) in enumerate(zip(Jl, Ju)):
        if (jl, ju) in a1:
            g[i], n[i] = a1[(jl, ju)]
        elif jl in a0:
            g[i], n[i] = a0[jl]
    return g, n


def test_exomol_driver_vs_oracle(O, tmp_path):
    L = synth.workload_list("voigt_1e6", scale=0.01)
    a0, a1 = broadeners(L.J)
    with open(tmp_path / "m.states", "w") as f:
        for i, (e, g, j) in enumerate(zip(L.E, L.g, L.J)):
            f.write(f"{i + 1:12d} {float(e)!r} {int(g):6d} {int(j):7d}\n")
    with open(tmp_path / "m.trans", "w") as f:
        for u, l, a in zip(L.upper, L.lower, L.A):
            f.write(f"{int(u) + 1} {int(l) + 1} {float(a)!r}\n")
    with open(tmp_path / "m.broad", "w") as f:
        for j, (g, n) in a0.items():
            f.write(f"a0 {g!r} {n!r} {j}\n")
        for (jl, ju), (g, n) in a1.items():
            f.write(f"a1 {g!r} {n!r} {jl} {ju}\n")
    tps = [(296.0, 1.0), (1200.0, 0.2)]
    N = 200_000
    cfg = "\n".join([
        "EXOMOL", "STATE-FILE m.states", "TRANS-FILES", "m.trans", "END", "BROADENERS",
        "m.broad", "END", "PROFILE VOIGT", f"N-POINTS {N}", "NU-RANGE 0.0 10000.0",
        "MEAN-MASS 18.0153", "GAMMA 0.07", "GAMMA-N 0.5", "CUTOFF 25", "T-REF 296",
        f"PRESSURE-SCALE {1 / 1.01325!r}", "CHUNK-LINES 1500", "TEMPERATURE-PRESSURE-GRID",
        *[f"{t} {p}" for t, p in tps], "END", ""])
    out, band, err = run_cli(cfg, tmp_path)
    assert "selected 10000, chunks 7" in err  # 10000 selected lines / 1500 per chunk
    # the same transitions bzip2-compressed (two concatenated streams): same bits
    import bz2
    raw = (tmp_path / "m.trans").read_bytes()
    cut = raw.index(b"\n", len(raw) // 2) + 1
    (tmp_path / "m.trans.bz2").write_bytes(bz2.compress(raw[:cut]) + bz2.compress(raw[cut:]))
    out_bz, _, _ = run_cli(cfg.replace("\nm.trans\n", "\nm.trans.bz2\n"), tmp_path)
    assert all(np.array_equal(out_bz[k], out[k]) for k in out)
    nu, E, A, g, _, _ = L.lines()
    gam, nex = gamma_n(L, a0, a1)
    keep = (nu >= -100.0) & (nu <= 10100.0) & (nu != 0)
    args = (nu[keep], E[keep], A[keep], g[keep], gam[keep], nex[keep])
    for k, (T, P) in enumerate(tps):
        Q = O.partition(L.E, L.g, T)
        p = O.params("voigt", T=T, P=P, Q=Q, cutoff=25.0, p_scale=1 / 1.01325)
        ref, _ = O.cross_section(p, 0.0, 10000.0 / N, N, *args, nthreads=8)
        e = np.max(np.abs(out[k] - ref)) / ref.max()
        bnd = abs(out[k].sum() - ref.sum()) / ref.sum()
        print(f"T={T} P={P}: peak-rel {e:.2e} band {bnd:.2e}", flush=True)
        assert e <= 1e-4 and bnd <= 1e-6
        if k == 0:
            ref0 = ref
    # the band-strength report (BaseProfile.cpp:250-267, 283-291): the summed
    # line intensity of the first (T,P) equals the oracle's to FP64 rounding,
    # and sum(sigma) * dnu the oracle's within the band tolerance
    T0 = tps[0][0]
    total = O.total_intensity(T0, O.partition(L.E, L.g, T0), *args[:4])
    summed, integ = band_sums(err)
    assert abs(summed - total) <= 1e-12 * total, (summed, total)
    want = ref0.sum() * (10000.0 / N)
    assert abs(integ - want) <= 1e-6 * want, (integ, want)
    a, b, pct = report_numbers(band)
    assert abs(a - summed) <= 5e-7 * summed and abs(b - integ) <= 5e-7 * integ
    assert abs(pct - (summed - integ) * 100.0 / summed) <= 5e-6


def parse_output(text):
    """Lines of gexs_xsec / OutputProfile text: ("band", numbers),
    ("head", text), ("blank", "") or ("pt", k, frequency text, sigma)."""
    rows = []
    for line in text.splitlines():
        if line.startswith("Our summed intensity"):
            rows.append(("band", report_numbers(line)))
        elif line.startswith("["):
            k, rest = line[1:].split("]", 1)
            f, sg = rest.split()
            rows.append(("pt", int(k), f, float(sg)))
        elif line.strip() == "":
            rows.append(("blank", ""))
        else:
            rows.append(("head", line))
    return rows


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_golden_text_output(O, tmp_path, precision):
    """gexs_xsec on a tiny ExoMol list (tests/golden/xsec_tiny, written by
    make_xsec_golden.py from the reference's own ComputeVoigtTP and
    compute_partition, compiled from its sources) under the reference CPU
    conventions.  The output has the reference's layout line for line
    (OutputProfile, BaseProfile.cpp:269-301): the report line, one header per
    (T,P), one "[i] nu sigma" line per point with the same frequency text;
    sigma within 1e-4 of the peak and 1e-6 in band strength (FP32C), or
    2e-8 of the peak (FP64: the 9 digits of the text); the summed line
    intensity as printed."""
    import shutil
    gold = os.path.join(ROOT, "tests", "golden", "xsec_tiny")
    for f in ("tiny.states", "tiny.trans"):
        shutil.copy(os.path.join(gold, f), tmp_path / f)
    cfg = open(os.path.join(gold, "input.txt")).read()
    if precision == "fp64":
        cfg = cfg.replace("PROFILE VOIGT", "PROFILE VOIGT\nPRECISION fp64")
    out, err = run_cli(cfg, tmp_path, raw=True)
    assert "selected 33," in err
    got = parse_output(out)
    want = parse_output(open(os.path.join(gold, "expected.txt")).read())
    assert len(got) == len(want)
    assert [r[0] for r in got] == [r[0] for r in want]
    for g, w in zip(got, want):
        if g[0] == "head":
            assert g == w
        elif g[0] == "pt":
            assert g[1:3] == w[1:3]
    for k in (0, 1):
        gs = np.array([r[3] for r in got if r[0] == "pt" and r[1] == k])
        ws = np.array([r[3] for r in want if r[0] == "pt" and r[1] == k])
        e = np.max(np.abs(gs - ws)) / ws.max()
        b = abs(gs.sum() - ws.sum()) / ws.sum()
        print(f"{precision} block {k}: peak-rel {e:.2e} band {b:.2e}", flush=True)
        if precision == "fp64":
            assert e <= 2e-8
        else:
            assert e <= 1e-4 and b <= 1e-6
    (a, b_, _), (wa, wb, _) = got[0][1], want[0][1]
    assert abs(a - wa) <= 5e-7 * wa  # the same FP64 sum over the same lines, at print precision
    assert abs(b_ - wb) <= (1e-6 if precision == "fp32c" else 1e-8) * wb + 5e-7 * wb


def par_record(nu, S, A, gair, gself, E, nair, gup):
    """One 160-character HITRAN record with the reader's field widths."""
    f = [f"{1:2d}", "1", f"{nu:12.6f}", f"{S:10.3E}", f"{A:10.3E}", f"{gair:5.4f}"[1:].rjust(5),
         f"{gself:5.3f}".rjust(5), f"{E:10.4f}", f"{nair:4.2f}", f"{0.0:8.6f}", " " * 15,
         " " * 15, " " * 15, " " * 15, "000000", " " * 12, " ", f"{gup:7.1f}", f"{gup:7.1f}"]
    rec = "".join(f)
    assert len(rec) == 160
    return rec


def test_hitran_driver_vs_oracle(O, tmp_path):
    rng = np.random.default_rng(11)
    n = 400
    nu = np.sort(rng.uniform(1000.0, 1100.0, n))
    E = rng.uniform(0.0, 3000.0, n)
    A = 10 ** rng.uniform(-3, 1, n)
    gair = rng.uniform(0.05, 0.09, n)
    gself = rng.uniform(0.2, 0.4, n)
    nair = rng.uniform(0.5, 0.8, n)
    gup = rng.integers(1, 30, n).astype(float)
    recs = [par_record(*row) for row in zip(nu, np.zeros(n), A, gair, gself, E, nair, gup)]
    (tmp_path / "x.par").write_text("\n".join(recs) + "\n")
    (tmp_path / "q.txt").write_text("200 150.0\n300 240.0\n1000 1300.0\n")
    # what the records hold after fixed-width formatting
    nu_r = np.array([float(r[3:15]) for r in recs])
    A_r = np.array([float(r[25:35]) for r in recs])
    ga_r = np.array([float(r[35:40]) for r in recs])
    gs_r = np.array([float(r[40:45]) for r in recs])
    E_r = np.array([float(r[45:55]) for r in recs])
    n_r = np.array([float(r[55:59]) for r in recs])
    g_r = np.array([float(r[146:153]) for r in recs]).astype(np.int32)
    mix, T, P, N = 0.8, 500.0, 0.5, 100_000
    cfg = "\n".join([
        "HITRAN", "TRANS-FILES", "x.par", "END", "PROFILE VOIGT", f"N-POINTS {N}",
        "NU-RANGE 990.0 1110.0", "MEAN-MASS 18.0153", f"HITRAN-MIXTURE-AIR {mix}", "CUTOFF 10",
        "PARTITION 240.0", "PARTITION-FILE q.txt", "TEMPERATURE-PRESSURE-GRID", f"{T} {P}", "END",
        ""])
    out, band, err = run_cli(cfg, tmp_path)
    Q = 240.0 + (1300.0 - 240.0) * (T - 300.0) / 700.0
    gam = ga_r * mix + gs_r * (1 - mix)
    p = O.params("voigt", T=T, P=P, Q=Q, cutoff=10.0, p_scale=1.0)
    ref, _ = O.cross_section(p, 990.0, 120.0 / N, N, nu_r, E_r, A_r, g_r, gam, n_r, nthreads=8)
    e = np.max(np.abs(out[0] - ref)) / ref.max()
    assert e <= 1e-4, e
    assert abs(out[0].sum() - ref.sum()) / ref.sum() <= 1e-6
