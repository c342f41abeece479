"""The ExoMol transition reader on plain and bzip2-compressed files (no GPU).

`gexs_xsec --lines` runs the driver's reader over the input's transition
files and prints each assembled line (nu, g', E'', A, gamma, n) exactly.  The
expected lines are built here from the same seeded list: nu = E_u - E_l in
FP64, g' of the upper state, E'' of the lower, A as written, gamma and n from
the broadener lookup (a1 by both J, then a0 by the lower J, else the
defaults; ExomolStateReader.cpp:158-167).
"""
import bz2
import os
import subprocess

import numpy as np
import pytest

from gexs_amd import synth
from tests.test_gpu_driver import broadeners, gamma_n

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gexs_amd", "gexs_xsec")


def record37(u, l, a):
    """One 37-byte .trans record (36 characters and the newline), the width
    the reference's bz2 reader takes (read_compress_trans.c:48)."""
    r = f"{u:12d} {l:12d} {a:10.4E}\n"
    assert len(r) == 37
    return r


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("exomol")
    L = synth.workload_list("voigt_1e6", scale=0.002)
    a0, a1 = broadeners(L.J)
    with open(d / "m.states", "w") as f:
        for i, (e, g, j) in enumerate(zip(L.E, L.g, L.J)):
            f.write(f"{i + 1:12d} {float(e)!r} {int(g):6d} {int(j):7d}\n")
    with open(d / "m.broad", "w") as f:
        for j, (g, n) in a0.items():
            f.write(f"a0 {g!r} {n!r} {j}\n")
        for (jl, ju), (g, n) in a1.items():
            f.write(f"a1 {g!r} {n!r} {jl} {ju}\n")
    text = "".join(record37(int(u) + 1, int(lo) + 1, float(a))
                   for u, lo, a in zip(L.upper, L.lower, L.A))
    raw = text.encode()
    (d / "m.trans").write_bytes(raw)
    (d / "one.trans.bz2").write_bytes(bz2.compress(raw))
    # three concatenated streams, cut in the middle of records
    cuts = [0, 37 * 500 + 11, 37 * 1300 + 29, len(raw)]
    (d / "multi.trans.bz2").write_bytes(b"".join(
        bz2.compress(raw[a:b]) for a, b in zip(cuts[:-1], cuts[1:])))
    return d, L, a0, a1


def dump(d, trans, expect_ok=True):
    cfg = "\n".join(["EXOMOL", "STATE-FILE m.states", "TRANS-FILES", *trans, "END",
                     "BROADENERS", "m.broad", "END", "GAMMA 0.07", "GAMMA-N 0.5",
                     "NU-RANGE 0.0 10000.0", "N-POINTS 1000", ""])
    r = subprocess.run([CLI, "--lines"], input=cfg, capture_output=True, text=True, cwd=d,
                       timeout=120)
    if not expect_ok:
        return r
    assert r.returncode == 0, r.stderr
    return np.array([[float(x) for x in ln.split()] for ln in r.stdout.splitlines()])


def expected(L, a0, a1):
    u, lo = L.upper.astype(np.int64), L.lower.astype(np.int64)
    A = np.array([float(f"{float(a):10.4E}") for a in L.A])
    gam, nex = gamma_n(L, a0, a1)
    return np.stack([L.E[u] - L.E[lo], L.g[u].astype(float), L.E[lo], A, gam, nex], axis=1)


def test_plain_37_byte_records(files):
    d, L, a0, a1 = files
    got = dump(d, ["m.trans"])
    assert got.shape == (L.n_trans, 6)
    assert np.array_equal(got, expected(L, a0, a1))


@pytest.mark.parametrize("name", ["one.trans.bz2", "multi.trans.bz2"])
def test_bzip2_equals_plain(files, name):
    d, L, a0, a1 = files
    got = dump(d, [name])
    assert np.array_equal(got, expected(L, a0, a1))


def test_several_files_in_order(files):
    d, L, a0, a1 = files
    got = dump(d, ["multi.trans.bz2", "m.trans"])
    exp = expected(L, a0, a1)
    assert np.array_equal(got, np.concatenate([exp, exp]))


def test_truncated_and_corrupt_bzip2_fail_loudly(files):
    d, L, _, _ = files
    data = (d / "one.trans.bz2").read_bytes()
    (d / "cut.trans.bz2").write_bytes(data[: len(data) // 2])
    r = dump(d, ["cut.trans.bz2"], expect_ok=False)
    assert r.returncode == 1 and "bzip2" in r.stderr
    # a flipped byte: bzip2 finds it by the block CRC at the latest (decoded
    # garbage before that point may fail to parse first); either way the run
    # ends with an error, never with a short list
    bad = bytearray(data)
    bad[len(bad) // 2] ^= 0xFF
    (d / "bad.trans.bz2").write_bytes(bytes(bad))
    r = dump(d, ["bad.trans.bz2"], expect_ok=False)
    assert r.returncode == 1
    assert "bzip2" in r.stderr or "unreadable transition record" in r.stderr


def test_hitran_160_char_records(tmp_path):
    """HITRAN .par records (HITRANStateReader.cpp:137-250): fields by the
    fixed widths 2,1,12,10,10,5,5,10,4,8,... (:3-18), gamma = g_air*mix +
    g_self*(1-mix) (P is applied later, once), and a record with A = 0 gets
    its A from the 296 K intensity with g' = 1 (:222-233).  The A-from-S
    records are checked by the inverse relation: the line strength that A
    gives at 296 K with Q(296) equals the record's S."""
    from tests.test_gpu_driver import par_record
    rng = np.random.default_rng(5)
    n = 300
    nu = np.sort(rng.uniform(500.0, 4000.0, n))
    S = 10 ** rng.uniform(-25, -19, n)
    A = 10 ** rng.uniform(-3, 1, n)
    A[::3] = 0.0
    rows = zip(nu, S, A, rng.uniform(0.05, 0.09, n), rng.uniform(0.2, 0.4, n),
               rng.uniform(0.0, 3000.0, n), rng.uniform(0.5, 0.8, n),
               rng.integers(1, 30, n).astype(float))
    recs = [par_record(*r) for r in rows]
    # a blank line between records is not a record
    (tmp_path / "x.par").write_text("\n".join(recs[:100]) + "\n\n" + "\n".join(recs[100:]) + "\n")
    mix, q296 = 0.7, 174.5
    cfg = "\n".join(["HITRAN", "TRANS-FILES", "x.par", "END", "PROFILE VOIGT", "N-POINTS 1000",
                     "NU-RANGE 400.0 4100.0", f"HITRAN-MIXTURE-AIR {mix}", f"PARTITION {q296}",
                     ""])
    r = subprocess.run([CLI, "--lines"], input=cfg, capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.array([[float(x) for x in ln.split()] for ln in r.stdout.splitlines()])
    assert got.shape == (n, 6)
    f = lambda a, b: np.array([float(rec[a:b]) for rec in recs])
    nu_r, S_r, A_r, ga, gs, E_r, n_r, g_r = (f(3, 15), f(15, 25), f(25, 35), f(35, 40),
                                             f(40, 45), f(45, 55), f(55, 59), f(146, 153))
    assert np.array_equal(got[:, 0], nu_r)
    assert np.array_equal(got[:, 2], E_r)
    assert np.array_equal(got[:, 4], ga * mix + gs * (1.0 - mix))
    assert np.array_equal(got[:, 5], n_r)
    has_a = A_r != 0.0
    assert np.array_equal(got[has_a, 3], A_r[has_a])
    assert np.array_equal(got[has_a, 1], g_r[has_a])
    assert np.all(got[~has_a, 1] == 1.0)
    beta = 6.6260693e-27 * 2.99792458e10 / (1.380658e-16 * 296.0)
    Aa, nn, EE = got[~has_a, 3], nu_r[~has_a], E_r[~has_a]
    S_back = Aa * np.exp(-beta * EE) * (1 - np.exp(-beta * nn)) / (
        8 * np.pi * 2.99792458e10 * nn * nn * q296)
    assert np.max(np.abs(S_back / S_r[~has_a] - 1)) < 1e-13
