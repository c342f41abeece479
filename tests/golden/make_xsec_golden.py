"""Writes tests/golden/xsec_tiny/: a tiny ExoMol list, the keyword input of
gexs_xsec for it, and the expected output text.

The expected cross-sections come from the reference's own CPU code,
OpenMPManager::ComputeVoigtTP with humlic (Managers/CPU/OpenMPManager.cpp:376-426,
compute/CPU/profiles.cpp:11-44), compiled from the reference sources into
oracle/_ref/libgexs_ref.so (``make -C oracle ref``; needs the reference tree,
so this script runs where it is present).  Q(T) comes from the reference's
compute_partition (Readers/Exomol/exomol_functions.cpp:50-73).  The input
therefore asks for the reference CPU conventions: PRESSURE-MODE ref_cpu
(gammaL without pressure, OpenMPManager.cpp:398), the default cutoff
min(25 P, 100) (:386) and WIDTH-CONVENTION ref_cpu (:404).

The text follows BaseProfile::OutputProfile (Profiles/BaseProfile.cpp:269-301):
the band-strength line (summed line intensity of the first (T,P), :250-267,
against sum(sigma) * dnu, :283-291), then per (T,P) a header and one
"[i] nu sigma" line per grid point.

    python tests/golden/make_xsec_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "xsec_tiny")
START, END, NPTS = 1000.0, 1005.0, 500
TPS = [(296.0, 1.0), (900.0, 0.3)]
MASS, GAMMA, GAMMA_N = 18.0153, 0.07, 0.5
HW = 100.0  # Voigt selection half width (VoigtProfile.h:15)
CM_COEF = 1.327209365e-12
PLANCK, VELLGT, BOLTZ = 6.6260693e-27, 2.99792458e10, 1.380658e-16


def make_list():
    """States and transitions: 24 lines inside the range, 6 in the wings
    outside it (975-1000 and 1005-1030 cm^-1: they reach the grid through
    their windows), 3 selected but too far to reach it (910-960 and
    1040-1090), 2 outside the selection (< 900 or > 1105) and one with
    E_u = E_l (nu = 0, never selected)."""
    rng = np.random.default_rng(20161203)
    nus = np.concatenate([rng.uniform(1000.2, 1004.8, 24), rng.uniform(975.0, 999.0, 3),
                          rng.uniform(1006.0, 1030.0, 3), [912.5, 958.25, 1077.0],
                          [850.0, 1150.0], [0.0]])
    El = rng.uniform(0.0, 3000.0, nus.shape[0])
    El = np.round(El, 6)
    nus = np.round(nus, 6)
    E = np.concatenate([El, El + nus])  # lower states, then upper states
    E = np.round(E, 6)
    g = rng.integers(1, 50, E.shape[0])
    J = rng.integers(0, 60, E.shape[0])
    n = nus.shape[0]
    lower = np.arange(n)
    upper = np.arange(n) + n
    upper[-1] = lower[-1]  # E_u = E_l
    A = 10.0 ** rng.uniform(-3.0, 1.0, n)
    order = rng.permutation(n)  # unsorted file, as ExoMol .trans files are in nu
    return E, g, J, upper[order], lower[order], A[order]


def main():
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libgexs_ref.so is not built (make -C oracle ref)")
    os.makedirs(OUT, exist_ok=True)
    E, g, J, up, lo, A = make_list()
    with open(os.path.join(OUT, "tiny.states"), "w") as f:
        for i, (e, gg, j) in enumerate(zip(E, g, J)):
            f.write(f"{i + 1:6d} {float(e)!r} {int(gg):4d} {int(j):4d}\n")
    with open(os.path.join(OUT, "tiny.trans"), "w") as f:
        for u, l, a in zip(up, lo, A):
            f.write(f"{int(u) + 1} {int(l) + 1} {float(a)!r}\n")
    cfg = ["EXOMOL", "STATE-FILE tiny.states", "TRANS-FILES", "tiny.trans", "END",
           "PROFILE VOIGT", f"N-POINTS {NPTS}", f"NU-RANGE {START} {END}",
           f"MEAN-MASS {MASS}", f"GAMMA {GAMMA}", f"GAMMA-N {GAMMA_N}", "T-REF 296",
           "PRESSURE-MODE ref_cpu", "WIDTH-CONVENTION ref_cpu", "TEMPERATURE-PRESSURE-GRID",
           *[f"{t} {p}" for t, p in TPS], "END", ""]
    with open(os.path.join(OUT, "input.txt"), "w") as f:
        f.write("\n".join(cfg))
    # the lines as the ExoMol reader assembles them (ExomolStateReader.cpp:158-162),
    # in file order, selected as BaseProfile.cpp:152-157 (as a filter)
    nu = E[up] - E[lo]
    El, gu = E[lo], g[up].astype(np.int32)
    keep = (nu >= START - HW) & (nu <= END + HW) & (nu != 0.0)
    nu, El, Ak, gu = nu[keep], El[keep], A[keep], gu[keep]
    gam, nex = np.full(nu.shape[0], GAMMA), np.full(nu.shape[0], GAMMA_N)
    dnu = (END - START) / NPTS
    freq = START + np.arange(NPTS) * dnu
    lines = []
    blocks = []
    for k, (T, P) in enumerate(TPS):
        Q = R.gxr_partition(E, g.astype(np.int32), E.shape[0], T)
        sig = O.ref_voigt_tp(START, dnu, NPTS, T, P, Q, MASS, 296.0, 1.0, nu, El, Ak, gu, gam,
                             nex, nthreads=1)
        blocks.append((T, P, Q, sig))
    T0, _, Q0, sig0 = blocks[0]
    beta = PLANCK * VELLGT / (BOLTZ * T0)
    total = float(np.sum(CM_COEF * Ak * gu * np.exp(-beta * El) * (1.0 - np.exp(-beta * nu)) /
                         (nu * nu * Q0)))
    cross = float(np.sum(sig0)) * dnu
    lines.append("Our summed intensity = %16.6E compared to the integrated intensity = %16.6E "
                 "with a difference of %12.6f %%" % (total, cross, (total - cross) * 100.0 / total))
    for k, (T, P, Q, sig) in enumerate(blocks):
        lines.append("-------T=%12.6f K--------P=%12.6f atm---------" % (T, P))
        lines.append("")
        for j in range(NPTS):
            lines.append("[%i] %12.6f %13.8E" % (k, freq[j], sig[j]))
        lines.append("")
    with open(os.path.join(OUT, "expected.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"{OUT}: {int(keep.sum())} selected lines, Q = {[b[2] for b in blocks]}")


if __name__ == "__main__":
    main()
