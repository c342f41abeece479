"""FP32 approximations the kernels use, restated in numpy float32 from the
constants in gexs_amd/csrc/profiles_device.h (read from the source, so a slip
in any constant fails here) and checked against scipy."""
import os
import re

import numpy as np
from scipy.special import erf, erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "gexs_amd", "csrc", "profiles_device.h")


def erfc_table_constants():
    src = open(SRC).read()
    n = int(re.search(r"constexpr int kErfcPieces = (\d+);", src).group(1))
    per = float(re.search(r"constexpr float kErfcPerUnit = ([0-9.]+)f;", src).group(1))
    return n, per


def erfc_pieces(n, per, swap=False):
    """FP32 coefficients (a0, a1, a2, a3) of each piece: the cubic Hermite
    interpolant of erfc and its derivative at the piece's ends."""
    k = np.arange(n, dtype=np.float64)
    e0, e1 = k / per - 8.0, (k + 1) / per - 8.0
    H0, H1 = erfc(e0), erfc(e1)
    s = -2.0 / np.sqrt(np.pi) / per
    D0, D1 = s * np.exp(-e0 * e0), s * np.exp(-e1 * e1)
    a2 = 3.0 * (H1 - H0) - 2.0 * D0 - D1
    a3 = 2.0 * (H0 - H1) + D0 + D1
    if swap:
        a2, a3 = a3, a2
    return np.stack([H0, D0, a2, a3], axis=1).astype(np.float32)


def fma32(a, b, c):
    # the float32 product is exact in float64; one rounding to float32 after
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def erfc_lookup_f32(e, tab, n, per):
    f = np.float32
    e = e.astype(f)
    v = np.clip(e * f(per), f(-n / 2), f(n / 2) - f(1.5e-5))  # exact scaling
    fl = np.floor(v)
    fr = (v - fl).astype(f)
    c = tab[fl.astype(np.int64) + n // 2]
    return fma32(fma32(fma32(c[:, 3], fr, c[:, 2]), fr, c[:, 1]), fr, c[:, 0])


def test_erfc_table_of_the_bin_averaged_gaussian():
    """The Gaussian profiles' erfc: cubic pieces over [-8, 8), looked up in
    FP32 as the kernel does (constants read from the source).  Within 2e-7
    of erfc everywhere (values up to 2, FP32 spacing 2.4e-7 there), 3e-8 where
    erfc < 0.25; the bin of a point, erfc(far edge) - erfc(near edge), within
    5e-7 of erf(near) - erf(far); swapped cubic coefficients fail."""
    n, per = erfc_table_constants()
    assert (n, per) == (512, 32.0)
    tab = erfc_pieces(n, per)
    e = np.concatenate([np.linspace(-12.0, 12.0, 600001), 10 ** np.linspace(-9, 1, 3000),
                        -(10 ** np.linspace(-9, 1, 3000))])
    got = erfc_lookup_f32(e, tab, n, per).astype(np.float64)
    ref = erfc(e.astype(np.float32).astype(np.float64))
    err = np.abs(got - ref)
    assert np.max(err) < 2e-7
    assert np.max(err[ref < 0.25]) < 3e-8
    rng = np.random.default_rng(22)
    A = 10 ** rng.uniform(-2, 1.5, 200000)     # points per width
    eb = rng.uniform(-7, 7, 200000)            # far edge of a point
    ea = eb + A                                # near edge: one bin up
    v = (erfc_lookup_f32(eb, tab, n, per).astype(np.float64)
         - erfc_lookup_f32(ea, tab, n, per).astype(np.float64))
    exact = erf(ea.astype(np.float32).astype(np.float64)) - erf(eb.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(v - exact)) < 5e-7
    bad = erfc_pieces(n, per, swap=True)
    assert np.max(np.abs(erfc_lookup_f32(e, bad, n, per).astype(np.float64) - ref)) > 1e-6


def quad_r2():
    src = open(SRC).read()
    return float(re.search(r"constexpr float kQuadR2 = ([0-9.e]+)f;", src).group(1))


def test_voigt_quad_far_rule_against_the_11_point_sum(O):
    """Voigt-Quad beyond the Zaghloul-Ali core (x^2 + y^2 >= 100): the oracle
    sums the reference's 11-point Gauss-Hermite rule in FP64.  The kernel takes
    the 4-point rule (Humlicek's region-II pairs) up to kQuadR2 and the one-term
    continued fraction (the region-I rational with 1/sqrt(pi)) beyond.  Both
    stay within 1.2e-6 of the oracle's value relative to it, also with the
    4-point pairs in FP32; the region-I rational at x^2 + y^2 >= 900 (a lower
    threshold) does not."""
    r2 = quad_r2()
    assert r2 == 1600.0
    rng = np.random.default_rng(21)
    n = 4000
    y = np.concatenate([10 ** rng.uniform(-6, 1.6, n), rng.uniform(0, 45, n)])
    rr = np.concatenate([rng.uniform(100.0, r2, n), 10 ** rng.uniform(2, np.log10(r2), n)])
    keep = rr > y * y
    y, rr = y[keep], rr[keep]
    x = np.sqrt(rr - y * y)
    ref = np.array([O.gh11(a, b) for a, b in zip(x, y)])

    t, w = np.polynomial.hermite.hermgauss(4)
    a2 = np.sort(t[t > 0] ** 2)
    c = 2 * w[t > 0][np.argsort(t[t > 0])] / np.pi

    def gh4(x, y, dt):
        x, y = dt(x), dt(y)
        x2 = x * x
        b = (x2 + y * y)[:, None] + a2.astype(dt)
        d = b * b - dt(4) * a2.astype(dt) * x2[:, None]
        return (y * np.sum(c.astype(dt) * b / d, axis=1)).astype(np.float64)

    assert np.max(np.abs(gh4(x, y, np.float64) - ref) / ref) < 2e-7
    assert np.max(np.abs(gh4(x, y, np.float32) - ref) / ref) < 1.2e-6

    def rational(x, y):
        s = x * x + y * y - 0.5
        return y * (s + 1.0) / (s * s + 2 * y * y) / np.sqrt(np.pi)

    yf = np.concatenate([10 ** rng.uniform(-6, 3, n), rng.uniform(0, 60, n)])
    xf = np.sqrt(np.maximum(r2 - yf * yf, 0.0)) + 10 ** rng.uniform(-6, 4, 2 * n)
    far = np.array([O.gh11(a, b) for a, b in zip(xf, yf)])
    assert np.max(np.abs(rational(xf, yf) - far) / far) < 1.2e-6
    xl = np.sqrt(np.maximum(900.0 - yf * yf, 0.0)) + 10 ** rng.uniform(-6, 0, 2 * n)
    low = np.array([O.gh11(a, b) for a, b in zip(xl, yf)])
    assert np.max(np.abs(rational(xl, yf) - low) / low) > 2e-6


def bin_gauss_reach(Af, Bf):
    """The reach line_record (gexs_amd/csrc/kernels.hip) gives a bin-Gauss
    line: ceil(((8 (1 + 1e-5) + |B|) / A) (1 + 1e-5) + 1/2), in FP64."""
    return np.ceil(((8.0 * (1.0 + 1e-5) + np.abs(Bf)) / Af) * (1.0 + 1e-5) + 0.5)


def clamped_edge_f32(c, j, A, B, h, n, per):
    """The kernel's table argument at the far edge of point j, clamped as
    erfc_lds clamps it: fma(cf - jf, 32 A, 32 (B - H)) in FP32."""
    f = np.float32
    A32 = (A * f(per)).astype(f)
    Bh32 = ((B - h).astype(f) * f(per)).astype(f)
    v = fma32((c.astype(f) - j.astype(f)), A32, Bh32)
    return np.clip(v, f(-n / 2), f(n / 2) - f(1.5e-5))


def test_bin_gauss_reach_adds_exact_zeros():
    """Beyond its reach K a bin-Gauss line adds exactly zero: at every point j
    with |c - j| >= K both edges of the bin (j's far edge and j - 1's) clamp
    to the same end of the erfc table, so the kernel's difference is 0 and the
    tiles may skip the line.  Checked in FP32 as the kernel computes the edges
    over widths from 0.01 to 30 points per width and offsets |B| <= A / 2; the
    bound is tight to within two points, and dropping its half-bin term lets
    nonzero bins through."""
    n, per = erfc_table_constants()
    f = np.float32
    rng = np.random.default_rng(5)
    m = 400000
    Af = 10 ** rng.uniform(-2.0, 1.5, m)
    Bf = rng.uniform(-0.5, 0.5, m) * Af
    A, B, h = Af.astype(f), Bf.astype(f), (0.5 * Af).astype(f)
    c = rng.integers(0, 1 << 23, m)
    K = bin_gauss_reach(Af, Bf).astype(np.int64)
    assert np.all(K >= 1)

    def same_clamp(j):
        return clamped_edge_f32(c, j, A, B, h, n, per) == clamped_edge_f32(c, j - 1, A, B, h, n, per)

    for d in range(0, 6):
        assert np.all(same_clamp(c - K - d)), d
        assert np.all(same_clamp(c + K + d)), d
    # tight: two points inside the reach some bins are not zero
    assert not np.all(same_clamp(c + K - 2)) and not np.all(same_clamp(c - K + 2))
    # without the half bin the bound fails
    K0 = np.ceil((8.0 + np.abs(Bf)) / Af).astype(np.int64)
    assert not (np.all(same_clamp(c - K0)) and np.all(same_clamp(c + K0)))


def far_expansion_f32(Kl, A, B, E2, d, T, nterms=16, flip=False):
    """The profile kernel's far-field expansion in FP32 (far_expand /
    far_add in gexs_amd/csrc/profile_tiles.h): the line's coefficients of
    Kl / ((Xc - A T tau)^2 + E2 - 1) by the two-term recurrence, Xc its x at
    the tile's centre, d = c - t_c (a half-integer)."""
    f = np.float32
    Kl, A, B, E2 = (np.asarray(v, f) for v in (Kl, A, B, E2))
    Xc = fma32(np.asarray(d, f), A, B)
    inv = (f(1) / fma32(Xc, Xc, (E2 - f(1)).astype(f))).astype(f)
    u = (A * f(T)).astype(f)
    al = ((f(2) * u).astype(f) * (Xc * inv).astype(f)).astype(f)
    be = (-(u * u).astype(f) * inv).astype(f)
    if flip:
        be = -be
    h = [(Kl * inv).astype(f)]
    h.append((al * h[0]).astype(f))
    for k in range(2, nterms):
        h.append(fma32(al, h[-1], (be * h[-2]).astype(f)))
    return np.stack(h, axis=1)


def horner_f32(C, tau):
    f = np.float32
    p = np.broadcast_to(C[..., -1:], C.shape[:-1] + (tau.shape[-1],)).astype(f)
    for k in range(C.shape[-1] - 2, -1, -1):
        p = fma32(p, tau.astype(f), np.broadcast_to(C[..., k:k + 1], p.shape))
    return p


def test_far_field_expansion_matches_the_one_term_form():
    """The far-field expansion the Voigt tiles use for far lines 4T + 2.5
    points or more from a tile's centre (T its half width): evaluated in
    FP32 as the kernel does, its value at every point of the tile is within
    5e-7 of the one-term form Kl / (x^2 + E2 - 1) computed exactly, over
    widths, offsets, Lorentz ratios y from 1e-2 to 1e2 and distances up to
    5000 points, at every tile width (the kernel's direct FP32 evaluation of
    the same form reaches 4.6e-7 on these lines: the FP32 rounding of the
    record's A, B and Kl); with the recurrence's second coefficient negated,
    or 10 terms instead of 16 (truncation up to 11 (0.256)^10 / 0.744^2 =
    2.4e-5), it is not."""
    rng = np.random.default_rng(9)
    for T in (32, 64, 128, 256):
        m = 20000
        A = 10 ** rng.uniform(-2.0, 1.0, m)
        B = rng.uniform(-0.5, 0.5, m) * A
        y = 10 ** rng.uniform(-2.0, 2.0, m)
        E2 = y * y - 0.5
        side = rng.choice([-1.0, 1.0], m)
        d = side * (4 * T + 2.5 + np.floor(rng.uniform(0, 1, m) ** 3 * 5000))
        tau = (np.arange(2 * T) - T + 0.5) / T
        x = (d[:, None] - T * tau[None, :]) * A[:, None] + B[:, None]
        s = x * x + E2[:, None]
        far = np.all(s * s >= (2 * y[:, None] ** 2 + 1) * 8388608.0, axis=1)
        sel = far
        assert np.sum(sel) > m // 4
        Kl = rng.uniform(0.5, 1.0, m)
        exact = Kl[sel, None] / (x[sel] ** 2 + (E2[sel, None] - 1.0))
        C = far_expansion_f32(Kl[sel], A[sel], B[sel], E2[sel], d[sel], T)
        got = horner_f32(C, np.broadcast_to(tau, (np.sum(sel), 2 * T))).astype(np.float64)
        err = np.max(np.abs(got - exact) / exact)
        assert err < 5e-7, (T, err)
        bad = horner_f32(far_expansion_f32(Kl[sel], A[sel], B[sel], E2[sel], d[sel], T, flip=True),
                         np.broadcast_to(tau, (np.sum(sel), 2 * T))).astype(np.float64)
        assert np.max(np.abs(bad - exact) / exact) > 1e-3
        few = horner_f32(far_expansion_f32(Kl[sel], A[sel], B[sel], E2[sel], d[sel], T, nterms=10),
                         np.broadcast_to(tau, (np.sum(sel), 2 * T))).astype(np.float64)
        assert np.max(np.abs(few - exact) / exact) > 5e-6
