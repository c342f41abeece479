"""FP32 approximations the kernels use, restated in numpy float32 from the
constants in gexs_amd/csrc/profiles_device.h (read from the source, so a slip
in any constant fails here) and checked against scipy."""
import os
import re

import numpy as np
from scipy.special import erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "gexs_amd", "csrc", "profiles_device.h")


def erfc_nonneg_coefficients():
    src = open(SRC).read()
    body = src[src.index("float erfc_nonneg(float z)"):]
    body = body[:body.index("}")]
    lead = re.search(r"float p = ([-0-9.e]+)f;", body).group(1)
    rest = re.findall(r"p = fmaf\(p, u, ([-0-9.e]+)f\);", body)
    return [float(lead)] + [float(c) for c in rest]  # highest degree first


def erfc_nonneg_f32(z, coef):
    f = np.float32
    z = z.astype(f)
    t = f(1) / (f(0.5) * z + f(1))
    u = f(2) * t - f(1)
    p = np.full_like(z, f(coef[0]))
    for c in coef[1:]:
        p = p * u + f(c)
    return t * np.exp(p - z * z)


def test_erfc_nonneg_accuracy():
    coef = erfc_nonneg_coefficients()
    assert len(coef) == 10
    z = np.concatenate([np.linspace(0.0, 12.0, 300001), 10 ** np.linspace(-9, -1, 3000)])
    got = erfc_nonneg_f32(z, coef).astype(np.float64)
    ref = erfc(z)
    assert np.max(np.abs(got - ref)) < 5e-7
    small = z <= 4.0
    assert np.max(np.abs(got - ref)[small] / ref[small]) < 3e-6
    assert erfc_nonneg_f32(np.array([0.0]), coef)[0] == np.float32(1.0) or \
        abs(erfc_nonneg_f32(np.array([0.0]), coef)[0] - 1.0) < 4e-7
    # a coefficient off in its fourth digit is visible
    bad = list(coef)
    bad[5] *= 1.001
    worse = erfc_nonneg_f32(z, bad).astype(np.float64)
    assert np.max(np.abs(worse - ref)) > 1e-6


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
