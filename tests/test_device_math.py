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
