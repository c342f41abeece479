"""The seeded line-list generator: determinism, distributions, golden digest."""
import hashlib
import json
import os

import numpy as np

from gexs_amd import synth

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "synth_digest.json")


def digest(L):
    h = hashlib.sha256()
    for a in (L.E, L.g, L.J, L.upper, L.lower, L.A, L.gamma, L.nexp):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_uniforms_counter_based():
    a = synth.uniforms(5, 2, 0, 1000)
    b = synth.uniforms(5, 2, 400, 600)
    assert np.array_equal(a[400:], b)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.03


def test_list_properties():
    L = synth.workload_list("voigt_1e6", scale=0.01)
    nu, E, A, g, gam, nex = L.lines()
    assert L.n_states == 500 and L.n_trans == 10000
    assert np.all(nu > 0) and np.all(nu <= 10000.0)
    assert L.E.min() >= 0 and L.E.max() <= 20000.0
    assert L.g.min() >= 1 and L.g.max() <= 50
    assert L.J.min() >= 0 and L.J.max() <= 60
    assert A.min() >= 1e-8 and A.max() <= 1e2
    assert gam.min() >= 0.02 and gam.max() <= 0.10
    assert nex.min() >= 0.3 and nex.max() <= 0.8
    # exponential energies: mean near 4000 (clipping at 5 sigma barely matters)
    assert abs(L.E.mean() - 4000.0) < 400.0


def test_golden_digest():
    want = json.load(open(GOLDEN))
    for name, spec in want.items():
        L = synth.workload_list(name, scale=spec["scale"])
        assert L.n_trans == spec["n_trans"]
        assert digest(L) == spec["sha256"], name
