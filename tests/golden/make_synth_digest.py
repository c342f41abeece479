"""Writes tests/golden/synth_digest.json: SHA-256 of small seeded lists.

The digests freeze the generator (SplitMix64 streams, inverse-CDF sampling,
acceptance rule) so that benchmark inputs stay identical across machines.
"""
import json
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))
from gexs_amd import synth  # noqa: E402
from tests.test_synth import digest  # noqa: E402

out = {}
for name, scale in (("doppler_2e5", 0.01), ("voigt_1e6", 0.002), ("voigt_tp_table_1e7", 0.0005)):
    L = synth.workload_list(name, scale=scale)
    out[name] = dict(scale=scale, n_trans=L.n_trans, sha256=digest(L))
json.dump(out, open(os.path.join(os.path.dirname(__file__), "synth_digest.json"), "w"), indent=1)
print(out)
