"""Seeded synthetic ExoMol-style line lists for the five benchmark workloads.

The generator is SplitMix64 used counter-style: draw ``i`` of stream ``k`` is
``mix64(seed_k + (i + 1) * GOLDEN)``, mapped to ``u = (x >> 11) * 2**-53`` in
[0, 1).  Every draw is therefore independent of how many came before it, so
the lists vectorise in numpy and are identical on every machine.  Sampling is
by inverse CDF only (no library distributions, whose algorithms differ between
implementations).

States  (ExoMol ``.states``: id, E, g, J; Readers/Exomol/exomol_functions.cpp:8-48)
    E = min(-mu * log1p(-u), E_max)   g = 1 + floor(50 u)   J = floor(61 u)
Transitions (ExoMol ``.trans``: upper, lower, A; ExomolStateReader.cpp:132-172)
    (upper, lower) uniform over states, kept iff E_u > E_l and nu = E_u - E_l
    <= nu_max; A = 10**(lo + (hi - lo) u); per-line Lorentz gamma_ref =
    0.02 + 0.08 u (cm^-1/atm) and temperature exponent n = 0.3 + 0.5 u.
Line assembly (ExomolStateReader.cpp:158-162): nu = E_u - E_l, E'' = E_l,
g' = g_u.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
BASE_SEED = 20161203

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# stream ids
_S_E, _S_G, _S_J, _S_UP, _S_LO, _S_A, _S_GAM, _S_N = range(8)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _stream_seed(seed: int, stream: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + np.uint64(stream + 1) * np.uint64(0xD1B54A32D192ED03)
        return _mix64(np.array([s], dtype=np.uint64))[0]


def uniforms(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """Draws start .. start+count-1 of one stream as float64 in [0, 1)."""
    base = _stream_seed(seed, stream)
    idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = _mix64(base + idx * GOLDEN)
    return (x >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class LineList:
    """Resident list: a state table plus transitions that index into it."""

    E: np.ndarray          # float64 [n_states]
    g: np.ndarray          # int32   [n_states]
    J: np.ndarray          # int32   [n_states]
    upper: np.ndarray      # uint32  [n_trans]
    lower: np.ndarray      # uint32  [n_trans]
    A: np.ndarray          # float64 or float32 [n_trans]
    gamma: np.ndarray      # float64 or float32 [n_trans]
    nexp: np.ndarray       # float64 or float32 [n_trans]
    meta: dict = field(default_factory=dict)

    @property
    def n_states(self) -> int:
        return int(self.E.shape[0])

    @property
    def n_trans(self) -> int:
        return int(self.upper.shape[0])

    def lines(self):
        """Assembled FP64 line arrays (nu, E'', A, g', gamma, n), the six
        arrays BaseManager::TransferVectors takes (BaseManager.h:26)."""
        u = self.upper.astype(np.int64)
        l = self.lower.astype(np.int64)
        nu = self.E[u] - self.E[l]
        return (nu, self.E[l].copy(), self.A.astype(np.float64), self.g[u].astype(np.int32),
                self.gamma.astype(np.float64), self.nexp.astype(np.float64))


def make_states(seed: int, n: int, mean_e: float, e_max: float):
    u = uniforms(seed, _S_E, 0, n)
    E = np.minimum(-mean_e * np.log1p(-u), e_max)
    g = (1 + np.floor(50.0 * uniforms(seed, _S_G, 0, n))).astype(np.int32)
    J = np.floor(61.0 * uniforms(seed, _S_J, 0, n)).astype(np.int32)
    return E, g, J


def make_list(seed: int, n_states: int, n_trans: int, mean_e: float, e_max: float,
              nu_max: float, a_lo: float, a_hi: float, f32: bool = False,
              block: int = 1 << 24) -> LineList:
    E, g, J = make_states(seed, n_states, mean_e, e_max)
    ups, los = [], []
    have = 0
    start = 0
    while have < n_trans:
        cnt = block
        up = np.floor(n_states * uniforms(seed, _S_UP, start, cnt)).astype(np.int64)
        lo = np.floor(n_states * uniforms(seed, _S_LO, start, cnt)).astype(np.int64)
        nu = E[up] - E[lo]
        keep = (nu > 0.0) & (nu <= nu_max)
        ups.append(up[keep].astype(np.uint32))
        los.append(lo[keep].astype(np.uint32))
        have += int(keep.sum())
        start += cnt
    upper = np.concatenate(ups)[:n_trans]
    lower = np.concatenate(los)[:n_trans]
    A = 10.0 ** (a_lo + (a_hi - a_lo) * uniforms(seed, _S_A, 0, n_trans))
    gam = 0.02 + 0.08 * uniforms(seed, _S_GAM, 0, n_trans)
    nex = 0.3 + 0.5 * uniforms(seed, _S_N, 0, n_trans)
    fdt = np.float32 if f32 else np.float64
    return LineList(E=E, g=g, J=J, upper=upper, lower=lower, A=A.astype(fdt),
                    gamma=gam.astype(fdt), nexp=nex.astype(fdt),
                    meta=dict(seed=seed, n_states=n_states, n_trans=n_trans, mean_e=mean_e,
                              e_max=e_max, nu_max=nu_max, a_lo=a_lo, a_hi=a_hi))


# The five workloads.  Pressures are in bar; gamma_ref is per atm, so the
# pressure scale is 1/1.01325.  T_ref = 296 K.
P_SCALE_BAR = 1.0 / 1.01325
T_REF = 296.0

WORKLOADS = {
    "doppler_2e5": dict(
        list=dict(n_states=10_000, n_trans=200_000, mean_e=4000.0, e_max=20000.0,
                  nu_max=10000.0, a_lo=-8.0, a_hi=2.0),
        profile="doppler", T=1000.0, P=None, mass=18.0153,
        grid=(0.0, 10000.0, 1_000_000), cutoff=1.0),
    "voigt_1e6": dict(
        list=dict(n_states=50_000, n_trans=1_000_000, mean_e=4000.0, e_max=20000.0,
                  nu_max=10000.0, a_lo=-8.0, a_hi=2.0),
        profile="voigt", T=296.0, P=1.0, mass=18.0153,
        grid=(0.0, 10000.0, 1_000_000), cutoff=25.0),
    "voigt_1e8_stream": dict(
        list=dict(n_states=2_000_000, n_trans=100_000_000, mean_e=8000.0, e_max=40000.0,
                  nu_max=40000.0, a_lo=-10.0, a_hi=3.0, f32=True),
        profile="voigt", T=2000.0, P=0.1, mass=28.0,
        grid=(100.0, 30000.0, 2_990_000), cutoff=25.0),
    "voigt_tp_table_1e7": dict(
        list=dict(n_states=200_000, n_trans=10_000_000, mean_e=8000.0, e_max=40000.0,
                  nu_max=40000.0, a_lo=-10.0, a_hi=3.0, f32=True),
        profile="voigt", T=None, P=None, mass=28.0,
        grid=(500.0, 15000.0, 1_450_000), cutoff=25.0),
}


def tp_table():
    """10 x 12 (T, P) table: T log-spaced 200-3000 K, P log-spaced 1e-5-100 bar."""
    T = 200.0 * 15.0 ** (np.arange(10) / 9.0)
    P = 1e-5 * 10.0 ** (7.0 * np.arange(12) / 11.0)
    return [(float(t), float(p)) for t in T for p in P]


def workload_list(name: str, scale: float = 1.0) -> LineList:
    """The workload's list; ``scale`` < 1 shrinks states and transitions
    proportionally (same distributions, same seed) for fast tests."""
    k = list(WORKLOADS).index(name)
    spec = dict(WORKLOADS[name]["list"])
    if scale != 1.0:
        spec["n_states"] = max(16, int(spec["n_states"] * scale))
        spec["n_trans"] = max(1, int(spec["n_trans"] * scale))
    return make_list(BASE_SEED + k, **spec)
