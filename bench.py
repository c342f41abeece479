"""Benchmark of the line-by-line Voigt cross-section path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload voigt_1e6] [--no-cpu]

A step is one cross-section: ExecuteCrossSection of the BaseManager drop-in
(FP64 per-line pre-pass, then the FP32-compensated tiled Voigt kernel) over the
resident line list of the workload, at its (T, P).  Inputs are resident in HBM
before the timed region starts.  value = N_eval * K / elapsed (line-point
evaluations per second, N_eval as defined by the reference CPU window).

The JSON line also carries
  roofline      the profile kernel's achieved FP32 rate (12 flop per Voigt
                evaluation, 16 per bin-averaged Gaussian) against 157.3 TFLOP/s,
                from HIP events around that kernel on its own stream;
  cpu_baseline  the reference CPU path (oracle/_ref, or the oracle port) timed
                on this host on a bounded sample of the same workload;
  fp64          the plain FP64 GPU path (one work-item per point) on the same
                input, and the speed-up of the FP32-compensated path over it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12  # MI355X FP32 vector peak, FLOP/s
FLOP_PER_EVAL = {"voigt": 12, "voigt_quad": 12, "doppler": 16, "gaussian": 16}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="voigt_1e6")
    ap.add_argument("--ppl", type=int, default=0, help="points per lane (0: tuned default)")
    ap.add_argument("--order", type=int, default=1)
    ap.add_argument("--split", type=int, default=1, help="waves per tile (1 or 4)")
    ap.add_argument("--persistent", type=int, default=0,
                    help="persistent workgroups of the profile kernel with --ppl (0: one per tile)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 5 times with a plain copy kernel "
                         "(the reference point of the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    return ap.parse_args()


class quiet_stdout:
    """Sends file descriptor 1 to /dev/null (the reference CPU library prints
    progress there; stdout carries only the JSON line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *exc):
        import ctypes
        sys.stdout.flush()
        ctypes.CDLL(None).fflush(None)  # C stdio buffers of the library
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def cpu_baseline(lst, spec, T, P, Q, target_s):
    """Reference CPU path on a contiguous nu slice of the list, all host cores,
    repeated until about target_s seconds of CPU work are timed."""
    from oracle import oracle as O
    nu, E, A, g, gam, nex = lst.lines()
    o = np.argsort(nu, kind="stable")
    nu, E, A, g, gam, nex = nu[o], E[o], A[o], g[o], gam[o], nex[o]
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(int(cores), 64))
    use_ref = O.ref() is not None
    p = O.params("voigt", T=T, P=P, Q=Q, mass=spec["mass"], cutoff=spec["cutoff"],
                 p_scale=1 / 1.01325)
    mid = nu.shape[0] // 3
    n_lines = min(nu.shape[0] - mid, 200_000)
    sl = slice(mid, mid + n_lines)
    ev_one = O.neval(p, start, dnu, N, nu[sl], E[sl], A[sl], g[sl])
    total_ev, total_t, reps = 0, 0.0, 0
    while total_t < target_s and reps < 100:
        t0 = time.perf_counter()
        if use_ref:
            O.ref_voigt_tp(start, dnu, N, T, P, Q, spec["mass"], 296.0, 1 / 1.01325, nu[sl],
                           E[sl], A[sl], g[sl], gam[sl], nex[sl], nthreads=cores, parallel=True)
        else:
            O.cross_section(p, start, dnu, N, nu[sl], E[sl], A[sl], g[sl], gam[sl], nex[sl],
                            nthreads=cores)
        total_t += time.perf_counter() - t0
        total_ev += ev_one
        reps += 1
    return {"value": total_ev / total_t, "unit": "evals/s", "cores": cores,
            "kind": "reference" if use_ref else "port",
            "sample": f"{n_lines} consecutive nu-sorted lines from the middle of the list on the "
                      f"full grid ({ev_one:.3e} evals) x {reps} runs = {total_ev:.3e} evals in "
                      f"{total_t:.1f} s"
                      + ("; OpenMPManager::ComputeVoigtTP + humlic compiled from the reference "
                         "sources (gammaL without pressure, as the reference CPU path)"
                         if use_ref else "; FP64 oracle port of ComputeVoigtTP")}


def traffic_from_profile(workload):
    """HBM bytes per launch of the profile kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (None if absent)."""
    f = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def calib_copy(mib):
    """A plain copy kernel of known size: torch's device-to-device copy of
    mib MiB, five times.  Each launch reads and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copied {mib} MiB x 5")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md)"}))
        sys.exit(2)
    if a.calib_copy:
        calib_copy(a.calib_copy)
        return
    from gexs_amd import synth
    from gexs_amd.manager import HipManager

    spec = synth.WORKLOADS[a.workload]
    prof = spec["profile"]
    T = spec["T"] if spec["T"] else 1000.0
    P = spec["P"] if spec["P"] is not None else 0.0
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    log(f"[bench] generating {a.workload}")
    lst = synth.workload_list(a.workload)
    log(f"[bench] {lst.n_trans} transitions over {lst.n_states} states")

    steps, warm = a.steps, a.warmup
    m = HipManager(prof)
    if a.ppl:
        m.set_tuning(a.ppl, a.order, a.split)
        m.set_launch(a.persistent, False)
    m.InitializeVectors(N, steps + warm)
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF, synth.P_SCALE_BAR,
                          cutoff=spec["cutoff"])
    m.TransferFreq(start + np.arange(N, dtype=np.float64) * dnu)
    t0 = time.perf_counter()
    n = m.load_list(lst)  # H2D, device assembly and nu sort; outside the timed region
    Q = float(m.partition([T])[0])
    log(f"[bench] list resident + sorted in {time.perf_counter() - t0:.2f} s, Q(T)={Q:.6g}")
    tuning = None
    if not a.no_autotune and not a.ppl:
        best, cands = m.autotune([(T, P, Q)], start, end, N)
        tuning = {"best": best, "candidates": cands}
        log(f"[bench] autotune: {best}")
        for c in sorted(cands, key=lambda c: c["kernel_ms"]):
            log(f"[bench]   candidate {c}")
    for _ in range(warm):
        m.ExecuteCrossSection(N, n, 0, T, P, Q)
    m.sync()
    n_eval = m.stats()["n_eval"]
    m.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.ExecuteCrossSection(N, n, 0, T, P, Q)
    m.sync()
    elapsed = time.perf_counter() - t0
    k_ms, tot_ms, cnt = m.timing()
    assert cnt == steps, (cnt, steps)
    st = m.stats()
    value = n_eval * steps / elapsed
    k_avg = k_ms / cnt
    flops = FLOP_PER_EVAL[prof]
    achieved = n_eval * flops / (k_avg * 1e-3)
    log(f"[bench] {steps} steps in {elapsed * 1e3:.1f} ms; kernel avg {k_avg:.3f} ms; "
        f"N_eval {n_eval:.4e}; {value:.4e} evals/s")

    # plain FP64 GPU path on the same input
    fp64 = None
    if a.fp64_steps > 0:
        m64 = HipManager(prof, precision="fp64")
        m64.InitializeVectors(N, a.fp64_steps + 1)
        m64.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF,
                                synth.P_SCALE_BAR, cutoff=spec["cutoff"])
        m64.TransferFreq(start + np.arange(N, dtype=np.float64) * dnu)
        m64.load_list(lst)
        m64.ExecuteCrossSection(N, n, 0, T, P, Q)
        m64.sync()
        t1 = time.perf_counter()
        for _ in range(a.fp64_steps):
            m64.ExecuteCrossSection(N, n, 0, T, P, Q)
        m64.sync()
        e64 = time.perf_counter() - t1
        v64 = n_eval * a.fp64_steps / e64
        fp64 = {"value": v64, "unit": "evals/s", "ms_per_step": e64 * 1e3 / a.fp64_steps,
                "speedup_fp32c_over_fp64": value / v64}
        log(f"[bench] plain FP64 path {v64:.4e} evals/s ({value / v64:.2f}x slower than FP32C)")
        m64.close()

    cpu = None
    if not a.no_cpu:
        log("[bench] timing the reference CPU path on a sample")
        with quiet_stdout():
            cpu = cpu_baseline(lst, spec, T, P, Q, a.cpu_seconds)
        log(f"[bench] cpu baseline {cpu['value']:.4e} evals/s on {cpu['cores']} cores")

    out = {
        "metric": "Voigt line-point evals/s",
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": elapsed * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded SplitMix64 ExoMol-style list (inverse-CDF sampling)",
        "config": {
            "workload": f"{a.workload}: {prof}, {lst.n_trans} transitions / {lst.n_states} "
                        f"states, T={T} K, P={P} bar, grid {start}-{end} cm-1 x {N} points, "
                        f"cutoff {spec['cutoff']} cm-1, FP64 pre-pass + FP32 Kahan-compensated "
                        f"profile sum",
            "n_eval_per_step": n_eval,
            "xsec_per_s": steps / elapsed,
            "tiles": st["n_tiles"],
            "tuning": tuning["best"] if tuning else {"points_per_lane": a.ppl or 4,
                                                     "waves_per_tile": a.split, "order": a.order,
                                                     "persistent_blocks": a.persistent},
        },
        "roofline": {
            # compute-bound: FP32 vector ALU work against the FP32 peak (the
            # kernel issues no matrix instructions)
            "bound": "mfma",
            "compute": "fp32 VALU",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32,
            "traffic": traffic_from_profile(a.workload),
            "kernel": "k_profile_tiles<VOIGT>",
            "kernel_ms_avg": k_avg,
            "flop_per_eval": flops,
        },
        "cpu_baseline": cpu,
        "fp64_gpu_path": fp64,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
