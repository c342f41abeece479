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
    ap.add_argument("--profile", default=None,
                    help="profile of the headline workload (default: the workload's own)")
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
    ap.add_argument("--no-probe", action="store_true", help="skip the on-device VALU rate probe")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-table", action="store_true",
                    help="skip the batched (T,P) table (10^7 lines, 120 cross-sections)")
    ap.add_argument("--table-steps", type=int, default=2)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sweep")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the 10^8-transition hot list (T = 2000 K, P = 0.1 bar)")
    ap.add_argument("--no-profiles", action="store_true",
                    help="skip the Doppler (2e5 lines, cutoff 1 cm-1) and Voigt-Quad workloads")
    ap.add_argument("--serve-rates", default="10,25,50,100,200,500")
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
    """The committed rocprofv3 figures of this workload's profile kernel (HBM
    bytes per launch from FETCH_SIZE / WRITE_SIZE, VALU instructions per
    launch from SQ_INSTS_VALU, and the launch shape they were taken at), or
    None if absent."""
    f = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        return json.load(fh)


_LISTS = {}


def big_list(name):
    from gexs_amd import synth
    if name not in _LISTS:
        _LISTS[name] = synth.workload_list(name)
    return _LISTS[name]


def table_bench(a):
    """Batched (T,P) table: 120 cross-sections (10 T x 12 P) of the resident
    10^7-line list over 500-15000 cm^-1 at 0.01 cm^-1, Q(T) on the device,
    one graph-captured launch sequence per batch, results left in HBM.
    The plain FP64 path runs a 6-point subset of the same table."""
    from gexs_amd import synth
    from gexs_amd.manager import HipManager
    name = "voigt_tp_table_1e7"
    spec = synth.WORKLOADS[name]
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    lst = big_list(name)
    tps = synth.tp_table()
    sub6 = tps[5::20]
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 1000.0, 1.0, dnu, spec["mass"], 1.0, synth.T_REF,
                          synth.P_SCALE_BAR, cutoff=spec["cutoff"])
    m.load_list(lst)
    best, _ = m.autotune(sub6, start, end, N)
    log(f"[bench] table autotune on 6 (T,P): {best}")
    m.set_launch(best["persistent_blocks"], True)
    # results stay in HBM (out="device"); run_batch is synchronous
    m.run_batch(tps, start, end, N, out="device")  # captures the graph
    m.run_batch(tps, start, end, N, out="device")
    m.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(a.table_steps):
        m.run_batch(tps, start, end, N, out="device")
    el = time.perf_counter() - t0
    k_ms, _, cnt = m.timing()
    st = m.stats()
    n_eval = st["n_eval"]
    value = n_eval * a.table_steps / el
    log(f"[bench] table: {len(tps)} x-sections in {el * 1e3 / a.table_steps:.1f} ms per batch, "
        f"{n_eval:.4e} evals, {value:.4e} evals/s, device {k_ms / cnt:.1f} ms per batch")
    m.close()
    m64 = HipManager("voigt", precision="fp64")
    m64.InitializeConstants(100.0, 1000.0, 1.0, dnu, spec["mass"], 1.0, synth.T_REF,
                            synth.P_SCALE_BAR, cutoff=spec["cutoff"])
    m64.load_list(lst)
    out64 = np.empty((len(sub6), N))
    m64.run_batch(sub6[:1], start, end, N, out=out64[:1])
    t1 = time.perf_counter()
    m64.run_batch(sub6, start, end, N, out=out64)
    e64 = time.perf_counter() - t1
    v64 = m64.stats()["n_eval"] / e64
    m64.close()
    log(f"[bench] table FP64 path on 6 (T,P): {v64:.4e} evals/s")
    return {"workload": f"{name}: {lst.n_trans} transitions / {lst.n_states} states, 120 (T,P) "
                        f"(T 200-3000 K x P 1e-5-100 bar, log-spaced), grid {start}-{end} cm-1 x "
                        f"{N} points, cutoff {spec['cutoff']} cm-1",
            "value": value, "unit": "evals/s", "xsec_per_s": len(tps) * a.table_steps / el,
            "ms_per_batch": el * 1e3 / a.table_steps, "n_eval_per_batch": n_eval,
            "device_ms_per_batch": k_ms / cnt, "graph": True, "tuning": best,
            "fp64_gpu_path": {"value": v64, "unit": "evals/s", "sample": "6 of the 120 (T,P)",
                              "speedup_fp32c_over_fp64": value / v64}}


def stream_bench(a):
    """The 10^8-transition list (2e6 states, FP32 A) at T = 2000 K, P = 0.1 bar
    over 100-30000 cm^-1 at 0.01 cm^-1.  Streamed: the list stays in host
    memory and goes through a 2 GiB device budget in list-order chunks
    (copy, sort and pass of successive chunks overlapped), results left in
    HBM; the line's value is this rate, host copies included.  Beside it the
    same list resident in HBM in one pass, and the plain FP64 path."""
    from gexs_amd import synth
    from gexs_amd.manager import HipManager
    name = "voigt_1e8_stream"
    spec = synth.WORKLOADS[name]
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    T, P = spec["T"], spec["P"]
    budget = 2 << 30
    lst = big_list(name)
    m = HipManager("voigt")
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF, synth.P_SCALE_BAR,
                          cutoff=spec["cutoff"])
    t0 = time.perf_counter()
    m.load_list(lst)
    log(f"[bench] stream: {lst.n_trans} transitions resident + sorted in "
        f"{time.perf_counter() - t0:.2f} s")
    best, _ = m.autotune([(T, P)], start, end, N)
    steps = max(1, a.table_steps)
    # resident: one pass over the whole selected list
    m.run_batch([(T, P)], start, end, N, out="device")
    m.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.run_batch([(T, P)], start, end, N, out="device")
    el_res = time.perf_counter() - t0
    k_ms, _, cnt = m.timing()
    st_res = m.stats()
    v_res = st_res["n_eval"] * steps / el_res
    # streamed at the same launch shape
    m.stream_xsec(lst, [(T, P)], start, end, N, mem_budget=budget, out="device")
    walls, devs, parts = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = m.stream_xsec(lst, [(T, P)], start, end, N, mem_budget=budget, out="device")
        walls.append(st["wall_ms"])
        devs.append(st["device_ms"])
        parts.append([st["stage_ms"], st["copy_ms"], st["sort_ms"], st["pass_ms"]])
    el = time.perf_counter() - t0
    value = st["n_eval"] * steps / el
    assert st["n_eval"] == st_res["n_eval"], (st["n_eval"], st_res["n_eval"])
    m.close()
    h2d_gb = lst.n_trans * (4 + 4 + 3 * lst.A.dtype.itemsize) / 1e9
    m64 = HipManager("voigt", precision="fp64")
    m64.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF, synth.P_SCALE_BAR,
                            cutoff=spec["cutoff"])
    m64.load_list(lst)
    m64.run_batch([(T, P)], start, end, N, out="device")
    t1 = time.perf_counter()
    m64.run_batch([(T, P)], start, end, N, out="device")
    e64 = time.perf_counter() - t1
    v64 = m64.stats()["n_eval"] / e64
    m64.close()
    log(f"[bench] stream: {st['n_chunks']} chunks of {st['chunk']} transitions in "
        f"{st['device_bytes'] / 2**30:.2f} GiB: {el * 1e3 / steps:.1f} ms per pass "
        f"(device {np.mean(devs):.1f} ms), {value:.4e} evals/s, {h2d_gb / (el / steps):.1f} GB/s "
        f"of list from host memory")
    parts = np.mean(np.array(parts), axis=0)
    log(f"[bench] stream: per pass, summed over the chunks: host staging {parts[0]:.1f} ms, "
        f"DMA {parts[1]:.1f} ms, keys+sort+gather {parts[2]:.1f} ms, passes {parts[3]:.1f} ms")
    log(f"[bench] stream: resident one pass {el_res * 1e3 / steps:.1f} ms, {v_res:.4e} evals/s "
        f"(device {k_ms / cnt:.1f} ms); FP64 path {v64:.4e} evals/s")
    return {"workload": f"{name}: {lst.n_trans} transitions / {lst.n_states} states in host "
                        f"memory, streamed through a {budget >> 30} GiB device budget in "
                        f"list-order chunks, T={T} K, P={P} bar, grid {start}-{end} cm-1 x "
                        f"{N} points, cutoff {spec['cutoff']} cm-1",
            "value": value, "unit": "evals/s", "ms_per_pass": el * 1e3 / steps,
            "device_ms": float(np.mean(devs)), "n_eval": st["n_eval"], "n_lines": st["n_lines"],
            "stage_ms_summed": {"host_staging": parts[0], "dma": parts[1], "sort": parts[2],
                                "passes": parts[3]},
            "chunks": st["n_chunks"], "chunk_transitions": st["chunk"],
            "device_bytes": st["device_bytes"], "list_bytes_from_host": h2d_gb * 1e9,
            "tuning": best,
            "resident": {"value": v_res, "unit": "evals/s", "ms_per_pass": el_res * 1e3 / steps,
                         "device_ms": k_ms / cnt,
                         "note": "the whole selected list resident in HBM, one pass"},
            "fp64_gpu_path": {"value": v64, "unit": "evals/s", "speedup_fp32c_over_fp64": value / v64}}


def serve_requests(rng, n):
    """Requests of the serving workload: T ~ U[200, 3000] K, P log-uniform in
    [1e-5, 100] bar, a range of width log-uniform in [100, 5000] cm^-1 placed
    uniformly inside 500-15000 cm^-1, resolution in {0.01, 0.1, 1.0}."""
    out = []
    for _ in range(n):
        T = rng.uniform(200.0, 3000.0)
        P = 10.0 ** rng.uniform(-5.0, 2.0)
        w = float(np.exp(rng.uniform(np.log(100.0), np.log(5000.0))))
        lo = rng.uniform(500.0, 15000.0 - w)
        res = float(rng.choice([0.01, 0.1, 1.0]))
        w = round(w / res) * res
        out.append((T, P, lo, lo + w, res))
    return out


def serve_bench(a):
    """Poisson arrivals at each rate against the continuous-batching server
    over the resident 10^7-line list; latency = completion - arrival (host
    clock).  The one-request-at-a-time plain FP64 path is timed on the same
    requests and its FIFO queue is replayed on the same arrival times."""
    from gexs_amd import synth
    from gexs_amd.manager import HipManager
    name = "voigt_tp_table_1e7"
    spec = synth.WORKLOADS[name]
    lst = big_list(name)
    rates = [float(r) for r in a.serve_rates.split(",") if r]
    rng = np.random.default_rng(20240611)
    m = HipManager("voigt")
    m.InitializeConstants(100.0, 1000.0, 1.0, 0.01, spec["mass"], 1.0, synth.T_REF,
                          synth.P_SCALE_BAR, cutoff=spec["cutoff"])
    m.load_list(lst)
    m.set_tuning(4, 3, 4)  # the shape the autotuner picks on the single-(T,P) workload
    m.serve_start(max_inflight=4, mem_budget=24 << 30)
    for r in serve_requests(np.random.default_rng(1), 4):  # first-use costs
        m.wait(m.submit(*r))
    sweeps = []
    all_reqs = []
    for rate in rates:
        n = int(max(40, min(400, rate * 4.0)))
        reqs = serve_requests(rng, n)
        arr = np.cumsum(rng.exponential(1.0 / rate, n))
        lat = np.zeros(n)
        srv = np.zeros(n)  # submit -> completion as the server saw it
        dev = np.zeros(n)  # device time of the request's batch
        nev = np.zeros(n)  # line-point evaluations of the request
        pending = {}
        i = 0
        t0 = time.perf_counter()
        while i < n or pending:
            now = time.perf_counter() - t0
            while i < n and arr[i] <= now:
                pending[m.submit(*reqs[i])] = i
                i += 1
            for rid in list(pending):
                r = m.poll(rid)
                if r["status"] == 0:
                    k = pending.pop(rid)
                    lat[k] = time.perf_counter() - t0 - arr[k]
                    srv[k] = r["latency_ms"]
                    dev[k] = r["device_ms"]
                    nev[k] = r["n_eval"]
                    m.release(rid)
            # the client polls without sleeping while requests are out, and
            # sleeps to within 0.3 ms of the next arrival when none is (sleep
            # overshoots by ~0.1 ms, which would count as latency)
            if i < n and not pending:
                gap = arr[i] - (time.perf_counter() - t0) - 3e-4
                if gap > 0:
                    time.sleep(min(gap, 0.002))
        span = time.perf_counter() - t0
        sweeps.append({"rate": rate, "n": n, "req_per_s": n / span,
                       "p50_ms": float(np.percentile(lat, 50) * 1e3),
                       "p99_ms": float(np.percentile(lat, 99) * 1e3), "arrivals": arr,
                       "server_p50_ms": float(np.percentile(srv, 50)),
                       "server_p99_ms": float(np.percentile(srv, 99)),
                       "device_p50_ms": float(np.percentile(dev, 50)),
                       "device_p99_ms": float(np.percentile(dev, 99)),
                       "mean_points": float(np.mean([round((r[3] - r[2]) / r[4]) for r in reqs])),
                       "evals_per_s": float(nev.sum() / span),
                       # the slowest requests: points, N_eval, latency, server and device ms
                       "slowest": [[int(round((reqs[k][3] - reqs[k][2]) / reqs[k][4])),
                                    float(f"{nev[k]:.3g}"), round(float(lat[k] * 1e3), 2),
                                    round(float(srv[k]), 2), round(float(dev[k]), 2)]
                                   for k in np.argsort(lat)[::-1][:3]]})
        all_reqs.append(reqs)
        log(f"[bench] serve {rate:6.1f} req/s offered: {n / span:7.1f} req/s, "
            f"p50 {sweeps[-1]['p50_ms']:.2f} ms, p99 {sweeps[-1]['p99_ms']:.2f} ms (server "
            f"{sweeps[-1]['server_p50_ms']:.2f} / {sweeps[-1]['server_p99_ms']:.2f} ms, device "
            f"{sweeps[-1]['device_p50_ms']:.2f} / {sweeps[-1]['device_p99_ms']:.2f} ms), {sweeps[-1]['evals_per_s']:.3g} evals/s; slowest "
            f"[points, N_eval, latency, server, device ms]: {sweeps[-1]['slowest']}")
    m.serve_stop()
    m.close()
    # one request at a time on the plain FP64 path: service times, then the
    # FIFO queue on the same arrivals
    m64 = HipManager("voigt", precision="fp64")
    m64.InitializeConstants(100.0, 1000.0, 1.0, 0.01, spec["mass"], 1.0, synth.T_REF,
                            synth.P_SCALE_BAR, cutoff=spec["cutoff"])
    m64.load_list(lst)
    r0 = serve_requests(np.random.default_rng(1), 1)[0]
    m64.run_batch([(r0[0], r0[1])], r0[2], r0[3], int(round((r0[3] - r0[2]) / r0[4])))
    for sw, reqs in zip(sweeps, all_reqs):
        svc = []
        for (T, P, lo, hi, res) in reqs:
            npt = int(round((hi - lo) / res))
            t1 = time.perf_counter()
            m64.run_batch([(T, P)], lo, hi, npt)
            svc.append(time.perf_counter() - t1)
        arr = sw.pop("arrivals")
        fin, lat = 0.0, []
        for t_a, s in zip(arr, svc):
            fin = max(fin, t_a) + s
            lat.append(fin - t_a)
        sw["fp64_one_at_a_time"] = {"req_per_s": len(reqs) / max(fin, arr[-1]),
                                    "capacity_req_per_s": len(svc) / sum(svc),
                                    "p50_ms": float(np.percentile(lat, 50) * 1e3),
                                    "p99_ms": float(np.percentile(lat, 99) * 1e3)}
        log(f"[bench] fp64 one-at-a-time at {sw['rate']:6.1f} req/s: "
            f"p50 {sw['fp64_one_at_a_time']['p50_ms']:.2f} ms, "
            f"p99 {sw['fp64_one_at_a_time']['p99_ms']:.2f} ms, capacity "
            f"{sw['fp64_one_at_a_time']['capacity_req_per_s']:.1f} req/s")
    m64.close()
    return {"workload": f"{name} list resident; T U[200,3000] K, P log-U[1e-5,100] bar, range "
                        f"width log-U[100,5000] cm-1 in 500-15000, resolution in "
                        f"{{0.01, 0.1, 1.0}} cm-1, Poisson arrivals",
            "max_inflight": 4, "sweep": sweeps}


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


def valu_probe():
    """On-device VALU rates (gexs_amd/libgexs_peak.so): independent FP32 FMA
    chains (the vector rate this build can reach, packed FP32 being off),
    v_rcp_f32 chains, and the profile kernel's region-I sequence (six VALU
    instructions per point).  8192 workgroups of 256 threads, 1024 steps of
    16 chains per thread, 5 timed launches each."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "gexs_amd", "libgexs_peak.so"))
    lib.gexs_peak_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_float)]
    blocks, iters, reps = 8192, 1024, 5
    steps = float(blocks) * 256 * iters * 16  # chain steps (lane-instructions) per launch
    ms = {}
    for kind, name in ((0, "fma"), (1, "rcp"), (2, "region1")):
        t = ctypes.c_float()
        rc = lib.gexs_peak_run(kind, blocks, iters, reps, ctypes.byref(t))
        if rc != 0:
            raise RuntimeError(f"VALU probe {name} failed (hip error {rc})")
        ms[name] = t.value
    fma_rate = steps / (ms["fma"] * 1e-3)  # lane-FMAs/s
    res = {"fp32_fma_tflops": 2.0 * fma_rate / 1e12,
           "rcp_issue_cycles_per_fma": ms["rcp"] / ms["fma"],
           "region1_evals_per_s": steps / (ms["region1"] * 1e-3),
           "region1_issue_cycles_per_eval": ms["region1"] / ms["fma"],
           "probe": "8192 x 256 threads x 1024 steps x 16 independent chains"}
    log(f"[bench] VALU probe: FP32 FMA {res['fp32_fma_tflops']:.1f} TFLOP/s, rcp "
        f"{res['rcp_issue_cycles_per_fma']:.2f} FMA slots, region-I sequence "
        f"{res['region1_evals_per_s']:.4e} evals/s ({res['region1_issue_cycles_per_eval']:.2f} "
        f"FMA slots per evaluation)")
    return res


def execute_bench(a, workload, profile=None, steps=None, warm=None, fp64_steps=None,
                  tuned=None, log_candidates=False, alone=10):
    """ExecuteCrossSection over the resident list of `workload` (its profile,
    or `profile`): autotuned launch shape (or `tuned`, or --ppl), `warm`
    untimed then `steps` timed steps, HIP events around each profile kernel;
    then `alone` steps each followed by a synchronisation (the profile kernel
    with nothing beside it: consecutive executes of short kernels overlap in
    the timed steps, which lengthens each kernel's event span); then the plain
    FP64 path on the same input.  Returns the measurements and the manager
    inputs the headline also uses."""
    from gexs_amd import synth
    from gexs_amd.manager import HipManager
    spec = synth.WORKLOADS[workload]
    prof = profile or spec["profile"]
    steps = a.steps if steps is None else steps
    warm = a.warmup if warm is None else warm
    fp64_steps = a.fp64_steps if fp64_steps is None else fp64_steps
    T = spec["T"] if spec["T"] else 1000.0
    P = spec["P"] if spec["P"] is not None else 0.0
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    lst = big_list(workload)
    m = HipManager(prof)
    if tuned is not None:
        m.set_tuning(tuned["points_per_lane"], tuned["order"], tuned["waves_per_tile"])
        m.set_launch(tuned["persistent_blocks"], False)
    elif a.ppl:
        m.set_tuning(a.ppl, a.order, a.split)
        m.set_launch(a.persistent, False)
    m.InitializeVectors(N, steps + warm + alone)
    m.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF, synth.P_SCALE_BAR,
                          cutoff=spec["cutoff"])
    m.TransferFreq(start + np.arange(N, dtype=np.float64) * dnu)
    t0 = time.perf_counter()
    n = m.load_list(lst)  # H2D, device assembly and nu sort; outside the timed region
    Q = float(m.partition([T])[0])
    log(f"[bench] {workload} ({prof}): list resident + sorted in {time.perf_counter() - t0:.2f} s, "
        f"Q(T)={Q:.6g}")
    tuning = None
    if tuned is None and not a.no_autotune and not a.ppl:
        best, cands = m.autotune([(T, P, Q)], start, end, N)
        tuning = {"best": best, "candidates": cands}
        log(f"[bench] {workload} ({prof}) autotune: {best}")
        if log_candidates:
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
    k_ms, _, cnt = m.timing()
    assert cnt == steps, (cnt, steps)
    st = m.stats()
    m.timing(reset=True)
    for _ in range(alone):
        m.ExecuteCrossSection(N, n, 0, T, P, Q)
        m.sync()
    k_alone_ms, _, cnt_alone = m.timing()
    m.close()
    k_alone = k_alone_ms / cnt_alone if cnt_alone else None
    value = n_eval * steps / elapsed
    k_avg = k_ms / cnt
    flops = FLOP_PER_EVAL[prof]
    achieved = n_eval * flops / (k_avg * 1e-3)
    log(f"[bench] {workload} ({prof}): {steps} steps in {elapsed * 1e3:.1f} ms; kernel avg "
        f"{k_avg:.4f} ms; N_eval {n_eval:.4e}; {value:.4e} evals/s")
    fp64 = None
    if fp64_steps > 0:
        m64 = HipManager(prof, precision="fp64")
        m64.InitializeVectors(N, fp64_steps + 1)
        m64.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF,
                                synth.P_SCALE_BAR, cutoff=spec["cutoff"])
        m64.TransferFreq(start + np.arange(N, dtype=np.float64) * dnu)
        m64.load_list(lst)
        m64.ExecuteCrossSection(N, n, 0, T, P, Q)
        m64.sync()
        t1 = time.perf_counter()
        for _ in range(fp64_steps):
            m64.ExecuteCrossSection(N, n, 0, T, P, Q)
        m64.sync()
        e64 = time.perf_counter() - t1
        v64 = n_eval * fp64_steps / e64
        fp64 = {"value": v64, "unit": "evals/s", "ms_per_step": e64 * 1e3 / fp64_steps,
                "speedup_fp32c_over_fp64": value / v64}
        log(f"[bench] {workload} ({prof}): plain FP64 path {v64:.4e} evals/s (FP32C "
            f"{value / v64:.2f}x faster)")
        m64.close()
    used = tuning["best"] if tuning else (tuned or {
        "points_per_lane": a.ppl or 4, "waves_per_tile": a.split, "order": a.order,
        "persistent_blocks": a.persistent})
    return {"workload": workload, "profile": prof, "lst": lst, "T": T, "P": P, "Q": Q, "n": n,
            "spec": spec, "steps": steps, "warm": warm, "elapsed": elapsed, "value": value,
            "n_eval": n_eval, "k_avg": k_avg, "k_alone": k_alone, "flops": flops, "achieved": achieved,
            "tuning": tuning, "used": used, "tiles": st["n_tiles"], "fp64": fp64}


VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4  # wave64 FP32 VALU instructions/s: 1024 SIMDs, one per 4 cycles


def roofline_of(r, prof=None, probe=None):
    """The profile kernel's FP32 VALU rate against the MI355X FP32 vector peak
    (frac_unpacked: against half of it, the rate without packed FP32
    instructions, which this build does not emit).  With the committed
    counters of the same launch shape (prof), the HBM bytes per launch
    (traffic) and the fraction of the VALU issue rate the kernel reaches at
    this run's kernel time."""
    out = {"bound": "valu", "compute": "fp32 VALU", "achieved": r["achieved"] / 1e12,
           "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": r["achieved"] / PEAK_FP32,
           "frac_unpacked": r["achieved"] / (PEAK_FP32 / 2),
           "traffic": prof.get("hbm_bytes_per_launch") if prof else None,
           "kernel": f"k_profile_tiles<{r['profile'].upper()}>", "valu_probe": probe,
           "kernel_ms_avg": r["k_avg"], "flop_per_eval": r["flops"],
           "flop_note": "achieved counts the reference's flops per evaluation (12 Voigt, 16 "
                        "bin-Gauss) for every line-point pair in the windows; the kernel sums far "
                        "wings as a polynomial per tile and skips exact zeros, so frac can pass 1; "
                        "valu_issue_frac is the hardware's own measure"}
    if r.get("k_alone"):
        # the kernel with nothing beside it (each execute synchronised)
        out["kernel_ms_alone"] = r["k_alone"]
        out["frac_alone"] = r["n_eval"] * r["flops"] / (r["k_alone"] * 1e-3) / PEAK_FP32
    if prof and prof.get("valu_insts_per_launch"):
        same = all(prof["launch"].get(k) == r["used"].get(k)
                   for k in ("points_per_lane", "waves_per_tile"))
        out["counters_launch"] = prof["launch"]
        out["valu_insts_per_eval"] = prof["valu_insts_per_launch"] / (r["n_eval"] / 64)
        if same:
            out["valu_issue_frac"] = prof["valu_insts_per_launch"] / (r["k_avg"] * 1e-3) / VALU_ISSUE_PEAK
    return out


def side_workload(a, workload, profile=None, steps=None):
    """An extra execute-path workload for the JSON: throughput, roofline and
    the plain FP64 path beside it."""
    r = execute_bench(a, workload, profile=profile, steps=steps or max(5, a.steps // 2),
                      warm=max(1, a.warmup), fp64_steps=max(1, a.fp64_steps))
    sp = r["spec"]
    start, end, N = sp["grid"]
    return {"workload": f"{workload}: {r['profile']}, {r['lst'].n_trans} transitions / "
                        f"{r['lst'].n_states} states, T={r['T']} K, P={r['P']} bar, grid {start}-"
                        f"{end} cm-1 x {N} points, cutoff {sp['cutoff']} cm-1",
            "value": r["value"], "unit": "evals/s", "ms_per_step": r["elapsed"] * 1e3 / r["steps"],
            "steps": r["steps"], "n_eval_per_step": r["n_eval"], "tuning": r["used"],
            "roofline": roofline_of(r, traffic_from_profile(f"{workload}_{r['profile']}")),
            "fp64_gpu_path": r["fp64"]}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the cross-section path runs on one GPU here; --gpus must be "
                                   "1 (sharding by wavenumber range over GPUs is not built)"}))
        sys.exit(2)
    if a.calib_copy:
        calib_copy(a.calib_copy)
        return
    from gexs_amd import synth
    from gexs_amd.manager import HipManager

    log(f"[bench] generating {a.workload}")
    r = execute_bench(a, a.workload, profile=a.profile, log_candidates=True)
    spec, prof, lst = r["spec"], r["profile"], r["lst"]
    T, P, Q = r["T"], r["P"], r["Q"]
    start, end, N = spec["grid"]
    dnu = (end - start) / N
    steps, warm, elapsed, value = r["steps"], r["warm"], r["elapsed"], r["value"]
    n_eval, k_avg, achieved, flops = r["n_eval"], r["k_avg"], r["achieved"], r["flops"]
    tuning, fp64 = r["tuning"], r["fp64"]
    log(f"[bench] {lst.n_trans} transitions over {lst.n_states} states")

    # PCIe-inclusive: the BaseManager boundary hands over host buffers, so
    # time steps that also move the line arrays in (TransferVectors, which
    # sorts them on the device) and the grid out (TransferResults)
    nu_h, E_h, A_h, g_h, gam_h, nex_h = lst.lines()
    mp = HipManager(prof)
    if tuning:
        b = tuning["best"]
        mp.set_tuning(b["points_per_lane"], b["order"], b["waves_per_tile"])
        mp.set_launch(b["persistent_blocks"], False)
    elif a.ppl:
        mp.set_tuning(a.ppl, a.order, a.split)
        mp.set_launch(a.persistent, False)
    mp.InitializeVectors(N, 1)
    mp.InitializeConstants(100.0, T, 1.0, dnu, spec["mass"], P, synth.T_REF, synth.P_SCALE_BAR,
                           cutoff=spec["cutoff"])
    freq = start + np.arange(N, dtype=np.float64) * dnu
    mp.TransferFreq(freq)
    host_out = np.zeros(N)
    pcie_steps = 5
    for k in range(pcie_steps + 1):
        if k == 1:
            t2 = time.perf_counter()
        mp.TransferVectors(nu_h.shape[0], E_h, nu_h, A_h, g_h, gam_h, nex_h)
        mp.ExecuteCrossSection(N, nu_h.shape[0], 0, T, P, Q)
        mp.TransferResults(freq, host_out, N, 0, accumulate=False)
    e2 = time.perf_counter() - t2
    pcie = {"value": n_eval * pcie_steps / e2, "unit": "evals/s",
            "ms_per_step": e2 * 1e3 / pcie_steps,
            "step": "TransferVectors (host FP64 line arrays, device nu sort) + "
                    "ExecuteCrossSection + TransferResults (host FP64 grid)"}
    log(f"[bench] PCIe-inclusive step {pcie['ms_per_step']:.2f} ms, {pcie['value']:.4e} evals/s")
    mp.close()

    doppler = voigt_quad = None
    if not a.no_profiles:
        # ~30 us passes: enough of them for a timed region of ~10 ms
        doppler = side_workload(a, "doppler_2e5", steps=max(5, 15 * a.steps))
        voigt_quad = side_workload(a, "voigt_1e6", profile="voigt_quad")
    table = None if a.no_table else table_bench(a)
    served = None if a.no_serve else serve_bench(a)
    stream = None if a.no_stream else stream_bench(a)

    probe = None if a.no_probe else valu_probe()

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
                        f"cutoff {spec['cutoff']} cm-1, list values (E, nu, A, gamma, n) held "
                        f"in FP64 ({lst.A.dtype.name} in the input list), FP64 pre-pass + FP32 "
                        f"Kahan-compensated profile sum",
            "n_eval_per_step": n_eval,
            "xsec_per_s": steps / elapsed,
            "tiles": r["tiles"],
            "tuning": r["used"],
        },
        "roofline": roofline_of(r, traffic_from_profile(a.workload), probe),
        "cpu_baseline": cpu,
        "fp64_gpu_path": fp64,
        "pcie_inclusive": pcie,
        "doppler": doppler,
        "voigt_quad": voigt_quad,
        "batched": table,
        "served": served,
        "stream": stream,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
