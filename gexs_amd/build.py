"""Builds gexs_amd/libgexs_hip.so for gfx950 with hipcc (in-tree).

    python -m gexs_amd.build            # or gexs_amd.build.build()

All HIP sources are compiled with packed FP32 instructions disabled
(``-target-feature -packed-fp32-ops``); the FP64 window arithmetic relies on
explicit ``#pragma clang fp contract(off)`` regions, not on global flags.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgexs_hip.so")
PEAK_LIB = os.path.join(HERE, "libgexs_peak.so")

SOURCES = ["kernels.hip", "tiles_voigt.hip", "tiles_quad.hip", "tiles_gauss.hip", "sort.hip",
           "serve.hip", "capi.hip"]
HOST_SOURCES = ["HipManager.cpp", "host/readers.cpp", "host/cross_section.cpp"]
CLI = os.path.join(HERE, "gexs_xsec")
ARCH = os.environ.get("GEXS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gexs HIP library cannot be built")


def flags():
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall",
            "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _compile(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"{what} failed")


def build(verbose: bool = False) -> str:
    cc = hipcc()
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    jobs = []  # (command, description, object)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        jobs.append(([cc, *flags(), "-c", path, "-o", obj], f"hipcc on {src}", obj))
    cxx = shutil.which("g++") or "g++"
    hflags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"),
              "-I", CSRC]
    for src in HOST_SOURCES:
        obj = os.path.join(bdir, os.path.basename(src).replace(".cpp", ".o"))
        jobs.append(([cxx, *hflags, "-c", os.path.join(CSRC, src), "-o", obj], f"g++ on {src}",
                     obj))
    # the translation units are independent: compile them side by side
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        for f in [ex.submit(_compile, cmd, what, verbose) for cmd, what, _ in jobs]:
            f.result()
    objs = [obj for _, _, obj in jobs]
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread",
           "-l:libbz2.so.1"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libgexs_hip.so failed")
    os.replace(tmp, LIB)
    cmd = [cxx, *hflags, os.path.join(CSRC, "host", "xsec_main.cpp"), "-o", CLI + ".tmp",
           "-L", HERE, "-lgexs_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of gexs_xsec failed")
    os.replace(CLI + ".tmp", CLI)
    # the VALU rate probes bench.py runs beside the profile kernel
    obj = os.path.join(bdir, "peak.o")
    for cmd in ([cc, *flags(), "-c", os.path.join(CSRC, "peak.hip"), "-o", obj],
                [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", PEAK_LIB + ".tmp", obj]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("build of libgexs_peak.so failed")
    os.replace(PEAK_LIB + ".tmp", PEAK_LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
