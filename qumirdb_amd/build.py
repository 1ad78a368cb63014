"""Builds the HIP library libqdb_gpu.so for gfx950 in-tree.

Every source is compiled with hipcc (in parallel), then linked into
qumirdb_amd/libqdb_gpu.so. The packed-FP32 target feature is disabled for all
device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqdb_gpu.so")

SOURCES = ["runtime.hip", "filter_project.hip", "aggregate.hip", "aggregate_kw1.hip", "aggregate_kw2.hip",
           "aggregate_kw3.hip", "aggregate_kw4.hip", "join.hip", "primitives.hip", "exec.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("QDB_GPU_ARCH", "gfx950")
FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str) -> tuple[str, float, str]:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-Wall", "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    proc = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    return obj, time.time() - t0, out


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(_compile, SOURCES))
    objs = []
    for (obj, secs, out), src in zip(results, SOURCES):
        objs.append(obj)
        if verbose:
            print(f"[qdb_gpu build] {src}: {secs:.1f}s", flush=True)
            if out.strip():
                print(out, flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[qdb_gpu build] linked {LIB}", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
