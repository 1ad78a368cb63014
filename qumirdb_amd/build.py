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

SOURCES = ["runtime.hip", "filter_project.hip", "runtime_filter.hip", "shuffle.hip", "sort.hip", "staging.hip", "aggregate.hip", "aggregate_kw1.hip", "aggregate_kw2.hip",
           "aggregate_kw3.hip", "aggregate_kw4.hip", "join.hip", "primitives.hip", "exec.cpp", "jit.cpp", "autotune.cpp"]

# Device headers built into the library for kernels specialized at run time
# (name seen by the generated source's #include, path in the tree).
JIT_HEADERS = [("qdb_gpu.h", os.path.join(HERE, "..", "include", "qdb_gpu.h")),
               ("device.h", os.path.join(CSRC, "device.h")),
               ("aggregate_kernels.h", os.path.join(CSRC, "aggregate_kernels.h"))]
JIT_INC = os.path.join(CSRC, "jit_headers.inc")

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
        cmd = [HIPCC, "-O3", "-fPIC", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
               "-I" + os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include"), "-x", "c++", "-c",
               os.path.join(CSRC, src), "-o", obj]
    t0 = time.time()
    proc = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{out}")
    return obj, time.time() - t0, out


def write_jit_headers() -> None:
    """Writes csrc/jit_headers.inc: the device headers as string literals."""
    delim = "QDBJIT"
    names, texts = [], []
    for name, path in JIT_HEADERS:
        with open(path) as f:
            text = f.read()
        if f"){delim}\"" in text:
            raise RuntimeError(f"{path} contains the raw-string delimiter")
        names.append(f'"{name}"')
        texts.append(f'R"{delim}({text}){delim}"')
    out = ["// Generated by build.py from the device headers; do not edit.",
           f"static const int kJitHeaderCount = {len(names)};",
           "static const char* kJitHeaderNames[] = {" + ", ".join(names) + "};",
           "static const char* kJitHeaderTexts[] = {"] + [t + "," for t in texts] + ["};", ""]
    tmp = JIT_INC + ".tmp"
    with open(tmp, "w") as f:
        f.write("\n".join(out))
    os.replace(tmp, JIT_INC)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    write_jit_headers()
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
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
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
