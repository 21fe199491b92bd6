"""Builds libsf_align.so (HIP, gfx950) and libsf_workload.so (host C++) in-tree.

    python -m spearfish_amd.build          # both libraries, sources compiled in parallel

The libraries land next to this file so they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
# Slowest to compile first: the build runs them on a pool of 8.
HIP_SOURCES = ["affine_mix1.hip", "serve_mix.hip", "affine_pk.hip", "affine_mix.hip", "myers_mix.hip",
               "affine_pk64.hip", "affine.hip", "myers.hip", "runtime.hip", "serve.hip", "probe.hip"]
ALIGN_LIB = os.path.join(HERE, "libsf_align.so")
WORKLOAD_LIB = os.path.join(HERE, "libsf_workload.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build_align() -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        futs = []
        for src in HIP_SOURCES:
            obj = os.path.join(OBJ, src + ".o")
            objs.append(obj)
            futs.append(ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
        for f in futs:
            f.result()
    tmp = ALIGN_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, ALIGN_LIB)
    return ALIGN_LIB


def build_workload() -> str:
    tmp = WORKLOAD_LIB + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-o", tmp,
          os.path.join(CSRC, "workload.cpp")])
    os.replace(tmp, WORKLOAD_LIB)
    return WORKLOAD_LIB


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        a = ex.submit(build_align)
        w = ex.submit(build_workload)
        a.result()
        w.result()


if __name__ == "__main__":
    build_all()
    print(ALIGN_LIB)
    print(WORKLOAD_LIB)
