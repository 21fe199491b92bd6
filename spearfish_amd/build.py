"""Builds libsf_align.so (HIP, gfx950), libsf_diag.so (HIP diagnostics for tests and
measurements) and libsf_workload.so (host C++) in-tree.

    python -m spearfish_amd.build          # every library, sources compiled in parallel

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
HIP_SOURCES = ["affine_pk_pass.hip", "affine_pk4.hip", "affine_mix1.hip", "mix_persistent_affine1.hip", "serve_mix.hip", "affine_pk.hip", "affine_mix.hip",
               "mix_persistent_affine.hip", "myers_mix.hip", "mix_persistent_myers.hip", "noalloc.hip",
               "affine_pk64.hip", "affine.hip", "myers.hip", "runtime.hip", "serve.hip", "tune.hip",
               "affine_wide.hip", "tiebreak.hip", "myers_start.hip"]
DIAG_SOURCES = ["probe.hip"]
ALIGN_LIB = os.path.join(HERE, "libsf_align.so")
DIAG_LIB = os.path.join(HERE, "libsf_diag.so")
WORKLOAD_LIB = os.path.join(HERE, "libsf_workload.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def _build_hip(sources, lib) -> str:
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        futs = []
        for src in sources:
            obj = os.path.join(OBJ, src + ".o")
            objs.append(obj)
            futs.append(ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]))
        for f in futs:
            f.result()
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


def build_align() -> str:
    return _build_hip(HIP_SOURCES, ALIGN_LIB)


def build_diag() -> str:
    return _build_hip(DIAG_SOURCES, DIAG_LIB)


def build_workload() -> str:
    tmp = WORKLOAD_LIB + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-o", tmp,
          os.path.join(CSRC, "workload.cpp")])
    os.replace(tmp, WORKLOAD_LIB)
    return WORKLOAD_LIB


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        for f in [ex.submit(build_align), ex.submit(build_diag), ex.submit(build_workload)]:
            f.result()


if __name__ == "__main__":
    build_all()
    print(ALIGN_LIB)
    print(DIAG_LIB)
    print(WORKLOAD_LIB)
