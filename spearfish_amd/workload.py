"""Synthetic workloads of the benchmark configurations (libsf_workload.so, host C++).

    pairs, bands = generate(cfg, n, seed)   # cfg 1 parity set, 2 uniform, 3 uneven, 4 long

Definitions are in spearfish_amd/csrc/workload.cpp.  The data are synthetic by construction
(random ACGT targets, error-model queries); nothing is downloaded.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as N
from .align import PairSet

NAMES = {1: "parity-4096x200-1000bp-8pct-band64", 2: "uniform-1e6x1000bp-6pct-band128",
         3: "uneven-5e5-lognormal800-5to15pct-adaptive-band", 4: "long-2e5x2to10kb-12pct-band256"}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(N.WORKLOAD_LIB):
            raise N.NativeError(N.SF_ENODEV, f"{N.WORKLOAD_LIB} is missing: run `python -m spearfish_amd.build`")
        L = C.CDLL(N.WORKLOAD_LIB)
        vp, u64 = C.c_void_p, C.c_uint64
        L.sf_workload_generate.argtypes = [C.c_int, u64, u64, C.c_int, C.POINTER(vp)]
        L.sf_workload_generate.restype = C.c_int
        for f in ("sf_workload_size", "sf_workload_pat_bases", "sf_workload_txt_bases"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = u64
        L.sf_workload_export.argtypes = [vp] * 8
        L.sf_workload_free.argtypes = [vp]
        _lib = L
    return _lib


def generate(cfg: int, n: int, seed: int = 0, threads: int | None = None):
    """(PairSet of one-code-per-byte pairs, per-pair bands as uint64)."""
    L = lib()
    threads = threads or min(16, os.cpu_count() or 1)
    h = C.c_void_p()
    if L.sf_workload_generate(cfg, seed, n, threads, C.byref(h)) != 0:
        raise ValueError(f"workload {cfg}")
    try:
        pb, tb = L.sf_workload_pat_bases(h), L.sf_workload_txt_bases(h)
        pat = np.empty(max(pb, 1), np.uint8)
        txt = np.empty(max(tb, 1), np.uint8)
        poff = np.empty(n, np.uint64)
        plen = np.empty(n, np.uint32)
        toff = np.empty(n, np.uint64)
        tlen = np.empty(n, np.uint32)
        band = np.empty(n, np.uint64)
        L.sf_workload_export(h, pat.ctypes.data, poff.ctypes.data, plen.ctypes.data, txt.ctypes.data,
                             toff.ctypes.data, tlen.ctypes.data, band.ctypes.data)
    finally:
        L.sf_workload_free(h)
    return PairSet(pat, poff, plen, txt, toff, tlen), band
