"""The serving loop (sf_server_*): continuous batching of reads over a transcript store kept in
HBM, admission by HBM budget, per-read latency.  See include/sf_align.h.

    store, reads = generate_serve_workload(n_reads=...)
    with Server(store, band=64) as s:
        ticket = s.submit(reads, i); s.wait(ticket)
        s.replay(reads, rate=20000)  # Poisson arrivals
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .align import Affine, Mode


@dataclass
class Store:
    words: np.ndarray      # 2-bit packed transcripts (uint32)
    bases: int
    tx_off: np.ndarray     # per transcript
    tx_len: np.ndarray
    tx_gene: np.ndarray


@dataclass
class Reads:
    codes: np.ndarray      # one code per byte, reads back to back
    read_off: np.ndarray
    read_len: np.ndarray
    nseg: np.ndarray
    seg0: np.ndarray       # first segment of each read
    seg_read_off: np.ndarray
    seg_read_len: np.ndarray
    seg_tx_off: np.ndarray
    seg_tx_len: np.ndarray

    def __len__(self):
        return len(self.read_len)


def _wl():
    from .workload import lib
    L = lib()
    if not hasattr(L, "_serve_ready"):
        vp, u64, u32 = C.c_void_p, C.c_uint64, C.c_uint32
        L.sf_txome_generate.argtypes = [u64, u32, u32, C.POINTER(vp)]
        L.sf_txome_generate.restype = C.c_int
        L.sf_txome_bases.argtypes = [vp]
        L.sf_txome_bases.restype = u64
        L.sf_txome_count.argtypes = [vp]
        L.sf_txome_count.restype = u32
        L.sf_txome_export.argtypes = [vp] * 5
        L.sf_txome_free.argtypes = [vp]
        L.sf_reads_generate.argtypes = [vp, u64, u64, C.POINTER(vp)]
        L.sf_reads_generate.restype = C.c_int
        L.sf_reads_bases.argtypes = [vp]
        L.sf_reads_bases.restype = u64
        L.sf_reads_segments.argtypes = [vp]
        L.sf_reads_segments.restype = u64
        L.sf_reads_export.argtypes = [vp] * 9
        L.sf_reads_free.argtypes = [vp]
        L._serve_ready = True
    return L


def generate_serve_workload(n_reads: int, seed: int = 0, genes: int = 5000, transcripts: int = 20000):
    """A synthetic transcriptome (genes with exon skipping and alternative 3' sites) and reads
    with their isoform segments; definitions in spearfish_amd/csrc/workload.cpp."""
    L = _wl()
    t = C.c_void_p()
    if L.sf_txome_generate(seed, genes, transcripts, C.byref(t)) != 0:
        raise ValueError("transcriptome parameters")
    try:
        nb, nt = L.sf_txome_bases(t), L.sf_txome_count(t)
        words = np.empty((nb + 15) // 16, np.uint32)
        off, ln, gene = np.empty(nt, np.uint64), np.empty(nt, np.uint32), np.empty(nt, np.uint32)
        L.sf_txome_export(t, words.ctypes.data, off.ctypes.data, ln.ctypes.data, gene.ctypes.data)
        store = Store(words, nb, off, ln, gene)
        r = C.c_void_p()
        L.sf_reads_generate(t, seed, n_reads, C.byref(r))
        try:
            rb, ns = L.sf_reads_bases(r), L.sf_reads_segments(r)
            codes = np.empty(max(rb, 1), np.uint8)
            roff, rlen, nseg = np.empty(n_reads, np.uint64), np.empty(n_reads, np.uint32), np.empty(n_reads, np.uint32)
            sro, srl, sto, stl = (np.empty(max(ns, 1), np.uint32), np.empty(max(ns, 1), np.uint32),
                                  np.empty(max(ns, 1), np.uint64), np.empty(max(ns, 1), np.uint32))
            L.sf_reads_export(r, codes.ctypes.data, roff.ctypes.data, rlen.ctypes.data, nseg.ctypes.data,
                              sro.ctypes.data, srl.ctypes.data, sto.ctypes.data, stl.ctypes.data)
        finally:
            L.sf_reads_free(r)
    finally:
        L.sf_txome_free(t)
    seg0 = np.zeros(n_reads, np.uint64)
    if n_reads:
        seg0[1:] = np.cumsum(nseg[:-1], dtype=np.uint64)
    return store, Reads(codes, roff, rlen, nseg, seg0, sro[:ns], srl[:ns], sto[:ns], stl[:ns])


class Results:
    """Per-segment outputs for a set of reads (caller-owned buffers the server writes into)."""

    def __init__(self, reads: Reads):
        ns = len(reads.seg_read_len)
        self.dist = np.full(ns, 0xFFFFFFFF, np.uint32)
        self.end = np.full(ns, 0xFFFFFFFF, np.uint32)
        self.score = np.zeros(ns, np.int32)
        self.score_col = np.zeros(ns, np.uint32)


def requests(reads: Reads, res: Results, idx=None):
    """A ctypes array of sf_read_request for reads `idx` (all by default), outputs into `res`."""
    idx = np.arange(len(reads)) if idx is None else np.asarray(idx)
    arr = (N.ReadRequest * max(len(idx), 1))()
    base = reads.codes.ctypes.data
    for q, i in enumerate(idx):
        s0 = int(reads.seg0[i])
        arr[q] = N.ReadRequest(
            base + int(reads.read_off[i]), int(reads.read_len[i]), int(reads.nseg[i]),
            reads.seg_read_off.ctypes.data + 4 * s0, reads.seg_read_len.ctypes.data + 4 * s0,
            reads.seg_tx_off.ctypes.data + 8 * s0, reads.seg_tx_len.ctypes.data + 4 * s0,
            res.dist.ctypes.data + 4 * s0, res.end.ctypes.data + 4 * s0, res.score.ctypes.data + 4 * s0,
            res.score_col.ctypes.data + 4 * s0)
    return arr


class Server:
    def __init__(self, store: Store, hbm_budget_bytes: int = 256 << 20, slots: int = 4, max_reads_per_batch: int = 0,
                 myers_mode: Mode = Mode.Infix, affine_mode: Mode = Mode.Global, sc: Affine = Affine.ONT,
                 band: int = 64, batch_wait_us: int = 0):
        self.lib = N.load()
        self.store = store
        cfg = N.ServerConfig(hbm_budget_bytes, slots, max_reads_per_batch, int(myers_mode), int(affine_mode), sc.c(),
                             band, batch_wait_us)
        h = C.c_void_p()
        N.check(self.lib.sf_server_create(store.words.ctypes.data, store.bases, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.sf_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, req: N.ReadRequest) -> int:
        t = C.c_uint64()
        N.check(self.lib.sf_server_submit(self.h, C.byref(req), C.byref(t)))
        return t.value

    def submit_status(self, req: N.ReadRequest) -> int:
        t = C.c_uint64()
        return self.lib.sf_server_submit(self.h, C.byref(req), C.byref(t))

    def wait(self, ticket: int, timeout: float = -1.0) -> int:
        return self.lib.sf_server_wait(self.h, ticket, timeout)

    def cancel(self, ticket: int) -> int:
        return self.lib.sf_server_cancel(self.h, ticket)

    def stats(self, reset: bool = False) -> N.ServerStats:
        s = N.ServerStats()
        N.check(self.lib.sf_server_get_stats(self.h, C.byref(s), int(reset)))
        return s

    def replay(self, reqs, arrivals: np.ndarray) -> float:
        """Submits reqs[i] at arrivals[i] seconds; returns the wall time (first arrival to last
        completion).  Requests refused at admission are counted in self.refused."""
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        wall = C.c_double()
        ref = C.c_uint64()
        N.check(self.lib.sf_server_replay(self.h, reqs, arr.ctypes.data_as(C.POINTER(C.c_double)), len(arr),
                                          C.byref(wall), C.byref(ref)))
        self.refused = ref.value
        return wall.value


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))
