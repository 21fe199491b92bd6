"""Benchmark of the alignment hot path on one MI355X.

Workload (default): 1,000,000 synthetic pairs of 1,000 bp targets with queries at 6% ONT-like
error, band 128 -- the uniform roofline batch.  One step aligns every pair as the reference's
kernels do (scores only, one launch per kernel shape):
  1. Myers semi-global (Infix) edit distance with the end column, k unbounded;
  2. banded affine Global score, ONT scoring (2, 4, 4, 2), band 128.
The same step with the affine traceback (CIGAR, this project's addition) is timed after it and
reported under "with_cigar".  Inputs are packed 2-bit into HBM before the timed region; results
stay in HBM.

value = DP cells of a step (sum of m*n for Myers plus the affine band cells, counted as the
reference's DP_WORK does) / wall time per step, in GCUPS.

    python bench.py [--steps K] [--warmup W] [--pairs N] [--no-cpu] [--no-extra]
    python bench.py --workload uneven|long|serve|tiebreak   # one of the other workloads alone

Prints one JSON line.  By default the line also carries the other workloads under
"other_workloads" (about a minute more): the uneven batch with its CPU baseline and the ratio to
it, the long-segment batch, the serving loop, and the tie-break's call mix.

Other workloads (same JSON shape, their own config):
  uneven  5e5 pairs, targets LogNormal(median 800 bp, sigma 0.8) clipped to 100..20,000, 5-15%
          error, band max(32, 1.2|m - n| + 32) capped at 512; Myers Infix + affine Global scores.
  long    2e5 pairs, queries 2-10 kb, targets 1.05x, 12% error; Myers Infix distance, end and start,
          then affine Global band 256 with CIGAR on text[start..end).
  serve   20,000-transcript synthetic transcriptome, reads of median 1.2 kb at 8% error with 1-16
          isoform segments, Poisson arrivals at 5k, 20k and 50k reads/s; reads/s and p50/p99
          latency with continuous batching next to one read per launch.
  tiebreak  the tie-break's calls (read ends and interior stretches against each isoform's text)
          of 20,000 reads of the serving transcriptome, all in one call over the device-resident
          packed store; candidate alignments/s, next to one call per segment and the CPU port.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "GCUPS (bit-exact) batched vs per-pair on 1 GPU; serve reads/s + p50/p99 latency"
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4 cycles at 2.4 GHz,
# in lane-instructions (x 64): 39.3 T/s.  Packed 16-bit instructions do two lane-ops each.
VALU_ISSUE_PEAK_T = 256 * 4 * 64 / 4 * 2.4e9 / 1e12
VALU16_PEAK_TOPS = 2 * VALU_ISSUE_PEAK_T
HBM_PEAK_GBS = 8000.0
# The reference's i32 recurrence, per cell (F: 2 sub + 2 max; substitution: compare + select;
# x: add + max; E/H: 2 max + 1 sub); the traceback adds 4 compares.
AFFINE_OPS_PER_CELL = 11
AFFINE_TB_OPS_PER_CELL = 11 + 4
# Issued VALU wave-instructions per DP cell on the uniform workload, from rocprofv3 --pmc
# SQ_INSTS_VALU runs of this script (profiles/valu_issue_uniform_1e6x1000_band128.md): the
# score step's affine phase (its 4-lane packed bins and the LP64 bin) and the traceback DP.
VALU_INSTS_PER_CELL = {"score": 1.62788e10 / 2.46731e11, "traceback": 5.4447e10 / 2.46731e11}
# HBM bytes per launch of the dominant kernels on the uniform workload, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE runs of this script, scaled by the same counters on a 1 GiB copy
# (FETCH x2.0, WRITE x1.0); profiles/affine_uniform_1e6x1000_band128_traffic.md.  Score: the
# step's affine phase; traceback: the average of the DP's four launches a step.
TRAFFIC_BYTES = {"score": 729_798_181, "traceback": 49_847_008_272}
# Issued VALU wave-instructions of one extension pass of the tie-break workload (its four
# affine_pk16_pass_kernel launches, 16-bit packed, 8 lanes a pair), from a rocprofv3 --pmc
# SQ_INSTS_VALU run of this script (profiles/tiebreak_20k_reads_passes.md).
TB_EXT_VALU_INSTS = 5.2471e8 + 1.9663e8 + 9.3688e6 + 5.0622e5


# Issue-rate probe (sf_probe_issue): waves of 32 independent VALU instructions of one encoding.
PROBE_KINDS = {"vop2_i32": 0, "vop3p_pk16": 1, "dpp": 2, "vop3_i32": 3}
# The probe workload's further kinds: which instruction streams the SIMD issues two a quad-cycle.
PROBE_PAIRING_KINDS = {"bitop3_alone": 4, "alignbit_alone": 5, "or7_alignbit1": 6, "or31_alignbit1": 7,
                       "add_co_addc_pairs": 8}


def issue_probe(blocks=8192, iters=20000, kinds=PROBE_KINDS):
    """Measured VALU issue rates on this GPU, in T lane-instructions/s, per encoding: two-operand
    32-bit ops (VOP2), packed 16-bit ops (VOP3P, the affine kernels'), DPP moves, three-operand
    32-bit ops (VOP3).  About 35 ms of GPU time per kind."""
    import ctypes as C
    from spearfish_amd import _native as N
    lib = N.load_diag()
    ms = C.c_float()
    out = {}
    for name, kind in kinds.items():
        N.check_diag(lib.sf_probe_issue(kind, iters, blocks, C.byref(ms)))
        out[name] = round(blocks * 4 * iters * 32 * 64 / (ms.value * 1e-3) / 1e12, 2)
    return out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-pairs", type=int, default=0, help="CPU baseline sample size (0: auto)")
    ap.add_argument("--per-pair", type=int, default=256, help="pairs for the per-pair-launch comparison")
    ap.add_argument("--workload", default="uniform", choices=["uniform", "uneven", "long", "serve", "tiebreak", "probe"])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-replays", type=int, default=0,
                    help="serve: replay the 50k reads/s row this many more times in the same process "
                         "and report each replay's p99")
    ap.add_argument("--no-extra", action="store_true", help="the uniform workload only")
    ap.add_argument("--tune-out", default=None, help="write the autotuner's measurements here (uneven)")
    ap.add_argument("--no-schedules", action="store_true",
                    help="uneven: skip the schedule comparison and the autotuner (for profiled runs)")
    return ap.parse_args()


def host_threads():
    """The host cores this process may run on (its CPU affinity), capped by OMP_NUM_THREADS where
    the environment sets a share: the CPU baseline runs on all of them."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(share))) if share.isdigit() and int(share) > 0 else max(1, n)


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def cpu_baseline(pairs, bands, gpu, threads):
    """The reference's CPU path restated in C on a bounded sample, all host threads: Myers as the
    scalar distance_with (the distinct patterns leave batch_distances one text a call), the affine
    pass as the x86 dispatch runs it -- 16-bit AVX-512 rows where `narrow` holds, 32-bit AVX-512
    rows otherwise -- or the scalar rows on a host without AVX-512."""
    from oracle import sf_oracle as O
    simd = O.avx512_available()
    n = len(pairs)
    cells = (pairs.pat_len.astype(np.float64) * pairs.txt_len).sum()
    t0 = time.perf_counter()
    d = O.batch_myers(pairs.pat, pairs.pat_off, pairs.pat_len, pairs.txt, pairs.txt_off, pairs.txt_len, O.INFIX,
                      O.NONE, threads)
    t1 = time.perf_counter()
    s, _ = O.batch_affine(pairs.pat, pairs.pat_off, pairs.pat_len, pairs.txt, pairs.txt_off, pairs.txt_len,
                          O.GLOBAL, O.ONT, bands, threads, avx512=simd)
    t2 = time.perf_counter()
    acells = sum(O.affine_cells(int(pairs.pat_len[i]), int(pairs.txt_len[i]), O.GLOBAL, int(bands[i]))
                 for i in range(n))
    exact = bool(np.array_equal(d, gpu["dist"][:n]) and np.array_equal(s, gpu["score"][:n]))
    return {"cells": cells + acells, "seconds": t2 - t0, "myers_s": t1 - t0, "affine_s": t2 - t1,
            "bit_exact": exact, "pairs": n, "rows": "avx512 (i16 where narrow holds, else i32)" if simd else "scalar"}


def ukkonen_rows(batch, steps=3):
    """The workload's Myers pass with a threshold (the reference's capped calls: Mapper::banded
    aligns later texts at k = best + cap, crates/map/src/lib.rs:851-911) and the reversed pass that
    finds `start`, over Ukkonen's band and over every block (sf_set_myers_ukk): device ms a pass,
    and whether both give the same results."""
    from spearfish_amd import _native as N
    from spearfish_amd import align as A
    lib = N.load()
    rows = []
    for k, locate in ((A.U32_MAX, 2), (100, 1), (30, 1)):
        row = {"k": None if k == A.U32_MAX else k,
               "pass": "infix distance+end" + ("+start" if locate == 2 else "")}
        res = []
        for on in (0, 1):
            N.check(lib.sf_set_myers_ukk(on))
            batch.myers(A.Mode.Infix, k, locate)
            batch.sync()
            batch.reset_timing()
            for _ in range(steps):
                batch.myers(A.Mode.Infix, k, locate)
            batch.sync()
            tm = batch.timing()
            key = "band" if on else "every_block"
            row[f"{key}_ms"] = round(tm.myers_ms / steps, 2)
            if locate == 2:
                row[f"{key}_start_ms"] = round(tm.locate_ms / steps, 2)
            res.append(batch.myers_results(locate))
        row["bit_exact"] = all(np.array_equal(x, y) for x, y in zip(res[0], res[1]) if x is not None)
        log(f"ukkonen: {row}")
        rows.append(row)
    N.check(lib.sf_set_myers_ukk(1))
    return rows


def per_pair_launch_gcups(pairs, bands, k):
    """The same step launched one pair at a time (one batch of one per call)."""
    from spearfish_amd import align as A
    from spearfish_amd.align import PairSet
    sub = PairSet(pairs.pat, pairs.pat_off[:k], pairs.pat_len[:k], pairs.txt, pairs.txt_off[:k], pairs.txt_len[:k])
    b = A.Batch(sub)
    cells = 0
    singles = []
    for i in range(k):
        one = PairSet(pairs.pat, pairs.pat_off[i:i + 1], pairs.pat_len[i:i + 1], pairs.txt, pairs.txt_off[i:i + 1],
                      pairs.txt_len[i:i + 1])
        singles.append(A.Batch(one))
    b.close()
    t0 = time.perf_counter()
    for i, s in enumerate(singles):
        s.myers(A.Mode.Infix, locate=1)
        s.affine(A.Mode.Global, A.Affine.ONT, bands=bands[i:i + 1], cigar=False)
        s.sync()
    dt = time.perf_counter() - t0
    for i, s in enumerate(singles):
        t = s.timing()
        cells += t.myers_cells + t.affine_cells
        s.close()
    return cells / dt / 1e9


def bench_uniform(args):
    from spearfish_amd import align as A
    from spearfish_amd import workload as W

    t0 = time.perf_counter()
    pairs, bands = W.generate(2, args.pairs, args.seed)
    log(f"generated {len(pairs)} pairs in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    batch = A.Batch(pairs)
    log(f"packed and uploaded in {time.perf_counter() - t0:.1f}s")

    def run(cigar, steps, warmup, label):
        for w in range(warmup):
            batch.myers(A.Mode.Infix, locate=1)
            batch.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=cigar)
            batch.sync()
            log(f"{label}: warmup {w + 1}/{warmup} done")
        batch.reset_timing()
        batch.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.myers(A.Mode.Infix, locate=1)
            batch.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=cigar)
        batch.sync()
        wall = time.perf_counter() - t0
        tm = batch.timing()
        log(f"{label}: timed {steps} steps in {wall:.3f}s")
        return wall, tm

    def roofline(tm, phase, kernel, ops, launches_per_step):
        # The unit is one launch of the phase's dominant shape (the score phase's bins run as one
        # launch each on the batch's streams, so the phase is the unit there).  `achieved` is the
        # issued VALU lane-instructions a second: measured instructions per cell x cells / time;
        # `reference_ops` restates it at the reference's 11 i32 ops a cell (the packed kernels
        # issue fewer, so that figure can pass the packed-i16 peak).
        launches = max(launches_per_step * args.steps, 1)
        per_launch_ms = tm.affine_ms / launches
        cells_per_launch = tm.affine_cells / launches
        secs = per_launch_ms * 1e-3
        ipc = VALU_INSTS_PER_CELL[phase]
        achieved = cells_per_launch * ipc * 64 / secs / 1e12
        traffic = TRAFFIC_BYTES.get(phase)
        # The packed kernels issue VOP3P instructions: the probe's rate for that encoding is the
        # measured ceiling beside the nominal one.
        measured = probe["vop3p_pk16"]
        return {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 3),
                "peak": round(VALU_ISSUE_PEAK_T, 1),
                "unit": "T VALU lane-instructions/s issued (SQ_INSTS_VALU per cell x cells / launch time)",
                "frac": round(achieved / VALU_ISSUE_PEAK_T, 4), "traffic": traffic,
                "peak_measured_vop3p": measured, "frac_of_measured": round(achieved / measured, 4),
                "valu_insts_per_cell": round(ipc, 5), "cells_per_launch": int(cells_per_launch),
                "avg_launch_ms": round(per_launch_ms, 3),
                "hbm_gbs": round(traffic / secs / 1e9, 1) if traffic else None,
                "reference_ops": {"ops_per_cell": ops,
                                  "tops": round(cells_per_launch * ops / secs / 1e12, 3),
                                  "packed_i16_peak_tops": round(VALU16_PEAK_TOPS, 1)}}

    probe = issue_probe()
    log(f"issue probe (T lane-instructions/s): {probe}")
    wall, tm = run(False, args.steps, args.warmup, "scores")
    cells = tm.myers_cells + tm.affine_cells
    value = cells / wall / 1e9
    ms_per_step = wall / args.steps * 1e3
    dist, _, _ = batch.myers_results(0)
    res = batch.affine_results(cigar=False)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 bit-vectors (Myers) + u16x2 packed, exact (affine)",
        "data": "synthetic",
        "config": {"workload": "uniform-1e6x1000bp-6pct-band128", "pairs": len(pairs), "target_bp": 1000,
                   "error": "6% (3% sub, 1.5% ins, 1.5% del, geometric indels p=0.7)", "band": 128,
                   "step": "myers infix distance+end, affine global ONT band 128 score"},
        "myers_gcups": round(tm.myers_cells / (tm.myers_ms * 1e-3) / 1e9, 1) if tm.myers_ms else None,
        "affine_gcups": round(tm.affine_cells / (tm.affine_ms * 1e-3) / 1e9, 1) if tm.affine_ms else None,
        "kernel_ms_per_step": {"myers": round(tm.myers_ms / args.steps, 2),
                               "affine": round(tm.affine_ms / args.steps, 2)},
        "roofline": roofline(tm, "score", "affine_mix_kernel: affine_pk16_kernel<4,33..36,true> bodies (4 "
                             "lanes x 66, 68, 70 or 72 slots a pair, 16 pairs a wave) and the 64-lane bin in one "
                             "launch; the score step's affine phase", AFFINE_OPS_PER_CELL, 1),
        "issue_probe_t": probe,
    }
    wall_c, tm_c = run(True, args.steps, 1, "with cigar")
    out["with_cigar"] = {
        "value": round((tm_c.myers_cells + tm_c.affine_cells) / wall_c / 1e9, 2), "unit": "GCUPS",
        "ms_per_step": round(wall_c / args.steps * 1e3, 2),
        "affine_gcups": round(tm_c.affine_cells / ((tm_c.affine_ms + tm_c.walk_ms) * 1e-3) / 1e9, 1),
        "kernel_ms_per_step": {"myers": round(tm_c.myers_ms / args.steps, 2),
                               "affine_tb": round(tm_c.affine_ms / args.steps, 2),
                               "walk_compact": round(tm_c.walk_ms / args.steps, 2)},
        "roofline": roofline(tm_c, "traceback", "affine_pk16_tb_kernel<9> (one traceback chunk)",
                             AFFINE_TB_OPS_PER_CELL,
                             # every chunk is one launch of the dominant class plus, for this
                             # workload, one small LP64 launch a step for the widest pairs
                             max(tm_c.affine_launches // max(args.steps, 1) - 1, 1)),
    }
    # The same step synchronised one at a time: the median step time (robust to a slow outlier).
    times = []
    for _ in range(max(args.steps, 5)):
        t0 = time.perf_counter()
        batch.myers(A.Mode.Infix, locate=1)
        batch.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=False)
        batch.sync()
        times.append(time.perf_counter() - t0)
    out["ms_per_step_median"] = round(sorted(times)[len(times) // 2] * 1e3, 2)
    out["steps_for_median"] = len(times)
    out["ukkonen"] = ukkonen_rows(batch)
    # End to end from host buffers, never `value`: 2-bit packing and upload, the launch plans, both
    # kernels, and the distances, ends, scores and columns copied back.
    batch.close()
    t0 = time.perf_counter()
    b2 = A.Batch(pairs)
    b2.myers(A.Mode.Infix, locate=1)
    b2.affine(A.Mode.Global, A.Affine.ONT, bands=bands, cigar=False)
    b2.myers_results(0)
    b2.affine_results(cigar=False)
    e2e = time.perf_counter() - t0
    b2.close()
    out["end_to_end"] = {"gcups": round(cells / args.steps / e2e / 1e9, 2), "ms": round(e2e * 1e3, 1),
                         "includes": "host 2-bit packing, upload, launch plans, one step, results to host"}
    if args.per_pair:
        try:
            out["per_pair_launch_gcups"] = round(per_pair_launch_gcups(pairs, bands, args.per_pair), 3)
        except Exception as e:  # report, don't hide
            out["per_pair_launch_gcups"] = f"error: {e}"
    if not args.no_cpu:
        threads = host_threads()
        # The whole workload: about 5 s of host work at 16 threads with the AVX-512 rows.
        k = args.cpu_pairs or len(pairs)
        from spearfish_amd.align import PairSet
        sub = PairSet(pairs.pat, pairs.pat_off[:k], pairs.pat_len[:k], pairs.txt, pairs.txt_off[:k],
                      pairs.txt_len[:k])
        cb = cpu_baseline(sub, bands[:k], {"dist": dist, "score": res["score"]}, threads)
        out["cpu_baseline"] = {"value": round(cb["cells"] / cb["seconds"] / 1e9, 3), "unit": "GCUPS", "cores": threads,
                               "kind": "port",
                               "sample": f"{k} pairs (the whole workload), scores only: C restatement of "
                                         f"distance_with (scalar) + run_core with {cb['rows']} rows, "
                                         f"{cb['seconds']:.1f}s (Myers {cb['myers_s']:.1f}s, affine "
                                         f"{cb['affine_s']:.1f}s)",
                               "sample_bit_exact_vs_gpu": cb["bit_exact"]}
    batch.close()
    return out


def common(args, value, unit, ms, dtype, config, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": config}


def bench_uneven(args, steps=None, warmup=None):
    from spearfish_amd import align as A
    from spearfish_amd import workload as W
    steps = steps or args.steps
    warmup = args.warmup if warmup is None else warmup
    n = args.pairs if (args.pairs != 1_000_000 and args.workload == "uneven") else 500_000
    pairs, bands = W.generate(3, n, args.seed)
    log(f"generated {n} uneven pairs")
    b = A.Batch(pairs)
    for w in range(warmup):
        b.myers(A.Mode.Infix, locate=1)
        b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
        b.sync()
        log(f"uneven: warmup {w + 1} done")
    b.reset_timing()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.myers(A.Mode.Infix, locate=1)
        b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)
    b.sync()
    wall = time.perf_counter() - t0
    tm = b.timing()
    log(f"uneven: timed {steps} steps in {wall:.3f}s")
    cells = tm.myers_cells + tm.affine_cells
    out = common(args, round(cells / wall / 1e9, 2), "GCUPS", round(wall / steps * 1e3, 2), "int32",
                 {"workload": "uneven-5e5-lognormal800-5to15pct-adaptive-band", "pairs": n,
                  "step": "myers infix distance+end, affine global ONT adaptive band score"})
    out["steps"], out["warmup"] = steps, warmup
    out["myers_gcups"] = round(tm.myers_cells / (tm.myers_ms * 1e-3) / 1e9, 1)
    out["affine_gcups"] = round(tm.affine_cells / (tm.affine_ms * 1e-3) / 1e9, 1)
    out["kernel_ms_per_step"] = {"myers": round(tm.myers_ms / steps, 2), "affine": round(tm.affine_ms / steps, 2)}
    out["affine_launches_per_step"] = tm.affine_launches // steps
    if not args.no_schedules:
        out["schedules"] = uneven_schedules(b, bands, steps, args.tune_out)
    if not args.no_cpu:
        dist, _, _ = b.myers_results(0)
        res = b.affine_results()
        threads = host_threads()
        k = args.cpu_pairs or len(pairs)  # the whole workload: about 4 s at 16 threads with AVX-512 rows
        from spearfish_amd.align import PairSet
        sub = PairSet(pairs.pat, pairs.pat_off[:k], pairs.pat_len[:k], pairs.txt, pairs.txt_off[:k], pairs.txt_len[:k])
        cb = cpu_baseline(sub, bands[:k], {"dist": dist, "score": res["score"]}, threads)
        out["cpu_baseline"] = {"value": round(cb["cells"] / cb["seconds"] / 1e9, 3), "unit": "GCUPS", "cores": threads,
                               "kind": "port", "sample": f"{k} pairs (the whole workload), scores only: C restatement of "
                                                         f"distance_with (scalar) + run_core with {cb['rows']} rows, "
                                                         f"{cb['seconds']:.1f}s",
                               "sample_bit_exact_vs_gpu": cb["bit_exact"]}
        # The target: at least 20x the all-host-cores CPU path on this workload.
        out["gpu_vs_cpu"] = {"ratio": round(out["value"] / out["cpu_baseline"]["value"], 1), "target": 20.0}
    b.close()
    return out


def uneven_schedules(b, bands, steps, tune_out=None):
    """The same step under each schedule, ms a step (wall time over `steps` steps after a warm-up):
    launches of one workgroup a block or persistent launches over work queues, each queued eagerly
    or replayed from a captured graph.  Every schedule's results are checked against the first."""
    from spearfish_amd import _native as N
    from spearfish_amd import align as A
    lib = N.load()

    def step():
        b.myers(A.Mode.Infix, locate=1)
        b.affine(A.Mode.Global, A.Affine.ONT, bands=bands)

    def res():
        d, e, _ = b.myers_results(1)
        r = b.affine_results()
        return d, e, r["score"], r["score_col"]

    out, ref = {}, None
    try:
        for pers in (0, 1, 3):
            N.check(lib.sf_set_persistent(pers))
            name = {0: "per_block", 1: "persistent_myers", 3: "persistent"}[pers]
            step()
            b.sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                step()
            b.sync()
            out[f"{name}_eager_ms"] = round((time.perf_counter() - t0) / steps * 1e3, 2)
            r = res()
            ref = ref or r
            ok = all(np.array_equal(x, y) for x, y in zip(r, ref))
            b.capture_begin()
            step()
            b.capture_end()
            b.replay(1)
            b.sync()
            t0 = time.perf_counter()
            b.replay(steps)
            b.sync()
            out[f"{name}_graph_ms"] = round((time.perf_counter() - t0) / steps * 1e3, 2)
            ok = ok and all(np.array_equal(x, y) for x, y in zip(res(), ref))
            out[f"{name}_bit_exact"] = ok
            log(f"uneven schedules: {out}")
        # The autotuner: measure the candidate classes of this batch's pairs, then the same step
        # planned with the measurements.
        N.check(lib.sf_set_persistent(0))
        N.check(lib.sf_tune_reset())
        t0 = time.perf_counter()
        b.autotune(A.Mode.Global, A.Affine.ONT, bands=bands)
        out["autotune_s"] = round(time.perf_counter() - t0, 2)
        out["autotune_classes"] = int(lib.sf_tune_classes())
        step()
        b.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        b.sync()
        out["tuned_eager_ms"] = round((time.perf_counter() - t0) / steps * 1e3, 2)
        out["tuned_bit_exact"] = all(np.array_equal(x, y) for x, y in zip(res(), ref))
        b.capture_begin()
        step()
        b.capture_end()
        b.replay(1)
        b.sync()
        t0 = time.perf_counter()
        b.replay(steps)
        b.sync()
        out["tuned_graph_ms"] = round((time.perf_counter() - t0) / steps * 1e3, 2)
        out["tuned_bit_exact"] = out["tuned_bit_exact"] and all(np.array_equal(x, y) for x, y in zip(res(), ref))
        if tune_out:
            N.check(lib.sf_tune_save(tune_out.encode()))
        log(f"uneven schedules: {out}")
    finally:
        N.check(lib.sf_set_persistent(0))
        N.check(lib.sf_tune_reset())
    return out


def bench_long(args, steps=None, warmup=None):
    from spearfish_amd import align as A
    from spearfish_amd import workload as W
    steps = steps or args.steps
    warmup = args.warmup if warmup is None else warmup
    n = args.pairs if (args.pairs != 1_000_000 and args.workload == "long") else 200_000
    pairs, _ = W.generate(4, n, args.seed)
    log(f"generated {n} long pairs")
    b = A.Batch(pairs)

    def step():
        b.use_located(False)
        b.myers(A.Mode.Infix, locate=2)
        b.use_located(True)
        b.affine(A.Mode.Global, A.Affine.ONT, band=256, cigar=True)

    for w in range(warmup):
        step()
        b.sync()
        log(f"long: warmup {w + 1} done")
    b.reset_timing()
    t0 = time.perf_counter()
    for s in range(steps):
        step()
        log(f"long: step {s + 1} queued")
    b.sync()
    wall = time.perf_counter() - t0
    tm = b.timing()
    cells = tm.myers_cells + tm.affine_cells
    out = common(args, round(cells / wall / 1e9, 2), "GCUPS", round(wall / steps * 1e3, 2), "int32",
                 {"workload": "long-2e5x2to10kb-12pct-band256", "pairs": n,
                  "step": "myers infix distance+end+start, affine global ONT band 256 + CIGAR on located window"})
    out["steps"], out["warmup"] = steps, warmup
    out["kernel_ms_per_step"] = {"myers": round(tm.myers_ms / steps, 2), "myers_start": round(tm.locate_ms / steps, 2),
                                 "affine_tb": round(tm.affine_ms / steps, 2), "walk": round(tm.walk_ms / steps, 2)}
    out["myers_gcups"] = round(tm.myers_cells / ((tm.myers_ms) * 1e-3) / 1e9, 1)
    b.close()
    return out


def bench_serve(args):
    from spearfish_amd import serve as S
    rates = [5000, 20000, 50000]
    dur = args.serve_seconds
    n_all = int(max(rates) * dur)
    store, reads = S.generate_serve_workload(n_all, seed=args.seed)
    log(f"transcriptome {store.bases} bases, {n_all} reads, {int(reads.nseg.sum())} segments")
    res = S.Results(reads)
    reqs = S.requests(reads, res)
    rows = []
    for rate in rates:
        n = int(rate * dur)
        arr = S.poisson_arrivals(n, rate, seed=rate)
        with S.Server(store, slots=4, band=64) as s:
            s.replay(reqs, S.poisson_arrivals(min(n, 2000), rate, seed=1))  # warm-up
            s.stats(reset=True)
            wall = s.replay(reqs, arr)
            st = s.stats()
            refused = s.refused
        row = {"rate": rate, "mode": "continuous batching", "reads": n, "refused": refused,
               "reads_per_s": round(st.reads_done / wall, 1),
               "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
               "mean_reads_per_batch": round(st.mean_reads_per_batch, 1), "gcups": round(st.cells / wall / 1e9, 1)}
        log(str(row))
        rows.append(row)
        npr = min(n, 4000)
        with S.Server(store, slots=1, max_reads_per_batch=1, band=64) as s:
            s.replay(reqs, S.poisson_arrivals(200, rate, seed=1))
            s.stats(reset=True)
            wall = s.replay(reqs, arr[:npr])
            st = s.stats()
        row = {"rate": rate, "mode": "one read per launch", "reads": npr, "reads_per_s": round(st.reads_done / wall, 1),
               "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3), "mean_reads_per_batch": 1.0}
        log(str(row))
        rows.append(row)
    # Saturation: every read offered at once, so the server runs full batches back to back; its
    # rate is the capacity the Poisson rows stay under (their latencies are the service's).
    import numpy as np
    with S.Server(store, slots=4, band=64) as s:
        s.replay(reqs, S.poisson_arrivals(2000, 50000, seed=1))  # warm-up
        s.stats(reset=True)
        wall = s.replay(reqs, np.zeros(n_all))
        st = s.stats()
        refused = s.refused
    sat = {"rate": "all at once", "mode": "continuous batching", "reads": n_all, "refused": refused,
           "reads_per_s": round(st.reads_done / wall, 1), "p50_ms": round(st.p50_ms, 3),
           "p99_ms": round(st.p99_ms, 3), "mean_reads_per_batch": round(st.mean_reads_per_batch, 1),
           "gcups": round(st.cells / wall / 1e9, 1)}
    log(str(sat))
    rows.append(sat)
    # The top rate replayed again and again in one process (one server, a warm-up first): the
    # spread of p99 across replays, which a rare host pause would show as an outlier.
    replays = []
    if args.serve_replays > 0:
        rate = max(rates)
        n = int(rate * dur)
        with S.Server(store, slots=4, band=64) as s:
            s.replay(reqs, S.poisson_arrivals(2000, rate, seed=1))
            for k in range(args.serve_replays):
                s.stats(reset=True)
                wall = s.replay(reqs, S.poisson_arrivals(n, rate, seed=100 + k))
                st = s.stats()
                replays.append({"reads_per_s": round(st.reads_done / wall, 1), "p50_ms": round(st.p50_ms, 3),
                                "p99_ms": round(st.p99_ms, 3), "max_ms": round(st.max_ms, 3)})
                log(f"replay {k + 1}/{args.serve_replays}: {replays[-1]}")
    top = [r for r in rows if r["mode"] == "continuous batching" and r["rate"] == max(rates)][-1]
    out = common(args, top["reads_per_s"], "reads/s", None, "int32",
                 {"workload": "serve-20k-transcripts-poisson-5k-20k-50k", "band": 64,
                  "per_segment": "myers infix distance+end, affine global ONT band 64 score",
                  "server": "4 slots, 256 MiB HBM budget"})
    out["steps"], out["warmup"] = None, None
    out["serve"] = rows
    if replays:
        out["replays_at_top_rate"] = replays
    return out


def bench_tiebreak(args, steps=None, warmup=None):
    """The tie-break's call mix: every `seg` call (crates/map/src/lib.rs:1307-1492) of 20,000 reads
    of the serving workload's transcriptome -- left ends, interior stretches and right ends against
    each isoform's text, windows and cuts as `segmented` makes them -- in one sf_tiebreak_segments
    call a step over the device-resident packed store (ONT scoring, the reference's tie-break
    defaults: band 16, X-drop 50, end bonus 5, bonus within the cut)."""
    from spearfish_amd import serve as S
    from spearfish_amd import tiebreak as T
    steps = steps or args.steps
    warmup = args.warmup if warmup is None else warmup
    n_reads = 20000
    store, reads = S.generate_serve_workload(n_reads, seed=args.seed)
    tb = T.TieBreak()
    segs, coff, clen = T.calls_from_isoform_segments(store.tx_off, store.tx_len, reads, tb)
    words = store.words if len(store.words) % 2 == 0 else np.append(store.words, np.uint32(0))
    words64 = words.view(np.uint64)  # the index's u64 words: the same bytes as 16-base u32 words
    ps = T.PackedStore(words64, store.bases)
    log(f"tiebreak: {n_reads} reads, {len(segs)} calls, {len(coff)} candidates")
    # The calls' inputs made resident once (reads and candidates uploaded, distinct texts found,
    # passes planned); a step runs every call's passes and reads the results back.
    t0 = time.perf_counter()
    batch = T.TbBatch(ps, reads.codes, segs, coff, clen)
    setup_s = time.perf_counter() - t0
    for w in range(warmup):
        batch.run(tb)
        batch.results()
    batch.timing(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(tb)
        out, rep = batch.results()
    wall = time.perf_counter() - t0
    tm = batch.timing()
    batch.close()
    distinct = int((rep == np.arange(len(rep)) - np.repeat(segs["cand0"], segs["ncand"])).sum())
    log(f"tiebreak: timed {steps} steps in {wall:.3f}s")
    rate = len(coff) * steps / wall
    # The one-shot call (host buffers in and out: uploads, dedup and planning every time).
    T.segments_arrays(ps, reads.codes, segs, coff, clen, tb)
    t0 = time.perf_counter()
    for _ in range(steps):
        one, _ = T.segments_arrays(ps, reads.codes, segs, coff, clen, tb)
    wall1 = time.perf_counter() - t0
    assert np.array_equal(one, out)
    res = common(args, round(rate, 1), "candidate alignments/s", round(wall / steps * 1e3, 2), "int32",
                 {"workload": "tiebreak-20k-reads-serve-transcriptome", "reads": n_reads, "calls": len(segs),
                  "candidates": len(coff), "distinct_texts": distinct,
                  "per_call": "seg: dedup by packed equality, left ends reversed, Prefix ends "
                              "affine_extend_sorted_cut_bonus, interior affine Global; ONT, band 16, xdrop 50, "
                              "end bonus 5"})
    res["steps"], res["warmup"] = steps, warmup
    # The dominant phase: the extension pass (affine_pk16_pass_kernel launches side by side, HIP events
    # around the phase), bound by VALU issue.
    ext_s = tm.ext_ms / steps * 1e-3
    achieved = TB_EXT_VALU_INSTS * 64 / ext_s / 1e12
    res["roofline"] = {"bound": "valu", "kernel": "affine_pk16_pass_kernel<8, 5..8, true> (the extension pass, 16-bit packed, 8 lanes a pair)",
                       "achieved": round(achieved, 3), "peak": round(VALU_ISSUE_PEAK_T, 1),
                       "unit": "T VALU lane-instructions/s issued (SQ_INSTS_VALU of a pass / pass time)",
                       "frac": round(achieved / VALU_ISSUE_PEAK_T, 4), "traffic": None,
                       "avg_pass_ms": round(tm.ext_ms / steps, 3), "sorted_ms": round(tm.sorted_ms / steps, 3),
                       "ext_cells_full_band": int(tm.ext_cells // steps),
                       "ext_gcups_full_band": round(tm.ext_cells / steps / ext_s / 1e9, 1)}
    res["calls_per_s"] = round(len(segs) * steps / wall, 1)
    res["reads_per_s"] = round(n_reads * steps / wall, 1)
    res["setup_ms"] = round(setup_s * 1e3, 2)
    res["one_shot"] = {"ms_per_call": round(wall1 / steps * 1e3, 2),
                       "candidate_alignments_per_s": round(len(coff) * steps / wall1, 1),
                       "note": "sf_tiebreak_segments from host buffers: reads and candidates uploaded, "
                               "distinct texts found and passes planned in every call"}
    # The same calls one `seg` call per library call, as the reference makes them.
    k1 = min(len(segs), 400)
    t0 = time.perf_counter()
    for i in range(k1):
        a, n = int(segs["cand0"][i]), int(segs["ncand"][i])
        one = segs[i:i + 1].copy()
        one["cand0"] = 0
        T.segments_arrays(ps, reads.codes, one, coff[a:a + n], clen[a:a + n], tb)
    w1 = time.perf_counter() - t0
    res["one_call_per_seg"] = {"calls": k1, "calls_per_s": round(k1 / w1, 1),
                               "candidate_alignments_per_s": round(int(segs["ncand"][:k1].sum()) / w1, 1)}
    if not args.no_cpu:
        from oracle import sf_oracle as O
        threads = host_threads()
        simd = O.avx512_available()
        kr = 4000  # reads in the sample: about 10 s of host work
        ks = int(np.searchsorted(segs["read_off"], reads.read_off[kr])) if kr < n_reads else len(segs)
        kc = int(segs["cand0"][ks]) if ks < len(segs) else len(coff)
        sc = (tb.affine.mat, tb.affine.mis, tb.affine.open, tb.affine.ext)
        t0 = time.perf_counter()
        cpu, crep = O.tiebreak_batch(words64, store.bases, reads.codes, segs[:ks], coff[:kc], clen[:kc], sc, tb.band,
                                     tb.end_bonus, tb.xdrop, threads=threads, avx512=simd)
        cs = time.perf_counter() - t0
        exact = bool(np.array_equal(cpu[:, 0], out["whole"][:kc]) and np.array_equal(cpu[:, 1], out["cut"][:kc])
                     and np.array_equal(cpu[:, 2], out["cols"][:kc]) and np.array_equal(crep, rep[:kc]))
        res["cpu_baseline"] = {"value": round(kc / cs, 1), "unit": "candidate alignments/s", "cores": threads,
                               "kind": "port",
                               "sample": f"the calls of the first {kr} reads ({ks} calls, {kc} candidates), {cs:.1f}s: "
                                         f"C restatement of seg with "
                                         + ("the x86 rows (32-bit AVX-512 rows for the X-drop passes, 16-bit "
                                            "where narrow holds for the others)" if simd else "scalar rows"),
                               "sample_bit_exact_vs_gpu": exact}
        res["gpu_vs_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (--gpus {args.gpus} requested)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "probe":
        # The issue-rate probe alone (the denominators of the VALU rooflines).
        print(json.dumps({"issue_probe_t_lane_instructions_per_s": issue_probe(kinds={**PROBE_KINDS, **PROBE_PAIRING_KINDS}),
                          "nominal_one_wave_instruction_per_4_cycles": round(VALU_ISSUE_PEAK_T, 1)}))
        return
    if args.workload != "uniform":
        print(json.dumps({"uneven": bench_uneven, "long": bench_long, "serve": bench_serve,
                          "tiebreak": bench_tiebreak}[args.workload](args)))
        return
    out = bench_uniform(args)
    if not args.no_extra:
        # The other workloads at bounded step counts (about a minute together), each in its own
        # record; `value` stays the uniform workload's.
        extra = {}
        for name, fn in (("uneven", lambda: bench_uneven(args, steps=5, warmup=1)),
                         ("long", lambda: bench_long(args, steps=2, warmup=1)),
                         ("serve", lambda: bench_serve(args)),
                         ("tiebreak", lambda: bench_tiebreak(args, steps=5, warmup=1))):
            t0 = time.perf_counter()
            try:
                r = fn()
                for key in ("metric", "n_gpus", "scaling", "vs_baseline", "data"):
                    r.pop(key, None)
                extra[name] = r
            except Exception as e:  # report, don't hide
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"{name}: done in {time.perf_counter() - t0:.1f}s")
        out["other_workloads"] = extra
    print(json.dumps(out))


if __name__ == "__main__":
    main()
