// affine_wide.hip -- the banded affine pass for bands wider than the register-resident kernels hold
// (more than 2,048 cells a row: wide bands, and Infix passes over long texts, which are unbanded).
//
// Computes crates/map/src/align.rs:2446-2666 `run_core` (mode, band laid out by geom, clip with
// X-drop and best cell, cut) or align.rs:783-852 `local_band_core` (diagonals dlo..=dhi, end
// bonuses, floor 0 except in the last row, read codes >= 4 matching nothing), the same values cell
// for cell as the oracle's row passes.  One pair per workgroup of 256 threads.  A row is kept in
// diagonal slots s = j - i + below (as the register-resident kernels keep it), two rows at a time
// (the previous and the current, H and F) in HBM scratch, and the workgroup sweeps a row in chunks
// of 2,048 slots, 8 consecutive slots a thread:
//   F = max(H_up - go, F_up - ext, NEG)          up: slot s + 1 of the previous row
//   x = max(H_diag + score, F, floor)            diag: slot s of the previous row
//   E = max over earlier slots s' of the row of (x(s') - go - ext (s - 1 - s'))
//   H = max(x, E)
// with E as an exclusive prefix maximum of x(s') + ext s' (64-bit: the bias grows with the width)
// across the thread's slots, the workgroup (wave scans, then the four waves through LDS) and the
// chunks (a carry).  Column 0 takes only F (and the floor), and seeds the row's horizontal gaps
// (align.rs:2564-2572, 800-806).  Slots outside a row's band hold NEG, and slot W (one past the
// widest row) is NEG in both rows: the reference's right sentinel.  The row's maximum with its
// first column, and its maximum up to the cut, are reduced over the workgroup after each row for
// the clip, the X-drop stop, the cut, and the last row's score.
#include <stdint.h>

#include "internal.hpp"

using namespace sf;

namespace {

constexpr int kThreads = 256, kSlots = 8, kChunk = kThreads * kSlots;
constexpr int64_t kNone64 = INT64_MIN / 2;

struct Best { // a row maximum and its first column (ties: the smaller column)
    int32_t v;
    int64_t j;
};

__device__ __forceinline__ Best better(Best a, Best b) {
    return (b.v > a.v || (b.v == a.v && b.j < a.j)) ? b : a;
}

// Workgroup reductions through LDS (wave shuffles, then the four waves).
__device__ Best reduce_best(Best x, Best *lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        Best y{__shfl_xor(x.v, o), __shfl_xor(x.j, o)};
        x = better(x, y);
    }
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    Best r = lds[0];
#pragma unroll
    for (int q = 1; q < kThreads / 64; ++q) r = better(r, lds[q]);
    __syncthreads();
    return r;
}

__device__ int32_t reduce_max(int32_t x, int32_t *lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    int32_t r = lds[0];
#pragma unroll
    for (int q = 1; q < kThreads / 64; ++q) r = max(r, lds[q]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ int32_t sat_sub32(int32_t a, int32_t b) {
    const int64_t r = (int64_t)a - b;
    return r < INT32_MIN ? INT32_MIN : (int32_t)r;
}
__device__ __forceinline__ int32_t sat_add32(int32_t a, int32_t b) {
    const int64_t r = (int64_t)a + b;
    return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}

__global__ __launch_bounds__(kThreads) void affine_wide_kernel(WideArgs a) {
    const uint32_t w = blockIdx.x;
    if (w >= a.p.count) return; // (whole workgroups)
    const uint32_t pi = a.p.order[w];
    const PairDesc dsc = a.p.desc[w];
    const int64_t m = dsc.m, n = dsc.n;
    const int64_t below = (int64_t)(int32_t)dsc.below, above = (int64_t)(int32_t)dsc.above;
    const int64_t W = below + above + 1; // slots of the widest row
    const bool local = a.local != 0, infix = a.p.mode == kInfix;
    const int32_t mat = a.p.mat, mis = a.p.mis, ext = a.p.ext, go = a.p.open + a.p.ext;
    const SeqRef P = seqref(a.p.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.p.pv.seq, dsc.toff, (uint32_t)n);
    const int tid = threadIdx.x;
    __shared__ int64_t scan_lds[2][kThreads / 64];
    __shared__ Best best_lds[kThreads / 64];
    __shared__ int32_t max_lds[kThreads / 64];

    if (local && (m == 0 || W <= 0)) { // align.rs:792-794: nothing to align scores 0
        if (tid == 0) a.p.best[pi] = 0;
        return;
    }
    int32_t *Hp = a.rows + (uint64_t)w * 4 * a.stride, *Fp = Hp + a.stride, *Hc = Fp + a.stride, *Fc = Hc + a.stride;
    // Row 0 (align.rs:2527-2537; local: the start bonus over the diagonals, align.rs:796-799).
    const int64_t hi0 = above < n ? above : n, lo0 = -below > 0 ? -below : 0;
    for (int64_t s = tid; s <= W; s += kThreads) {
        const int64_t j = s - below;
        int32_t h = kNeg;
        if (s < W && j >= lo0 && j <= hi0)
            h = local ? a.p.bonus0 : ((j == 0 || infix) ? 0 : -(a.p.open + ext * (int32_t)j));
        Hp[s] = h;
        Fp[s] = kNeg;
        if (s == W) Hc[s] = Fc[s] = kNeg; // the right sentinel of both rows
    }
    __syncthreads();
    int32_t clip_best = 0, cut_best = 0, lbest = 0;
    uint32_t clip_i = 0, clip_j = 0;
    const uint64_t cut = a.p.cut ? a.p.cut[pi] : UINT64_MAX;
    Best last{kNeg, 0};     // the last row's maximum and first column
    int32_t last_cut = kNeg; // the last row's maximum up to the cut
    int64_t llo = lo0, lhi = hi0;
    for (int64_t i = 1; i <= m; ++i) {
        const int64_t lo = i - below > 0 ? i - below : 0, hi = i + above < n ? i + above : n;
        if (!local && lo > hi) { // align.rs:2543-2556: the band has passed the text's end
            if (tid == 0) {
                if (a.p.out) a.p.out[pi] = DpOut{kNeg, 0, clip_best, clip_i, clip_j, (uint32_t)n, (uint32_t)i, cut_best, kNeg};
                else a.score[pi] = kNeg, a.score_col[pi] = 0;
            }
            return;
        }
        const uint32_t pc = code_at(P, (int32_t)(i - 1));
        bool none = false; // local: a read code >= 4 matches nothing (align.rs:731-734)
        if (local && a.p.pinv) {
            const uint64_t bit = dsc.poff + (uint64_t)(i - 1);
            none = (a.p.pinv[bit >> 5] >> (bit & 31)) & 1u;
        }
        const int32_t floor = local ? (i < m ? 0 : kNeg) : kNeg;
        int64_t carry = kNone64;
        Best tb{kNeg, INT64_MAX};
        int32_t tcut = kNeg;
        for (int64_t c0 = 0; c0 < W; c0 += kChunk) {
            const int64_t s0 = c0 + (int64_t)tid * kSlots;
            int32_t X[kSlots], Fv[kSlots];
            int64_t inc[kSlots];
            int64_t run = kNone64;
#pragma unroll
            for (int q = 0; q < kSlots; ++q) {
                const int64_t s = s0 + q, j = i - below + s;
                const bool valid = s < W && j >= lo && j <= hi;
                int32_t x = kNeg, f = kNeg;
                if (valid) {
                    f = (int32_t)max(max((int64_t)Hp[s + 1] - go, (int64_t)Fp[s + 1] - ext), (int64_t)kNeg);
                    if (j == 0) {
                        x = max(f, floor);
                    } else {
                        const bool match = !none && code_at(T, (int32_t)(j - 1)) == pc;
                        const int32_t d = Hp[s] + (match ? mat : -mis);
                        x = max(max(d, f), floor);
                    }
                    run = max(run, (int64_t)x + (int64_t)ext * s);
                }
                X[q] = x;
                Fv[q] = f;
                inc[q] = run;
            }
            // Exclusive scan of the threads' maxima over the workgroup, then the chunks' carry.
            int64_t v = run;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int64_t o = __shfl_up(v, d, 64);
                if ((tid & 63) >= d) v = max(v, o);
            }
            const int par = (int)((c0 / kChunk) & 1);
            if ((tid & 63) == 63) scan_lds[par][tid >> 6] = v;
            __syncthreads();
            int64_t excl = carry;
            for (int q = 0; q < (tid >> 6); ++q) excl = max(excl, scan_lds[par][q]);
            const int64_t up = __shfl_up(v, 1, 64);
            if ((tid & 63) > 0) excl = max(excl, up);
            for (int q = 0; q < kThreads / 64; ++q) carry = max(carry, scan_lds[par][q]);
#pragma unroll
            for (int q = 0; q < kSlots; ++q) {
                const int64_t s = s0 + q, j = i - below + s;
                const bool valid = s < W && j >= lo && j <= hi;
                const int64_t eb = q ? max(excl, inc[q - 1]) : excl;
                int32_t h = kNeg;
                if (valid) {
                    h = X[q];
                    if (j > 0 && eb > kNone64) h = (int32_t)max((int64_t)h, eb - go + ext - (int64_t)ext * s);
                    tb = better(tb, Best{h, j});
                    if ((uint64_t)j <= cut) tcut = max(tcut, h);
                }
                if (s < W) {
                    Hc[s] = h;
                    Fc[s] = valid ? Fv[q] : kNeg;
                }
            }
        }
        __syncthreads(); // the row is written before it is read as the previous one
        const Best rb = reduce_best(tb, best_lds);
        const int32_t rc = reduce_max(tcut, max_lds);
        const int32_t row_max = lo <= hi ? rb.v : kNeg;
        if (local) {
            lbest = max(lbest, row_max);
            if (i == m) lbest = max(lbest, sat_add32(row_max, a.p.bonus1));
        } else {
            if (cut != UINT64_MAX && (uint64_t)lo <= cut) cut_best = max(cut_best, rc);
            if (a.p.has_clip) {
                if (row_max > clip_best) {
                    clip_best = row_max;
                    if (a.p.clip_cell) clip_i = (uint32_t)i, clip_j = (uint32_t)rb.j;
                }
                if (row_max < sat_sub32(clip_best, a.p.xdrop)) { // align.rs:2613-2627
                    if (tid == 0) {
                        if (a.p.out)
                            a.p.out[pi] = DpOut{kNeg, 0, clip_best, clip_i, clip_j, (uint32_t)hi, (uint32_t)i, cut_best, kNeg};
                        else a.score[pi] = kNeg, a.score_col[pi] = 0;
                    }
                    return;
                }
            }
        }
        last = rb;
        last_cut = (cut != UINT64_MAX && (uint64_t)lo <= cut) ? rc : kNeg;
        llo = lo, lhi = hi;
        int32_t *t = Hp;
        Hp = Hc, Hc = t;
        t = Fp, Fp = Fc, Fc = t;
    }
    if (local) {
        if (tid == 0) a.p.best[pi] = lbest;
        return;
    }
    if (m == 0) { // the last row is row 0 (only its values were written)
        Best b0{kNeg, INT64_MAX};
        int32_t c0 = kNeg;
        for (int64_t j = llo + tid; j <= lhi; j += kThreads) {
            const int32_t h = Hp[j + below];
            b0 = better(b0, Best{h, j});
            if ((uint64_t)j <= cut) c0 = max(c0, h);
        }
        last = reduce_best(b0, best_lds);
        const int32_t rc = reduce_max(c0, max_lds);
        last_cut = (cut != UINT64_MAX && (uint64_t)llo <= cut) ? rc : kNeg;
    }
    if (tid != 0) return;
    int32_t score;
    uint32_t col;
    if (a.p.mode == kGlobal) { // align.rs:2643: H[m][n]
        const int64_t sN = n - m + below;
        score = (sN >= 0 && sN < W) ? Hp[sN] : kNeg;
        col = (uint32_t)n;
    } else { // align.rs:2644-2648: the last row's maximum, first column
        score = last.v;
        col = (uint32_t)last.j;
    }
    if (a.p.out) a.p.out[pi] = DpOut{score, col, clip_best, clip_i, clip_j, (uint32_t)lhi, (uint32_t)m, cut_best, last_cut};
    else a.score[pi] = score, a.score_col[pi] = col;
}

} // namespace

hipError_t sf_launch_affine_wide(const WideArgs &a, hipStream_t s) {
    if (a.p.count == 0) return hipSuccess;
    hipLaunchKernelGGL(affine_wide_kernel, dim3(a.p.count), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}
