// affine.hip -- launches of the 32-bit banded affine DP (affine_wave.hpp), the passes with per-row
// bookkeeping, the traceback walk and the CIGAR compaction.
#include "affine_wave.hpp"

#include <atomic>
#include <cstdlib>

namespace {

// ---- passes with per-row bookkeeping ----
//
// `affine_dp` with a clip (X-drop stop, best cell) or a cut, and `local_band_ends`, need each
// row's maximum over its band cells (and for the best cell its first column), so these passes add
// a group reduction per row to the 32-bit recurrence of affine_wave_kernel (same layout, bias and
// out-of-band handling).  The reductions are DPP row rotations, so every lane of a group ends with
// the group's value.

constexpr int kRowRor1 = 0x121, kRowRor2 = 0x122, kRowRor4 = 0x124, kRowRor8 = 0x128;

// (A rotation has no invalid source lane, so bound_ctrl changes nothing; set, it lets the max take
// the DPP operand directly.)
template <int LP>
__device__ __forceinline__ int32_t group_max(int32_t x) {
    x = imax(x, __builtin_amdgcn_update_dpp(x, x, kRowRor8, 0xf, 0xf, true));
    x = imax(x, __builtin_amdgcn_update_dpp(x, x, kRowRor4, 0xf, 0xf, true));
    x = imax(x, __builtin_amdgcn_update_dpp(x, x, kRowRor2, 0xf, 0xf, true));
    x = imax(x, __builtin_amdgcn_update_dpp(x, x, kRowRor1, 0xf, 0xf, true));
    if constexpr (LP == 64) {
        x = imax(x, __shfl_xor(x, 16));
        x = imax(x, __shfl_xor(x, 32));
    }
    return x;
}

template <int LP>
__device__ __forceinline__ int32_t group_min(int32_t x) {
    x = imin(x, __builtin_amdgcn_update_dpp(x, x, kRowRor8, 0xf, 0xf, true));
    x = imin(x, __builtin_amdgcn_update_dpp(x, x, kRowRor4, 0xf, 0xf, true));
    x = imin(x, __builtin_amdgcn_update_dpp(x, x, kRowRor2, 0xf, 0xf, true));
    x = imin(x, __builtin_amdgcn_update_dpp(x, x, kRowRor1, 0xf, 0xf, true));
    if constexpr (LP == 64) {
        x = imin(x, __shfl_xor(x, 16));
        x = imin(x, __shfl_xor(x, 32));
    }
    return x;
}

__device__ __forceinline__ int32_t sat_add32(int32_t a, int32_t b) {
    const int64_t r = (int64_t)a + b;
    return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}

template <int LP, int K, bool LOCAL>
__global__ __launch_bounds__(256) void affine_pass_kernel(PassArgs a) {
    constexpr bool ONE = LP == 64;
    constexpr int SHL = ONE ? kWaveShl1 : kRowShl1;
    constexpr int BLK = 16;
    constexpr int NWIN = (BLK + K - 1 + 15) / 16;
    constexpr int NY = (K + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int l = lane % LP;
    const uint32_t w = (blockIdx.x * 256u + threadIdx.x) / LP;
    const bool valid = w < a.count;
    if (ONE && !valid) return; // whole waves only
    const uint32_t pi = valid ? a.order[w] : 0u;
    const PairDesc dsc = valid ? a.desc[w] : PairDesc{0, 0, 0, 0, 0, 0};
    const int32_t m = (int32_t)dsc.m, n = (int32_t)dsc.n;
    const int32_t below = (int32_t)dsc.below, above = (int32_t)dsc.above; // local: -dlo, dhi
    const int32_t Wd = below + above + 1;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, (uint32_t)n);
    const int32_t mat = a.mat, mis = a.mis, ext = a.ext;
    const uint32_t dlt = (uint32_t)-a.open;
    const uint32_t scm = (uint32_t)(mat + 2 * ext);
    const int32_t ndl = -(mat + mis);
    const bool infix = a.mode == kInfix;
    const int32_t s0 = l * K;

    uint32_t V[K], Fb[K], Ub[K], dltv[K];
    // A slot's value is V - (row bias) - ext * s; slots past the band's width (cells outside it,
    // computed like band cells) also lose 2^30, which puts them below every band cell (scores stay
    // far from NEG), so a row's maximum needs no mask there.
    int32_t cpen[K];
    {
        const int32_t lo0 = below < 0 ? -below : 0; // row 0 covers [max(0, dlo), min(n, dhi)] (local)
        const int32_t hi0 = above < n ? above : n;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int32_t s = s0 + q;
            const int32_t j = s - below;
            int32_t h = kNeg;
            if (j >= lo0 && j <= hi0) h = LOCAL ? a.bonus0 : ((j == 0 || infix) ? 0 : -(a.open + ext * j));
            V[q] = (uint32_t)(h == kNeg ? kNeg : h + ext * s) + kOff;
            Fb[q] = (uint32_t)kNeg + kOff;
            Ub[q] = V[q] + dlt;
            dltv[q] = s >= Wd ? dlt - (1u << 30) : dlt;
            cpen[q] = ext * s + (s >= Wd ? (1 << 30) : 0);
        }
    }
    // Rows past n + below have no band cell (the band ran off the text).
    const int32_t last = valid ? (m < n + below ? m : n + below) : 0;
    int32_t rows = last;
    if constexpr (!ONE) {
#pragma unroll
        for (int o = 32; o >= LP; o >>= 1) rows = imax(rows, __shfl_xor(rows, o));
        rows = __builtin_amdgcn_readfirstlane(rows);
    }
    const int32_t xdrop = a.xdrop;
    // cut = usize::MAX is "no cut"; any other cut past the text covers the whole row.
    const uint64_t cut_in = (!LOCAL && valid && a.cut) ? a.cut[pi] : UINT64_MAX;
    const int64_t cut = cut_in == UINT64_MAX ? -1 : (int64_t)(cut_in < (uint64_t)INT32_MAX ? cut_in : INT32_MAX);
    int32_t clip_best = 0, cut_best = 0, best = 0;
    int32_t ci = 0, cj = 0;
    bool xstop = false;
    int32_t stop_row = 0, stop_read = 0;
    for (int32_t i0 = 1; i0 <= rows; i0 += BLK) {
        uint32_t pw = codes16(P, i0 - 1);
        uint32_t inv = 0; // local: rows whose read code matches nothing
        if (LOCAL && a.pinv) {
            const uint64_t b0 = dsc.poff + (uint64_t)(i0 - 1);
            const uint32_t lo32 = a.pinv[b0 >> 5], hi32 = a.pinv[(b0 >> 5) + 1];
            inv = __builtin_amdgcn_alignbit(hi32, lo32, (uint32_t)(b0 & 31));
        }
        uint32_t win[NWIN];
        const int32_t pos = i0 - below + s0 - 1;
#pragma unroll
        for (int k = 0; k < NWIN; ++k) win[k] = codes16(T, pos + 16 * k);
        const int32_t cnt = rows - i0 + 1 < BLK ? rows - i0 + 1 : BLK;
        for (int32_t r = 0; r < cnt; ++r) {
            const int32_t i = i0 + r;
            const uint32_t pcw = pw;
            pw >>= 2;
            const bool inv_row = (inv >> r) & 1u;
            if (i > last || xstop) continue; // this pair's rows are done
            const uint32_t prep = rep16(pcw, 0);
            uint32_t y[NY];
#pragma unroll
            for (int k = 0; k < NY; ++k) y[k] = inv_row ? 0xFFFFFFFFu : mismatch_bits(win[k], prep);
            // Band cells of row i: slots [slo, shi] (columns 0..n).
            const int32_t slo = below - i > 0 ? below - i : 0;
            const int32_t shi = n - i + below < Wd - 1 ? n - i + below : Wd - 1;
            const uint32_t rowb = (uint32_t)(2 * ext * i) + kOff; // bias of value 0 at slot 0 of row i
            const uint32_t Un = dpp0<SHL>(Ub[0]);
            const uint32_t Fn = dpp0<SHL>(Fb[0]);
            uint32_t X[K], G[K], FQ[K];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const uint32_t uu = q < K - 1 ? Ub[q + 1] : Un;
                const uint32_t fu = q < K - 1 ? Fb[q + 1] : Fn;
                const uint32_t mism = __builtin_amdgcn_ubfe(y[q / 16], 2 * (q % 16) + 1, 1);
                const uint32_t f = umax(uu, fu);
                const uint32_t d = V[q] + (uint32_t)__mul24((int)mism, ndl) + scm;
                uint32_t x = umax(d, f);
                if constexpr (LOCAL) {
                    // Local: a band cell never drops below 0 (align.rs:800-806), except in the last row.
                    const int32_t s = s0 + q;
                    if (i < m && s >= slo && s <= shi) x = umax(x, rowb + (uint32_t)(ext * s));
                }
                X[q] = x;
                FQ[q] = f;
                const uint32_t g = x + dlt;
                G[q] = q ? umax(G[q - 1], g) : g;
            }
            const uint32_t C = group_excl_umax<LP>(G[K - 1]);
#pragma unroll
            for (int q = 0; q < K; ++q) {
                V[q] = umax(X[q], q ? umax(C, G[q - 1]) : C);
                Fb[q] = FQ[q];
                Ub[q] = V[q] + dltv[q];
            }
            // The row's maximum over its band cells (align.rs:2599-2628 `row_max`): masked only in
            // rows whose band starts left of column 0 or ends past column n.
            int32_t hq[K];
            int32_t lm = INT32_MIN;
#pragma unroll
            for (int q = 0; q < K; ++q) hq[q] = (int32_t)(V[q] - rowb) - cpen[q];
            if (slo > 0 || shi < Wd - 1) {
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const int32_t s = s0 + q;
                    if (s >= slo && s <= shi) lm = imax(lm, hq[q]);
                }
            } else {
#pragma unroll
                for (int q = 0; q < K; ++q) lm = imax(lm, hq[q]);
            }
            const int32_t rmax = group_max<LP>(lm);
            if constexpr (LOCAL) {
                best = imax(best, rmax);
                if (i == m) best = imax(best, sat_add32(rmax, a.bonus1));
            } else {
                const int32_t lo = i - below > 0 ? i - below : 0;
                const int32_t hi = i + above < n ? i + above : n;
                if (cut >= 0 && lo <= cut) {
                    int32_t rc = rmax;
                    if (hi > cut) {
                        const int32_t sc_hi = (int32_t)cut - i + below;
                        int32_t lc = INT32_MIN;
#pragma unroll
                        for (int q = 0; q < K; ++q) {
                            const int32_t s = s0 + q;
                            if (s >= slo && s <= sc_hi) lc = imax(lc, hq[q]);
                        }
                        rc = group_max<LP>(lc);
                    }
                    cut_best = imax(cut_best, rc);
                }
                if (a.has_clip) {
                    if (rmax > clip_best) {
                        clip_best = rmax;
                        if (a.clip_cell) {
                            int32_t ls = INT32_MAX;
#pragma unroll
                            for (int q = 0; q < K; ++q) {
                                const int32_t s = s0 + q;
                                if (s >= slo && s <= shi && hq[q] == rmax) ls = imin(ls, s);
                            }
                            const int32_t sm = group_min<LP>(ls);
                            ci = i;
                            cj = sm + i - below;
                        }
                    }
                    // X-drop: the whole row is more than xdrop below the best so far.
                    const int64_t thr = (int64_t)clip_best - xdrop;
                    const int64_t sat = thr < INT32_MIN ? INT32_MIN : (thr > INT32_MAX ? INT32_MAX : thr);
                    if ((int64_t)rmax < sat) {
                        xstop = true;
                        stop_row = i;
                        stop_read = hi;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NWIN; ++k) win[k] = k + 1 < NWIN ? __builtin_amdgcn_alignbit(win[k + 1], win[k], 2) : win[k] >> 2;
        }
    }
    if constexpr (LOCAL) {
        if (valid && l == 0) a.best[pi] = (m == 0 || Wd <= 0) ? 0 : best;
        return;
    } else {
        // End of pass (align.rs:2643-2665): score and column, the cut score, bases read.
        const int32_t rowb = 2 * ext * m;
        const int32_t plo = m - below > 0 ? m - below : 0;
        const int32_t phi = m + above < n ? m + above : n;
        const int32_t slo = plo - (m - below), shi = phi - (m - below);
        int32_t score;
        uint32_t col;
        if (a.mode == kGlobal) {
            // H[m][n]; NEG when (m, n) lies outside the last row's band (a band laid out for
            // another text length).
            const int32_t sN = n - m + below;
            const bool inb = sN >= 0 && sN <= Wd - 1;
            const int owner = inb ? sN / K : 0, qq = inb ? sN % K : 0;
            uint32_t v = V[0];
#pragma unroll
            for (int q = 1; q < K; ++q) v = q == qq ? V[q] : v;
            const uint32_t vo = ONE ? (uint32_t)__builtin_amdgcn_readlane((int)v, owner)
                                    : (uint32_t)__shfl((int)v, (lane - l) + owner);
            score = inb ? (int32_t)(vo - kOff) - ext * sN - rowb : kNeg;
            col = (uint32_t)n;
        } else {
            int32_t bv = INT32_MIN, bs = INT32_MAX;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const int32_t s = s0 + q;
                const int32_t h = (int32_t)(V[q] - kOff) - ext * s - rowb;
                if (s >= slo && s <= shi && h > bv) {
                    bv = h;
                    bs = s;
                }
            }
            const int32_t mx = group_max<LP>(bv);
            const int32_t sm = group_min<LP>(bv == mx ? bs : INT32_MAX);
            score = mx;
            col = (uint32_t)(sm + (m - below));
        }
        int32_t cut_score = kNeg;
        if (cut >= 0 && plo <= cut) {
            const int32_t top = phi < cut ? phi : (int32_t)cut;
            const int32_t stop_s = top - (m - below);
            int32_t lc = INT32_MIN;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const int32_t s = s0 + q;
                const int32_t h = (int32_t)(V[q] - kOff) - ext * s - rowb;
                if (s >= slo && s <= stop_s) lc = imax(lc, h);
            }
            cut_score = imax(kNeg, group_max<LP>(lc));
        }
        if (valid && l == 0) {
            DpOut o;
            o.clip_best = clip_best;
            o.clip_i = (uint32_t)ci;
            o.clip_j = (uint32_t)cj;
            o.cut_best = cut_best;
            if (xstop) {
                o.score = kNeg, o.score_col = 0, o.text_read = (uint32_t)stop_read, o.last_row = (uint32_t)stop_row;
                o.cut_score = kNeg;
            } else if (last < m) {
                o.score = kNeg, o.score_col = 0, o.text_read = (uint32_t)n, o.last_row = (uint32_t)(last + 1);
                o.cut_score = kNeg;
            } else {
                o.score = score, o.score_col = col, o.text_read = (uint32_t)phi, o.last_row = (uint32_t)m;
                o.cut_score = cut_score;
            }
            a.out[pi] = o;
        }
    }
}

// One lane per pair: walks the traceback bits from the end cell back to the start, emitting
// BAM ops last to first (rules in the header comment of affine_wave_kernel and the oracle).
// Each step reads one u32 -- the word of the cell's lane for that row, which holds 8 slots of
// the lane -- so the walk keeps the 16-row block of its current lane-word in LDS (one 64-byte
// line), the pattern word of that block and the two text words around the cell's column in
// registers.  A lane steps on while its next cell is inside that data (diagonal, vertical and
// horizontal moves stay inside it until the path crosses to another lane-word or block); once
// every lane is done or needs new data, the lanes that need it load it together, so the wave
// waits on memory once per refill round rather than once per step of any lane.
constexpr int kWalkRows = 16;

__global__ __launch_bounds__(256) void walk_kernel(WalkArgs a) {
    // The walk's waves share SIMDs with the next chunk's DP waves, which issue every cycle they
    // can; raised priority lets the walk's one dependent chain issue first.
    __builtin_amdgcn_s_setprio(3);
    __shared__ uint32_t cache[256 * (kWalkRows + 1)];
    uint32_t *my = cache + threadIdx.x * (kWalkRows + 1);
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= a.count) return;
    const uint32_t pi = a.order[t];
    uint32_t *out = a.runs + a.runs_off[pi];
    if (a.score[pi] == kNeg) {
        a.nruns[pi] = 0;
        a.tstart[pi] = 0;
        return;
    }
    const PairDesc dsc = a.desc[t];
    const int32_t m = (int32_t)dsc.m, n = (int32_t)dsc.n, below = (int32_t)dsc.below;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, (uint32_t)n);
    const int32_t K = a.K, LP = a.LP;
    const int32_t nwl = (4 * K + 31) / 32, slots = LP * K;
    const uint32_t *tb = a.tb + a.tb_off[t];
    int32_t i = m, j = a.mode == kGlobal ? n : (int32_t)a.score_col[pi];
    // The cell's slot s = j - i + below as lane L and slot-in-lane q: a diagonal move keeps it, a
    // vertical one adds 1, a horizontal one subtracts 1.
    int32_t s = j - i + below;
    int32_t L = s >= 0 ? s / K : -1, q = s - L * K;
    int state = 0; // 0 H, 1 F (vertical), 2 E (horizontal)
    uint32_t nr = 0, cur_op = 0, cur_len = 0, tstart = 0;
    auto emit = [&](uint32_t op, uint32_t len) {
        if (len == 0) return;
        if (cur_len && op == cur_op) {
            cur_len += len;
        } else {
            if (cur_len) out[nr++] = (cur_len << 4) | cur_op;
            cur_op = op;
            cur_len = len;
        }
    };
    bool bad = false, done = false;
    int32_t cL = -1, cwd = -1, cb = -1;         // the lane-word and 16-row block held in LDS
    int32_t pwi = -1, twi = -2;                 // the pattern word (of block cb) and text words twi, twi - 1
    uint32_t pwv = 0, tw0 = 0, tw1 = 0;
    for (;;) {
        // Refill round: the line of the current cell, and for the H state its codes.
        if (state != 0 || (i > 0 && j > 0)) {
            if (i < 1 || i > m || s < 0 || s >= slots) {
                bad = true;
                break;
            }
            const int32_t wd = q >> 3, b = (i - 1) >> 4;
            const bool line = L != cL || wd != cwd || b != cb;
            if (line) {
                const uint4 *line = reinterpret_cast<const uint4 *>(tb + ((uint64_t)(b * nwl + wd) * LP + L) * 16);
                uint4 v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = line[k];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    my[4 * k] = v[k].x;
                    my[4 * k + 1] = v[k].y;
                    my[4 * k + 2] = v[k].z;
                    my[4 * k + 3] = v[k].w;
                }
                cL = L;
                cwd = wd;
                cb = b;
            }
            if (state == 0) {
                if (((i - 1) >> 4) != pwi) {
                    pwi = (i - 1) >> 4;
                    pwv = codes16(P, pwi * 16);
                }
                // Re-centred on the cell's word whenever a line is loaded too (same round): the
                // next 16 diagonal steps then stay inside the two words, and a text refill does
                // not cost a round of its own between two line refills.
                const int32_t tw = (j - 1) >> 4;
                if (tw != twi && (tw != twi - 1 || line)) {
                    twi = tw;
                    tw0 = codes16(T, tw * 16);
                    tw1 = tw > 0 ? codes16(T, (tw - 1) * 16) : 0u;
                }
            }
        }
        // Steps while this lane's data holds, as arithmetic on the state rather than branches
        // (the lanes of a wave sit in different states), with one exit: per step one nibble, at
        // most one run written.
        bool fin, oob;
        for (;;) {
            fin = state == 0 && (i == 0 || j == 0);
            oob = !fin && (i < 1 || i > m || s < 0 || s >= slots); // off the band: see the refill
            const int32_t wd = q >> 3, b = (i - 1) >> 4, tw = (j - 1) >> 4;
            const bool have = !fin && !oob && L == cL && wd == cwd && b == cb;
            const int32_t nq = min(K - (wd << 3), 8);
            const uint32_t w = (my[(i - 1) & 15] >> (4 * (nq - 1 - (q & 7)))) & 15u;
            const bool h = state == 0, diag = h && (w & 8u) == 0;
            const bool codes = b == pwi && (tw == twi || tw == twi - 1);
            if (!have || (diag && !codes)) break; // done, or a refill first
            const uint32_t pc = (pwv >> (2 * ((i - 1) & 15))) & 3u;
            const uint32_t tc = ((tw == twi ? tw0 : tw1) >> (2 * ((j - 1) & 15))) & 3u;
            // The step: H diagonal (= or X, i - 1, j - 1), H into F or E (no move), F (I, i - 1,
            // slot + 1, stays in F on fx), E (D, j - 1, slot - 1, stays in E on ex).
            const uint32_t op = diag ? (pc == tc ? 7u : 8u) : (h ? 0u : (uint32_t)state);
            const int32_t di = (diag || state == 1) ? 1 : 0, dj = (diag || state == 2) ? 1 : 0;
            const int32_t ds = state == 1 ? 1 : (state == 2 ? -1 : 0);
            const int hs = (w & 8u) == 0 ? 0 : ((w & 4u) ? 2 : 1);
            state = h ? hs : (state == 1 ? ((w & 2u) ? 1 : 0) : ((w & 1u) ? 2 : 0));
            const bool same = op == cur_op && cur_len != 0;
            if (op != 0 && !same && cur_len != 0) out[nr++] = (cur_len << 4) | cur_op;
            cur_len = op != 0 ? (same ? cur_len + 1 : 1u) : cur_len;
            cur_op = op != 0 ? op : cur_op;
            i -= di;
            j -= dj;
            s += ds;
            q += ds;
            L += q == K ? 1 : (q < 0 ? -1 : 0);
            q = q == K ? 0 : (q < 0 ? K - 1 : q);
        }
        if (fin) {
            if (i == 0) {
                if (a.mode == kInfix) tstart = (uint32_t)j;
                else emit(2, (uint32_t)j);
            } else {
                emit(1, (uint32_t)i);
            }
            done = true;
        } else if (oob) {
            bad = true;
            done = true;
        }
        if (done) break;
    }
    if (cur_len) out[nr++] = (cur_len << 4) | cur_op;
    if (bad) {
        nr = 0;
        tstart = kNone; // reported to the host as a device error
    }
    a.nruns[pi] = nr;
    a.tstart[pi] = tstart;
}

// 16 lanes per pair: each pass copies 16 consecutive runs (one 64-byte line each way), reversed.
__global__ __launch_bounds__(256) void compact_kernel(const uint32_t *order, uint32_t count, const uint32_t *runs,
                                                      const uint64_t *runs_off, const uint32_t *nruns,
                                                      const uint64_t *out_off, uint32_t *out) {
    const uint64_t g = ((uint64_t)blockIdx.x * 256u + threadIdx.x) >> 4;
    const uint32_t l = threadIdx.x & 15;
    if (g >= count) return;
    const uint32_t pi = order[g];
    const uint32_t nr = nruns[pi];
    const uint32_t *src = runs + runs_off[pi];
    uint32_t *dst = out + out_off[pi];
    for (uint32_t r = l; r < nr; r += 16) dst[r] = src[nr - 1 - r];
}

} // namespace

// Shape class for a band of `width` cells per row: 16 lanes per pair (four pairs per wave) up to
// the tuning limit (default 16 * 32 slots), where the per-row lane scan is shared by four pairs,
// then 64 lanes per pair.
static std::atomic<uint64_t> g_lp16_max{[] {
    const char *e = getenv("SF_AFFINE_LP16_MAX");
    return e ? (uint64_t)strtoull(e, nullptr, 10) : (uint64_t)16 * 32;
}()};

uint64_t sf_affine_lp16_max() { return g_lp16_max.load(std::memory_order_relaxed); }

// Slots per lane of the DP kernels' classes: every count up to 16, then steps that keep the unused
// slots of a class under about 1/9 (a pair uses the smallest class that holds its band).
static const int kClasses16[] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16,
                                 18, 20, 22, 24, 26, 28, 30, 32};
static const int kClasses64[] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 24, 28, 32};
// The passes with bookkeeping (small calls) use fewer classes.
static const int kClassesPass[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32};

int sf_affine_class(uint64_t width, bool lp16) {
    if (lp16 && width <= 16 * 32)
        for (int c : kClasses16)
            if (width <= (uint64_t)c * 16) return 16000 + c;
    for (int c : kClasses64)
        if (width <= (uint64_t)c * 64) return c;
    return 0;
}

int sf_affine_pass_class(uint64_t width, bool lp16) {
    if (lp16 && width <= 16 * 32)
        for (int c : kClassesPass)
            if (width <= (uint64_t)c * 16) return 16000 + c;
    for (int c : kClassesPass)
        if (width <= (uint64_t)c * 64) return c;
    return 0;
}

int sf_affine_k_for(uint64_t width) { return sf_affine_class(width, width <= sf_affine_lp16_max()); }

int sf_affine_pk16_class(int K, int32_t mat, int32_t mis, int32_t open, int32_t ext, uint32_t m, uint32_t below,
                         bool tb, int lanes) {
    if (mis != 2 * ext || (lanes != 16 && lanes != 32 && lanes != 64)) return 0;
    const int kp = K + (K & 1); // two slots a register
    if (kp > 32) return 0; // (the 64-lane classes' slot counts, 2 .. 20, 24, 28, 32, are all instantiated)
    if (lanes == 32 && (!tb || kp < 18)) return 0; // 32 lanes: traceback, 18 .. 32 slots a lane
    const int64_t dm = (int64_t)mat + mis;
    // The traceback kernel's offset is `open` higher (no band term saturates).
    const int64_t off = (int64_t)open * (tb ? 2 : 1) + (dm + ext) * below + 2;
    const int64_t top = off + dm * m + (int64_t)ext * lanes * kp;
    if (!tb && top < 0x7C00) return lanes * 1000 + 200 + kp; // the score kernel's three-term max
    return top <= 65535 ? lanes * 1000 + 100 + kp : 0;
}

int sf_affine_pk8_class(uint64_t width, int32_t mat, int32_t mis, int32_t open, int32_t ext, uint32_t m,
                        uint32_t below) {
    if (mis != 2 * ext || width > 8 * 36) return 0;
    const int k2 = (int)((width + 15) / 16); // registers a lane: 8 lanes x 2 slots each
    const int64_t dm = (int64_t)mat + mis;
    const int64_t off = (int64_t)open + (dm + ext) * below + 2;
    const int64_t top = off + dm * m + (int64_t)ext * 16 * k2;
    return top < 0x7C00 ? 8200 + 2 * k2 : 0;
}

int sf_affine_pk4_class(uint64_t width, int32_t mat, int32_t mis, int32_t open, int32_t ext, uint32_t m,
                        uint32_t below) {
    if (mis != 2 * ext) return 0;
    int kp = 0; // slots a lane: the instantiated 4-lane shapes
    for (int c : {36, 40, 48, 56, 64, 66, 68, 70, 72})
        if (width <= 4u * (uint64_t)c) {
            kp = c;
            break;
        }
    if (!kp) return 0;
    const int64_t dm = (int64_t)mat + mis;
    const int64_t off = (int64_t)open + (dm + ext) * below + 2;
    const int64_t top = off + dm * m + (int64_t)ext * 4 * kp;
    return top < 0x7C00 ? 4200 + kp : 0;
}

int sf_affine_pass_pk_class(uint64_t width, int lanes, int32_t mat, int32_t mis, int32_t open, int32_t ext,
                            uint32_t m, uint32_t below) {
    if (mis != 2 * ext || (lanes != 8 && lanes != 16)) return 0;
    int k2 = 0; // registers a lane: the instantiated packed pass shapes (2 slots a register)
    for (int c : {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16})
        if ((lanes == 16 || c >= 2) && width <= 2u * (uint64_t)lanes * c) {
            k2 = c;
            break;
        }
    if (!k2) return 0;
    const int64_t dm = (int64_t)mat + mis;
    const int64_t off = (int64_t)open + (dm + ext) * below + 2;
    const int64_t top = off + dm * m + (int64_t)ext * 2 * lanes * k2;
    if (top < 0x7C00) return lanes * 1000 + 200 + 2 * k2;
    return top <= 65535 ? lanes * 1000 + 100 + 2 * k2 : 0;
}

static std::atomic<int> g_pass_packed{[] {
    const char *e = getenv("SF_AFFINE_PASS_PACKED");
    const int v = e ? atoi(e) : 1;
    return v < 0 ? 0 : (v > 2 ? 2 : v);
}()};

int sf_affine_pass_packed() { return g_pass_packed.load(std::memory_order_relaxed); }
void sf_affine_set_pass_packed(int policy) { g_pass_packed.store(policy, std::memory_order_relaxed); }

static std::atomic<int> g_lp8_policy{[] {
    const char *e = getenv("SF_AFFINE_LP8");
    const int v = e ? atoi(e) : 0;
    return v < -1 ? -1 : (v > 2 ? 2 : v);
}()};

int sf_affine_lp8_policy() { return g_lp8_policy.load(std::memory_order_relaxed); }
void sf_affine_set_lp8_policy(int policy) { g_lp8_policy.store(policy, std::memory_order_relaxed); }

extern "C" int sf_set_affine_lp16_max(uint64_t slots) {
    g_lp16_max.store(slots, std::memory_order_relaxed);
    return 0;
}

hipError_t sf_launch_affine(const AffineArgs &a, int cls, bool tb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    const int LP = aff_lp(cls);
    dim3 grid((unsigned)(((uint64_t)a.count * LP + 255) / 256)), block(256);
    if (aff_packed(cls)) return sf_launch_affine_packed(a, cls, tb, s);
    switch (cls) {
#define CASE(LL, KK)                                                                             \
    case (LL == 64 ? 0 : LL * 1000) + KK:                                                       \
        if (tb) hipLaunchKernelGGL((affine_wave_kernel<LL, KK, true>), grid, block, 0, s, a);   \
        else hipLaunchKernelGGL((affine_wave_kernel<LL, KK, false>), grid, block, 0, s, a);     \
        break;
#define CASES(LL)                                                                                \
    CASE(LL, 1) CASE(LL, 2) CASE(LL, 3) CASE(LL, 4) CASE(LL, 5) CASE(LL, 6) CASE(LL, 7) CASE(LL, 8) \
    CASE(LL, 9) CASE(LL, 10) CASE(LL, 11) CASE(LL, 12) CASE(LL, 13) CASE(LL, 14) CASE(LL, 15)        \
    CASE(LL, 16) CASE(LL, 18) CASE(LL, 20) CASE(LL, 24) CASE(LL, 28) CASE(LL, 32)
        CASES(16)
        CASES(64)
        CASE(16, 22) CASE(16, 26) CASE(16, 30)
#undef CASES
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

namespace {
__global__ __launch_bounds__(256) void desc_kernel(uint64_t n, const uint32_t *order, PairsView pv, const uint64_t *ba,
                                                   PairDesc *desc) {
    const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= n) return;
    const uint32_t pi = order[w];
    const uint64_t g = ba ? ba[pi] : 0;
    desc[w] = PairDesc{pv.poff[pi], pv.toff[pi], pv.plen[pi], pv.tlen[pi], (uint32_t)g, (uint32_t)(g >> 32)};
}
} // namespace

hipError_t sf_launch_desc(uint64_t n, const uint32_t *order, const PairsView &pv, const uint64_t *ba, PairDesc *desc,
                          hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(desc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, order, pv, ba, desc);
    return hipGetLastError();
}

hipError_t sf_launch_pass(const PassArgs &a, int cls, bool local, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    if (aff_packed(cls)) return local ? hipErrorInvalidValue : sf_launch_pass_packed(a, cls, s);
    const int LP = aff_lp(cls);
    dim3 grid((unsigned)(((uint64_t)a.count * LP + 255) / 256)), block(256);
    switch (cls) {
#define CASE(LL, KK)                                                                             \
    case (LL == 64 ? 0 : LL * 1000) + KK:                                                       \
        if (local) hipLaunchKernelGGL((affine_pass_kernel<LL, KK, true>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((affine_pass_kernel<LL, KK, false>), grid, block, 0, s, a);      \
        break;
#define CASES(LL)                                                                                \
    CASE(LL, 1) CASE(LL, 2) CASE(LL, 3) CASE(LL, 4) CASE(LL, 5) CASE(LL, 6) CASE(LL, 7) CASE(LL, 8) \
    CASE(LL, 10) CASE(LL, 12) CASE(LL, 14) CASE(LL, 16) CASE(LL, 20) CASE(LL, 24) CASE(LL, 28) CASE(LL, 32)
        CASES(16)
        CASES(64)
#undef CASES
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t sf_launch_walk(const WalkArgs &a, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL(walk_kernel, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t sf_launch_compact(const uint32_t *order, uint32_t count, const uint32_t *runs, const uint64_t *runs_off,
                             const uint32_t *nruns, const uint64_t *out_off, uint32_t *out, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)(((uint64_t)count * 16 + 255) / 256)), dim3(256), 0, s, order,
                       count, runs, runs_off, nruns, out_off, out);
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_affine() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&desc_kernel));
}
