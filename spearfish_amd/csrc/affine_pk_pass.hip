// affine_pk_pass.hip -- the passes with per-row bookkeeping (affine_dp with a clip, X-drop stop, best
// cell, cut or band geometry: align.rs:2580-2665 `run_core`) in the 16-bit packed layout of
// affine_pk.hpp, 8 or 16 lanes a pair.  The reference runs these passes in 16 bits where its `narrow`
// bound holds (align.rs:2381-2426, `run16_avx512`); here the packed kernel's own bound decides
// (checked on the host), and the 32-bit pass (affine.hip) takes every other pair.
//
// Rows are the packed score kernel's (pk16_row).  After each row:
//   * its maximum over its band cells.  A slot's key is v + ext (SL - 1 - s) (SL = the pair's 16
//     (or 8) x 2 K2 slots): v = H + ext s + mis i + off, so the key is H + mis i + off + ext (SL - 1), one
//     offset for the whole row, positive for every band cell (v >= dm * below + 2 there) and within
//     16 bits under the same bound as v (off + dm m + ext SL).  Slots past the band's width take key
//     0 (one v_pk_mad: v * 1 + add, or v * 0 + 0), and only rows whose band reaches left of column
//     0 or past column n (or past the cut, for the cut's maximum) mask further slots.
//   * the X-drop test, the best row and its first best column, the cut's best, as the 32-bit pass
//     keeps them (same order of tests, 32-bit arithmetic on the unbiased maximum).
// A pair stops at its X-drop row: its lanes are masked off from then on.  Every cross-lane move
// stays inside a pair's lanes (a DPP row, or half of one for 8 lanes; the moves that could read the
// other half take the floor), so a stopped pair beside a running one changes nothing.
#include "affine_pk.hpp"

namespace {

constexpr int kRowRor1 = 0x121, kRowRor2 = 0x122, kRowRor4 = 0x124, kRowRor8 = 0x128;
constexpr int kHalfMirror = 0x141, kQuadXor1 = 0xB1, kQuadXor2 = 0x4E; // lane i <-> 7 - i; quad [1,0,3,2]; [2,3,0,1]

// Max / min of one value a lane over a pair's LP lanes, every lane ending with it: 16 lanes (one
// DPP row) by rotations; 8 lanes (half a row) by the half-row mirror, then swaps inside quads.
template <int LP, typename Op>
__device__ __forceinline__ int32_t group_reduce(int32_t x, Op op) {
    if constexpr (LP == 16) {
        x = op(x, __builtin_amdgcn_update_dpp(x, x, kRowRor8, 0xf, 0xf, true));
        x = op(x, __builtin_amdgcn_update_dpp(x, x, kRowRor4, 0xf, 0xf, true));
        x = op(x, __builtin_amdgcn_update_dpp(x, x, kRowRor2, 0xf, 0xf, true));
        return op(x, __builtin_amdgcn_update_dpp(x, x, kRowRor1, 0xf, 0xf, true));
    } else {
        static_assert(LP == 8, "8 or 16 lanes a pair");
        x = op(x, __builtin_amdgcn_update_dpp(x, x, kHalfMirror, 0xf, 0xf, true));
        x = op(x, __builtin_amdgcn_update_dpp(x, x, kQuadXor1, 0xf, 0xf, true));
        return op(x, __builtin_amdgcn_update_dpp(x, x, kQuadXor2, 0xf, 0xf, true));
    }
}
template <int LP>
__device__ __forceinline__ uint32_t group_umax(uint32_t x) {
    return (uint32_t)group_reduce<LP>((int32_t)x, [](int32_t a, int32_t b) { return (int32_t)umax((uint32_t)a, (uint32_t)b); });
}
template <int LP>
__device__ __forceinline__ int32_t group_imax(int32_t x) {
    return group_reduce<LP>(x, [](int32_t a, int32_t b) { return imax(a, b); });
}
template <int LP>
__device__ __forceinline__ int32_t group_imin(int32_t x) {
    return group_reduce<LP>(x, [](int32_t a, int32_t b) { return imin(a, b); });
}

__device__ __forceinline__ u16x2 pmin(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ u16x2 splat(uint32_t v) { return (u16x2)((uint16_t)v); }
// Max of a register's two halves.
__device__ __forceinline__ uint32_t hmax(u16x2 v) {
    const uint32_t x = as1(v), lo = x & 0xFFFFu, hi = x >> 16;
    return lo > hi ? lo : hi;
}
__device__ __forceinline__ uint32_t hmin(u16x2 v) {
    const uint32_t x = as1(v), lo = x & 0xFFFFu, hi = x >> 16;
    return lo < hi ? lo : hi;
}

template <int LP, int K2, bool M3>
__global__ __launch_bounds__(256) void affine_pk16_pass_kernel(PassArgs a) {
    constexpr int S = 2 * K2;       // slots a lane
    constexpr int SL = LP * S;      // slots a pair
    constexpr int BLK = 16;         // rows a block: one pattern word
    constexpr int NF = K2 + BLK - 1;
    constexpr uint32_t kBig = 0x4000u; // above every slot number (SL <= 1,024)
    const int lane = threadIdx.x & 63;
    const int l = lane & (LP - 1);
    const uint32_t w = (blockIdx.x * 256u + threadIdx.x) / LP;
    const bool valid = w < a.count;
    const uint32_t pi = valid ? a.order[w] : 0u;
    const PairDesc dsc = valid ? a.desc[w] : PairDesc{0, 0, 0, 0, 0, 0};
    const int32_t m = (int32_t)dsc.m, n = (int32_t)dsc.n;
    const int32_t below = (int32_t)dsc.below, above = (int32_t)dsc.above;
    const int32_t Wd = below + above + 1;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, (uint32_t)n);
    const int32_t ext = a.ext, open = a.open, mis = a.mis;
    const int32_t dm = a.mat + mis;
    const bool infix = a.mode == kInfix;
    const int32_t off = open + (dm + ext) * below + 2;
    const int32_t s0 = l * S;
    const u16x2 vdm = splat((uint32_t)dm);
    // 1 in both halves, hidden from the optimizer (min(x, 1) stays one packed min).
    uint32_t one1 = 0x00010001u;
    asm("" : "+s"(one1));
    const u16x2 one = as2(one1);

    u16x2 V[K2], U[K2], dl[K2];
    pk16_init<K2>(V, U, dl, s0, below, above, n, Wd, open, ext, infix, off);
    // Row-maximum keys: key = v * kmul + kadd; the slot numbers for the masks and the best column.
    u16x2 kmul[K2], kadd[K2], sidx[K2];
#pragma unroll
    for (int q = 0; q < K2; ++q) {
        uint32_t mu[2], ad[2], si[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int32_t s = s0 + q + h * K2;
            mu[h] = s < Wd ? 1u : 0u;
            ad[h] = s < Wd ? (uint32_t)(ext * (SL - 1 - s)) : 0u;
            si[h] = (uint32_t)s;
        }
        kmul[q] = as2(mu[0] | (mu[1] << 16));
        kadd[q] = as2(ad[0] | (ad[1] << 16));
        sidx[q] = as2(si[0] | (si[1] << 16));
    }
    const int32_t keyoff = ext * (SL - 1) + off; // key = H + mis i + keyoff
    // Rows past n + below have no band cell (the band ran off the text).
    const int32_t last = valid ? (m < n + below ? m : n + below) : 0;
    int32_t rows = last;
#pragma unroll
    for (int o = LP; o < 64; o <<= 1) rows = imax(rows, __shfl_xor(rows, o));
    rows = __builtin_amdgcn_readfirstlane(rows);
    const int32_t xdrop = a.xdrop;
    // cut = usize::MAX is "no cut"; any other cut past the text covers the whole row.
    const uint64_t cut_in = (valid && a.cut) ? a.cut[pi] : UINT64_MAX;
    const int64_t cut = cut_in == UINT64_MAX ? -1 : (int64_t)(cut_in < (uint64_t)INT32_MAX ? cut_in : INT32_MAX);
    int32_t clip_best = 0, cut_best = 0;
    int32_t ci = 0, cj = 0;
    bool xstop = false;
    int32_t stop_row = 0, stop_read = 0;

    // Masked keys: slots outside [lo_s, hi_s] take 0 (below every band cell's key).
    auto masked = [&](const u16x2 (&k)[K2], u16x2 (&out)[K2], int32_t lo_s, int32_t hi_s) {
        const u16x2 lov = splat((uint32_t)lo_s), hiv = splat((uint32_t)(hi_s + 1));
#pragma unroll
        for (int q = 0; q < K2; ++q) {
            const u16x2 in = pmin(pmin(psub(sidx[q] + one, lov), psub(hiv, sidx[q])), one);
            out[q] = k[q] * in;
        }
    };
    auto keys_max = [&](const u16x2 (&k)[K2]) {
        u16x2 t = k[0];
#pragma unroll
        for (int q = 1; q < K2; ++q) t = pmax(t, k[q]);
        return group_umax<LP>(hmax(t));
    };
    // The row's bookkeeping (align.rs:2599-2628), after row i of the recurrence.
    auto book = [&](const int32_t i) {
        const int32_t slo = below - i > 0 ? below - i : 0;
        const int32_t shi = n - i + below < Wd - 1 ? n - i + below : Wd - 1;
        const int32_t lo = i - below > 0 ? i - below : 0;
        const int32_t hi = i + above < n ? i + above : n;
        u16x2 key[K2];
#pragma unroll
        for (int q = 0; q < K2; ++q) key[q] = V[q] * kmul[q] + kadd[q];
        if (slo > 0 || shi < Wd - 1) masked(key, key, slo, shi);
        const uint32_t rk = keys_max(key);
        const int32_t rbias = mis * i + keyoff;
        const int32_t rmax = (int32_t)rk - rbias;
        if (cut >= 0 && lo <= cut) {
            int32_t rc = rmax;
            if (hi > cut) {
                u16x2 kc[K2];
                masked(key, kc, slo, (int32_t)cut - i + below);
                rc = (int32_t)keys_max(kc) - rbias;
            }
            cut_best = imax(cut_best, rc);
        }
        if (a.has_clip) {
            if (rmax > clip_best) {
                clip_best = rmax;
                if (a.clip_cell) {
                    // The first slot holding the row's maximum key (other slots, and masked ones,
                    // hold less).
                    const u16x2 rv = splat(rk), big = splat(kBig);
                    u16x2 cmin = big;
#pragma unroll
                    for (int q = 0; q < K2; ++q)
                        cmin = pmin(cmin, pmin(psub(rv, key[q]), one) * big + sidx[q]);
                    const int32_t sm = group_imin<LP>((int32_t)hmin(cmin));
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
    };

    uint32_t W[NF];
    text_fields<K2>(T, -below + s0, W);
    for (int32_t i0 = 1; i0 <= rows; i0 += BLK) {
        const uint32_t pw = codes16(P, i0 - 1);
        const int32_t posn = i0 + BLK - below + s0 - 1;
        const uint32_t nlo = codes16(T, posn + K2 - 1), nhi = codes16(T, posn + 2 * K2 - 1);
#pragma unroll
        for (int r = 0; r < BLK; ++r) {
            const int32_t i = i0 + r;
            if (i <= last && !xstop) { // this pair's rows go on
                pk16_row<LP, K2, M3>(V, U, dl, W, pw, r, l, vdm);
                book(i);
            }
        }
        text_fields_next<K2>(nlo, nhi, W);
    }

    // End of pass (align.rs:2643-2665): score and column, the cut score, bases read.
    int32_t sb = s0; // the lane's first slot, opaque (its slot numbers are computed here)
    asm volatile("" : "+v"(sb));
    const int32_t rowb = mis * m + off; // h = v - ext s - rowb in row m
    const int32_t plo = m - below > 0 ? m - below : 0;
    const int32_t phi = m + above < n ? m + above : n;
    const int32_t slo = plo - (m - below), shi = phi - (m - below);
    auto slot_h = [&](int wi) { // H of this lane's slot sb + wi in row m (wi static)
        const uint32_t x = as1(V[wi % K2]);
        return (int32_t)(wi < K2 ? (x & 0xFFFFu) : (x >> 16)) - ext * (sb + wi) - rowb;
    };
    int32_t score;
    uint32_t col;
    if (a.mode == kGlobal) {
        // H[m][n]; NEG when (m, n) lies outside the last row's band.
        const int32_t sN = n - m + below;
        const bool inb = sN >= 0 && sN <= Wd - 1;
        const int32_t owner = inb ? sN / S : 0, wn = inb ? sN - owner * S : 0;
        int32_t v = 0;
#pragma unroll
        for (int wi = 0; wi < S; ++wi) v = wn == wi ? slot_h(wi) : v;
        const int32_t vo = __shfl(v, (lane - l) + owner);
        score = inb ? vo : kNeg;
        col = (uint32_t)n;
    } else {
        int32_t bv = INT32_MIN, bs = INT32_MAX;
#pragma unroll
        for (int wi = 0; wi < S; ++wi) {
            const int32_t s = sb + wi;
            const int32_t h = slot_h(wi);
            if (s >= slo && s <= shi && h > bv) {
                bv = h;
                bs = s;
            }
        }
        const int32_t mx = group_imax<LP>(bv);
        const int32_t sm = group_imin<LP>(bv == mx ? bs : INT32_MAX);
        score = mx;
        col = (uint32_t)(sm + (m - below));
    }
    int32_t cut_score = kNeg;
    if (cut >= 0 && plo <= cut) {
        const int32_t top = phi < cut ? phi : (int32_t)cut;
        const int32_t stop_s = top - (m - below);
        int32_t lc = INT32_MIN;
#pragma unroll
        for (int wi = 0; wi < S; ++wi) {
            const int32_t s = sb + wi;
            if (s >= slo && s <= stop_s) lc = imax(lc, slot_h(wi));
        }
        cut_score = imax(kNeg, group_imax<LP>(lc));
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

} // namespace

hipError_t sf_launch_pass_packed(const PassArgs &a, int cls, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    const int LP = aff_lp(cls);
    if ((LP != 8 && LP != 16) || !aff_packed(cls)) return hipErrorInvalidValue;
    const bool m3 = cls % 1000 >= 200;
    dim3 grid((unsigned)(((uint64_t)a.count * LP + 255) / 256)), block(256);
    switch (aff_k(cls) * 100 + LP) {
#define PKP(LL, KK)                                                                                 \
    case 2 * KK * 100 + LL:                                                                         \
        if (m3) hipLaunchKernelGGL((affine_pk16_pass_kernel<LL, KK, true>), grid, block, 0, s, a);  \
        else hipLaunchKernelGGL((affine_pk16_pass_kernel<LL, KK, false>), grid, block, 0, s, a);    \
        break;
        PKP(16, 1) PKP(16, 2) PKP(16, 3) PKP(16, 4) PKP(16, 5) PKP(16, 6) PKP(16, 7) PKP(16, 8) PKP(16, 10)
        PKP(16, 12) PKP(16, 14) PKP(16, 16)
        PKP(8, 2) PKP(8, 3) PKP(8, 4) PKP(8, 5) PKP(8, 6) PKP(8, 7) PKP(8, 8) PKP(8, 10) PKP(8, 12) PKP(8, 14)
        PKP(8, 16)
#undef PKP
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
