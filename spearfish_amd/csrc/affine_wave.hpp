// affine_wave.hpp -- the 32-bit banded affine-gap (Gotoh) DP kernel on gfx950, one pair per wave
// or four (16 lanes a pair), with optional traceback bits.
//
// Computes crates/map/src/align.rs:2446-2666 `run_core` (geom = n, no clip, no cut): the same
// H in every band cell, hence the same score and score column.
//
// Layout.  Row i covers text columns [i - below, i + above] clipped to [0, n] (align.rs:
// 2468-2481).  The wave holds one row in diagonal coordinates: slot s = j - i + below, lane L
// owning slots L*K .. L*K + K - 1, so the band never moves across lanes.  From row i-1 to row i
// the vertical neighbour of slot s is slot s + 1 (a register, or lane L + 1's first slot via
// DPP wave_shl:1) and the diagonal neighbour is slot s itself.
//
// Values are kept biased by B(i, s) = ext*s + 2*ext*i:  Hb = H + B, Fb = F + B, xb = x + B.
// The reference's passes (align.rs:1018-1072) then become, per cell,
//   Fb = max(Ub_up, Fb_up)            with Ub = Hb + (ext - go), made by the row above
//   xb = max(Hb_diag + s + 2 ext, Fb)
//   Hb = max(xb, max{xb of the slots left of s in the row} + (ext - go))
// because a horizontal gap from slot s' to s costs go + ext*(s - 1 - s'), which the bias turns
// into the constant go - ext; the running max over the row is an in-lane max plus one exclusive
// DPP max-scan across the 64 lanes.  With open >= 0 this is Gotoh's recurrence (align.rs:
// 1060-1063); every value is exact i32.  The reference clamps F at NEG (align.rs:1053); that
// clamp only acts where both vertical terms are below NEG, where H comes from the diagonal, so
// it changes no H and is dropped.  Registers hold each value plus 2^31 (mod 2^32): unsigned max
// then orders them as signed values, and DPP's zero fill (lanes with no source) is below all.
// The diagonal term is Hb_diag + (mat + 2 ext) - (mat + mis) * [text code != pattern code], the
// mismatch bit taken from one xor of the lane's text codes with the row's pattern code.
//
// Cells outside the band.  The reference reads NEG just right of each row's band (align.rs:
// 1047-1048) and starts the horizontal gap of a row beginning at lo > 0 from NEG.  Here the slots
// s >= Wd right of the widest row make their up-values 2^30 lower, so the rightmost band cell's
// vertical terms are NEG-level, as in the reference; a row starting at lo > 0 starts at slot 0,
// whose scan carry is below every value; slots left of column 0 (top rows) and right of column n
// (bottom rows) hold values that real band values dominate and that no band cell reads as a
// vertical or diagonal neighbour.  The host rejects scorings for which band values could come
// within reach of NEG.
//
// Traceback bits (own definition; the reference has no traceback), per cell
//   nd: H != Hdiag + s,  ne: H != x,  fx: F > Hup - go,  ex: E > Hleft - go,
// accumulated per lane, 8 slots per u32 (slot q's nibble nd,ne,fx,ex from high to low, the first
// slot highest), NWL = ceil(4K / 32) words per lane and row, stored in 16-row blocks: row i, word
// wd of lane L at ((b * NWL + wd) * LP + L) * 16 + (i - 1) % 16 with b = (i - 1) / 16.
// The walk reads them back under the oracle's tie rules.
#pragma once
#include "internal.hpp"

using namespace sf;

namespace {

// acc = 2 * acc + (a != b): the compare's VCC is the carry-in of one add.
__device__ __forceinline__ uint32_t push_ne(uint32_t acc, int32_t a, int32_t b) {
    asm("v_cmp_ne_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc" : "+v"(acc) : "v"(a), "v"(b) : "vcc");
    return acc;
}

// acc = 2 * acc + (a > b) for values in the offset-unsigned domain.
__device__ __forceinline__ uint32_t push_ugt(uint32_t acc, uint32_t a, uint32_t b) {
    asm("v_cmp_gt_u32_e32 vcc, %1, %2\n\tv_addc_co_u32_e32 %0, vcc, %0, %0, vcc" : "+v"(acc) : "v"(a), "v"(b) : "vcc");
    return acc;
}

constexpr uint32_t kOff = 0x80000000u; // offset-unsigned domain: v = x + 2^31 (mod 2^32)

// The 2-bit code at bits sh, sh+1 of w, replicated into all 16 fields of a word: a byte of four
// copies, broadcast to the word's four bytes by one byte permute (no 32-bit multiply).
__device__ __forceinline__ uint32_t rep16(uint32_t w, uint32_t sh) {
    const uint32_t c4 = __umul24(__builtin_amdgcn_ubfe(w, sh, 2), 0x55u); // the code in 4 fields: one byte
    return __builtin_amdgcn_perm(0u, c4, 0u);                            // that byte in all 4 bytes
}

// Bit 2f+1 of the result is set where field f of `win` differs from field f of `rep`:
// t = win ^ rep, then (t << 1) | t in one v_lshl_or.
__device__ __forceinline__ uint32_t mismatch_bits(uint32_t win, uint32_t rep) {
    const uint32_t t = win ^ rep;
    uint32_t y;
    asm("v_lshl_or_b32 %0, %1, 1, %1" : "=v"(y) : "v"(t));
    return y;
}

// LP lanes per pair, K slots per lane: 64 / LP pairs per wave.  With LP = 16 every lane move
// stays inside a DPP row, so the four pairs of a wave never see each other; each runs its own
// rows (a pair whose rows are done sits out the wave's remaining ones) and its own geometry.
// The kernel's body for workgroup `blk` of a launch (a mixed launch runs several classes).
template <int LP, int K, bool TB>
__device__ __forceinline__ void affine_wave_body(const AffineArgs &a, uint32_t blk) {
    constexpr bool ONE = LP == 64; // one pair per wave: pair values are wave-uniform
    constexpr int SHL = ONE ? kWaveShl1 : kRowShl1;
    constexpr int SHR = ONE ? kWaveShr1 : kRowShr1;
    constexpr int NWL = kTbWords<K>;
    constexpr int BLK = 16;                       // rows per block: one pattern word
    constexpr int NWIN = (BLK + K - 1 + 15) / 16; // text window words: codes of the block's slots
    constexpr int NY = (K + 15) / 16;             // mismatch words per row
    const int lane = threadIdx.x & 63;
    const int l = lane % LP;
    const uint32_t w = (blk * 256u + threadIdx.x) / LP; // pair slot in the launch order
    const bool valid = w < a.count;
    if (ONE && !valid) return; // whole waves only
    auto uni = [](int32_t v) { return ONE ? __builtin_amdgcn_readfirstlane(v) : v; };
    const uint32_t pi = valid ? (uint32_t)uni((int32_t)a.order[w]) : 0u;
    const PairDesc dsc = valid ? a.desc[w] : PairDesc{0, 0, 0, 0, 0, 0};
    const int32_t m = uni((int32_t)dsc.m);
    const int32_t n = uni((int32_t)dsc.n);
    const int32_t below = uni((int32_t)dsc.below);
    const int32_t above = uni((int32_t)dsc.above);
    const int32_t Wd = below + above + 1;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, (uint32_t)n);
    const int32_t mat = a.mat, mis = a.mis, ext = a.ext;
    const uint32_t dlt = (uint32_t)-a.open; // ext - go
    const uint32_t scm = (uint32_t)(mat + 2 * ext);
    const int32_t ndl = -(mat + mis); // scx - scm
    const bool infix = a.mode == kInfix;
    const int32_t s0 = l * K;

    // Values live as v = Hb + 2^31 (mod 2^32), so unsigned max orders them as signed and DPP's
    // zero fill is below everything.
    uint32_t V[K], Fb[K], Ub[K];
    const int32_t hi0 = above < n ? above : n;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int32_t s = s0 + q;
        const int32_t j = s - below;
        const int32_t h = (j >= 0 && j <= hi0) ? ((j == 0 || infix) ? 0 : -(a.open + ext * j)) : kNeg;
        V[q] = (uint32_t)(h == kNeg ? kNeg : h + ext * s) + kOff;
        Fb[q] = (uint32_t)kNeg + kOff;
        Ub[q] = V[q] + dlt;
    }
    // Slots s >= Wd lie right of the band.  Their up-values Ub carry an extra -2^30, so every
    // up-value above the band stays below NEG (|Hb| < 2^28 over all LP * K slots, checked on
    // the host) and F there is exactly NEG: the last band slot sees the same NEG up-neighbour
    // as align.rs:2548-2556 gives it, with no per-row reset.
    uint32_t dltv[K];
#pragma unroll
    for (int q = 0; q < K; ++q) dltv[q] = s0 + q >= Wd ? dlt - (1u << 30) : dlt;
    // Row i has no band cell once i - below > n (align.rs:2543-2556); only a band laid out for a
    // longer text than this one gets there.
    const int32_t last = valid ? (m < n + below ? m : n + below) : 0;
    int32_t rows = last; // the wave's row count
    if constexpr (!ONE) {
#pragma unroll
        for (int o = 32; o >= LP; o >>= 1) rows = imax(rows, __shfl_xor(rows, o));
        rows = __builtin_amdgcn_readfirstlane(rows);
    }
    // Traceback: a lane stages its words of the current 16-row block in LDS (17-word stride:
    // conflict-free) and then writes each lane-word's 16 rows as one 64-byte line, the unit the
    // walk reads back.  Layout per pair: word ((b * NWL + wd) * LP + l) * 16 + (i - 1) % 16 for
    // row i in block b = (i - 1) / 16.
    __shared__ uint32_t tbs[TB ? 4 * NWL * 64 * 17 : 1];
    uint32_t *stage = tbs + ((threadIdx.x >> 6) * NWL * 64 + lane) * 17;
    uint32_t *tbp = nullptr;
    if constexpr (TB) tbp = a.tb + (valid ? a.tb_off[w] : 0) + l * 16;
    for (int32_t i0 = 1; i0 <= rows; i0 += BLK) {
        uint32_t pw = (uint32_t)uni((int32_t)codes16(P, i0 - 1));
        uint32_t win[NWIN];
        const int32_t pos = i0 - below + s0 - 1; // text base of this lane's first slot in row i0
#pragma unroll
        for (int k = 0; k < NWIN; ++k) win[k] = codes16(T, pos + 16 * k);
        const int32_t cnt = rows - i0 + 1 < BLK ? rows - i0 + 1 : BLK;
        for (int32_t r = 0; r < cnt; ++r) {
            const uint32_t pcw = pw;
            pw >>= 2;
            if (!ONE && i0 + r > last) continue; // this pair's rows are done
            // The row's pattern code in every 2-bit field.
            const uint32_t prep = rep16(pcw, 0);
            // Bit 2q+1 of y: slot q's text code differs from the pattern code.
            uint32_t y[NY];
#pragma unroll
            for (int k = 0; k < NY; ++k) y[k] = mismatch_bits(win[k], prep);
            const uint32_t Un = dpp0<SHL>(Ub[0]);
            const uint32_t Fn = dpp0<SHL>(Fb[0]);
            uint32_t X[K], FQ[K], G[K], D[K], Uu[K];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const uint32_t uu = q < K - 1 ? Ub[q + 1] : Un;
                const uint32_t fu = q < K - 1 ? Fb[q + 1] : Fn;
                const uint32_t mism = __builtin_amdgcn_ubfe(y[q / 16], 2 * (q % 16) + 1, 1);
                const uint32_t f = umax(uu, fu);
                Fb[q] = f; // in place: slot q - 1 has read Fb[q] already
                const uint32_t d = V[q] + (uint32_t)__mul24((int)mism, ndl) + scm;
                const uint32_t x = umax(d, f);
                X[q] = x;
                FQ[q] = f;
                D[q] = d;
                Uu[q] = uu;
                const uint32_t g = x + dlt;
                G[q] = q ? umax(G[q - 1], g) : g;
            }
            const uint32_t C = group_excl_umax<LP>(G[K - 1]);
            uint32_t E[K];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                E[q] = q ? umax(C, G[q - 1]) : C; // biased horizontal-gap value entering slot q
                V[q] = umax(X[q], E[q]);
            }
#pragma unroll
            for (int q = 0; q < K; ++q) Ub[q] = V[q] + dltv[q];
            if constexpr (TB) {
                const uint32_t ul0 = dpp0<SHR>(Ub[K - 1]);
#pragma unroll
                for (int wd = 0; wd < NWL; ++wd) {
                    uint32_t acc = 0;
#pragma unroll
                    for (int q = wd * 8; q < K && q < wd * 8 + 8; ++q) {
                        const uint32_t ul = q ? Ub[q - 1] : ul0;
                        acc = push_ne(acc, (int32_t)V[q], (int32_t)D[q]);
                        acc = push_ne(acc, (int32_t)V[q], (int32_t)X[q]);
                        acc = push_ugt(acc, FQ[q], Uu[q]);
                        acc = push_ugt(acc, E[q], ul);
                    }
                    stage[wd * 64 * 17 + r] = acc;
                }
            }
#pragma unroll
            for (int k = 0; k < NWIN; ++k) win[k] = k + 1 < NWIN ? __builtin_amdgcn_alignbit(win[k + 1], win[k], 2) : win[k] >> 2;
        }
        if constexpr (TB) {
            if (valid && i0 <= last) {
                uint32_t *blk = tbp + (uint32_t)((i0 - 1) >> 4) * (NWL * LP * 16);
#pragma unroll
                for (int wd = 0; wd < NWL; ++wd) {
                    const uint32_t *src = stage + wd * 64 * 17;
                    uint4 *dst = reinterpret_cast<uint4 *>(blk + wd * LP * 16);
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        dst[k] = make_uint4(src[4 * k], src[4 * k + 1], src[4 * k + 2], src[4 * k + 3]);
                }
            }
        }
    }
    const bool stopped = last < m;
    int32_t score;
    uint32_t col;
    const int32_t rowbias = 2 * ext * m;
    if (a.mode == kGlobal) {
        const int32_t sN = n - m + below; // slot of (m, n)
        const int owner = sN / K, qq = sN % K;
        uint32_t v = V[0];
#pragma unroll
        for (int q = 1; q < K; ++q) v = q == qq ? V[q] : v;
        const uint32_t vo = ONE ? (uint32_t)__builtin_amdgcn_readlane((int)v, owner)
                                : (uint32_t)__shfl((int)v, (lane - l) + owner);
        score = (int32_t)(vo - kOff) - ext * sN - rowbias;
        col = (uint32_t)n;
    } else {
        // align.rs:2643-2650: max of the last row, first argmax.
        const int32_t b = m - below;
        const int32_t lo = b > 0 ? b : 0;
        const int32_t hi = (b + Wd - 1) < n ? (b + Wd - 1) : n;
        const int32_t slo = lo - b, shi = hi - b;
        int32_t bv = INT32_MIN;
        int32_t bs = INT32_MAX;
        // The lane's first slot, made opaque: its slot numbers are then computed here rather than
        // kept (or spilled) across the row loop from the start-up code.
        int32_t s0r = s0;
        asm volatile("" : "+v"(s0r));
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int32_t s = s0r + q;
            const int32_t h = (int32_t)(V[q] - kOff) - ext * s - rowbias;
            if (s >= slo && s <= shi && h > bv) {
                bv = h;
                bs = s;
            }
        }
#pragma unroll
        for (int off = LP / 2; off > 0; off >>= 1) {
            const int32_t ov = __shfl_xor(bv, off), os = __shfl_xor(bs, off);
            if (ov > bv || (ov == bv && os < bs)) {
                bv = ov;
                bs = os;
            }
        }
        score = bv;
        col = (uint32_t)(b + bs);
    }
    if (stopped) {
        score = kNeg;
        col = 0;
    }
    if (valid && l == 0) {
        a.score[pi] = score;
        if (a.score_col) a.score_col[pi] = col;
    }
}

template <int LP, int K, bool TB>
__global__ __launch_bounds__(256) void affine_wave_kernel(AffineArgs a) {
    affine_wave_body<LP, K, TB>(a, blockIdx.x);
}

} // namespace
