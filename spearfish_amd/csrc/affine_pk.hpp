// affine_pk.hpp -- the 16-bit packed banded affine kernels (scores; scores + traceback bits) on
// gfx950, for scorings with mis = 2 ext: templates shared by affine_pk.hip (8 and 16 lanes a pair)
// and affine_pk64.hip (64 lanes), which instantiate and launch them.
#pragma once
#include "internal.hpp"

using namespace sf;

namespace {

// ---- 16-bit packed kernel: scores only, 16 lanes per pair ----
//
// For scorings with mis = 2 ext (ONT, HIFI) a second bias, Hb = H + ext*s + mis*i, makes both gap
// extensions and a mismatch step free: a match adds dm = mat + mis, a gap open subtracts `open`,
// nothing else changes a value.  Band cells then satisfy -open - ext*below <= Hb <= dm*m + ext*s
// (a path along the main diagonal and one gap reaches every band cell; H <= mat*i), and cells
// outside the band -- left of column 0, which only exist in the first `below` rows -- start at
// the floor and gain at most dm a row.  Stored as Hb + off with off = open + (dm + ext)*below + 2
// every value is exact in unsigned 16 bits whenever off + dm*m + ext*(slots) <= 65535 (checked on
// the host), out-of-band values stay below every band value, and the saturating subtractions only
// ever floor out-of-band values.  This is the reference's narrow i16 pass (align.rs:2381-2426) on
// the device, with its own bound.
//
// Each lane holds 2*K2 slots, slot s0 + r in the low half of register r and s0 + K2 + r in the
// high half, so the vertical neighbour of register r is register r + 1 in both halves; the text
// window pairs the codes the same way.
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 as2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as1(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 pmax(u16x2 a, u16x2 b) { return __builtin_elementwise_max(a, b); }
__device__ __forceinline__ u16x2 psub(u16x2 a, u16x2 b) { return __builtin_elementwise_sub_sat(a, b); }
// Max of three values below 0x7C00 in one op: read as half floats they are positive and finite,
// and positive halves order as their bit patterns (denormals are kept in this mode), so the packed
// half-float maximum3 is the integer maximum.
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ u16x2 pmax3(u16x2 a, u16x2 b, u16x2 c) {
    const h16x2 x = __builtin_bit_cast(h16x2, a), y = __builtin_bit_cast(h16x2, b), z = __builtin_bit_cast(h16x2, c);
    return __builtin_bit_cast(u16x2, __builtin_elementwise_maximum(x, __builtin_elementwise_maximum(y, z)));
}

// The packed kernels' result: H(m, n) (Global) or the last row's maximum and its first column,
// unbiased, or NEG for a pair whose band ended before the last row.
template <int LP, int K2>
__device__ __forceinline__ void pk16_result(const AffineArgs &a, const u16x2 (&V)[K2], bool valid, uint32_t pi,
                                            int lane, int l, int32_t m, int32_t n, int32_t below, int32_t Wd,
                                            int32_t last, int32_t off) {
    constexpr int S = 2 * K2;
    // The lane's first slot, made opaque: its slot numbers s0 + wi are then computed here rather
    // than kept (or spilled) across the row loop from the start-up code.
    int32_t s0 = l * S;
    asm volatile("" : "+v"(s0));
    const int32_t ext = a.ext;
    const bool stopped = last < m;
    int32_t score;
    uint32_t col;
    const int32_t rowbias = a.mis * m + off;
    auto slot_value = [&](int32_t wi) { // this lane's value at its slot s0 + wi
        uint32_t v = 0;
#pragma unroll
        for (int q = 0; q < K2; ++q) {
            const uint32_t x = as1(V[q]);
            v = wi == q ? (x & 0xFFFFu) : (wi == q + K2 ? (x >> 16) : v);
        }
        return v;
    };
    if (a.mode == kGlobal) {
        const int32_t sN = n - m + below; // slot of (m, n)
        const int32_t owner = sN / S, wi = sN - owner * S;
        const uint32_t v = slot_value(wi);
        const uint32_t vo = (uint32_t)__shfl((int)v, (lane - l) + owner);
        score = (int32_t)vo - ext * sN - rowbias;
        col = (uint32_t)n;
    } else {
        // align.rs:2643-2650: max of the last row, first argmax.
        const int32_t b = m - below;
        const int32_t lo = b > 0 ? b : 0;
        const int32_t hi = (b + Wd - 1) < n ? (b + Wd - 1) : n;
        const int32_t slo = lo - b, shi = hi - b;
        int32_t bv = INT32_MIN;
        int32_t bs = INT32_MAX;
#pragma unroll
        for (int wi = 0; wi < S; ++wi) {
            const int32_t s = s0 + wi;
            const uint32_t x = as1(V[wi % K2]);
            const int32_t v = (int32_t)(wi < K2 ? (x & 0xFFFFu) : (x >> 16));
            const int32_t h = v - ext * s - rowbias;
            if (s >= slo && s <= shi && h > bv) {
                bv = h;
                bs = s;
            }
        }
#pragma unroll
        for (int o = LP / 2; o > 0; o >>= 1) {
            const int32_t ov = __shfl_xor(bv, o), os = __shfl_xor(bs, o);
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

// The text fields of a 16-row block whose register 0 low slot reads text position `pos` in the
// block's first row: field f holds the codes at positions pos + f (low half) and pos + K2 + f
// (high half) as byte selectors (bytes 0 and 2; 0x0c selects a zero byte), so that one byte
// permute against a row's one-hot pattern code gives a register's two match bits.  Row r of the
// block reads field q + r in register q.
template <int K2>
__device__ __forceinline__ void text_fields(const SeqRef &T, int32_t pos, uint32_t (&W)[K2 + 15]) {
    constexpr int NF = K2 + 15;
    constexpr int NWP = (NF + 7) / 8; // paired text words holding them
    uint32_t wp[NWP];
#pragma unroll
    for (int k = 0; k < NWP; ++k)
        wp[k] = (codes16(T, pos + 8 * k) & 0xFFFFu) | (codes16(T, pos + K2 + 8 * k) << 16);
#pragma unroll
    for (int f = 0; f < NF; ++f)
        W[f] = ((wp[f / 8] >> (2 * (f % 8))) & 0x00030003u) | 0x0c000c00u;
}

// The fields of the next block (16 positions on): field f takes field f + 16, and the top 16
// fields come from `lo` and `hi`, the 16 codes at the next block's positions pos + K2 - 1 and
// pos + 2 K2 - 1 (loaded a block ahead, so the loads' latency hides behind the rows).
template <int K2>
__device__ __forceinline__ void text_fields_next(uint32_t lo, uint32_t hi, uint32_t (&W)[K2 + 15]) {
    constexpr int NF = K2 + 15;
#pragma unroll
    for (int f = 0; f < NF - 16; ++f) W[f] = W[f + 16];
    // Paired words as text_fields builds them: 8 low-half codes under 8 high-half codes.
    const uint32_t wa = __builtin_amdgcn_perm(hi, lo, 0x05040100u);
    const uint32_t wb = __builtin_amdgcn_perm(hi, lo, 0x07060302u);
#pragma unroll
    for (int k = 0; k < 16; ++k)
        W[NF - 16 + k] = (((k < 8 ? wa : wb) >> (2 * (k % 8))) & 0x00030003u) | 0x0c000c00u;
}

// Row 0 of a pair's band in the packed layout: V (H + bias), U (the up-values the next row reads)
// and dl (`open`, or the floor right of the band).  G = X - open (X = max(diagonal, F): H before
// the row's horizontal gaps) is both the row's gap-open term for E and, as U = max(G, F), the next
// row's vertical term: F never opens from E, a horizontal gap turning straight into a vertical
// one.  With mis = 2 ext that turn never scores more than cutting its corner: a horizontal gap of
// a and a vertical gap of b cost 2 open + ext (a + b); a - 1, a diagonal step and b - 1 cost at
// most 2 open + ext (a + b - 2) + mis, the same, over cells the first path used.  So every H (and
// the score) is unchanged.  Right of the band G is floored, as the up-values there
// (align.rs:1047-1048).
template <int K2>
__device__ __forceinline__ void pk16_init(u16x2 (&V)[K2], u16x2 (&U)[K2], u16x2 (&dl)[K2], int32_t s0, int32_t below,
                                          int32_t above, int32_t n, int32_t Wd, int32_t open, int32_t ext,
                                          bool infix, int32_t off) {
    const int32_t hi0 = above < n ? above : n;
#pragma unroll
    for (int r = 0; r < K2; ++r) {
        uint32_t hv[2], dv[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int32_t s = s0 + r + h * K2;
            const int32_t j = s - below;
            const bool in = j >= 0 && j <= hi0;
            hv[h] = in ? (uint32_t)(((j == 0 || infix) ? 0 : -(open + ext * j)) + ext * s + off) : 0u;
            dv[h] = s >= Wd ? 0xFFFFu : (uint32_t)open;
        }
        V[r] = as2(hv[0] | (hv[1] << 16));
        dl[r] = as2(dv[0] | (dv[1] << 16));
        U[r] = psub(V[r], dl[r]); // row 0: the boundary's H opens the gaps, as the reference's
    }
}

// One row of the packed recurrence: row r of the block whose 16 pattern codes are `pw` (r static
// after unrolling, so every text field index is a constant), lane l of its group of LP lanes.
template <int LP, int K2, bool M3>
__device__ __forceinline__ void pk16_row(u16x2 (&V)[K2], u16x2 (&U)[K2], const u16x2 (&dl)[K2],
                                         const uint32_t (&W)[K2 + 15], uint32_t pw, const int r, int l,
                                         u16x2 vdm) {
    constexpr int SHL = LP == 64 ? kWaveShl1 : kRowShl1; // next lane up
    const uint32_t onehot = 1u << (__builtin_amdgcn_ubfe(pw, 2 * r, 2) << 3);
    // Register K2 - 1: low half reads slot s0 + K2 (register 0, high half), high half reads slot
    // s0 + 2 K2 (the next lane's register 0, low half).  (With 4 or 8 lanes a pair the row shift
    // hands a group's top lane the next group's value: that lane takes the floor.)
    const uint32_t u0 = as1(U[0]);
    uint32_t un = dpp0<SHL>(u0);
    if (LP < 16) un = l == LP - 1 ? 0u : un;
    const u16x2 fK = as2(__builtin_amdgcn_alignbit(un, u0, 16));
    u16x2 X[K2], G[K2];
#pragma unroll
    for (int q = 0; q < K2; ++q) {
        const uint32_t mt = __builtin_amdgcn_perm(0u, onehot, W[q + r]); // match bits
        const u16x2 d = V[q] + as2(mt) * vdm;
        const u16x2 f = q < K2 - 1 ? U[q + 1] : fK; // F of this row
        X[q] = pmax(d, f);
        G[q] = psub(X[q], dl[q]);
        U[q] = pmax(G[q], f); // in place: register q - 1 has read U[q] already
    }
    // Horizontal gaps: Pre[q] = max(G[0 .. q-1]), the lane's own prefix, is built while the group
    // scan runs; then every register takes max(X, Pre, gaps entering the lane) with no chain
    // after the scan.
    u16x2 Pre[K2];
    Pre[1 % K2] = G[0];
#pragma unroll
    for (int q = 2; q < K2; ++q) Pre[q] = pmax(Pre[q - 1], G[q - 1]);
    const u16x2 tot = K2 > 1 ? pmax(Pre[K2 - 1], G[K2 - 1]) : G[0];
    const uint32_t tt = as1(tot);
    const uint32_t tlo = tt & 0xFFFFu, thi = tt >> 16;
    const uint32_t C = group_excl_umax<LP>(tlo > thi ? tlo : thi);
    const u16x2 E = as2(C | ((C > tlo ? C : tlo) << 16)); // gap values entering register 0
    V[0] = pmax(X[0], E);
#pragma unroll
    for (int q = 1; q < K2; ++q) V[q] = M3 ? pmax3(X[q], Pre[q], E) : pmax(X[q], pmax(Pre[q], E));
}

// LP lanes per pair: 16 (four pairs a wave) or 8 (eight pairs a wave: the per-row work that every
// lane repeats -- scan, neighbour move, pattern code -- is shared by twice the slots).  M3: every
// value is below 0x7C00 (checked on the host), so H takes its three terms in one max.
// The kernel's body for workgroup `blk` of a launch (a mixed launch runs several classes).
template <int LP, int K2, bool M3>
__device__ __forceinline__ void affine_pk16_body(const AffineArgs &a, uint32_t blk) {
    static_assert(LP == 4 || LP == 8 || LP == 16 || LP == 64, "4, 8, 16 or 64 lanes a pair");
    constexpr int S = 2 * K2;                     // slots per lane
    constexpr int BLK = 16;                       // rows per block: one pattern word
    constexpr int NF = K2 + BLK - 1;              // text fields of a block: q + r, q < K2, r < 16
    const int lane = threadIdx.x & 63;
    const int l = lane & (LP - 1);
    const uint32_t w = (blk * 256u + threadIdx.x) / LP;
    const bool valid = w < a.count;
    const uint32_t pi = valid ? a.order[w] : 0u;
    const PairDesc dsc = valid ? a.desc[w] : PairDesc{0, 0, 0, 0, 0, 0};
    const int32_t m = (int32_t)dsc.m;
    const int32_t n = (int32_t)dsc.n;
    const int32_t below = (int32_t)dsc.below;
    const int32_t above = (int32_t)dsc.above;
    const int32_t Wd = below + above + 1;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, (uint32_t)n);
    const int32_t ext = a.ext, open = a.open, mis = a.mis;
    const int32_t dm = a.mat + mis;
    const bool infix = a.mode == kInfix;
    const int32_t off = open + (dm + ext) * below + 2;
    const int32_t s0 = l * S;
    const u16x2 vdm = (u16x2)((uint16_t)dm);

    u16x2 V[K2], U[K2], dl[K2];
    pk16_init<K2>(V, U, dl, s0, below, above, n, Wd, open, ext, infix, off);
    const int32_t last = valid ? (m < n + below ? m : n + below) : 0;
    int32_t rows = last, rows_all = last; // the wave's longest pair's rows, and rows every pair has
#pragma unroll
    for (int o = LP; o < 64; o <<= 1) {
        rows = imax(rows, __shfl_xor(rows, o));
        rows_all = imin(rows_all, __shfl_xor(rows_all, o));
    }
    rows = __builtin_amdgcn_readfirstlane(rows);
    rows_all = __builtin_amdgcn_readfirstlane(rows_all);
    // The text window slides 16 positions a block (register 0's low slot reads position
    // i0 - below + s0 - 1 in row i0): the first block loads every field, each later one only
    // its top 16.  Positions outside the text give unspecified codes, here as in a fresh load:
    // they only reach cells right of column n or left of column 0, whose values never flow back
    // into the band (see the bias note above).
    uint32_t W[NF];
    text_fields<K2>(T, -below + s0, W);
    uint32_t pw = 0; // the block's 16 pattern codes
    auto row = [&](const int r) { pk16_row<LP, K2, M3>(V, U, dl, W, pw, r, l, vdm); };
    // A block's pattern codes, and the next block's new text codes (loaded a block ahead).
    auto block_loads = [&](int32_t i0, uint32_t &nlo, uint32_t &nhi) {
        pw = codes16(P, i0 - 1);
        const int32_t posn = i0 + BLK - below + s0 - 1;
        nlo = codes16(T, posn + K2 - 1);
        nhi = codes16(T, posn + 2 * K2 - 1);
    };
    int32_t i0 = 1;
    // Blocks whose 16 rows every pair of the wave has: no per-row guard, one loop body, so the
    // rows' values stay in fixed registers across blocks.
    for (; i0 + BLK - 1 <= rows_all; i0 += BLK) {
        uint32_t nlo, nhi;
        block_loads(i0, nlo, nhi);
#pragma unroll
        for (int r = 0; r < BLK; ++r) row(r);
        text_fields_next<K2>(nlo, nhi, W);
    }
    // The wave's last blocks: each pair stops at its own last row.
    for (; i0 <= rows; i0 += BLK) {
        uint32_t nlo, nhi;
        block_loads(i0, nlo, nhi);
#pragma unroll
        for (int r = 0; r < BLK; ++r)
            if (i0 + r <= last) row(r);
        text_fields_next<K2>(nlo, nhi, W);
    }
    pk16_result<LP, K2>(a, V, valid, pi, lane, l, m, n, below, Wd, last, off);
}

template <int LP, int K2, bool M3>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LP == 4 ? 2 : 1))) void affine_pk16_kernel(AffineArgs a) {
    affine_pk16_body<LP, K2, M3>(a, blockIdx.x);
}

// A value from the next lane up (down) inside the pair's group of LP lanes, 0 past the group's
// last (first) lane.  32 lanes: the wave shift, with the move across the two groups' boundary
// (lanes 31 and 32) replaced by 0.
template <int LP>
__device__ __forceinline__ uint32_t from_up(uint32_t x) {
    if constexpr (LP == 64) return dpp0<kWaveShl1>(x);
    else if constexpr (LP == 32) {
        const uint32_t y = dpp0<kWaveShl1>(x);
        return (threadIdx.x & 31) == 31 ? 0u : y;
    } else return dpp0<kRowShl1>(x);
}

template <int LP>
__device__ __forceinline__ uint32_t from_down(uint32_t x) {
    if constexpr (LP == 64) return dpp0<kWaveShr1>(x);
    else if constexpr (LP == 32) {
        const uint32_t y = dpp0<kWaveShr1>(x);
        return (threadIdx.x & 31) == 0 ? 0u : y;
    } else return dpp0<kRowShr1>(x);
}

// Traceback variant of the packed kernel (16, 32 or 64 lanes a pair).  Same layout and bias; the
// offset is raised by `open` so that no band cell's up-, left- or gap term saturates, and F opens
// from H as the reference's does (the traceback reads F).  Per cell it writes the 32-bit kernel's bits nd, ne, fx, ex, each
// a saturating difference clamped to 1, in the 32-bit kernel's layout for K = 2*K2 slots a lane
// (slot q < K2 is register q's low half, slot K2 + q register q's high half), so the same walk
// reads them back.  Exact whenever off + dm*m + ext*slots <= 65535 (checked on the host).
template <int LP, int K2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(K2 == 9 ? 3 : 1))) void affine_pk16_tb_kernel(AffineArgs a) {
    static_assert(LP == 16 || LP == 32 || LP == 64, "16, 32 or 64 lanes a pair");
    constexpr int S = 2 * K2;
    constexpr int BLK = 16;
    constexpr int NF = K2 + BLK - 1;
    constexpr int NWL = kTbWords<S>;   // words per lane and row
    constexpr int NACC = (K2 + 3) / 4; // nibble accumulators, four registers each
    const int lane = threadIdx.x & 63;
    const int l = lane & (LP - 1);
    const uint32_t w = (blockIdx.x * 256u + threadIdx.x) / LP;
    const bool valid = w < a.count;
    const uint32_t pi = valid ? a.order[w] : 0u;
    const PairDesc dsc = valid ? a.desc[w] : PairDesc{0, 0, 0, 0, 0, 0};
    const int32_t m = (int32_t)dsc.m;
    const int32_t n = (int32_t)dsc.n;
    const int32_t below = (int32_t)dsc.below;
    const int32_t above = (int32_t)dsc.above;
    const int32_t Wd = below + above + 1;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, (uint32_t)n);
    const int32_t ext = a.ext, open = a.open;
    const int32_t dm = a.mat + a.mis;
    const bool infix = a.mode == kInfix;
    const int32_t off = 2 * open + (dm + ext) * below + 2;
    const int32_t s0 = l * S;
    const u16x2 vopen = (u16x2)((uint16_t)open), vdm = (u16x2)((uint16_t)dm);
    // 1 and 2 in both halves, hidden from the optimizer: min(a -sat b, 1) is a bit as one packed
    // min and 2 * acc + bit one packed multiply-add; seen as constants they are rewritten into
    // compares and selects, and a shift and an add.
    uint32_t one1 = 0x00010001u, two1 = 0x00020002u;
    asm("" : "+s"(one1), "+s"(two1));
    const u16x2 one = as2(one1), two = as2(two1);

    u16x2 V[K2], Fb[K2], Ub[K2], dl[K2];
    const int32_t hi0 = above < n ? above : n;
#pragma unroll
    for (int r = 0; r < K2; ++r) {
        uint32_t hv[2], dv[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int32_t s = s0 + r + h * K2;
            const int32_t j = s - below;
            const bool in = j >= 0 && j <= hi0;
            hv[h] = in ? (uint32_t)(((j == 0 || infix) ? 0 : -(open + ext * j)) + ext * s + off) : 0u;
            dv[h] = s >= Wd ? 0xFFFFu : (uint32_t)open; // right of the band: up-values floored
        }
        V[r] = as2(hv[0] | (hv[1] << 16));
        Fb[r] = as2(0u);
        dl[r] = as2(dv[0] | (dv[1] << 16));
        Ub[r] = psub(V[r], dl[r]);
    }
    const int32_t last = valid ? (m < n + below ? m : n + below) : 0;
    int32_t rows = last, rows_all = last; // the wave's longest pair's rows, and rows every pair has
#pragma unroll
    for (int o = LP; o < 64; o <<= 1) {
        rows = imax(rows, __shfl_xor(rows, o));
        rows_all = imin(rows_all, __shfl_xor(rows_all, o));
    }
    rows = __builtin_amdgcn_readfirstlane(rows);
    rows_all = __builtin_amdgcn_readfirstlane(rows_all);
    // Staging as in the 32-bit kernel: a lane's words of the current 16-row block in LDS, then
    // each lane-word's 16 rows written as one 64-byte line.
    __shared__ uint32_t tbs[4 * NWL * 64 * 17];
    uint32_t *stage = tbs + ((threadIdx.x >> 6) * NWL * 64 + lane) * 17;
    uint32_t *tbp = a.tb + (valid ? a.tb_off[w] : 0) + l * 16;
    // The text window slides 16 positions a block, as in the score kernel.
    uint32_t W[NF];
    text_fields<K2>(T, -below + s0, W);
    uint32_t pw = 0; // the block's 16 pattern codes
    auto row = [&](const int r) {
        const uint32_t onehot = 1u << (__builtin_amdgcn_ubfe(pw, 2 * r, 2) << 3);
        const uint32_t un = from_up<LP>(as1(Ub[0]));
        const uint32_t fn = from_up<LP>(as1(Fb[0]));
        const u16x2 uuK = as2(__builtin_amdgcn_alignbit(un, as1(Ub[0]), 16));
        const u16x2 fuK = as2(__builtin_amdgcn_alignbit(fn, as1(Fb[0]), 16));
        u16x2 X[K2], G[K2], D[K2], BF[K2];
#pragma unroll
        for (int q = 0; q < K2; ++q) {
            const u16x2 uu = q < K2 - 1 ? Ub[q + 1] : uuK;
            const u16x2 fu = q < K2 - 1 ? Fb[q + 1] : fuK;
            const uint32_t mt = __builtin_amdgcn_perm(0u, onehot, W[q + r]);
            D[q] = V[q] + as2(mt) * vdm;
            BF[q] = __builtin_elementwise_min(psub(fu, uu), one); // fx: F extends the gap above
            const u16x2 f = pmax(uu, fu);
            Fb[q] = f; // in place: register q - 1 has read Fb[q] already
            X[q] = pmax(D[q], f);
            G[q] = psub(X[q], vopen);
        }
        u16x2 Pre[K2];
        Pre[1 % K2] = G[0];
#pragma unroll
        for (int q = 2; q < K2; ++q) Pre[q] = pmax(Pre[q - 1], G[q - 1]);
        const u16x2 tot = K2 > 1 ? pmax(Pre[K2 - 1], G[K2 - 1]) : G[0];
        const uint32_t tt = as1(tot);
        const uint32_t tlo = tt & 0xFFFFu, thi = tt >> 16;
        const uint32_t C = group_excl_umax<LP>(tlo > thi ? tlo : thi);
        const u16x2 E = as2(C | ((C > tlo ? C : tlo) << 16));
        // The last register first: its new Ub holds the left neighbours of register 0 (slot
        // s0 - 1, the previous lane's last slot, in the low half; slot s0 + K2 - 1 in the
        // high half).  Then each register's new value and its four bits in one pass, so that
        // D, X and the gap term of a register die with it (fewer live registers, more waves).
        const u16x2 EQL = K2 > 1 ? pmax(Pre[K2 - 1], E) : E; // the horizontal gap entering
        V[K2 - 1] = pmax(X[K2 - 1], EQL);
        Ub[K2 - 1] = psub(V[K2 - 1], dl[K2 - 1]);
        const uint32_t ulp = from_down<LP>(as1(Ub[K2 - 1]));
        const u16x2 ul0 = as2(__builtin_amdgcn_alignbit(as1(Ub[K2 - 1]), ulp, 16));
        u16x2 acc[NACC];
#pragma unroll
        for (int g = 0; g < NACC; ++g) acc[g] = as2(0u);
#pragma unroll
        for (int q = 0; q < K2; ++q) {
            const u16x2 ul = q ? Ub[q - 1] : ul0;
            u16x2 EQq = EQL;
            if (q < K2 - 1) {
                EQq = q ? pmax(Pre[q], E) : E;
                V[q] = pmax(X[q], EQq);
                Ub[q] = psub(V[q], dl[q]);
            }
            const u16x2 nd = __builtin_elementwise_min(psub(V[q], D[q]), one);
            const u16x2 ne = __builtin_elementwise_min(psub(V[q], X[q]), one);
            const u16x2 ex = __builtin_elementwise_min(psub(EQq, ul), one);
            u16x2 A = acc[q / 4];
            A = A * two + nd;
            A = A * two + ne;
            A = A * two + BF[q];
            A = A * two + ex;
            acc[q / 4] = A;
        }
        // The row's nibbles in slot order -- the low halves' registers, then the high
        // halves' -- cut into words of 8 slots, the first slot highest; the last word holds
        // the remaining slots in its low bits.
        uint64_t buf = 0;
        int nb = 0, wd = 0;
        auto push = [&](uint32_t v, int bits) {
            buf = (buf << bits) | v;
            nb += bits;
            if (nb >= 32) {
                stage[wd * 64 * 17 + r] = (uint32_t)(buf >> (nb - 32));
                ++wd;
                nb -= 32;
                buf &= (1ull << nb) - 1;
            }
        };
#pragma unroll
        for (int g = 0; g < NACC; ++g) {
            const int cnt = K2 - 4 * g < 4 ? K2 - 4 * g : 4;
            push(as1(acc[g]) & 0xFFFFu, 4 * cnt);
        }
#pragma unroll
        for (int g = 0; g < NACC; ++g) {
            const int cnt = K2 - 4 * g < 4 ? K2 - 4 * g : 4;
            push(as1(acc[g]) >> 16, 4 * cnt);
        }
        if (nb > 0) stage[wd * 64 * 17 + r] = (uint32_t)buf;
    };
    // A block's pattern codes, and the next block's new text codes (loaded a block ahead).
    auto block_loads = [&](int32_t i0, uint32_t &nlo, uint32_t &nhi) {
        pw = codes16(P, i0 - 1);
        const int32_t posn = i0 + BLK - below + s0 - 1;
        nlo = codes16(T, posn + K2 - 1);
        nhi = codes16(T, posn + 2 * K2 - 1);
    };
    // The block's staged words out: each lane-word's 16 rows as one 64-byte line.
    auto block_store = [&](int32_t i0) {
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
    };
    int32_t i0 = 1;
    // Blocks whose 16 rows every pair of the wave has (one loop body, no per-row guard), then the
    // wave's last blocks, where each pair stops at its own last row.
    for (; i0 + BLK - 1 <= rows_all; i0 += BLK) {
        uint32_t nlo, nhi;
        block_loads(i0, nlo, nhi);
#pragma unroll
        for (int r = 0; r < BLK; ++r) row(r);
        block_store(i0);
        text_fields_next<K2>(nlo, nhi, W);
    }
    for (; i0 <= rows; i0 += BLK) {
        uint32_t nlo, nhi;
        block_loads(i0, nlo, nhi);
#pragma unroll
        for (int r = 0; r < BLK; ++r)
            if (i0 + r <= last) row(r);
        block_store(i0);
        text_fields_next<K2>(nlo, nhi, W);
    }
    pk16_result<LP, K2>(a, V, valid, pi, lane, l, m, n, below, Wd, last, off);
}

} // namespace
