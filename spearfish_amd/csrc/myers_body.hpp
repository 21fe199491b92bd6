// myers_body.hpp -- the Myers/Hyyro bit-parallel edit distance kernel bodies on gfx950: one pair
// per lane (myers_lane_body) or one pair per G lanes in a pipeline over the text
// (myers_pipe_body).  A body runs as workgroup `blk` of its own launch; the kernels of
// myers.hip and the mixed launch of myers_mix.hip are thin wrappers around them.
//
// Computes what crates/map/src/align.rs:92-168 `distance_with` returns.  That function's
// Ukkonen cut-off only skips blocks whose rows all score above k; every cell it keeps is exact
// when <= k and an upper bound otherwise, so its result is exactly "the distance if <= k, else
// None" (the reference's own property tests check it against a full DP, align.rs:2899-2906,
// 3343-3371).  This kernel computes the full column of blocks for every text base and applies k
// at the end: the same result for every k.
//
// Each lane holds its pattern as two bit-planes per 64-row block (bit 0 and bit 1 of the codes),
// so a text base c selects its match mask as (P1 ^ m1) & (P2 ^ m2) with m1/m2 all-ones or zero
// from c; the vertical deltas pv/mv per block stay in registers (MAXB blocks, unrolled).  Blocks
// below every lane's last block use bit 63 for the horizontal delta out; each lane's last block
// takes its own row.  Bits above a pattern's last row never reach rows below it (carries and
// shifts move upward), so the last block needs no mask.
#pragma once
#include "internal.hpp"

#include <cstdint>
#include <type_traits>

namespace {

using namespace sf;


__device__ __forceinline__ void load_planes(const SeqRef &P, uint32_t m, int b, bool rev, uint64_t &p1,
                                            uint64_t &p2) {
    uint32_t w[4];
    if (!rev) {
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = codes16(P, 64 * (int64_t)b + 16 * q);
    } else {
        // Row i of the reversed pattern is P[m - 1 - i]: rows 64b.. are codes s.. reversed,
        // s = m - 64 - 64b.
        int64_t s = (int64_t)m - 64 - 64 * (int64_t)b;
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = reverse2(codes16(P, s + 16 * (3 - q)));
    }
    uint32_t lo1 = even16(w[0]) | (even16(w[1]) << 16), hi1 = even16(w[2]) | (even16(w[3]) << 16);
    uint32_t lo2 = even16(w[0] >> 1) | (even16(w[1] >> 1) << 16), hi2 = even16(w[2] >> 1) | (even16(w[3] >> 1) << 16);
    p1 = (uint64_t)lo1 | ((uint64_t)hi1 << 32);
    p2 = (uint64_t)lo2 | ((uint64_t)hi2 << 32);
}

__device__ __forceinline__ int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }

// Three-input bitwise ops (v_bitop3_b32): the immediate is the truth table, bit 4a + 2b + c, so a
// function's table is the function applied to the three columns below.
template <uint32_t TT>
__device__ __forceinline__ uint32_t lop3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, TT);
}
constexpr uint32_t kColA = 0xF0u, kColB = 0xCCu, kColC = 0xAAu;
constexpr uint32_t kXorAnd = ((kColA ^ kColB) & kColC) & 0xFFu; // (a ^ b) & c
constexpr uint32_t kXorOr = ((kColA ^ kColB) | kColC) & 0xFFu;  // (a ^ b) | c
constexpr uint32_t kOrNor = (kColA | ~(kColB | kColC)) & 0xFFu; // a | ~(b | c)

// One text base through one 64-row block (Hyyro's step, align.rs:51-66) on 32-bit halves (lo,
// hi), so no value needs an aligned register pair.  M1/M2 are all-ones or zero from the base's
// code bits; hp/hn carry the horizontal delta in (0 or 1 each) and leave with the delta out of
// row `bit` of the block.
struct Blk {
    uint32_t p1l, p1h, p2l, p2h; // pattern bit-planes
    uint32_t vpl, vph, vml, vmh; // vertical deltas +1 / -1
};

template <bool LAST63>
__device__ __forceinline__ void block_step(Blk &k, uint32_t M1, uint32_t M2, uint32_t &hp, uint32_t &hn, unsigned bit) {
    const uint32_t eql = lop3<kXorAnd>(k.p1l, M1, k.p2l ^ M2);
    const uint32_t eqh = lop3<kXorAnd>(k.p1h, M1, k.p2h ^ M2);
    const uint32_t xvl = eql | k.vml, xvh = eqh | k.vmh;
    const uint32_t enl = eql | hn, enh = eqh;
    // (en & pv) + pv over 64 bits: one add with carry out, one with carry in.
    const uint32_t al = enl & k.vpl, ah = enh & k.vph;
    const uint32_t suml = al + k.vpl;
    const uint32_t sumh = ah + k.vph + (suml < al ? 1u : 0u);
    const uint32_t xhl = lop3<kXorOr>(suml, k.vpl, enl), xhh = lop3<kXorOr>(sumh, k.vph, enh);
    const uint32_t hpl = lop3<kOrNor>(k.vml, xhl, k.vpl), hph = lop3<kOrNor>(k.vmh, xhh, k.vph); // +1 out
    const uint32_t hnl = k.vpl & xhl, hnh = k.vph & xhh;                                         // -1 out
    uint32_t op, om;
    if (LAST63) {
        op = hph >> 31;
        om = hnh >> 31;
    } else {
        const bool lo = bit < 32;
        op = __builtin_amdgcn_ubfe(lo ? hpl : hph, bit & 31u, 1);
        om = __builtin_amdgcn_ubfe(lo ? hnl : hnh, bit & 31u, 1);
    }
    const uint32_t spl = (hpl << 1) | hp, sph = __builtin_amdgcn_alignbit(hph, hpl, 31);
    const uint32_t sml = (hnl << 1) | hn, smh = __builtin_amdgcn_alignbit(hnh, hnl, 31);
    k.vpl = lop3<kOrNor>(sml, xvl, spl);
    k.vph = lop3<kOrNor>(smh, xvh, sph);
    k.vml = spl & xvl;
    k.vmh = sph & xvh;
    hp = op;
    hn = om;
}

// block_step<true> (the delta out of row 63 into hp/hn) that also ORs the horizontal deltas of the
// rows selected by (ml, mh) into (ap, an): with the lane's last-row bit as the mask in its last block
// and zero elsewhere, ap/an end nonzero exactly where that row's delta out is +1 / -1.
__device__ __forceinline__ void block_step_cap(Blk &k, uint32_t M1, uint32_t M2, uint32_t &hp, uint32_t &hn,
                                               uint32_t ml, uint32_t mh, uint32_t &ap, uint32_t &an) {
    const uint32_t eql = lop3<kXorAnd>(k.p1l, M1, k.p2l ^ M2);
    const uint32_t eqh = lop3<kXorAnd>(k.p1h, M1, k.p2h ^ M2);
    const uint32_t xvl = eql | k.vml, xvh = eqh | k.vmh;
    const uint32_t enl = eql | hn, enh = eqh;
    const uint32_t al = enl & k.vpl, ah = enh & k.vph;
    const uint32_t suml = al + k.vpl;
    const uint32_t sumh = ah + k.vph + (suml < al ? 1u : 0u);
    const uint32_t xhl = lop3<kXorOr>(suml, k.vpl, enl), xhh = lop3<kXorOr>(sumh, k.vph, enh);
    const uint32_t hpl = lop3<kOrNor>(k.vml, xhl, k.vpl), hph = lop3<kOrNor>(k.vmh, xhh, k.vph);
    const uint32_t hnl = k.vpl & xhl, hnh = k.vph & xhh;
    ap |= (hpl & ml) | (hph & mh);
    an |= (hnl & ml) | (hnh & mh);
    const uint32_t spl = (hpl << 1) | hp, sph = __builtin_amdgcn_alignbit(hph, hpl, 31);
    const uint32_t sml = (hnl << 1) | hn, smh = __builtin_amdgcn_alignbit(hnh, hnl, 31);
    k.vpl = lop3<kOrNor>(sml, xvl, spl);
    k.vph = lop3<kOrNor>(smh, xvh, sph);
    k.vml = spl & xvl;
    k.vmh = sph & xvh;
    hp = hph >> 31;
    hn = hnh >> 31;
}

// The columns of a lane pass over the wave's band of Ukkonen's cut-off (align.rs:113-149), in the
// add-only form: the wave computes blocks 0..L, and after a column it adds block L + 1 when some
// live lane holding that block has its band's bottom row (row 64 (L + 1)) at or below its
// threshold kk.  A cell scoring <= kk has its up-left neighbour at <= kk (D[i][j] >= D[i-1][j-1]),
// so every such cell of the next column lies in the band, and a block added after column j has no
// cell <= kk in column j: its untouched start state (every vertical delta +1, rows counting up from
// the row above) is an upper bound there, as the reference's added blocks are (align.rs:129-131).
// Computed cells are therefore exact where they score <= kk and upper bounds elsewhere: the
// distance when <= kk, its first end column, and above kk otherwise, exactly as the full column
// gives them.  Blocks are never dropped (a wider band than the reference's, still exact), so no
// block's state is ever reset, and the band only grows: each width L runs as its own loop
// (ukk_phase<L>, its L + 1 block steps unrolled, no per-block test) until the band grows.
//
// Lanes of a class hold LB + 1 .. MAXB blocks (ukk_lb): blocks below LB step as in every lane's
// interior; from LB on, each lane also takes the deltas out of its own last row (`hlast`), and
// blocks past a lane's last run on its zero bit-planes, ignored (rows below a pattern never reach
// it).  `sc` is the score at the bottom row of block min(L, lastb).
template <int MAXB>
constexpr int ukk_lb() { return MAXB <= 2 ? MAXB - 1 : (MAXB <= 4 ? 2 : (MAXB <= 8 ? 4 : MAXB - 4)); }

struct UkkLane {
    int nbw, lastb;        // blocks held (0 without work), the pattern's last block
    unsigned last_bit;
    uint32_t bl, bh;       // the last row's bit in the low and high word of its block
    int32_t k;             // the threshold
    int32_t sc, best, gres;
    uint32_t bestj, n, top;
    bool rev, global;
    SeqRef T;
    int nb_hi;             // wave-uniform: the most blocks a lane holds
    int32_t nmax;          // wave-uniform: the longest text
};

template <int MAXB, int L>
__device__ __forceinline__ int32_t ukk_phase(UkkLane &u, Blk (&K)[MAXB], int32_t j, uint32_t &tw) {
    constexpr int LB = ukk_lb<MAXB>();
    for (; j < u.nmax; ++j) {
        const bool live = j < (int32_t)u.n; // (a lane past its text runs on clamped codes; its results are final)
        if ((j & 15) == 0) tw = u.rev ? reverse2(codes16(u.T, (int64_t)u.n - 16 - (int64_t)j)) : codes16(u.T, j);
        const uint32_t c = tw & 3;
        tw >>= 2;
        const uint32_t M1 = (c & 1) ? 0u : ~0u;
        const uint32_t M2 = (c & 2) ? 0u : ~0u;
        uint32_t hp = u.top, hn = 0, ap = 0, an = 0;
#pragma unroll
        for (int b = 0; b <= L; ++b) {
            if (b < LB) {
                block_step<true>(K[b], M1, M2, hp, hn, 63u);
            } else {
                // Bit 63 on to the next block, and the deltas out of the lane's own last row.
                const bool own = b == u.lastb;
                block_step_cap(K[b], M1, M2, hp, hn, own ? u.bl : 0u, own ? u.bh : 0u, ap, an);
            }
        }
        // Out of block min(L, lastb): the lane's own last row, or the band's bottom row.
        u.sc += u.lastb <= L ? (int32_t)(ap != 0) - (int32_t)(an != 0) : (int32_t)hp - (int32_t)hn;
        int32_t nl = L;
        if constexpr (L + 1 < MAXB) {
            if (L + 1 < u.nb_hi) {
                const bool holds = L + 1 < u.nbw;
                if (__ballot(live && holds && u.sc <= u.k)) {
                    nl = L + 1;
                    if (holds) u.sc += L + 1 == u.lastb ? (int32_t)u.last_bit + 1 : 64;
                }
            }
        }
        if (live && nl >= u.lastb) { // the pattern's last row is in the lane's band: sc = D[m][j + 1]
            if (!u.global) {
                if (u.sc < u.best) {
                    u.best = u.sc;
                    u.bestj = j + 1;
                }
            } else if (j + 1 == (int32_t)u.n) {
                u.gres = u.sc;
            }
        }
        if (nl != L) return j + 1;
    }
    return j;
}

// The phases from width L on, entered at width L0 (the band only grows, one block at a time).
template <int MAXB, int L = 0>
__device__ __forceinline__ void ukk_phases(UkkLane &u, Blk (&K)[MAXB], int32_t &j, uint32_t &tw, int L0) {
    if constexpr (L < MAXB) {
        if (L >= L0 && j < u.nmax) j = ukk_phase<MAXB, L>(u, K, j, tw);
        ukk_phases<MAXB, L + 1>(u, K, j, tw, L0);
    }
}

// Returns false (nothing done) for a wave holding a pattern below its class's block counts.
template <int MAXB>
__device__ __forceinline__ bool myers_lane_ukk_columns(Blk (&K)[MAXB], uint32_t m, uint32_t n, int nb, int lastb,
                                                       unsigned last_bit, bool work, bool rev, int mode, uint32_t top,
                                                       const SeqRef &T, uint32_t kk, int32_t &best, uint32_t &bestj) {
    if (__ballot(work && lastb < ukk_lb<MAXB>())) return false;
    UkkLane u;
    u.nbw = work ? nb : 0;
    u.lastb = work ? lastb : MAXB; // (no work: holds nothing, never in a band)
    u.last_bit = last_bit;
    u.bl = last_bit < 32 ? 1u << last_bit : 0u;
    u.bh = last_bit < 32 ? 0u : 1u << (last_bit - 32);
    u.k = (int32_t)imin64((int64_t)kk, (int64_t)INT32_MAX - 64);
    u.best = best;
    u.bestj = bestj;
    u.gres = (mode == kGlobal && n == 0) ? (int32_t)m : INT32_MAX;
    u.n = n;
    u.top = top;
    u.rev = rev;
    u.global = mode == kGlobal;
    u.T = T;
    // Column 0: blocks 0 .. ceil((kk + 1) / 64) - 1 (align.rs:110), the wave's widest.
    int L0 = work ? (int)imin64((int64_t)nb, ((int64_t)kk + 64) / 64) - 1 : -1;
    int nb_hi = u.nbw, nmax = (int)n;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        L0 = imax(L0, __shfl_xor(L0, off));
        nb_hi = imax(nb_hi, __shfl_xor(nb_hi, off));
        nmax = imax(nmax, __shfl_xor(nmax, off));
    }
    // Wave-uniform; the columns run to the longest text, each lane live while inside its own (as
    // align.rs:237-238 runs eight texts).
    L0 = __builtin_amdgcn_readfirstlane(L0);
    u.nb_hi = __builtin_amdgcn_readfirstlane(nb_hi);
    u.nmax = __builtin_amdgcn_readfirstlane(nmax);
    u.sc = !work ? 0 : (L0 >= lastb ? (int32_t)m : 64 * (L0 + 1));
    uint32_t tw = 0;
    int32_t j = 0;
    if (L0 >= 0) ukk_phases<MAXB>(u, K, j, tw, L0);
    best = u.best;
    bestj = u.bestj;
    if (mode == kGlobal) {
        best = u.gres;
        bestj = n;
    }
    return true;
}

// UKK: the column runs only the blocks of the wave's Ukkonen band (myers_lane_ukk_columns), for
// passes with a threshold below the patterns' lengths.
template <int MAXB, bool UKK = false>
__device__ __forceinline__ void myers_lane_body(const MyersArgs &a, uint32_t blk) {
    const uint32_t t = blk * 256u + threadIdx.x;
    const bool valid = t < a.count;
    const uint32_t pi = valid ? a.order[t] : 0;
    const PairDesc dsc = valid ? a.desc[t] : PairDesc{0, 0, 0, 0, 0, 0};
    uint32_t m = dsc.m;
    const uint32_t tlen = dsc.n;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, tlen);
    const bool rev = a.reverse != 0;
    int mode = a.mode;
    uint32_t n = tlen;
    bool work = valid;
    if (valid && rev) {
        // The reversed Prefix pass runs only for Infix pairs with a distance and a pattern.
        mode = kPrefix;
        const uint32_t d = a.dist[pi];
        if (d == kNone || m == 0) work = false;
        else n = a.end[pi];
    }
    if (valid && !rev && m == 0) {
        // align.rs:96-99: the empty pattern is n (Global) or 0 away.
        const uint32_t d = mode == kGlobal ? tlen : 0;
        const bool ok = d <= a.thresh(pi);
        a.dist[pi] = ok ? d : kNone;
        if (a.end) a.end[pi] = ok ? (mode == kGlobal ? tlen : 0) : kNone;
        if (a.start) a.start[pi] = ok ? 0 : kNone;
        work = false;
    }
    if (!work) {
        m = 0;
        n = 0;
    }
    const int nb = work ? (int)((m + 63) / 64) : MAXB;
    // The fewest blocks of any lane in the wave: blocks below nb_lo - 1 are nobody's last.
    int nb_lo = nb;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nb_lo = imin(nb_lo, __shfl_xor(nb_lo, off));
    const int lastb = nb - 1;
    const unsigned last_bit = work ? (m - 1) & 63u : 63u;

    Blk K[MAXB];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        uint64_t p1 = 0, p2 = 0;
        if (b < nb && work) load_planes(P, m, b, rev, p1, p2);
        K[b] = Blk{(uint32_t)p1, (uint32_t)(p1 >> 32), (uint32_t)p2, (uint32_t)(p2 >> 32), ~0u, ~0u, 0u, 0u};
    }
    const uint32_t top = mode == kInfix ? 0 : 1;
    const bool full = nb_lo == MAXB; // every lane of the wave has all MAXB blocks
    int32_t score = (int32_t)m; // D[m][0]
    int32_t best = (int32_t)m;
    uint32_t bestj = 0;
    uint32_t tw = 0;
    bool ukk_done = false;
    if constexpr (UKK) {
        // The threshold: the call's k clamped to m + n (align.rs:100), or for the reversed pass
        // the pair's distance (the reversed Prefix pass from the end column reaches exactly it).
        const uint32_t kk = !work ? 0u : (rev ? a.dist[pi] : (uint32_t)imin64((int64_t)a.thresh(pi), (int64_t)m + n));
        ukk_done = myers_lane_ukk_columns<MAXB>(K, m, n, nb, lastb, last_bit, work, rev, mode, top, T, kk, best, bestj);
    }
    if (!ukk_done)
    for (uint32_t j = 0; j < n; ++j) {
        if ((j & 15) == 0) {
            tw = rev ? reverse2(codes16(T, (int64_t)n - 16 - (int64_t)j)) : codes16(T, j);
        }
        const uint32_t c = tw & 3;
        tw >>= 2;
        const uint32_t M1 = (c & 1) ? 0u : ~0u;
        const uint32_t M2 = (c & 2) ? 0u : ~0u;
        uint32_t hp = top, hn = 0;
        if (full) {
#pragma unroll
            for (int b = 0; b < MAXB - 1; ++b) block_step<true>(K[b], M1, M2, hp, hn, 63u);
            block_step<false>(K[MAXB - 1], M1, M2, hp, hn, last_bit);
        } else {
#pragma unroll
            for (int b = 0; b < MAXB; ++b) {
                if (b < nb_lo - 1)
                    block_step<true>(K[b], M1, M2, hp, hn, 63u);
                else if (b < nb)
                    block_step<false>(K[b], M1, M2, hp, hn, b == lastb ? last_bit : 63u);
            }
        }
        score += (int32_t)hp - (int32_t)hn;
        if (score < best) {
            best = score;
            bestj = j + 1;
        }
    }
    if (!work) return;
    if (!rev) {
        if (mode == kGlobal && !ukk_done) {
            best = score;
            bestj = n;
        }
        const bool ok = (uint32_t)best <= a.thresh(pi);
        a.dist[pi] = ok ? (uint32_t)best : kNone;
        if (a.end) a.end[pi] = ok ? bestj : kNone;
        if (a.start && (mode != kInfix || !ok)) a.start[pi] = ok ? 0 : kNone;
    } else {
        a.start[pi] = n - bestj;
    }
}

// Long patterns: G lanes per pair (G = 16: four pairs per wave, one DPP row each; G = 32: two
// pairs; G = 64: one pair per wave), lane l holding R consecutive 64-row blocks.  The lanes form a pipeline over the
// text in steps of CB columns: at step t lane l works columns CB (t - l) .. CB (t - l) + CB - 1,
// taking each column's horizontal delta out of lane l - 1's last block (computed one step
// earlier) and the columns' text codes, both moved by DPP row_shr:1 / wave_shr:1; lane 0 takes
// `top` and the next codes.  A step runs its CB columns block by block (block r's column c needs
// block r's column c - 1 and block r - 1's column c), so the fetch, the lane moves, the guards
// and the loop control are paid once per CB columns.  Blocks are placed so that the pattern's
// last block is the last slot of lane G - 1 (positions below off = G*R - nb hold no rows and pass
// the deltas through), so one slot per lane needs the per-pair last-row bit.  Columns past a
// text's end (in its last step) run on unspecified codes and are left out of the score.
//
// ST: one stripe of a pattern longer than G * R blocks (MyersArgs::stripe): the pipeline holds the
// stripe's blocks, lane 0 takes each column's delta from the stripe above (hcol_in; the first
// stripe takes `top`), the last lane writes its deltas out (hcol_out) unless the stripe holds the
// pattern's last block, which keeps the score as above.
template <int G, int R, bool ST = false>
__device__ __forceinline__ void myers_pipe_body(const MyersArgs &a, uint32_t blk) {
    // 16 lanes: one DPP row each; 32 and 64: the wave shift (with 32, lane 32 -- the second pair's
    // lane 0 -- keeps its own deltas, as lane 0 does).
    constexpr int SHR = G == 16 ? kRowShr1 : kWaveShr1;
    constexpr int CB = 4;       // text columns a lane takes per step
    constexpr int SW = 16 / CB; // steps per fetched text word
    const int lane = threadIdx.x & 63;
    const int l = lane % G;
    const uint32_t grp = (blk * 256u + threadIdx.x) / G;
    const bool valid = grp < a.count;
    const uint32_t pi = valid ? a.order[grp] : 0;
    const PairDesc dsc = valid ? a.desc[grp] : PairDesc{0, 0, 0, 0, 0, 0};
    const uint32_t m = dsc.m;
    const uint32_t tlen = dsc.n;
    const SeqRef P = seqref(a.pv.seq, dsc.poff, m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, tlen);
    const bool rev = a.reverse != 0;
    int mode = a.mode;
    uint32_t n = tlen;
    bool work = valid && m > 0;
    if (work && rev) {
        mode = kPrefix;
        if (a.dist[pi] == kNone) work = false;
        else n = a.end[pi];
    }
    // Stripes: this launch's blocks [b0, b0 + nb) of the pair's pattern (all of them without ST).
    bool st_last = true;
    int b0 = 0;
    const uint8_t *hcin = nullptr;
    uint8_t *hcout = nullptr;
    if constexpr (ST) {
        const int nb_all = work ? (int)((m + 63) / 64) : 0;
        const int nst = (nb_all + G * R - 1) / (G * R);
        st_last = (int)a.stripe == nst - 1;
        if ((int)a.stripe >= nst) work = false; // (its stripes are done)
        b0 = (int)a.stripe * G * R;
        if (a.stripe > 0) hcin = a.hcol_in + (uint64_t)grp * a.hcol_stride;
        if (!st_last) hcout = a.hcol_out + (uint64_t)grp * a.hcol_stride;
    }
    if (!work) n = 0;
    int nb = work ? (int)((m + 63) / 64) : 0;
    if constexpr (ST) nb = !work ? 0 : (!st_last ? G * R : nb - b0);
    const int off = G * R - nb; // >= 0: the launcher's class holds every block
    const unsigned last_bit = work ? (m - 1) & 63u : 63u;
    const unsigned bitR = (l == G - 1 && st_last) ? last_bit : 63u; // row of the delta out of slot R - 1
    Blk K[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        uint64_t p1 = 0, p2 = 0;
        const int b = l * R + r - off;
        if (work && b >= 0) load_planes(P, m, ST ? b0 + b : b, rev, p1, p2);
        K[r] = Blk{(uint32_t)p1, (uint32_t)(p1 >> 32), (uint32_t)p2, (uint32_t)(p2 >> 32), ~0u, ~0u, 0u, 0u};
    }
    // Steps: the wave runs until its longest text has passed every lane.  From step G on, up to
    // the shortest text's last whole step, every lane of the wave has CB valid columns, so those
    // steps run without guards; the first G steps (lanes starting) and the rest (texts ending)
    // check each lane and column.
    int32_t steps = work ? (int32_t)((n + CB - 1) / CB) + G - 1 : 0;
    int32_t full = work ? (int32_t)(n / CB) : INT32_MAX;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        steps = imax(steps, __shfl_xor(steps, o));
        full = imin(full, __shfl_xor(full, o));
    }
    steps = __builtin_amdgcn_readfirstlane(steps); // wave-uniform: scalar loops
    full = __builtin_amdgcn_readfirstlane(full < steps ? full : steps); // (no work in the wave: steps = 0)
    const int32_t fill = steps < G ? steps : G;
    const int32_t steady = full > fill ? fill + (full - fill) / SW * SW : fill;
    int32_t score = (int32_t)m, best = (int32_t)m;
    uint32_t bestj = 0; // as CB t + c + 1 (wave-uniform per column), the column j + 1 plus CB l
    uint32_t codes = 0;
    // hin/nin: the deltas into this lane's first block, per column of the step, moved in from the
    // lane above by DPP into the same registers, so lane 0 of a pair keeps its `top` (+1 for
    // Prefix and Global, 0 for Infix) and 0.  hp/hn: the deltas out of this lane's last block.
    uint32_t hin[CB], nin[CB], hp[CB], hn[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        hin[c] = mode == kInfix ? 0u : 1u;
        nin[c] = 0;
        hp[c] = hn[c] = 0;
    }
    // The 4 codes of the step's columns, for lane 0.
    auto fetch4 = [&](int32_t t) -> uint32_t {
        if (!work) return 0u;
        return rev ? reverse2(codes16(T, (int32_t)n - 4 - CB * t)) >> 24 : codes16(T, CB * t) & 0xFFu;
    };
    auto step = [&](uint32_t fresh, int32_t t, auto guarded) {
        constexpr bool GD = decltype(guarded)::value;
        codes = (uint32_t)__builtin_amdgcn_update_dpp((int32_t)codes, (int32_t)codes, SHR, 0xf, 0xf, false);
        codes = l == 0 ? fresh : codes;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp((int32_t)hin[c], (int32_t)hp[c], SHR, 0xf, 0xf, false);
            const uint32_t un = (uint32_t)__builtin_amdgcn_update_dpp((int32_t)nin[c], (int32_t)hn[c], SHR, 0xf, 0xf, false);
            hin[c] = (G == 32 && l == 0) ? hin[c] : up;
            nin[c] = (G == 32 && l == 0) ? nin[c] : un;
        }
        const int32_t j0 = CB * (t - l);
        if constexpr (ST) {
            // Lane 0: the step's four columns' deltas from the stripe above (one aligned word).
            if (hcin && l == 0 && j0 < (int32_t)n) {
                const uint32_t v = *reinterpret_cast<const uint32_t *>(hcin + j0);
#pragma unroll
                for (int c = 0; c < CB; ++c) {
                    hin[c] = (v >> (8 * c)) & 1u;
                    nin[c] = (v >> (8 * c + 1)) & 1u;
                }
            }
        }
        if (!GD || (work && j0 >= 0 && j0 < (int32_t)n)) {
            // Match masks: all-ones where the column's code bit is 0 (block_step's M1/M2).
            const uint32_t nc = ~codes;
            uint32_t M1[CB], M2[CB], xp[CB], xn[CB];
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                M1[c] = (uint32_t)__builtin_amdgcn_sbfe((int32_t)nc, 2 * c, 1);
                M2[c] = (uint32_t)__builtin_amdgcn_sbfe((int32_t)nc, 2 * c + 1, 1);
                xp[c] = hin[c];
                xn[c] = nin[c];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (l * R + r >= off) {
#pragma unroll
                    for (int c = 0; c < CB; ++c) {
                        if (r < R - 1) block_step<true>(K[r], M1[c], M2[c], xp[c], xn[c], 63u);
                        else block_step<false>(K[r], M1[c], M2[c], xp[c], xn[c], bitR);
                    }
                }
            }
            if constexpr (ST) {
                // The deltas out of the stripe's last block, for the stripe below.
                if (hcout && l == G - 1 && j0 >= 0 && j0 < (int32_t)n) {
                    uint32_t v = 0;
#pragma unroll
                    for (int c = 0; c < CB; ++c) v |= (xp[c] | (xn[c] << 1)) << (8 * c);
                    *reinterpret_cast<uint32_t *>(hcout + j0) = v; // (the row holds n rounded up to 4)
                }
            }
            // The score row (lane G - 1's last slot): the pattern's last row.
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                hp[c] = xp[c];
                hn[c] = xn[c];
                const int32_t nsc = score + (int32_t)xp[c] - (int32_t)xn[c];
                const bool in = !GD || j0 + c < (int32_t)n;
                score = in ? nsc : score;
                const bool lower = nsc < best;
                bestj = (in && lower) ? (uint32_t)(CB * t + c + 1) : bestj;
                best = in ? imin(best, nsc) : best;
            }
        }
    };
    std::integral_constant<bool, true> guarded;
    std::integral_constant<bool, false> bare;
    int32_t t = 0;
    for (; t < fill; ++t) step(fetch4(t), t, guarded);
    for (; t < steady; t += SW) {
        // 16 codes: the next SW steps' columns (lane 0's; the other lanes load the same word).
        const uint32_t tw = !work ? 0u : (rev ? reverse2(codes16(T, (int32_t)n - 16 - CB * t)) : codes16(T, CB * t));
#pragma unroll
        for (int u = 0; u < SW; ++u) step((tw >> (2 * CB * u)) & 0xFFu, t + u, bare);
    }
    for (; t < steps; ++t) step(fetch4(t), t, guarded);
    if (valid && !rev && m == 0) {
        const uint32_t d = mode == kGlobal ? tlen : 0;
        const bool ok = d <= a.thresh(pi);
        if (l == 0) {
            a.dist[pi] = ok ? d : kNone;
            if (a.end) a.end[pi] = ok ? (mode == kGlobal ? tlen : 0) : kNone;
            if (a.start) a.start[pi] = ok ? 0 : kNone;
        }
        return;
    }
    if (!work || l != G - 1) return;
    if constexpr (ST) {
        if (!st_last) return;
    }
    if (bestj) bestj -= CB * l;
    if (!rev) {
        if (mode == kGlobal) {
            best = score;
            bestj = n;
        }
        const bool ok = (uint32_t)best <= a.thresh(pi);
        a.dist[pi] = ok ? (uint32_t)best : kNone;
        if (a.end) a.end[pi] = ok ? bestj : kNone;
        if (a.start && (mode != kInfix || !ok)) a.start[pi] = ok ? 0 : kNone;
    } else {
        a.start[pi] = n - bestj;
    }
}

} // namespace
