// internal.hpp -- shared between the kernel translation units and the runtime of libsf_align.so.
//
// Device layout of a batch (HBM):
//   seq   : all sequences 2-bit packed, 16 bases per u32 (base k at bits 2*(k%16) of word k/16),
//           each sequence starting on a word boundary when the batch is made.
//   poff/toff (u64) : base offset of pair i's pattern / text in `seq`;  plen/tlen (u32): lengths.
//                     (Base offsets let a pair's text be restricted to any window in place.)
//   order (u32)     : pair indices grouped by kernel shape class ("bins"); a launch covers one bin.
// Loads past a sequence's last word are clamped to it, so no padding words are needed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sf {

constexpr int32_t kNeg = INT32_MIN / 4;       // align.rs:693
// Largest scoring value (mat, mis, open, ext): the 32-bit kernels take the mismatch step
// -(mat + mis) as a signed 24-bit multiplier; per pair, sf::scores_in_range bounds the scores.
constexpr int32_t kScoreMax = (1 << 22) - 1;
constexpr int32_t kScanIdent = INT32_MIN / 2; // below every value a cell or gap can take
constexpr uint32_t kNone = 0xffffffffu;
constexpr int kInfix = 0, kPrefix = 1, kGlobal = 2;

struct PairsView {
    const uint32_t *seq;
    const uint64_t *poff;
    const uint64_t *toff;
    const uint32_t *plen;
    const uint32_t *tlen;
};

// One pair as a launch reads it, stored in launch order so that a launch reads its pairs' sizes
// and offsets sequentially (32 bytes a pair) rather than gathering them by pair index.
struct PairDesc {
    uint64_t poff, toff; // base offsets of the pattern and the text
    uint32_t m, n;       // their lengths
    uint32_t below, above; // affine band geometry (0 for Myers launches)
};

struct MyersArgs {
    PairsView pv;
    const PairDesc *desc; // per order slot of the launch
    const uint32_t *order;
    uint32_t count;
    int mode;
    uint32_t k;
    const uint32_t *kvec; // a threshold per pair (by pair index) in place of k, or null
    int reverse; // 0: forward pass (dist, end); 1: reversed Prefix pass for `start`
    uint32_t *dist;
    uint32_t *end;
    uint32_t *start;
    // Patterns longer than the largest class (kMyersStriped): launch `stripe` runs blocks
    // [stripe * 1536, (stripe + 1) * 1536) of each pair's pattern; a column's horizontal delta into
    // the stripe's first block comes from hcol_in (the stripe above's), the delta out of its last
    // block goes to hcol_out, one byte a column (bit 0: +1, bit 1: -1) at launch slot * hcol_stride.
    uint32_t stripe;
    const uint8_t *hcol_in;
    uint8_t *hcol_out;
    uint64_t hcol_stride;
    __device__ uint32_t thresh(uint32_t pi) const { return kvec ? kvec[pi] : k; }
};

// The start pass of patterns over 1,536 bases, stripe by stripe over the diagonal band
// (myers_start.hip), over launch order slots [first, first + count): row u of each stripe buffer
// belongs to slot first + u and holds a stripe's bottom-row deltas, 2 bits a column, hc_words a
// row (ping-pong between hc0 and hc1).
struct MyersStartArgs {
    PairsView pv;
    const PairDesc *desc;
    const uint32_t *order;
    uint32_t first, count;
    const uint32_t *dist, *end;
    uint32_t *start;
    uint32_t *hc0, *hc1;
    uint64_t hc_words;
};

struct AffineArgs {
    PairsView pv;
    const PairDesc *desc; // per order slot of the launch
    const uint32_t *order;
    uint32_t count;
    int mode;
    int32_t mat, mis, open, ext;
    int32_t *score;
    uint32_t *score_col;
    uint32_t *tb;           // traceback bits (chunk base); null for score only
    const uint64_t *tb_off; // per order slot of the launch: word offset into tb
};

// One launch over several 64-lane score classes: segment i holds count[i] pairs of class cls[i]
// at launch-order offset off[i] and starts at workgroup blk0[i] (4 pairs a workgroup); blocks is
// the launch's total.
//
// queue (null: one workgroup a block index, blocks workgroups): two device words, a cursor and a
// count of workgroups done, both zero before the launch.  The launch is then persistent: as many
// workgroups as the chip holds at once, each taking the next block index from the cursor (segments
// in the launch's order: the longest chains first) until none is left; the last workgroup to finish
// sets both words back to zero, so the next launch (or graph replay) finds them ready.
struct MixArgs {
    static constexpr uint32_t kMax = 48;
    uint32_t nseg, blocks;
    int32_t cls[kMax];
    uint32_t off[kMax], count[kMax], blk0[kMax];
    uint32_t *queue;
};

// The block indices of a mixed launch, as MixArgs::queue describes:
//     for (uint32_t b = mix_next(q, true); b < mx.blocks; b = mix_next(q, false)) body(b);
//     mix_done(q);
// Without a queue that is the workgroup's own index, once.  With one, each workgroup takes indices
// from the cursor until they run out; the bodies hold no workgroup barrier, so lanes and waves that
// end a block early simply wait at the next index's barrier.
__device__ __forceinline__ uint32_t mix_next(uint32_t *queue, bool first) {
    if (!queue) return first ? blockIdx.x : 0xffffffffu;
    __shared__ uint32_t next;
    if (threadIdx.x == 0) next = atomicAdd(queue, 1u);
    __syncthreads();
    // (The same in every lane: readfirstlane lets the compiler keep it, and all it indexes, scalar.)
    const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)next);
    __syncthreads(); // every thread has read `next` before thread 0 takes the next index
    return b;
}

// Every workgroup has taken its last index before it counts itself done, so the last one to count
// can reset the cursor: no index is taken after it.
__device__ __forceinline__ void mix_done(uint32_t *queue) {
    if (queue && threadIdx.x == 0 && atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
        atomicExch(&queue[0], 0u);
        atomicExch(&queue[1], 0u);
    }
}

// A pass with per-row bookkeeping: align.rs:2446-2666 run_core with clip (X-drop, best cell) and
// cut tracking, or (local) align.rs:783-852 local_band_core.  Band geometry comes in the pair
// records (local: below = -dlo, above = dhi, as two's complement).
struct DpOut { // align.rs:2111-2129 `Dp` (the usize fields fit 32 bits: sequences are < 2^30)
    int32_t score;
    uint32_t score_col;
    int32_t clip_best;
    uint32_t clip_i, clip_j;
    uint32_t text_read, last_row;
    int32_t cut_best, cut_score;
};

struct PassArgs {
    PairsView pv;
    const PairDesc *desc;
    const uint32_t *order;
    uint32_t count;
    int mode;
    int32_t mat, mis, open, ext;
    int has_clip, clip_cell;
    int32_t xdrop;
    const uint64_t *cut;     // per pair (null: no cut)
    DpOut *out;              // per pair (run_core passes)
    int32_t bonus0, bonus1;  // local: end bonuses for the read's start and end
    const uint32_t *pinv;    // local: bit (base offset) set where a pattern code is >= 4 (null: none)
    int32_t *best;           // per pair (local passes)
};

// A wide-band pass (affine_wide.hip): the pass arguments (run_core bookkeeping into p.out, or local
// passes into p.best), or with p.out and p.best null a score run into score / score_col; rows:
// scratch of 4 * stride int32 per pair of the launch (H and F of the previous and the current row,
// each at least the widest band plus one slot).
struct WideArgs {
    PassArgs p;
    int local;
    int32_t *score;
    uint32_t *score_col;
    int32_t *rows;
    uint64_t stride;
};

struct WalkArgs {
    PairsView pv;
    const PairDesc *desc; // per order slot of the launch
    const uint32_t *order;
    uint32_t count;
    int mode;
    int K, LP; // the DP's shape: K slots per lane, LP lanes per pair
    const int32_t *score;
    const uint32_t *score_col;
    const uint32_t *tb;
    const uint64_t *tb_off;  // per order slot
    uint32_t *runs;          // reversed runs per pair, at runs_off[pair]
    const uint64_t *runs_off; // per pair
    uint32_t *nruns;          // per pair
    uint32_t *tstart;         // per pair
};

// ---- device helpers ----

// One sequence inside the packed words: code k of the sequence is code (base + k) of `seq`.
// Word loads are clamped to the sequence's own words, so reads just outside it stay in bounds
// (their codes are unspecified) and nothing needs padding.  Positions are 32-bit (sequences are
// under 2^30 bases); only the first word's address is 64-bit.
struct SeqRef {
    const uint32_t *w; // first word holding the sequence (word 0 of `seq` for an empty one)
    int32_t lead;      // position of code 0 inside that word
    int32_t last;      // index of the last word holding the sequence, relative to `w`
};

__device__ __forceinline__ SeqRef seqref(const uint32_t *seq, uint64_t base, uint32_t len) {
    SeqRef r;
    r.w = len ? seq + (base >> 4) : seq;
    r.lead = len ? (int32_t)(base & 15) : 0;
    r.last = len ? (int32_t)(((base & 15) + len - 1) >> 4) : 0;
    return r;
}

__device__ __forceinline__ uint32_t ldw(const SeqRef &r, int32_t i) {
    return r.w[i < 0 ? 0 : (i > r.last ? r.last : i)];
}

// 16 codes starting at sequence position s (any integer): code (s + k) at bits 2k.  Positions
// outside the sequence give unspecified codes.
__device__ __forceinline__ uint32_t codes16(const SeqRef &r, int32_t s) {
    const int32_t abs = r.lead + s;
    const int32_t i = abs >> 4; // floor for negative positions
    const uint32_t a = ldw(r, i), b = ldw(r, i + 1);
    return __builtin_amdgcn_alignbit(b, a, (uint32_t)(abs & 15) * 2);
}

__device__ __forceinline__ uint32_t code_at(const SeqRef &r, int32_t k) {
    const int32_t abs = r.lead + k;
    return (ldw(r, abs >> 4) >> (2 * (abs & 15))) & 3u;
}

// Reverses the order of the 16 two-bit codes of a word.
__device__ __forceinline__ uint32_t reverse2(uint32_t x) {
    x = __builtin_bitreverse32(x);
    return ((x >> 1) & 0x55555555u) | ((x << 1) & 0xAAAAAAAAu);
}

// The 16 even bits of x, compacted to the low half.
__device__ __forceinline__ uint32_t even16(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

__device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
__device__ __forceinline__ int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

// DPP lane moves (gfx9 encodings).  Lanes without a source keep `fill`.
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ int32_t dpp(int32_t fill, int32_t v) {
    return __builtin_amdgcn_update_dpp(fill, v, CTRL, ROWMASK, 0xf, false);
}
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118, kRowShl1 = 0x101;
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

// A DPP move whose lanes without a source read 0 (bound_ctrl): the minimum of unsigned values.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp0(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}

__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// Exclusive max-scan of unsigned values across the 64 lanes of a wave, lane 0 getting 0.  DPP's
// zero fill of lanes without a source is the identity, so each step is one v_max_u32 with a DPP
// operand.
__device__ __forceinline__ uint32_t wave_excl_umax(uint32_t x) {
    x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr1, 0xf, 0xf, true));
    x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr2, 0xf, 0xf, true));
    x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr4, 0xf, 0xf, true));
    x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr8, 0xf, 0xf, true));
    x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast15, 0xa, 0xf, true));
    x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast31, 0xc, 0xf, true));
    return dpp0<kWaveShr1>(x);
}

// Traceback words per lane per row: 4 bits per slot.
template <int K>
constexpr int kTbWords = (4 * K + 31) / 32;

// Exclusive max-scan of unsigned values over each group of LP lanes (16: one DPP row; 32: two;
// 64: the wave); the group's first lane gets 0.
template <int LP>
__device__ __forceinline__ uint32_t group_excl_umax(uint32_t x) {
    if constexpr (LP == 64) {
        return wave_excl_umax(x);
    } else if constexpr (LP == 4) {
        // Four groups share a DPP row: the input moves up a lane inside its group (the group's
        // first lane takes 0), then is scanned inside the quad (quad permutes [0,0,1,2] and
        // [0,1,0,1]; a lane reading itself is harmless under max).
        uint32_t y = dpp0<kRowShr1>(x);
        y = (threadIdx.x & 3) ? y : 0u;
        y = umax(y, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)y, 0x90, 0xf, 0xf, false));
        return umax(y, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)y, 0x44, 0xf, 0xf, false));
    } else if constexpr (LP == 8) {
        // Two groups share a DPP row.  The input moves up a lane inside its group (the group's
        // first lane takes 0), is scanned inside quads (quad permutes [0,0,1,2] and [0,1,0,1]: a
        // lane reading itself is harmless under max), and the upper quad of each group takes the
        // lower quad's total by a row shift of 4 written to banks 1 and 3 only.
        uint32_t y = dpp0<kRowShr1>(x);
        y = (threadIdx.x & 7) ? y : 0u;
        y = umax(y, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)y, 0x90, 0xf, 0xf, false));
        y = umax(y, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)y, 0x44, 0xf, 0xf, false));
        const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)y, 0xff, 0xf, 0xf, false);
        return umax(y, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)t, kRowShr4, 0xf, 0xa, false));
    } else if constexpr (LP == 32) {
        // Two DPP rows: the rows' inclusive scans, then rows 1 and 3 take rows 0 and 2's totals
        // (row_bcast:15 written to those rows only), then the exclusive shift, whose move across
        // lanes 31 -> 32 (the next group) the group's first lane replaces by 0.
        x = umax(x, dpp0<kRowShr1>(x));
        x = umax(x, dpp0<kRowShr2>(x));
        x = umax(x, dpp0<kRowShr4>(x));
        x = umax(x, dpp0<kRowShr8>(x));
        x = umax(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast15, 0xa, 0xf, true));
        const uint32_t y = dpp0<kWaveShr1>(x);
        return (threadIdx.x & 31) ? y : 0u;
    } else {
        static_assert(LP == 16, "groups of 8, 16, 32 or 64 lanes");
        x = umax(x, dpp0<kRowShr1>(x));
        x = umax(x, dpp0<kRowShr2>(x));
        x = umax(x, dpp0<kRowShr4>(x));
        x = umax(x, dpp0<kRowShr8>(x));
        return dpp0<kRowShr1>(x);
    }
}


// Exclusive max-scan across the 64 lanes of a wave, lane 0 getting INT32_MIN.  Signed values are
// mapped to unsigned order (x ^ 0x80000000) so that DPP's zero fill of lanes without a source is
// the identity and each step is one v_max_u32 with a DPP operand.
__device__ __forceinline__ int32_t wave_excl_max_min(int32_t v) {
    uint32_t x = (uint32_t)v ^ 0x80000000u;
    auto step = [](uint32_t y, uint32_t z) { return y > z ? y : z; };
    x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr1, 0xf, 0xf, true));
    x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr2, 0xf, 0xf, true));
    x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr4, 0xf, 0xf, true));
    x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowShr8, 0xf, 0xf, true));
    x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast15, 0xa, 0xf, true));
    x = step(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kRowBcast31, 0xc, 0xf, true));
    x = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kWaveShr1, 0xf, 0xf, true);
    return (int32_t)(x ^ 0x80000000u);
}

// Exclusive max-scan across the 64 lanes of a wave (lane 0 gets `ident`).
__device__ __forceinline__ int32_t wave_excl_max(int32_t v, int32_t ident) {
    int32_t x = v;
    x = imax(x, dpp<kRowShr1>(ident, x));
    x = imax(x, dpp<kRowShr2>(ident, x));
    x = imax(x, dpp<kRowShr4>(ident, x));
    x = imax(x, dpp<kRowShr8>(ident, x));
    x = imax(x, dpp<kRowBcast15, 0xa>(ident, x));
    x = imax(x, dpp<kRowBcast31, 0xc>(ident, x));
    return dpp<kWaveShr1>(ident, x);
}

} // namespace sf

// Affine shape classes: K slots per lane and LP lanes per pair, coded K for LP = 64 and
// 1000 * LP + K otherwise, plus 100 for the 16-bit packed kernel (scores only).  kAffWide is the
// kernel for bands over 2,048 cells a row (affine_wide.hip: one pair per 256-thread workgroup, the
// rows in scratch, any width up to kAffWideMax), counted as 256 lanes of 8 slots.
constexpr int kAffWide = 60000;
constexpr uint64_t kAffWideMax = 1ull << 26;
constexpr int aff_lp(int cls) { return cls == kAffWide ? 256 : (cls >= 1000 ? cls / 1000 : 64); }
constexpr int aff_k(int cls) { return cls == kAffWide ? 8 : (cls >= 1000 ? cls % 100 : cls); }
constexpr bool aff_packed(int cls) { return cls >= 1000 && cls % 1000 >= 100; }
constexpr uint64_t aff_slots(int cls) { return (uint64_t)aff_lp(cls) * (uint64_t)aff_k(cls); }
// Traceback words per row of one pair: 4 bits per slot, a lane's slots in ceil(4K / 32) words.
constexpr uint64_t aff_tb_words_per_row(int cls) { return (uint64_t)aff_lp(cls) * ((4 * aff_k(cls) + 31) / 32); }

// Launchers (defined in the kernel TUs).  Each returns hipSuccess or the launch error.
hipError_t sf_launch_myers(const sf::MyersArgs &a, int maxb, hipStream_t s);
hipError_t sf_launch_affine(const sf::AffineArgs &a, int K, bool tb, hipStream_t s);
uint64_t sf_affine_lp16_max(); // the current 16-lane limit (slots), part of the shape choice
int sf_affine_class(uint64_t width, bool lp16);      // 16 lanes per pair if allowed and it fits, else 64
int sf_affine_pass_class(uint64_t width, bool lp16); // the same for the passes with bookkeeping
// The 16-bit packed pass class on `lanes` lanes (8 or 16; lanes * 1000 + 100 + slots a lane, + 200
// with the three-term max) for a band of `width` slots, or 0 (mis != 2 ext, wider than the shape's
// 32 slots a lane, or past the range).
int sf_affine_pass_pk_class(uint64_t width, int lanes, int32_t mat, int32_t mis, int32_t open, int32_t ext,
                            uint32_t m, uint32_t below);
// Passes with bookkeeping: 0 the 32-bit kernel throughout; 1 (default) the packed kernel where its
// range holds, on 8 lanes where the band fits 256 slots, else 16; 2 the packed kernel on 16 lanes.
int sf_affine_pass_packed();
void sf_affine_set_pass_packed(int policy);
// The 16-bit packed class for a pair of the 32-bit class with `lanes` (16 or 64) lanes and K slots
// per lane, or 0 where its exact range does not hold (the scoring must have mis = 2 ext); tb: for
// the traceback kernel.
int sf_affine_pk16_class(int K, int32_t mat, int32_t mis, int32_t open, int32_t ext, uint32_t m, uint32_t below,
                         bool tb, int lanes);
// When score-only runs use the 4- and 8-lane packed kernels: 0 by estimate, 1 always 8 lanes (never
// 4), 2 always 4 lanes first, -1 neither.
int sf_affine_lp8_policy();
void sf_affine_set_lp8_policy(int policy);
// The 4-lane packed score class (4200 + slots a lane, 36..72) for a band of `width` slots, or 0.
int sf_affine_pk4_class(uint64_t width, int32_t mat, int32_t mis, int32_t open, int32_t ext, uint32_t m,
                        uint32_t below);
// The 8-lane packed class (scores, three-term max) for a band of `width` slots, or 0.
int sf_affine_pk8_class(uint64_t width, int32_t mat, int32_t mis, int32_t open, int32_t ext, uint32_t m,
                        uint32_t below);
// The packed classes (aff_packed): the 16-bit score and traceback kernels.
hipError_t sf_launch_affine_packed(const sf::AffineArgs &a, int cls, bool tb, hipStream_t s);
hipError_t sf_launch_affine_packed64(const sf::AffineArgs &a, int cls, bool tb, hipStream_t s);
hipError_t sf_launch_affine_packed4(const sf::AffineArgs &a, int cls, bool tb, hipStream_t s);
// Score-only 64-lane classes that can share one launch, and that launch.
bool sf_affine_mixable(int cls);
hipError_t sf_launch_affine_mix(const sf::AffineArgs &a, const sf::MixArgs &mx, hipStream_t s);
// Score classes sharing one launch by register group (affine_mix.hip, affine_mix1.hip): the group
// of a class (1-2; 0 for a class with launches of its own); workgroups a class takes for `count`
// pairs; the launch of one group's segments.
int sf_affine_mix_group(int cls);
uint32_t sf_affine_blocks_for(int cls, uint32_t count);
hipError_t sf_launch_affine_mix_group(const sf::AffineArgs &a, const sf::MixArgs &mx, int group, hipStream_t s);
hipError_t sf_launch_affine_mix1(const sf::AffineArgs &a, const sf::MixArgs &mx, hipStream_t s);
// A small serving batch's Myers and affine class runs in one launch (serve_mix.hip): the classes
// it holds, and the launch (affine segments first, offsets into aa.order / ma.order).
bool sf_serve_mix_affine(int cls);
bool sf_serve_mix_myers(int cls);
hipError_t sf_launch_serve_mix(const sf::AffineArgs &aa, const sf::MixArgs &amx, const sf::MyersArgs &ma,
                               const sf::MixArgs &mmx, hipStream_t s);
// The runtime loads a file's code object when one of its kernels is first used, which holds up the
// launching thread for milliseconds.  Each of these loads one file's object (a kernel's attributes
// need it); sf_load_code_objects calls them all, for callers that must not pause later.
hipError_t sf_load_code_myers();
hipError_t sf_load_code_myers_mix();
hipError_t sf_load_code_affine();
hipError_t sf_load_code_affine_pk();
hipError_t sf_load_code_affine_pk64();
hipError_t sf_load_code_affine_mix();
hipError_t sf_load_code_affine_mix1();
hipError_t sf_load_code_serve_mix();
hipError_t sf_load_code_objects();
// Myers classes sharing one launch (myers_mix.hip): the register group of a class (1-3; 0 for a
// class with launches of its own); workgroups a class takes for `count` pairs; the launch of one
// group's segments (as for sf_launch_affine_mix: offsets into a.order / a.desc).
int sf_myers_mix_group(int cls);
uint32_t sf_myers_blocks_for(int cls, uint32_t count);
hipError_t sf_launch_myers_mix(const sf::MyersArgs &a, const sf::MixArgs &mx, int group, hipStream_t s);
hipError_t sf_launch_walk(const sf::WalkArgs &a, hipStream_t s);
hipError_t sf_launch_affine_wide(const sf::WideArgs &a, hipStream_t s);
// cls: a 32-bit affine class (16 or 64 lanes); local selects local_band_core.
hipError_t sf_launch_pass(const sf::PassArgs &a, int cls, bool local, hipStream_t s);
hipError_t sf_launch_pass_packed(const sf::PassArgs &a, int cls, hipStream_t s); // affine_pk_pass.hip
// desc[w] of pair order[w], w < n (ba may be null: no band).
hipError_t sf_launch_desc(uint64_t n, const uint32_t *order, const sf::PairsView &pv, const uint64_t *ba,
                          sf::PairDesc *desc, hipStream_t s);
hipError_t sf_launch_compact(const uint32_t *order, uint32_t count, const uint32_t *runs, const uint64_t *runs_off,
                             const uint32_t *nruns, const uint64_t *out_off, uint32_t *out, hipStream_t s);

// Workgroups of 256 threads of `kernel` the device holds at once (its occupancy x compute units),
// at most `blocks`: the grid of a persistent mixed launch (cached per kernel).
uint32_t sf_persistent_grid(const void *kernel, uint32_t blocks);
// Which mixed launches run persistent (MixArgs::queue; bit 0 Myers runs, bit 1 score runs):
// sf_set_persistent / SF_PERSISTENT.
int sf_persistent_policy();
// The persistent forms of the mixed launches (mix_persistent_*.hip; mx.queue set, segments checked
// by the callers above).
hipError_t sf_launch_myers_mix_persistent(const sf::MyersArgs &a, const sf::MixArgs &mx, int group, hipStream_t s);
hipError_t sf_launch_affine_mix_persistent(const sf::AffineArgs &a, const sf::MixArgs &mx, hipStream_t s);
hipError_t sf_launch_affine_mix1_persistent(const sf::AffineArgs &a, const sf::MixArgs &mx, hipStream_t s);

// A batch (sf_batch, include/sf_align.h) over sequences already on the device in the 2-bit word
// layout: pair i's pattern at base pbase[i] of `seq`, its text at tbase[i].  The batch does not own
// `seq`: it must stay allocated and unchanged until the batch is destroyed.  Lengths are under 2^30.
struct sf_batch;
int sf_batch_create_resident(const uint32_t *seq, uint64_t n, const uint32_t *plen, const uint32_t *tlen,
                             const uint64_t *pbase, const uint64_t *tbase, sf_batch **out);

// Supported shape classes.
int sf_myers_maxb_for(uint32_t nblocks);     // 0 if unsupported
// Patterns over 1,536 blocks (98,304 bases): the 64-lane, 24-block pipeline run stripe by stripe
// (myers_pipe_body<64, 24, true>), up to kMyersStripedMax blocks.
constexpr int kMyersStriped = 99000;
constexpr uint32_t kMyersStripedBlocks = 64 * 24;
constexpr uint32_t kMyersStripedMax = 1u << 24;
hipError_t sf_launch_myers_striped(const sf::MyersArgs &a, hipStream_t s);
// Lane classes (2..24 blocks) over the wave's Ukkonen band (myers_lane_ukk_kernel): for a finite k
// below the patterns' lengths, and for the reversed pass, whose threshold is the distance.
bool sf_myers_ukk_class(int cls);
hipError_t sf_launch_myers_start_stripe(const sf::MyersStartArgs &a, hipStream_t s);
hipError_t sf_launch_myers_ukk(const sf::MyersArgs &a, int maxb, hipStream_t s);
// The 64-lane pipeline class with the shorter chain that holds a 32-lane class's patterns.
int sf_myers_wide_class(int cls);
int sf_affine_k_for(uint64_t width);         // 0 if unsupported

namespace sf {

// Exact-equivalent band geometry of a pair (m, n): (below, above) of align.rs:2468-2481, with any
// band that keeps every row at [0, n] (Infix, usize::MAX, or band >= max(m, n)) as (m, n).
inline void band_geom(uint32_t m, uint32_t n, int mode, uint64_t band, uint64_t *below, uint64_t *above) {
    const uint64_t mx = m > n ? m : n;
    if (mode == kInfix || band == UINT64_MAX || band >= mx) {
        *below = m;
        *above = n;
    } else {
        *below = band + (m > n ? m - n : 0);
        *above = band + (n > m ? n - m : 0);
    }
}

// The band of a pass laid out for a text of `geom` bases (align.rs:2464-2474): below = band +
// (m - geom)+, above = band + (geom - m)+; the whole matrix for Infix, band = usize::MAX, or a band
// that covers every row anyway.
inline void band_geom_for(uint32_t m, uint32_t n, uint64_t geom, int mode, uint64_t band, uint64_t *below,
                          uint64_t *above) {
    if (mode == kInfix || band >= (uint64_t)m + n) { // (band = usize::MAX included)
        *below = m;
        *above = n;
        return;
    }
    const uint64_t bl = band + (m > geom ? m - geom : 0), ab = band + (geom > m ? geom - m : 0);
    if (bl >= m && ab >= n) {
        *below = m;
        *above = n;
    } else {
        *below = bl;
        *above = ab;
    }
}

// Cells of a pass as the reference's DP_WORK counts them (align.rs:2573-2578): the sum over rows
// r = 1..R, R = min(m, n + below), of min(n, r + above) - max(0, r - below) + 1, in closed form.
inline uint64_t band_cells(uint32_t m, uint32_t n, uint64_t below, uint64_t above) {
    const uint64_t R = m < (uint64_t)n + below ? m : (uint64_t)n + below;
    const uint64_t c = n > above ? ((uint64_t)n - above < R ? (uint64_t)n - above : R) : 0; // rows with r + above < n
    const uint64_t d = R > below ? R - below : 0;                                           // rows with r > below
    return R + c * (c + 1) / 2 + c * above + (R - c) * n - d * (d + 1) / 2;
}

// Bound on the biased values Hb = H + ext (s + 2 i) over `width` slots: |H| <= (mat + mis + ext)
// (m + n + 2) + 2 open, and the bias adds at most ext (width + 2 m).  Under 2^28 the band stays
// far above NEG = -2^29 and slots right of the band, lowered by 2^30, stay below it.
inline bool scores_in_range(int32_t mat, int32_t mis, int32_t open, int32_t ext, uint64_t m, uint64_t n,
                            uint64_t width) {
    const uint64_t bound = (uint64_t)(mat + mis + ext) * (m + n + 2) + 2ull * (uint64_t)open +
                           (uint64_t)ext * (width + 2 * m);
    return bound < (1ull << 28);
}

} // namespace sf
