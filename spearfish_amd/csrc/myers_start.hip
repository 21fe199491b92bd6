// myers_start.hip -- the reversed Prefix pass that finds `start` for patterns over 1,536 bases,
// stripe by stripe over Ukkonen's diagonal band.
//
// The pass (the edlib convention, as the plain kernels run it: the pattern reversed against the
// text up to the forward pass's end column, reversed, in Prefix mode) reaches exactly the pair's
// distance d, and `start` is the text length minus its first column reaching it.  In Prefix mode
// D[i][j] >= |i - j|, so every cell at or below d lies in the diagonal band |i - j| <= d.
//
// Stripe s of a pattern is rows (1024 s, 1024 s + rows] of the reversed pattern (16 blocks of 64
// rows, fewer in the last stripe), run by one lane (the lane kernel's block step and bit-planes)
// over the columns of the band that meet those rows:
// jlo = max(1, 1024 s + 1 - d) .. jhi = min(n, 1024 s + rows + d).  The cells outside that window
// score above d, so they are replaced by upper bounds: the stripe starts at column jlo - 1 with
// every vertical delta +1 below the row above (D[i][j] <= D[i-1][j] + 1), and takes +1 as the
// horizontal delta into its first row where the stripe above has no column (D[i][j] <= D[i][j-1]
// + 1).  Upper bounds in give upper bounds out, and a cell at or below d has its optimal
// predecessor at or below d, inside the windows, so every cell at or below d is exact.  The last
// stripe's bottom row is the pattern's last row: its first column at d is `start`'s.
//
// Between stripes: the horizontal deltas out of a stripe's bottom row, 2 bits a column (+1, -1),
// 16 columns a word at their absolute column (a stripe reads exactly the window the stripe above
// wrote), and the absolute score of that row at the next stripe's column jlo - 1 (the anchor of the
// next stripe's scores).  A window is 1,024 + 2d columns against the whole text for the pipeline;
// the kernel holds 16 blocks of state a lane (two waves a SIMD).
#include "myers_body.hpp"

using namespace sf;

namespace {

constexpr int kStartBlocks = 16; // blocks a stripe
constexpr int64_t kStartRows = 64 * kStartBlocks;

// 16 two-bit entries from column c (0-based) of a row of `words` words: entry c + k at bits 2k.
// Entries outside the row's words are unspecified.
__device__ __forceinline__ uint32_t bits16(const uint32_t *row, uint64_t words, int64_t c) {
    const int64_t w = c >> 4;
    const int64_t last = (int64_t)words - 1;
    const uint32_t a = row[w < 0 ? 0 : (w > last ? last : w)];
    const uint32_t b = row[w + 1 < 0 ? 0 : (w + 1 > last ? last : w + 1)];
    return __builtin_amdgcn_alignbit(b, a, (uint32_t)(c & 15) * 2);
}

// One lane a pair, its stripes one after another (each stripe reads the bottom-row deltas the lane
// wrote for the stripe above, and keeps that row's anchor in a register), the longest patterns'
// workgroups first so that their longer chains start early.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void myers_start_stripe_kernel(MyersStartArgs a) {
    constexpr int MAXB = kStartBlocks;
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const bool valid = t < a.count;
    const uint32_t u = valid ? a.count - 1 - t : 0; // row of the stripe buffers: the longest first
    const uint32_t slot = a.first + u;              // launch order slot
    const uint32_t pi = valid ? a.order[slot] : 0;
    const PairDesc dsc = valid ? a.desc[slot] : PairDesc{0, 0, 0, 0, 0, 0};
    const int32_t m = (int32_t)dsc.m;
    const uint32_t d = valid ? a.dist[pi] : kNone;
    const bool work = valid && m > 0 && d != kNone;
    // Columns and rows fit 31 bits (texts and patterns of the batch are below 2^31 bases).
    const int32_t np = work ? (int32_t)a.end[pi] : 0; // the reversed text's length
    const int32_t dd = work ? (int32_t)d : 0;
    const int32_t stripes = work ? (m + kStartRows - 1) / kStartRows : 0;
    int32_t smax = stripes;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) smax = imax(smax, __shfl_xor(smax, off));
    smax = __builtin_amdgcn_readfirstlane(smax);
    const SeqRef P = seqref(a.pv.seq, dsc.poff, (uint32_t)m);
    const SeqRef T = seqref(a.pv.seq, dsc.toff, dsc.n);
    const uint64_t row_off = (uint64_t)u * a.hc_words;
    int32_t anchor = 0; // the stripe above's bottom-row score at this stripe's jlo - 1
    int32_t best = INT32_MAX, bestj = 0;
    for (int32_t st = 0; st < smax; ++st) {
        const bool on = st < stripes; // this lane's pattern has stripe st
        const int32_t r0 = kStartRows * st;
        const int32_t rows = on && m - r0 < kStartRows ? m - r0 : kStartRows;
        const bool last = on && st == stripes - 1;
        const int nb = on ? (rows + 63) / 64 : MAXB;
        const int lastb = nb - 1;
        const unsigned last_bit = (unsigned)((rows - 1) & 63);
        // The window [jlo, jhi] (1-based columns), as counts from jlo: the columns the stripe
        // above wrote (k <= kin; it starts at or before jlo), and the next stripe's anchor
        // column (k == kan; only when that column is at least 1).
        const int32_t jlo = imax(1, r0 + 1 - dd);
        const int32_t jhi = imin(np, r0 + rows + dd);
        const int32_t cols = on ? imax(0, jhi - jlo + 1) : 0;
        const int32_t kin = imin(np, r0 + dd) - jlo;
        const int32_t kan = r0 + kStartRows - dd - jlo;
        const bool has_in = on && st > 0;
        const uint32_t *hin = ((st & 1) ? a.hc1 : a.hc0) + row_off;
        uint32_t *hout = ((st & 1) ? a.hc0 : a.hc1) + row_off;
        int nb_lo = nb, cmax = cols;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            nb_lo = imin(nb_lo, __shfl_xor(nb_lo, off));
            cmax = imax(cmax, __shfl_xor(cmax, off));
        }
        nb_lo = __builtin_amdgcn_readfirstlane(nb_lo);
        cmax = __builtin_amdgcn_readfirstlane(cmax);
        Blk K[MAXB];
        const int gb0 = r0 / 64; // the stripe's first block of the reversed pattern
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            uint64_t p1 = 0, p2 = 0;
            if (on && b < nb) load_planes(P, (uint32_t)m, gb0 + b, true, p1, p2);
            K[b] = Blk{(uint32_t)p1, (uint32_t)(p1 >> 32), (uint32_t)p2, (uint32_t)(p2 >> 32), ~0u, ~0u, 0u, 0u};
        }
        // The bottom row's score at column jlo - 1: exact at column 0 (D[i][0] = i), else the
        // anchor plus this stripe's rows at +1 each.
        int32_t B = jlo == 1 ? r0 + rows : anchor + rows;
        if (last && jlo == 1) {
            best = m; // D[m][0]
            bestj = 0;
        }
        int32_t tpos = np - 16 - (jlo - 1); // text position of the next 16 reversed columns
        uint32_t tw = 0, hw = 0, ow = 0;
        for (int32_t k = 0; k < cmax; ++k) {
            const bool live = k < cols;
            const int32_t c = jlo - 1 + k; // 0-based column of the reversed text
            if ((k & 15) == 0) {
                tw = reverse2(codes16(T, tpos)); // columns c .. c + 15
                tpos -= 16;
                if (has_in) hw = bits16(hin, a.hc_words, c);
            }
            const uint32_t code = tw & 3;
            tw >>= 2;
            const uint32_t M1 = (code & 1) ? 0u : ~0u;
            const uint32_t M2 = (code & 2) ? 0u : ~0u;
            // Prefix: D[0][j] = j; and +1 where the stripe above has no column.
            const bool in = has_in && k <= kin;
            uint32_t hp = in ? (hw & 1u) : 1u, hn = in ? ((hw >> 1) & 1u) : 0u;
            hw >>= 2;
            if (nb_lo == MAXB) {
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
            B += (int32_t)hp - (int32_t)hn; // the bottom row at column c + 1
            if (live) {
                if (last) {
                    if (B < best) {
                        best = B;
                        bestj = c + 1;
                    }
                } else {
                    ow |= (hp | (hn << 1)) << (2 * (c & 15));
                    if ((c & 15) == 15 || k == cols - 1) {
                        hout[c >> 4] = ow;
                        ow = 0;
                    }
                    if (k == kan) anchor = B;
                }
            }
        }
    }
    if (work) a.start[pi] = (uint32_t)(np - bestj);
}

} // namespace

hipError_t sf_launch_myers_start_stripe(const MyersStartArgs &a, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL(myers_start_stripe_kernel, dim3((a.count + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}
