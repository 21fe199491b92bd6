// myers.hip -- Myers/Hyyro bit-parallel edit distance on gfx950, one pair per lane.
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
#include "internal.hpp"

using namespace sf;

namespace {

__device__ __forceinline__ void load_planes(const uint32_t *P, int64_t pnw, uint32_t m, int b, bool rev,
                                            uint64_t &p1, uint64_t &p2) {
    uint32_t w[4];
    if (!rev) {
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = ldw(P, (int64_t)b * 4 + q, pnw);
    } else {
        // Row i of the reversed pattern is P[m - 1 - i]: rows 64b.. are codes s.. reversed,
        // s = m - 64 - 64b.
        int64_t s = (int64_t)m - 64 - 64 * (int64_t)b;
#pragma unroll
        for (int q = 0; q < 4; ++q) w[q] = reverse2(codes16(P, s + 16 * (3 - q), pnw));
    }
    uint32_t lo1 = even16(w[0]) | (even16(w[1]) << 16), hi1 = even16(w[2]) | (even16(w[3]) << 16);
    uint32_t lo2 = even16(w[0] >> 1) | (even16(w[1] >> 1) << 16), hi2 = even16(w[2] >> 1) | (even16(w[3] >> 1) << 16);
    p1 = (uint64_t)lo1 | ((uint64_t)hi1 << 32);
    p2 = (uint64_t)lo2 | ((uint64_t)hi2 << 32);
}

template <int MAXB>
__global__ __launch_bounds__(256) void myers_lane_kernel(MyersArgs a) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const bool valid = t < a.count;
    const uint32_t pi = valid ? a.order[t] : 0;
    uint32_t m = valid ? a.pv.plen[pi] : 0;
    const uint32_t tlen = valid ? a.pv.tlen[pi] : 0;
    const uint32_t *P = a.pv.seq + (valid ? a.pv.poff[pi] : 0);
    const uint32_t *T = a.pv.seq + (valid ? a.pv.toff[pi] : 0);
    const int64_t pnw = (m + 15) / 16, tnw = (tlen + 15) / 16;
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
        const bool ok = d <= a.k;
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

    uint64_t P1[MAXB], P2[MAXB], PV[MAXB], MV[MAXB];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        P1[b] = P2[b] = 0;
        if (b < nb && work) load_planes(P, pnw, m, b, rev, P1[b], P2[b]);
        PV[b] = ~0ull;
        MV[b] = 0;
    }
    const uint64_t top = mode == kInfix ? 0 : 1;
    int32_t score = (int32_t)m; // D[m][0]
    int32_t best = (int32_t)m;
    uint32_t bestj = 0;
    uint32_t tw = 0;
    for (uint32_t j = 0; j < n; ++j) {
        if ((j & 15) == 0) {
            tw = rev ? reverse2(codes16(T, (int64_t)n - 16 - (int64_t)j, tnw)) : ldw(T, j >> 4, tnw);
        }
        const uint32_t c = tw & 3;
        tw >>= 2;
        const uint64_t M1 = (c & 1) ? 0ull : ~0ull;
        const uint64_t M2 = (c & 2) ? 0ull : ~0ull;
        uint64_t hp = top, hn = 0;
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            if (b < nb_lo - 1) {
                const uint64_t eq = (P1[b] ^ M1) & (P2[b] ^ M2);
                const uint64_t p = PV[b], mm = MV[b];
                const uint64_t xv = eq | mm;
                const uint64_t eqn = eq | hn;
                const uint64_t xh = (((eqn & p) + p) ^ p) | eqn;
                uint64_t ph = mm | ~(xh | p);
                uint64_t mh = p & xh;
                const uint64_t op = ph >> 63, om = mh >> 63;
                ph = (ph << 1) | hp;
                mh = (mh << 1) | hn;
                PV[b] = mh | ~(xv | ph);
                MV[b] = ph & xv;
                hp = op;
                hn = om;
            } else if (b < nb) {
                const uint64_t eq = (P1[b] ^ M1) & (P2[b] ^ M2);
                const uint64_t p = PV[b], mm = MV[b];
                const uint64_t xv = eq | mm;
                const uint64_t eqn = eq | hn;
                const uint64_t xh = (((eqn & p) + p) ^ p) | eqn;
                uint64_t ph = mm | ~(xh | p);
                uint64_t mh = p & xh;
                const unsigned bit = b == lastb ? last_bit : 63u;
                const uint64_t op = (ph >> bit) & 1, om = (mh >> bit) & 1;
                ph = (ph << 1) | hp;
                mh = (mh << 1) | hn;
                PV[b] = mh | ~(xv | ph);
                MV[b] = ph & xv;
                hp = op;
                hn = om;
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
        if (mode == kGlobal) {
            best = score;
            bestj = n;
        }
        const bool ok = (uint32_t)best <= a.k;
        a.dist[pi] = ok ? (uint32_t)best : kNone;
        if (a.end) a.end[pi] = ok ? bestj : kNone;
        if (a.start && (mode != kInfix || !ok)) a.start[pi] = ok ? 0 : kNone;
    } else {
        a.start[pi] = n - bestj;
    }
}

__global__ void probe_kernel(int32_t *out) {
    // Lane-movement check of the DPP encodings the kernels rely on: 5 x 64 values.
    const int32_t lane = threadIdx.x;
    out[lane] = dpp<kWaveShl1>(-1, lane);
    out[64 + lane] = dpp<kWaveShr1>(-1, lane);
    out[128 + lane] = dpp<kRowShr1>(-1, lane);
    out[192 + lane] = wave_excl_max(lane * 3 % 17, -100);
    out[256 + lane] = __shfl_xor(lane, 5);
}

} // namespace

int sf_myers_maxb_for(uint32_t nblocks) {
    static const int cls[] = {1, 2, 4, 8, 12, 16, 20, 24, 32};
    for (int c : cls)
        if (nblocks <= (uint32_t)c) return c;
    return 0;
}

hipError_t sf_launch_myers(const MyersArgs &a, int maxb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    dim3 grid((a.count + 255) / 256), block(256);
    switch (maxb) {
#define CASE(B) \
    case B: hipLaunchKernelGGL(myers_lane_kernel<B>, grid, block, 0, s, a); break;
        CASE(1) CASE(2) CASE(4) CASE(8) CASE(12) CASE(16) CASE(20) CASE(24) CASE(32)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t sf_launch_probe(int32_t *out, hipStream_t s) {
    hipLaunchKernelGGL(probe_kernel, dim3(1), dim3(64), 0, s, out);
    return hipGetLastError();
}
