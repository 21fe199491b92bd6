// affine.hip -- banded affine-gap (Gotoh) DP on gfx950, one pair per wave, plus the traceback.
//
// Computes crates/map/src/align.rs:2446-2666 `run_core` (geom = n, no clip, no cut): the same
// H and F in every band cell, hence the same score and score column.
//
// Layout.  Row i covers text columns [i - below, i + above] clipped to [0, n] (align.rs:
// 2468-2481).  The wave holds one row in diagonal coordinates: slot s = j - i + below, lane L
// owning slots L*K .. L*K + K - 1, so the band never moves across lanes.  From row i-1 to row i
// the vertical neighbour of slot s is slot s + 1 (a register, or lane L + 1's first slot via
// DPP wave_shl:1) and the diagonal neighbour is slot s itself.
//
// A row is the reference's three passes (align.rs:1018-1072) fused: per cell
//   F = max(Hup - go, Fup - ext, NEG),  x = max(Hdiag + s, F),  g = x - go + ext*(slot + 1),
// then E at slot s = (max of g over slots < s) - ext*slot, as an in-lane running max and one
// exclusive DPP max-scan across the 64 lanes, and H = max(x, E).  With open >= 0 this is
// Gotoh's recurrence (align.rs:1060-1063), and all values are exact i32.
//
// Cells outside the band.  The reference reads a NEG sentinel just right of each row's band
// (align.rs:1047-1048) and starts the horizontal gap of a row beginning at lo > 0 from NEG.
// Here: slots >= Wd (lanes' padding) are forced to H = NEG every row, which is the sentinel the
// next row reads at its right edge (their F then stays NEG); a row starting at lo > 0 starts at
// slot 0, whose scan carry is below NEG; slots left of column 0 (top rows) and right of column
// n (bottom rows) hold values that stay within NEG + mat * m of NEG and are never read by a band
// cell except as a horizontal-gap source or a diagonal that real band values dominate (the host
// checks the score bound that guarantees it, as the reference's own `narrow` does for 16 bits).
//
// Traceback bits (own definition; the reference has no traceback): per cell, ballots of
//   nd: H != Hdiag + s,  ne: H != x,  fx: F > Hup - go,  ex: E > Hleft - go,
// written as 64-lane masks: row i, slot-in-lane q, lane half h, bit b at word
// (i-1)*8K + q*8 + h*4 + b.  The walk reads them back under the oracle's tie rules.
#include "internal.hpp"

using namespace sf;

namespace {

template <int K, bool TB>
__global__ __launch_bounds__(256) void affine_wave_kernel(AffineArgs a) {
    constexpr int NW = (8 * K + 63) / 64;
    const int lane = threadIdx.x & 63;
    const uint32_t w = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (w >= a.count) return; // whole waves only
    const uint32_t pi = a.order[w];
    const int64_t m = a.pv.plen[pi], n = a.pv.tlen[pi];
    const uint64_t ba = a.below_above[pi];
    const int64_t below = (int64_t)(uint32_t)ba, above = (int64_t)(ba >> 32);
    const int64_t Wd = below + above + 1;
    const uint32_t *P = a.pv.seq + a.pv.poff[pi];
    const uint32_t *T = a.pv.seq + a.pv.toff[pi];
    const int64_t pnw = (m + 15) / 16, tnw = (n + 15) / 16;
    const int32_t mat = a.mat, mis = a.mis, ext = a.ext, go = a.open + a.ext;
    const bool infix = a.mode == kInfix;

    int32_t Hp[K], Fp[K], GC[K], ES[K];
    bool PAD[K];
    const int64_t hi0 = above < n ? above : n;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int64_t s = (int64_t)lane * K + q;
        const int64_t j = s - below;
        Hp[q] = (j >= 0 && j <= hi0) ? ((j == 0 || infix) ? 0 : -(a.open + ext * (int32_t)j)) : kNeg;
        Fp[q] = kNeg;
        GC[q] = ext * (int32_t)(s + 1) - go;
        ES[q] = ext * (int32_t)s;
        PAD[q] = s >= Wd;
    }
    uint64_t win = 0;
    int avail = 0;
    uint32_t pw = 0;
    uint32_t *tbp = nullptr;
    if constexpr (TB) tbp = a.tb + a.tb_off[w];
    bool stopped = false;
    for (int64_t i = 1; i <= m; ++i) {
        const int64_t b = i - below;
        const int64_t lo = b > 0 ? b : 0;
        const int64_t hi = (b + Wd - 1) < n ? (b + Wd - 1) : n;
        if (lo > hi) {
            stopped = true; // align.rs:2543-2556
            break;
        }
        if (((i - 1) & 15) == 0) pw = codes16(P, i - 1, pnw);
        const uint32_t pc = __builtin_amdgcn_readfirstlane(pw & 3u);
        pw >>= 2;
        if (avail < K) {
            const int64_t pos = b + (int64_t)lane * K - 1; // text base of this lane's first slot
            win = (uint64_t)codes16(T, pos, tnw) | ((uint64_t)codes16(T, pos + 16, tnw) << 32);
            avail = 32;
        }
        const int32_t Hn = dpp<kWaveShl1>(kNeg, Hp[0]);
        const int32_t Fn = dpp<kWaveShl1>(kNeg, Fp[0]);
        int32_t X[K], FQ[K], G[K];
        int32_t D[K], U[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int32_t hu = q < K - 1 ? Hp[q + 1] : Hn;
            const int32_t fu = q < K - 1 ? Fp[q + 1] : Fn;
            const uint32_t code = (uint32_t)(win >> (2 * q)) & 3u;
            const int32_t sc = code == pc ? mat : -mis;
            const int32_t u = hu - go;
            const int32_t f = imax(imax(u, fu - ext), kNeg);
            const int32_t d = Hp[q] + sc;
            const int32_t x = imax(d, f);
            X[q] = x;
            FQ[q] = f;
            D[q] = d;
            U[q] = u;
            const int32_t g = x + GC[q];
            G[q] = q ? imax(G[q - 1], g) : g;
        }
        const int32_t C = wave_excl_max(G[K - 1], kScanIdent);
        int32_t H[K], E[K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int32_t e = (q ? imax(C, G[q - 1]) : C) - ES[q];
            const int32_t h = imax(X[q], e);
            E[q] = e;
            H[q] = PAD[q] ? kNeg : h;
        }
        if constexpr (TB) {
            const int32_t hl0 = dpp<kWaveShr1>(kNeg, H[K - 1]);
            int32_t out[NW];
#pragma unroll
            for (int r = 0; r < NW; ++r) out[r] = 0;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const int32_t hl = q ? H[q - 1] : hl0;
                const uint64_t bits[4] = {
                    __builtin_amdgcn_ballot_w64(H[q] != D[q]),
                    __builtin_amdgcn_ballot_w64(H[q] != X[q]),
                    __builtin_amdgcn_ballot_w64(FQ[q] > U[q]),
                    __builtin_amdgcn_ballot_w64(E[q] > hl - go),
                };
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int bt = 0; bt < 4; ++bt) {
                        const int idx = q * 8 + h * 4 + bt;
                        const uint32_t v = (uint32_t)(bits[bt] >> (32 * h));
                        out[idx / 64] = lane == idx % 64 ? (int32_t)v : out[idx / 64];
                    }
            }
            uint32_t *row = tbp + (uint64_t)(i - 1) * (8 * K);
#pragma unroll
            for (int r = 0; r < NW; ++r)
                if (r * 64 + lane < 8 * K) row[r * 64 + lane] = (uint32_t)out[r];
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            Hp[q] = H[q];
            Fp[q] = FQ[q];
        }
        win >>= 2;
        --avail;
    }
    int32_t score;
    uint32_t col;
    if (stopped) {
        score = kNeg;
        col = 0;
    } else if (a.mode == kGlobal) {
        const int64_t sN = n - m + below; // slot of (m, n)
        const int owner = (int)(sN / K), qq = (int)(sN % K);
        int32_t v = Hp[0];
#pragma unroll
        for (int q = 1; q < K; ++q) v = q == qq ? Hp[q] : v;
        score = __builtin_amdgcn_readlane(v, owner);
        col = (uint32_t)n;
    } else {
        // align.rs:2643-2650: max of the last row, first argmax.
        const int64_t b = m - below;
        const int64_t lo = b > 0 ? b : 0;
        const int64_t hi = (b + Wd - 1) < n ? (b + Wd - 1) : n;
        const int64_t slo = lo - b, shi = hi - b;
        int32_t bv = INT32_MIN;
        int32_t bs = INT32_MAX;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int64_t s = (int64_t)lane * K + q;
            if (s >= slo && s <= shi && Hp[q] > bv) {
                bv = Hp[q];
                bs = (int32_t)s;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int32_t ov = __shfl_xor(bv, off), os = __shfl_xor(bs, off);
            if (ov > bv || (ov == bv && os < bs)) {
                bv = ov;
                bs = os;
            }
        }
        score = bv;
        col = (uint32_t)(b + bs);
    }
    if (lane == 0) {
        a.score[pi] = score;
        if (a.score_col) a.score_col[pi] = col;
    }
}

// One lane per pair: walks the traceback bits from the end cell back to the start, emitting
// BAM ops last to first (rules in the header comment of affine_wave_kernel and the oracle).
__global__ __launch_bounds__(256) void walk_kernel(WalkArgs a) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= a.count) return;
    const uint32_t pi = a.order[t];
    uint32_t *out = a.runs + a.runs_off[pi];
    if (a.score[pi] == kNeg) {
        a.nruns[pi] = 0;
        a.tstart[pi] = 0;
        return;
    }
    const int64_t m = a.pv.plen[pi], n = a.pv.tlen[pi];
    const int64_t below = (int64_t)(uint32_t)a.below_above[pi];
    const uint32_t *P = a.pv.seq + a.pv.poff[pi];
    const uint32_t *T = a.pv.seq + a.pv.toff[pi];
    const int64_t pnw = (m + 15) / 16, tnw = (n + 15) / 16;
    const int K = a.K;
    const uint32_t *tb = a.tb + a.tb_off[t];
    int64_t i = m, j = a.mode == kGlobal ? n : (int64_t)a.score_col[pi];
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
    bool bad = false;
    auto bits_at = [&](int64_t ii, int64_t jj) -> uint4 {
        const int64_t s = jj - (ii - below);
        if (ii < 1 || ii > m || s < 0 || s >= 64 * (int64_t)K) {
            // Off the band: cannot happen on a path the DP produced; stop rather than read out of
            // the pair's bits.
            bad = true;
            return make_uint4(0, 0, 0, 0);
        }
        const int64_t L = s / K, q = s % K;
        const uint4 *p = reinterpret_cast<const uint4 *>(tb + (uint64_t)(ii - 1) * (8 * K) + q * 8 + (L >> 5) * 4);
        return *p;
    };
    for (;;) {
        if (bad) break;
        if (state == 0) {
            if (i == 0) {
                if (a.mode == kInfix) tstart = (uint32_t)j;
                else emit(2, (uint32_t)j);
                break;
            }
            if (j == 0) {
                emit(1, (uint32_t)i);
                break;
            }
            const uint4 w = bits_at(i, j);
            const uint32_t bit = (uint32_t)((j - (i - below)) / K) & 31u;
            if (((w.x >> bit) & 1u) == 0) {
                const uint32_t pc = (ldw(P, (i - 1) >> 4, pnw) >> (2 * ((i - 1) & 15))) & 3u;
                const uint32_t tc = (ldw(T, (j - 1) >> 4, tnw) >> (2 * ((j - 1) & 15))) & 3u;
                emit(pc == tc ? 7u : 8u, 1);
                --i;
                --j;
            } else if (((w.y >> bit) & 1u) == 0) {
                state = 1;
            } else {
                state = 2;
            }
        } else if (state == 1) {
            emit(1, 1);
            const uint4 w = bits_at(i, j);
            const uint32_t bit = (uint32_t)((j - (i - below)) / K) & 31u;
            state = ((w.z >> bit) & 1u) ? 1 : 0;
            --i;
        } else {
            emit(2, 1);
            const uint4 w = bits_at(i, j);
            const uint32_t bit = (uint32_t)((j - (i - below)) / K) & 31u;
            state = ((w.w >> bit) & 1u) ? 2 : 0;
            --j;
        }
    }
    if (cur_len) out[nr++] = (cur_len << 4) | cur_op;
    if (bad) {
        nr = 0;
        tstart = kNone; // reported to the host as a device error
    }
    a.nruns[pi] = nr;
    a.tstart[pi] = tstart;
}

__global__ __launch_bounds__(256) void compact_kernel(const uint32_t *order, uint32_t count, const uint32_t *runs,
                                                      const uint64_t *runs_off, const uint32_t *nruns,
                                                      const uint64_t *out_off, uint32_t *out) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= count) return;
    const uint32_t pi = order[t];
    const uint32_t nr = nruns[pi];
    const uint32_t *src = runs + runs_off[pi];
    uint32_t *dst = out + out_off[pi];
    for (uint32_t r = 0; r < nr; ++r) dst[r] = src[nr - 1 - r];
}

} // namespace

int sf_affine_k_for(uint64_t width) {
    static const int cls[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32};
    for (int c : cls)
        if (width <= (uint64_t)c * 64) return c;
    return 0;
}

hipError_t sf_launch_affine(const AffineArgs &a, int K, bool tb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    dim3 grid((a.count + 3) / 4), block(256);
    switch (K) {
#define CASE(KK)                                                                                 \
    case KK:                                                                                     \
        if (tb) hipLaunchKernelGGL((affine_wave_kernel<KK, true>), grid, block, 0, s, a);       \
        else hipLaunchKernelGGL((affine_wave_kernel<KK, false>), grid, block, 0, s, a);         \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(10) CASE(12) CASE(14) CASE(16)
        CASE(20) CASE(24) CASE(28) CASE(32)
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
    hipLaunchKernelGGL(compact_kernel, dim3((count + 255) / 256), dim3(256), 0, s, order, count, runs, runs_off, nruns,
                       out_off, out);
    return hipGetLastError();
}
