// myers.hip -- Myers/Hyyro bit-parallel edit distance on gfx950: the kernels (bodies in
// myers_body.hpp), the shape classes and their launchers.
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
#include "myers_body.hpp"

using namespace sf;

namespace {

// MAXB = 16 (the 1 kb patterns of a uniform batch) in 168 registers, three waves a SIMD instead
// of two at 174: two values a column come back from scratch, and the pass is still faster.
template <int MAXB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MAXB == 16 ? 3 : 1))) void myers_lane_kernel(MyersArgs a) {
    myers_lane_body<MAXB>(a, blockIdx.x);
}

// The lane kernel over the wave's Ukkonen band, for passes whose threshold is below the patterns'
// lengths (a finite k, and the reversed pass for `start`, whose threshold is the distance).
template <int MAXB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MAXB >= 12 && MAXB <= 16 ? 2 : 1))) void myers_lane_ukk_kernel(MyersArgs a) {
    myers_lane_body<MAXB, true>(a, blockIdx.x);
}

template <int G, int R>
__global__ __launch_bounds__(256) void myers_pipe_kernel(MyersArgs a) {
    myers_pipe_body<G, R>(a, blockIdx.x);
}

// One stripe of 1,536 blocks of patterns longer than the largest class.
__global__ __launch_bounds__(256) void myers_striped_kernel(MyersArgs a) {
    myers_pipe_body<64, 24, true>(a, blockIdx.x);
}

} // namespace

// Shape classes: lane kernel MAXB (1..24; more planes would spill), then pipelined classes
// encoded as 1000 * G + R: 16 lanes per pair up to 96 blocks, then 64 lanes per pair, which keeps
// the per-lane chain of a long pattern (its critical path) at most 24 blocks, with 32-lane classes
// of 160 and 224 blocks between 64 x 2, 64 x 3 and 64 x 4 (less padding for a class that fills
// the chip; the batch moves a 32-lane bin too small for that to 64 lanes, sf_myers_wide_class).
int sf_myers_maxb_for(uint32_t nblocks) {
    static const int lane_cls[] = {1, 2, 4, 8, 12, 16, 20, 24};
    for (int c : lane_cls)
        if (nblocks <= (uint32_t)c) return c;
    static const int pipe16[] = {2, 3, 4, 5, 6};
    for (int r : pipe16)
        if (nblocks <= (uint32_t)(16 * r)) return 16000 + r;
    static const int pipe[] = {64002, 32005, 64003, 32007, 64004, 64005, 64006, 64008,
                               64010, 64012, 64016, 64020, 64024};
    for (int c : pipe)
        if (nblocks <= (uint32_t)((c / 1000) * (c % 1000))) return c;
    return 0;
}

int sf_myers_wide_class(int cls) { return cls / 1000 == 32 ? 64000 + (cls % 1000 + 1) / 2 : cls; }

hipError_t sf_launch_myers(const MyersArgs &a, int maxb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    dim3 grid((a.count + 255) / 256), block(256);
    if (maxb >= 16000) {
        const int G = maxb / 1000, R = maxb % 1000;
        const uint64_t threads = (uint64_t)a.count * G;
        dim3 pg((unsigned)((threads + 255) / 256));
        switch (maxb) {
#define PCASE(GG, RR) \
    case GG * 1000 + RR: hipLaunchKernelGGL((myers_pipe_kernel<GG, RR>), pg, block, 0, s, a); break;
            // (16, 1): one block a lane, for patterns of up to 16 blocks when a segment's chain, not
            // the issue rate, sets the time (small serving batches); sf_myers_maxb_for never picks it.
            PCASE(16, 1) PCASE(16, 2) PCASE(16, 3) PCASE(16, 4) PCASE(16, 5) PCASE(16, 6) PCASE(32, 5) PCASE(32, 7)
            PCASE(64, 2) PCASE(64, 3)
            PCASE(64, 4) PCASE(64, 5) PCASE(64, 6) PCASE(64, 8) PCASE(64, 10) PCASE(64, 12) PCASE(64, 16)
            PCASE(64, 20) PCASE(64, 24)
#undef PCASE
        default:
            return hipErrorInvalidValue;
        }
        (void)G;
        (void)R;
        return hipGetLastError();
    }
    switch (maxb) {
#define CASE(B) \
    case B: hipLaunchKernelGGL(myers_lane_kernel<B>, grid, block, 0, s, a); break;
        CASE(1) CASE(2) CASE(4) CASE(8) CASE(12) CASE(16) CASE(20) CASE(24)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}


bool sf_myers_ukk_class(int cls) { return cls >= 2 && cls <= 24; }

hipError_t sf_launch_myers_ukk(const MyersArgs &a, int maxb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    dim3 grid((a.count + 255) / 256), block(256);
    switch (maxb) {
#define CASE(B) \
    case B: hipLaunchKernelGGL(myers_lane_ukk_kernel<B>, grid, block, 0, s, a); break;
        CASE(2) CASE(4) CASE(8) CASE(12) CASE(16) CASE(20) CASE(24)
#undef CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t sf_launch_myers_striped(const MyersArgs &a, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    const uint64_t threads = (uint64_t)a.count * 64;
    hipLaunchKernelGGL(myers_striped_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_myers() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&myers_lane_kernel<1>));
}
