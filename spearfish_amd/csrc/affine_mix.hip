// affine_mix.hip -- one launch for several shape classes of a score-only run: the 64-lane classes,
// the widest 16-bit packed 8- and 16-lane ones and the 4-lane ones (two waves a SIMD);
// affine_mix1.hip holds the narrower 8- and 16-lane classes.
//
// Pairs whose rows are long run one pair a wave, and a class holding only a few of them has too few
// waves to fill the chip: its time is its longest pair's chain of rows.  Launched one class after
// another such classes add up their chains; here each workgroup finds its class (its segment of
// the launch) and runs that class's kernel body, so they all run side by side.
#include "mix_blocks.hpp"

namespace {


__global__ __launch_bounds__(256) void affine_mix_kernel(AffineArgs a, MixArgs mx) {
    affine_mix_block(a, mx, blockIdx.x);
}

} // namespace

bool sf_affine_mixable(int cls) {
    if (cls < 1000) {
        for (int k : {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 24, 28, 32})
            if (cls == k) return true;
        return false;
    }
    if (aff_lp(cls) != 64 || !aff_packed(cls)) return false;
    const int kp = aff_k(cls);
    return kp >= 2 && kp % 2 == 0 && (kp <= 20 || kp == 24 || kp == 28 || kp == 32);
}

// Register groups of the score classes that share a launch: 1 for the packed 8- and 16-lane
// classes of at most 13 registers of slot pairs a lane (at most 168 VGPRs: three waves a SIMD;
// affine_mix1.hip), 2 for the wider ones, the 4-lane ones and the 64-lane classes
// (affine_mix_kernel, two waves), 0 for classes that launch alone (the 32-bit 16-lane kernels).
int sf_affine_mix_group(int cls) {
    if (!aff_packed(cls)) return sf_affine_mixable(cls) ? 2 : 0;
    const int lp = aff_lp(cls), kp = aff_k(cls), k2 = kp / 2;
    const bool m3 = cls % 1000 >= 200;
    if (kp % 2 || k2 < 1) return 0;
    if (lp == 4) return m3 ? 2 : 0;
    if (lp == 8) return !m3 ? 0 : (k2 <= 13 ? 1 : (k2 <= 18 ? 2 : 0));
    if (lp == 16) return k2 <= 13 ? 1 : (k2 <= 16 ? 2 : 0);
    return sf_affine_mixable(cls) ? 2 : 0;
}

uint32_t sf_affine_blocks_for(int cls, uint32_t count) {
    return (uint32_t)(((uint64_t)count * aff_lp(cls) + 255) / 256);
}

static bool mix_segments_ok(const MixArgs &mx, int group) {
    if (mx.nseg == 0 || mx.nseg > MixArgs::kMax) return false;
    for (uint32_t i = 0; i < mx.nseg; ++i)
        if (sf_affine_mix_group(mx.cls[i]) != group ||
            (uint64_t)mx.blk0[i] + sf_affine_blocks_for(mx.cls[i], mx.count[i]) !=
                (i + 1 < mx.nseg ? mx.blk0[i + 1] : mx.blocks))
            return false;
    return true;
}

hipError_t sf_launch_affine_mix_group(const AffineArgs &a, const MixArgs &mx, int group, hipStream_t s) {
    if ((group != 1 && group != 2) || !mix_segments_ok(mx, group)) return hipErrorInvalidValue;
    if (mx.blocks == 0) return hipSuccess;
    if (group == 1) return sf_launch_affine_mix1(a, mx, s);
    if (mx.queue) return sf_launch_affine_mix_persistent(a, mx, s);
    hipLaunchKernelGGL(affine_mix_kernel, dim3(mx.blocks), dim3(256), 0, s, a, mx);
    return hipGetLastError();
}

hipError_t sf_launch_affine_mix(const AffineArgs &a, const MixArgs &mx, hipStream_t s) {
    if (mx.nseg == 0 || mx.nseg > MixArgs::kMax) return hipErrorInvalidValue;
    for (uint32_t i = 0; i < mx.nseg; ++i)
        if (!sf_affine_mixable(mx.cls[i]) ||
            (uint64_t)mx.blk0[i] + ((uint64_t)mx.count[i] * 64 + 255) / 256 != (i + 1 < mx.nseg ? mx.blk0[i + 1] : mx.blocks))
            return hipErrorInvalidValue;
    if (mx.blocks == 0) return hipSuccess;
    if (mx.queue) return sf_launch_affine_mix_persistent(a, mx, s);
    hipLaunchKernelGGL(affine_mix_kernel, dim3(mx.blocks), dim3(256), 0, s, a, mx);
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_affine_mix() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&affine_mix_kernel));
}
