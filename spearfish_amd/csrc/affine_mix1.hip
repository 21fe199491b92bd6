// affine_mix1.hip -- score runs: the 16-bit packed shape classes of 8 and 16 lanes a pair with at
// most 168 VGPRs a lane (three waves a SIMD) in one launch.
//
// A length-skewed batch spreads its pairs over a dozen shape classes; launched one after another,
// each ends in a tail while the chip empties.  Here each workgroup finds its class (its segment of
// the launch) and runs that class's kernel body, so the chip refills as waves end.  The wider
// classes and the 64-lane ones (long pairs, few waves) share affine_mix_kernel (affine_mix.hip,
// two waves a SIMD).  Segments begin on workgroup boundaries, so a wave never holds two classes.
#include "mix_blocks.hpp"

namespace {


__global__ __launch_bounds__(256) void affine_mix1_kernel(AffineArgs a, MixArgs mx) {
    affine_mix1_block(a, mx, blockIdx.x);
}

} // namespace

// Launched through sf_launch_affine_mix_group (affine_mix.hip), which checks the segments.
hipError_t sf_launch_affine_mix1(const AffineArgs &a, const MixArgs &mx, hipStream_t s) {
    if (mx.queue) return sf_launch_affine_mix1_persistent(a, mx, s);
    hipLaunchKernelGGL(affine_mix1_kernel, dim3(mx.blocks), dim3(256), 0, s, a, mx);
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_affine_mix1() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&affine_mix1_kernel));
}
