// affine_mix1.hip -- score runs: the 16-bit packed shape classes of 8 and 16 lanes a pair with at
// most 168 VGPRs a lane (three waves a SIMD) in one launch.
//
// A length-skewed batch spreads its pairs over a dozen shape classes; launched one after another,
// each ends in a tail while the chip empties.  Here each workgroup finds its class (its segment of
// the launch) and runs that class's kernel body, so the chip refills as waves end.  The wider
// classes and the 64-lane ones (long pairs, few waves) share affine_mix_kernel (affine_mix.hip,
// two waves a SIMD).  Segments begin on workgroup boundaries, so a wave never holds two classes.
#include "affine_pk.hpp"

namespace {

__global__ __launch_bounds__(256) void affine_mix1_kernel(AffineArgs a, MixArgs mx) {
    uint32_t s = 0;
    while (s + 1 < mx.nseg && blockIdx.x >= mx.blk0[s + 1]) ++s;
    AffineArgs x = a;
    x.order += mx.off[s];
    x.desc += mx.off[s];
    x.count = mx.count[s];
    const uint32_t blk = blockIdx.x - mx.blk0[s];
    switch (mx.cls[s]) {
#define P8(KK)                                                                                   \
    case 8200 + 2 * KK: affine_pk16_body<8, KK, true>(x, blk); break;
#define P16(KK)                                                                                  \
    case 16100 + 2 * KK: affine_pk16_body<16, KK, false>(x, blk); break;                          \
    case 16200 + 2 * KK: affine_pk16_body<16, KK, true>(x, blk); break;
        P8(1) P8(2) P8(3) P8(4) P8(5) P8(6) P8(7) P8(8) P8(9) P8(10) P8(11) P8(12) P8(13)
        P16(1) P16(2) P16(3) P16(4) P16(5) P16(6) P16(7) P16(8) P16(9) P16(10) P16(11) P16(12) P16(13)
#undef P8
#undef P16
    default:
        break; // the host launches only this group's classes (sf_affine_mix_group)
    }
}

} // namespace

// Launched through sf_launch_affine_mix_group (affine_mix.hip), which checks the segments.
hipError_t sf_launch_affine_mix1(const AffineArgs &a, const MixArgs &mx, hipStream_t s) {
    hipLaunchKernelGGL(affine_mix1_kernel, dim3(mx.blocks), dim3(256), 0, s, a, mx);
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_affine_mix1() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&affine_mix1_kernel));
}
