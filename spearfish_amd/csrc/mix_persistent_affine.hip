// mix_persistent_affine.hip -- the persistent form of affine_mix_kernel (affine_mix.hip): as many workgroups as the device holds at once
// (sf_persistent_grid), each taking block indices from the launch's work queue (MixArgs::queue: the
// classes' segments in the launch's order, the longest chains first) until none is left.  A kernel
// of its own, apart from the one-workgroup-a-block form, because the loop raises the register count
// (every class body's constants are hoisted above it): that cost stays out of the default launches.
#include "mix_blocks.hpp"

using namespace sf;

namespace {

// Held to the occupancy of affine_mix_kernel (2 waves a SIMD).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void affine_mixp_kernel(AffineArgs a,
                                                                                                 MixArgs mx) {
    uint32_t *const q = mx.queue;
    for (uint32_t b = mix_next(q, true); b < mx.blocks; b = mix_next(q, false)) affine_mix_block(a, mx, b);
    mix_done(q);
}

} // namespace

hipError_t sf_launch_affine_mix_persistent(const AffineArgs &a, const MixArgs &mx, hipStream_t s) {
    const uint32_t grid = sf_persistent_grid(reinterpret_cast<const void *>(&affine_mixp_kernel), mx.blocks);
    hipLaunchKernelGGL(affine_mixp_kernel, dim3(grid), dim3(256), 0, s, a, mx);
    return hipGetLastError();
}
