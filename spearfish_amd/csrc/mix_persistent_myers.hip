// mix_persistent_myers.hip -- the persistent form of the Myers mixed launches (myers_mix.hip): as many workgroups as the device holds at once
// (sf_persistent_grid), each taking block indices from the launch's work queue (MixArgs::queue: the
// classes' segments in the launch's order, the longest chains first) until none is left.  A kernel
// of its own, apart from the one-workgroup-a-block form, because the loop raises the register count
// (every class body's constants are hoisted above it): that cost stays out of the default launches.
#include "mix_blocks.hpp"

using namespace sf;

namespace {

// Held to the occupancy of the one-workgroup-a-block kernels (4, 3 and 2 waves a SIMD).
template <int GROUP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(GROUP == 1 ? 4 : GROUP == 2 ? 3 : 2)))
void myers_mixp_kernel(MyersArgs a, MixArgs mx) {
    uint32_t *const q = mx.queue;
    for (uint32_t b = mix_next(q, true); b < mx.blocks; b = mix_next(q, false)) myers_mix_block<GROUP>(a, mx, b);
    mix_done(q);
}

} // namespace

hipError_t sf_launch_myers_mix_persistent(const MyersArgs &a, const MixArgs &mx, int group, hipStream_t s) {
    const void *k = group == 1   ? reinterpret_cast<const void *>(&myers_mixp_kernel<1>)
                    : group == 2 ? reinterpret_cast<const void *>(&myers_mixp_kernel<2>)
                                 : reinterpret_cast<const void *>(&myers_mixp_kernel<3>);
    const uint32_t grid = sf_persistent_grid(k, mx.blocks);
    switch (group) {
    case 1: hipLaunchKernelGGL(myers_mixp_kernel<1>, dim3(grid), dim3(256), 0, s, a, mx); break;
    case 2: hipLaunchKernelGGL(myers_mixp_kernel<2>, dim3(grid), dim3(256), 0, s, a, mx); break;
    default: hipLaunchKernelGGL(myers_mixp_kernel<3>, dim3(grid), dim3(256), 0, s, a, mx); break;
    }
    return hipGetLastError();
}
