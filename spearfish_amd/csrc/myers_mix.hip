// myers_mix.hip -- Myers bins of several shape classes in one launch.
//
// A length-skewed batch spreads its pairs over many classes.  Launched one after another on a
// stream, each bin ends in a tail (its last waves finishing while the chip empties) and a bin of
// a few hundred waves of long patterns runs as long as its slowest pair's chain; streams put only
// as many bins side by side as the process has hardware queues.  Here each workgroup finds its
// class (its segment of the launch) and runs that class's body, so the chip refills as waves end.
// A kernel holds the registers of its widest body, so the classes form three groups of similar
// footprint (at most 108, 140 and 248 VGPRs: 4, 3 and 2 waves a SIMD), one kernel each, and a
// run takes at most three launches.  Segments begin on workgroup boundaries, so a wave never
// holds two classes.
#include "mix_blocks.hpp"

namespace {


template <int GROUP>
__global__ __launch_bounds__(256) void myers_mix_kernel(MyersArgs a, MixArgs mx) {
    myers_mix_block<GROUP>(a, mx, blockIdx.x);
}

} // namespace

// The pipeline classes past 64 x 6 blocks (patterns over 24,576 bases) hold 153-256 registers a
// lane and keep launches of their own.
int sf_myers_mix_group(int cls) {
    switch (cls) {
    case 1: case 2: case 4: case 8: case 16001: case 16002: case 16003: case 64002: case 64003:
        return 1;
    case 12: case 16004: case 16005: case 16006: case 32005: case 32007: case 64004: case 64005: case 64006:
        return 2;
    case 16: case 20: case 24:
        return 3;
    default:
        return 0;
    }
}

uint32_t sf_myers_blocks_for(int cls, uint32_t count) {
    const uint64_t lanes = cls >= 16000 ? (uint64_t)(cls / 1000) : 1;
    return (uint32_t)((count * lanes + 255) / 256);
}

hipError_t sf_launch_myers_mix(const MyersArgs &a, const MixArgs &mx, int group, hipStream_t s) {
    if (mx.nseg == 0 || mx.nseg > MixArgs::kMax || group < 1 || group > 3) return hipErrorInvalidValue;
    for (uint32_t i = 0; i < mx.nseg; ++i)
        if (sf_myers_mix_group(mx.cls[i]) != group ||
            (uint64_t)mx.blk0[i] + sf_myers_blocks_for(mx.cls[i], mx.count[i]) !=
                (i + 1 < mx.nseg ? mx.blk0[i + 1] : mx.blocks))
            return hipErrorInvalidValue;
    if (mx.blocks == 0) return hipSuccess;
    if (mx.queue) return sf_launch_myers_mix_persistent(a, mx, group, s);
    switch (group) {
    case 1: hipLaunchKernelGGL(myers_mix_kernel<1>, dim3(mx.blocks), dim3(256), 0, s, a, mx); break;
    case 2: hipLaunchKernelGGL(myers_mix_kernel<2>, dim3(mx.blocks), dim3(256), 0, s, a, mx); break;
    default: hipLaunchKernelGGL(myers_mix_kernel<3>, dim3(mx.blocks), dim3(256), 0, s, a, mx); break;
    }
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_myers_mix() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&myers_mix_kernel<1>));
}
