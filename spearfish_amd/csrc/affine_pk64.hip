// affine_pk64.hip -- instantiates and launches the 16-bit packed kernels (affine_pk.hpp) for 64
// lanes a pair (one pair a wave: bands too wide for 16 lanes, or few long pairs).
#include "affine_pk.hpp"

hipError_t sf_launch_affine_packed64(const AffineArgs &a, int cls, bool tb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    dim3 grid((unsigned)(((uint64_t)a.count * 64 + 255) / 256)), block(256);
    const bool m3 = cls % 1000 >= 200;
    if (tb && m3) return hipErrorInvalidValue;
    switch (aff_k(cls)) {
#define PK(KK)                                                                                   \
    case 2 * KK:                                                                                 \
        if (tb) hipLaunchKernelGGL((affine_pk16_tb_kernel<64, KK>), grid, block, 0, s, a);      \
        else if (m3) hipLaunchKernelGGL((affine_pk16_kernel<64, KK, true>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((affine_pk16_kernel<64, KK, false>), grid, block, 0, s, a);     \
        break;
        PK(1) PK(2) PK(3) PK(4) PK(5) PK(6) PK(7) PK(8) PK(9) PK(10) PK(12) PK(14) PK(16)
#undef PK
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_affine_pk64() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&affine_pk16_kernel<64, 1, false>));
}
