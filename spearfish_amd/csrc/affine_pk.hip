// affine_pk.hip -- instantiates and launches the 16-bit packed kernels (affine_pk.hpp) for 8, 16
// and (traceback) 32 lanes a pair; 64 lanes a pair are in affine_pk64.hip.
#include "affine_pk.hpp"

hipError_t sf_launch_affine_packed(const AffineArgs &a, int cls, bool tb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    const int LP = aff_lp(cls);
    dim3 grid((unsigned)(((uint64_t)a.count * LP + 255) / 256)), block(256);
    if (LP == 64) return sf_launch_affine_packed64(a, cls, tb, s);
    if (LP == 4) return sf_launch_affine_packed4(a, cls, tb, s);
    if (LP == 32) {
        // 32 lanes a pair (two pairs a wave): traceback runs of bands over 512 slots.
        if (!tb || cls % 1000 >= 200) return hipErrorInvalidValue;
        switch (aff_k(cls)) {
#define PK32(KK)                                                                                 \
    case 2 * KK: hipLaunchKernelGGL((affine_pk16_tb_kernel<32, KK>), grid, block, 0, s, a); break;
            PK32(9) PK32(10) PK32(11) PK32(12) PK32(13) PK32(14) PK32(15) PK32(16)
#undef PK32
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (LP == 8) {
        if (tb || cls % 1000 < 200) return hipErrorInvalidValue; // 8 lanes: scores, three-term max
        switch (aff_k(cls)) {
#define PK8(KK)                                                                                  \
    case 2 * KK: hipLaunchKernelGGL((affine_pk16_kernel<8, KK, true>), grid, block, 0, s, a); break;
            PK8(1) PK8(2) PK8(3) PK8(4) PK8(5) PK8(6) PK8(7) PK8(8) PK8(9) PK8(10) PK8(11) PK8(12) PK8(13)
            PK8(14) PK8(15) PK8(16) PK8(17) PK8(18)
#undef PK8
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    {
        const bool m3 = cls % 1000 >= 200;
        if (tb && m3) return hipErrorInvalidValue;
        switch (aff_k(cls)) {
#define PK(KK)                                                                                   \
    case 2 * KK:                                                                                 \
        if (tb) hipLaunchKernelGGL((affine_pk16_tb_kernel<16, KK>), grid, block, 0, s, a);      \
        else if (m3) hipLaunchKernelGGL((affine_pk16_kernel<16, KK, true>), grid, block, 0, s, a); \
        else hipLaunchKernelGGL((affine_pk16_kernel<16, KK, false>), grid, block, 0, s, a);     \
        break;
            PK(1) PK(2) PK(3) PK(4) PK(5) PK(6) PK(7) PK(8) PK(9) PK(10) PK(11) PK(12) PK(13) PK(14) PK(15)
            PK(16)
#undef PK
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_affine_pk() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&affine_pk16_kernel<16, 1, false>));
}
