// affine_pk4.hip -- instantiates and launches the 16-bit packed score kernel (affine_pk.hpp) for 4
// lanes a pair (sixteen pairs a wave): the per-row work every lane repeats -- the group scan, the
// neighbour move, the pattern code -- is shared by twice the slots of the 8-lane shape, at two
// waves a SIMD (the 8-lane kernel's registers doubled do not fit three).
#include "affine_pk.hpp"

hipError_t sf_launch_affine_packed4(const AffineArgs &a, int cls, bool tb, hipStream_t s) {
    if (a.count == 0) return hipSuccess;
    if (tb || cls % 1000 < 200) return hipErrorInvalidValue; // scores, three-term max
    dim3 grid((unsigned)(((uint64_t)a.count * 4 + 255) / 256)), block(256);
    switch (aff_k(cls)) {
#define PK4(KK) \
    case 2 * KK: hipLaunchKernelGGL((affine_pk16_kernel<4, KK, true>), grid, block, 0, s, a); break;
        PK4(18) PK4(20) PK4(24) PK4(28) PK4(32) PK4(33) PK4(34) PK4(35) PK4(36)
#undef PK4
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
