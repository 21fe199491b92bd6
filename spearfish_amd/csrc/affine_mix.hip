// affine_mix.hip -- one launch for several shape classes of a score-only run on 64 lanes a pair.
//
// Pairs whose rows are long run one pair a wave, and a class holding only a few of them has too few
// waves to fill the chip: its time is its longest pair's chain of rows.  Launched one class after
// another such classes add up their chains; here each workgroup finds its class (its segment of
// the launch) and runs that class's kernel body, so they all run side by side.
#include "affine_pk.hpp"
#include "affine_wave.hpp"

namespace {

__global__ __launch_bounds__(256) void affine_mix_kernel(AffineArgs a, MixArgs mx) {
    uint32_t s = 0;
    while (s + 1 < mx.nseg && blockIdx.x >= mx.blk0[s + 1]) ++s;
    AffineArgs x = a;
    x.order += mx.off[s];
    x.desc += mx.off[s];
    x.count = mx.count[s];
    const uint32_t blk = blockIdx.x - mx.blk0[s];
    switch (mx.cls[s]) {
#define W(KK)                                                                                    \
    case KK: affine_wave_body<64, KK, false>(x, blk); break;
        W(1) W(2) W(3) W(4) W(5) W(6) W(7) W(8) W(9) W(10) W(11) W(12) W(13) W(14) W(15) W(16) W(18) W(20)
        W(24) W(28) W(32)
#undef W
#define P(KK)                                                                                    \
    case 64100 + 2 * KK: affine_pk16_body<64, KK, false>(x, blk); break;                          \
    case 64200 + 2 * KK: affine_pk16_body<64, KK, true>(x, blk); break;
        P(1) P(2) P(3) P(4) P(5) P(6) P(7) P(8) P(9) P(10) P(12) P(14) P(16)
#undef P
    default:
        break; // the host launches only the classes above (sf_affine_mixable)
    }
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

hipError_t sf_launch_affine_mix(const AffineArgs &a, const MixArgs &mx, hipStream_t s) {
    if (mx.nseg == 0 || mx.nseg > MixArgs::kMax) return hipErrorInvalidValue;
    for (uint32_t i = 0; i < mx.nseg; ++i)
        if (!sf_affine_mixable(mx.cls[i]) ||
            (uint64_t)mx.blk0[i] + ((uint64_t)mx.count[i] * 64 + 255) / 256 != (i + 1 < mx.nseg ? mx.blk0[i + 1] : mx.blocks))
            return hipErrorInvalidValue;
    if (mx.blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(affine_mix_kernel, dim3(mx.blocks), dim3(256), 0, s, a, mx);
    return hipGetLastError();
}
