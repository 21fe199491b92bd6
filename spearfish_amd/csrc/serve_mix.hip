// serve_mix.hip -- a small serving batch's Myers and affine segments in one launch.
//
// A serving batch holds a few reads, so its time is its longest segment's chain.  The Myers and
// affine runs are independent: in one launch (the first workgroups the 16-bit packed affine
// classes at 64 lanes a pair, the rest the Myers classes of the first register group) they run
// side by side without a second stream, whose placement on the process's few hardware queues
// would decide whether they overlap.  A serving batch's segments nearly all fall in these classes
// (reads well under the 16-bit range, patterns under 3,072 bases); the host launches any other on
// its own kernel after this one (sf_serve_mix_affine, sf_serve_mix_myers).  Segments begin on
// workgroup boundaries, so a wave never holds two classes.
#include "affine_pk.hpp"
#include "myers_body.hpp"

namespace {

__global__ __launch_bounds__(256) void serve_mix_kernel(AffineArgs aa, MixArgs amx, MyersArgs ma, MixArgs mmx) {
    if (blockIdx.x < amx.blocks) {
        uint32_t s = 0;
        while (s + 1 < amx.nseg && blockIdx.x >= amx.blk0[s + 1]) ++s;
        AffineArgs x = aa;
        x.order += amx.off[s];
        x.desc += amx.off[s];
        x.count = amx.count[s];
        const uint32_t blk = blockIdx.x - amx.blk0[s];
        switch (amx.cls[s]) {
#define P(KK)                                                                                    \
    case 64100 + 2 * KK: affine_pk16_body<64, KK, false>(x, blk); break;                          \
    case 64200 + 2 * KK: affine_pk16_body<64, KK, true>(x, blk); break;
            P(1) P(2) P(3) P(4) P(5) P(6) P(7) P(8) P(9) P(10) P(12) P(14) P(16)
#undef P
        default:
            break; // the host launches only the classes above (sf_serve_mix_affine)
        }
        return;
    }
    const uint32_t b = blockIdx.x - amx.blocks;
    uint32_t s = 0;
    while (s + 1 < mmx.nseg && b >= mmx.blk0[s + 1]) ++s;
    MyersArgs x = ma;
    x.order += mmx.off[s];
    x.desc += mmx.off[s];
    x.count = mmx.count[s];
    const uint32_t blk = b - mmx.blk0[s];
    switch (mmx.cls[s]) {
#define L(B)                                                                                     \
    case B: myers_lane_body<B>(x, blk); break;
#define Q(GG, RR)                                                                                \
    case GG * 1000 + RR: myers_pipe_body<GG, RR>(x, blk); break;
        L(1) L(2) L(4) L(8) Q(16, 1) Q(16, 2) Q(16, 3) Q(64, 2) Q(64, 3)
#undef L
#undef Q
    default:
        break; // the host launches only the classes above (sf_serve_mix_myers)
    }
}

} // namespace

// The classes the combined launch holds: the 16-bit packed 64-lane affine ones, and Myers's first
// register group.
bool sf_serve_mix_affine(int cls) { return aff_lp(cls) == 64 && aff_packed(cls) && sf_affine_mixable(cls); }
bool sf_serve_mix_myers(int cls) { return sf_myers_mix_group(cls) == 1; }

hipError_t sf_launch_serve_mix(const AffineArgs &aa, const MixArgs &amx, const MyersArgs &ma, const MixArgs &mmx,
                               hipStream_t s) {
    if (amx.nseg > MixArgs::kMax || mmx.nseg > MixArgs::kMax) return hipErrorInvalidValue;
    for (uint32_t i = 0; i < amx.nseg; ++i)
        if (!sf_serve_mix_affine(amx.cls[i]) || (uint64_t)amx.blk0[i] + sf_affine_blocks_for(amx.cls[i], amx.count[i]) !=
                                                  (i + 1 < amx.nseg ? amx.blk0[i + 1] : amx.blocks))
            return hipErrorInvalidValue;
    for (uint32_t i = 0; i < mmx.nseg; ++i) {
        if (!sf_serve_mix_myers(mmx.cls[i]) || (uint64_t)mmx.blk0[i] + sf_myers_blocks_for(mmx.cls[i], mmx.count[i]) !=
                                      (i + 1 < mmx.nseg ? mmx.blk0[i + 1] : mmx.blocks))
            return hipErrorInvalidValue;
    }
    if ((amx.nseg == 0) != (amx.blocks == 0) || (mmx.nseg == 0) != (mmx.blocks == 0)) return hipErrorInvalidValue;
    const uint64_t blocks = (uint64_t)amx.blocks + mmx.blocks;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7FFFFFFFu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(serve_mix_kernel, dim3((unsigned)blocks), dim3(256), 0, s, aa, amx, ma, mmx);
    return hipGetLastError();
}

// Loads this file's code object now (see sf_load_code_objects).
hipError_t sf_load_code_serve_mix() {
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(&serve_mix_kernel));
}
