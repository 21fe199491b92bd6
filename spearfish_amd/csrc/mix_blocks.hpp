// mix_blocks.hpp -- one block index of a mixed launch (several shape classes in one launch): the
// workgroup finds its class's segment of the launch and runs that class's kernel body.  Shared by
// the launches of one workgroup per block index (myers_mix.hip, affine_mix.hip, affine_mix1.hip)
// and the persistent launches that take block indices from a work queue (mix_persistent_*.hip).
#pragma once
#include "affine_pk.hpp"
#include "affine_wave.hpp"
#include "myers_body.hpp"

namespace sf {

// Myers classes by register group: 1 (lane classes to 8 blocks, pipelines to 16 x 3 and 64 x 3),
// 2 (lane 12, pipelines to 16 x 6, 32 x 7 and 64 x 6), 3 (lane 16-24).
template <int GROUP>
__device__ __forceinline__ void myers_mix_block(const MyersArgs &a, const MixArgs &mx, uint32_t b) {
    uint32_t s = 0;
    while (s + 1 < mx.nseg && b >= mx.blk0[s + 1]) ++s;
    MyersArgs x = a;
    x.order += mx.off[s];
    x.desc += mx.off[s];
    x.count = mx.count[s];
    const uint32_t blk = b - mx.blk0[s];
#define L(B)                                                                                     \
    case B: myers_lane_body<B>(x, blk); break;
#define P(GG, RR)                                                                                \
    case GG * 1000 + RR: myers_pipe_body<GG, RR>(x, blk); break;
    if constexpr (GROUP == 1) {
        switch (mx.cls[s]) {
            L(1) L(2) L(4) L(8) P(16, 1) P(16, 2) P(16, 3) P(64, 2) P(64, 3)
        default:
            break; // the host launches only this group's classes (sf_myers_mix_group)
        }
    } else if constexpr (GROUP == 2) {
        switch (mx.cls[s]) {
            L(12) P(16, 4) P(16, 5) P(16, 6) P(32, 5) P(32, 7) P(64, 4) P(64, 5) P(64, 6)
        default:
            break;
        }
    } else {
        switch (mx.cls[s]) {
            L(16) L(20) L(24)
        default:
            break;
        }
    }
#undef L
#undef P
}

// Score classes of affine_mix_kernel: the 64-lane classes, the widest 8- and 16-lane packed ones and
// the 4-lane ones.
__device__ __forceinline__ void affine_mix_block(const AffineArgs &a, const MixArgs &mx, uint32_t b) {
    uint32_t s = 0;
    while (s + 1 < mx.nseg && b >= mx.blk0[s + 1]) ++s;
    AffineArgs x = a;
    x.order += mx.off[s];
    x.desc += mx.off[s];
    x.count = mx.count[s];
    const uint32_t blk = b - mx.blk0[s];
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
#define P8(KK)                                                                                   \
    case 8200 + 2 * KK: affine_pk16_body<8, KK, true>(x, blk); break;
#define P16(KK)                                                                                  \
    case 16100 + 2 * KK: affine_pk16_body<16, KK, false>(x, blk); break;                          \
    case 16200 + 2 * KK: affine_pk16_body<16, KK, true>(x, blk); break;
        P8(14) P8(15) P8(16) P8(17) P8(18) P16(14) P16(15) P16(16)
#undef P8
#undef P16
#define P4(KK)                                                                                   \
    case 4200 + 2 * KK: affine_pk16_body<4, KK, true>(x, blk); break;
        P4(18) P4(20) P4(24) P4(28) P4(32) P4(33) P4(34) P4(35) P4(36)
#undef P4
    default:
        break; // the host launches only the classes above (sf_affine_mix_group)
    }
}

// Score classes of affine_mix1_kernel: the packed 8- and 16-lane classes up to 13 registers of slot
// pairs a lane.
__device__ __forceinline__ void affine_mix1_block(const AffineArgs &a, const MixArgs &mx, uint32_t b) {
    uint32_t s = 0;
    while (s + 1 < mx.nseg && b >= mx.blk0[s + 1]) ++s;
    AffineArgs x = a;
    x.order += mx.off[s];
    x.desc += mx.off[s];
    x.count = mx.count[s];
    const uint32_t blk = b - mx.blk0[s];
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

} // namespace sf
