// probe.hip -- issue-rate probes: kernels with a known number of VALU instructions, to measure the
// chip's VALU issue rate per instruction kind (the denominator of the kernels' VALU rooflines).
//
// Each wave runs `iters` iterations of 32 VALU instructions: 8 independent register chains, each
// register written every 8 instructions, so no dependency or DPP hazard stalls the wave, and 4
// waves a workgroup.  Kinds: 0 = two-operand 32-bit integer ops (VOP2: v_add, v_xor, v_and, v_or),
// 1 = packed 16-bit ops of the affine kernels (VOP3P: v_pk_max_u16, v_pk_sub_u16, v_pk_mad_u16,
// v_pk_min_u16), 2 = DPP moves (v_mov_b32_dpp row_shr:1), 3 = three-operand 32-bit integer ops of
// the Myers block step (VOP3: v_bitop3, v_alignbit, v_lshl_or, v_add3).
#include "internal.hpp"

namespace {

#define SF_R8(OP) OP(0) OP(1) OP(2) OP(3) OP(4) OP(5) OP(6) OP(7)

template <int KIND>
__global__ __launch_bounds__(256) void issue_probe_kernel(uint32_t iters, uint32_t *out) {
    uint32_t r0 = threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6,
             r7 = r0 + 7;
    const uint32_t c = blockIdx.x | 1u, d = (blockIdx.x << 3) | 5u;
    for (uint32_t i = 0; i < iters; ++i) {
        if constexpr (KIND == 0) {
#define A1(K) "v_add_u32_e32 %" #K ", %8, %" #K "\n\t"
#define A2(K) "v_xor_b32_e32 %" #K ", %9, %" #K "\n\t"
#define A3(K) "v_and_b32_e32 %" #K ", %8, %" #K "\n\t"
#define A4(K) "v_or_b32_e32 %" #K ", %9, %" #K "\n\t"
            asm volatile(SF_R8(A1) SF_R8(A2) SF_R8(A3) SF_R8(A4)
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d));
#undef A1
#undef A2
#undef A3
#undef A4
        } else if constexpr (KIND == 3) {
#define D1(K) "v_bitop3_b32 %" #K ", %" #K ", %8, %9 bitop3:0x96\n\t"
#define D2(K) "v_alignbit_b32 %" #K ", %" #K ", %9, 7\n\t"
#define D3(K) "v_lshl_or_b32 %" #K ", %" #K ", 1, %8\n\t"
#define D4(K) "v_add3_u32 %" #K ", %" #K ", %8, %9\n\t"
            asm volatile(SF_R8(D1) SF_R8(D2) SF_R8(D3) SF_R8(D4)
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d));
#undef D1
#undef D2
#undef D3
#undef D4
        } else if constexpr (KIND == 1) {
#define B1(K) "v_pk_max_u16 %" #K ", %" #K ", %8\n\t"
#define B2(K) "v_pk_sub_u16 %" #K ", %" #K ", %9\n\t"
#define B3(K) "v_pk_mad_u16 %" #K ", %" #K ", %8, %9\n\t"
#define B4(K) "v_pk_min_u16 %" #K ", %" #K ", %9\n\t"
            asm volatile(SF_R8(B1) SF_R8(B2) SF_R8(B3) SF_R8(B4)
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d));
#undef B1
#undef B2
#undef B3
#undef B4
        } else {
#define C1(K) "v_mov_b32_dpp %" #K ", %" #K " row_shr:1 row_mask:0xf bank_mask:0xf\n\t"
            asm volatile(SF_R8(C1) SF_R8(C1) SF_R8(C1) SF_R8(C1)
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d));
#undef C1
        }
    }
    out[blockIdx.x * 256u + threadIdx.x] = r0 ^ r1 ^ r2 ^ r3 ^ r4 ^ r5 ^ r6 ^ r7;
}

#undef SF_R8

} // namespace

hipError_t sf_launch_issue_probe(int kind, uint32_t iters, uint32_t blocks, uint32_t *out, hipStream_t s) {
    switch (kind) {
    case 0: hipLaunchKernelGGL(issue_probe_kernel<0>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 1: hipLaunchKernelGGL(issue_probe_kernel<1>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 2: hipLaunchKernelGGL(issue_probe_kernel<2>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 3: hipLaunchKernelGGL(issue_probe_kernel<3>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
