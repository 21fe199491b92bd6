// probe.hip -- libsf_diag.so (include/sf_diag.h): diagnostics for tests and measurements, kept out
// of the alignment library.
//
// sf_selftest_lanes: a one-wave kernel that records the lane moves of the DPP and shuffle
// encodings the alignment kernels rely on.
//
// sf_probe_issue: issue-rate probes, kernels with a known number of VALU instructions, to measure
// the chip's VALU issue rate per instruction kind (the denominator of the kernels' VALU rooflines).
//
// Each wave runs `iters` iterations of 32 VALU instructions: 8 independent register chains, each
// register written every 8 instructions, so no dependency or DPP hazard stalls the wave, and 4
// waves a workgroup.  Kinds: 0 = two-operand 32-bit integer ops (VOP2: v_add, v_xor, v_and, v_or),
// 1 = packed 16-bit ops of the affine kernels (VOP3P: v_pk_max_u16, v_pk_sub_u16, v_pk_mad_u16,
// v_pk_min_u16), 2 = DPP moves (v_mov_b32_dpp row_shr:1), 3 = three-operand 32-bit integer ops of
// the Myers block step (VOP3: v_bitop3, v_alignbit, v_lshl_or, v_add3), 4 = v_bitop3 alone (three
// register operands), 5 = v_alignbit alone, 6 = seven v_or and one v_alignbit, 7 = thirty-one v_or
// and one v_alignbit, 8 = carry pairs (v_add_co_u32 then v_addc_co_u32 through VCC).  Kinds 4-8
// show when the SIMD issues two VALU instructions a quad-cycle: runs of two-operand and v_bitop3
// instructions only; one slower instruction among them costs several of their issue slots.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/sf_diag.h"
#include "internal.hpp"

using namespace sf;

namespace {

__global__ void lane_probe_kernel(int32_t *out) {
    // Lane-movement check of the DPP encodings the kernels rely on: 5 x 64 values.
    const int32_t lane = threadIdx.x;
    out[lane] = dpp<kWaveShl1>(-1, lane);
    out[64 + lane] = dpp<kWaveShr1>(-1, lane);
    out[128 + lane] = dpp<kRowShr1>(-1, lane);
    out[192 + lane] = wave_excl_max(lane * 3 % 17, -100);
    out[256 + lane] = __shfl_xor(lane, 5);
}

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
        } else if constexpr (KIND == 4) {
#define E1(K, A, B) "v_bitop3_b32 %" #K ", %" #K ", %" #A ", %" #B " bitop3:0x96\n\t"
#define E8 E1(0, 1, 2) E1(1, 2, 3) E1(2, 3, 4) E1(3, 4, 5) E1(4, 5, 6) E1(5, 6, 7) E1(6, 7, 0) E1(7, 0, 1)
            asm volatile(E8 E8 E8 E8
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d));
#undef E8
#undef E1
        } else if constexpr (KIND == 5) {
#define E1(K) "v_alignbit_b32 %" #K ", %" #K ", %9, 31\n\t"
            asm volatile(SF_R8(E1) SF_R8(E1) SF_R8(E1) SF_R8(E1)
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d));
#undef E1
        } else if constexpr (KIND == 6 || KIND == 7) {
#define F1(K, A) "v_or_b32_e32 %" #K ", %" #A ", %" #K "\n\t"
#define S1 "v_alignbit_b32 %0, %0, %9, 31\n\t"
#define F7 F1(1, 2) F1(2, 3) F1(3, 4) F1(4, 5) F1(5, 6) F1(6, 7) F1(7, 0)
#define F8 F1(0, 1) F7
            if constexpr (KIND == 6)
                asm volatile(F7 S1 F7 S1 F7 S1 F7 S1
                             : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                             : "v"(c), "v"(d));
            else
                asm volatile(F8 F8 F8 F7 S1
                             : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                             : "v"(c), "v"(d));
#undef F8
#undef F7
#undef S1
#undef F1
        } else if constexpr (KIND == 8) {
#define G1(K, A) "v_add_co_u32_e32 %" #K ", vcc, %" #K ", %8\n\ts_nop 0\n\tv_addc_co_u32_e32 %" #A ", vcc, %" #A ", %9, vcc\n\t"
#define G4 G1(0, 1) G1(2, 3) G1(4, 5) G1(6, 7)
            asm volatile(G4 G4 G4 G4
                         : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7)
                         : "v"(c), "v"(d)
                         : "vcc");
#undef G4
#undef G1
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

hipError_t launch_issue_probe(int kind, uint32_t iters, uint32_t blocks, uint32_t *out, hipStream_t s) {
    switch (kind) {
    case 0: hipLaunchKernelGGL(issue_probe_kernel<0>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 1: hipLaunchKernelGGL(issue_probe_kernel<1>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 2: hipLaunchKernelGGL(issue_probe_kernel<2>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 3: hipLaunchKernelGGL(issue_probe_kernel<3>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 4: hipLaunchKernelGGL(issue_probe_kernel<4>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 5: hipLaunchKernelGGL(issue_probe_kernel<5>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 6: hipLaunchKernelGGL(issue_probe_kernel<6>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 7: hipLaunchKernelGGL(issue_probe_kernel<7>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    case 8: hipLaunchKernelGGL(issue_probe_kernel<8>, dim3(blocks), dim3(256), 0, s, iters, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int ensure_gfx950() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(SF_DIAG_ENODEV, "no HIP device visible");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return fail(SF_DIAG_EDEVICE, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SF_DIAG_ENODEV, "device 0 is %s, this library is built for gfx950", prop.gcnArchName);
    if (hipSetDevice(0) != hipSuccess) return fail(SF_DIAG_EDEVICE, "hipSetDevice");
    return 0;
}

} // namespace

extern "C" {

const char *sf_diag_last_error(void) { return g_err.c_str(); }

int sf_selftest_lanes(int32_t *out) {
    if (!out) return fail(SF_DIAG_EINVAL, "out is null");
    if (const int st = ensure_gfx950()) return st;
    int32_t *d = nullptr;
    if (hipMalloc((void **)&d, 320 * 4) != hipSuccess) return fail(SF_DIAG_EDEVICE, "hipMalloc");
    hipLaunchKernelGGL(lane_probe_kernel, dim3(1), dim3(64), 0, nullptr, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, d, 320 * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SF_DIAG_EDEVICE, "lane probe: %s", hipGetErrorString(e));
    return 0;
}

int sf_probe_issue(int kind, uint32_t iters, uint32_t blocks, float *ms) {
    if (kind < 0 || kind > 8 || blocks == 0 || !ms) return fail(SF_DIAG_EINVAL, "probe: kind 0..8, blocks > 0, ms");
    if (const int st = ensure_gfx950()) return st;
    uint32_t *d = nullptr;
    if (hipMalloc((void **)&d, (uint64_t)blocks * 256 * 4) != hipSuccess) return fail(SF_DIAG_EDEVICE, "hipMalloc");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_issue_probe(kind, iters, blocks, d, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = launch_issue_probe(kind, iters, blocks, d, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(SF_DIAG_EDEVICE, "probe: %s", hipGetErrorString(e));
    return 0;
}

} // extern "C"
