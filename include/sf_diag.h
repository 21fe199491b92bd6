/*
 * sf_diag.h -- C ABI of libsf_diag.so: diagnostics for tests and measurements, kept out of the
 * alignment library (libsf_align.so, include/sf_align.h).  Neither call has a counterpart in
 * spearfish (crates/map/src/align.rs); nothing in the alignment path uses them.
 */
#ifndef SF_DIAG_H
#define SF_DIAG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_DIAG_EINVAL 1
#define SF_DIAG_EDEVICE 3
#define SF_DIAG_ENODEV 4

/* The calling thread's last error message. */
const char *sf_diag_last_error(void);

/* Runs a one-wave kernel that records the lane moves of the DPP and shuffle encodings the
 * alignment kernels rely on, 5 x 64 values into `out` (320 int32): wave_shl:1, wave_shr:1,
 * row_shr:1 (fill -1), an exclusive max scan of (lane * 3) % 17 (identity -100), shfl_xor 5. */
int sf_selftest_lanes(int32_t *out);

/* The VALU issue rate the kernels' rooflines are measured against.  Runs `blocks` workgroups of
 * 4 waves, each wave `iters` x 32 independent VALU instructions of one kind (0: two-operand 32-bit
 * integer ops; 1: packed 16-bit ops as in the affine kernels; 2: DPP lane moves; 3: three-operand
 * 32-bit integer ops as in the Myers step; 4: v_bitop3 alone; 5: v_alignbit alone; 6: seven v_or
 * and one v_alignbit; 7: thirty-one v_or and one v_alignbit; 8: v_add_co / v_addc pairs through
 * VCC, with the one wait state between them), once untimed, then once timed: *ms. */
int sf_probe_issue(int kind, uint32_t iters, uint32_t blocks, float *ms);

#ifdef __cplusplus
}
#endif
#endif
