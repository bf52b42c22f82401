/* measure.h -- libfpho_measure.so, measurement-only entry points used by
 * bench.py.  Not part of the product ABI (include/fpho.h); no forcepho
 * counterpart. */
#pragma once
#include <stdint.h>
#include "../include/fpho.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *fpho_measure_last_error(void);
/* VALU issue calibration of the current device.  kind 0 times independent
 * FP32 FMA chains, kind 1 FMA chains interleaved with v_exp_f32, kind 2 FMA
 * chains in the accumulate (v_fmac) encoding, kind 3 v_exp_f32 chains alone,
 * kinds 4 and 5 FP32 multiply and add chains, kind 6 FMA chains with
 * vector-register operands only, at 8 waves per SIMD on every CU.  Returns
 * the kernel time (ms, last of three runs), the wave64 VALU instructions it
 * issued and the device's shader clock (kHz).  Synchronous; allocates. */
int fpho_measure_issue_calibration(int kind, double *ms, double *wave_instructions,
                                   int *clock_khz);
/* Number of (pixel, image Gaussian) pairs inside the cutoff (exp argument
 * <= 36) for the given proposals (device), counted with the pixel pass's own
 * FP32 test over the batch's current plan.  Synchronous; allocates; rewrites
 * the batch's image-Gaussian workspace. */
int fpho_measure_batch_count_pairs(fpho_batch *b, const fpho_source *proposals_dev,
                                   int64_t *n_in);

#ifdef __cplusplus
}
#endif
