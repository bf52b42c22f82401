// calibrate.h -- measurement-only VALU issue-rate kernels (calibrate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace fpho {
// kind 0: 16 FP32 FMA chains per lane; kind 1: 8 FMA and 8 v_exp_f32 chains
// per lane; kind 2: 16 FMA chains in the accumulate (v_fmac) form; kind 3:
// 16 v_exp_f32 chains per lane; kinds 4 and 5: 16 FP32 multiply or add
// chains per lane; kind 6: 16 FMA chains with vector-register operands only.  256 threads per workgroup; writes
// blocks*256 floats to out.
// *instr_per_wave = VALU instructions of the measured loop per wave.
hipError_t launch_issue_calibration(int kind, int blocks, float *out, hipStream_t stream,
                                    int64_t *instr_per_wave);
}  // namespace fpho
