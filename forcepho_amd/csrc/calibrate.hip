// calibrate.hip -- measurement-only kernels: VALU issue rates of gfx950.
//
// The pixel pass is bound by VALU issue.  These kernels measure, on the
// device at hand, how many wave64 VALU instructions a SIMD issues per cycle
// for the two instruction types that dominate it: FP32 FMA (non-packed, as
// the build compiles with packed FP32 disabled) and FMA interleaved with
// v_exp_f32, FMA in its 4-byte accumulate encoding, v_exp_f32 alone (the
// exp-issue ceiling of the pixel pass), FP32 multiply and add alone, and FMA
// with vector-register operands only.  Each lane runs
// independent dependency chains so that issue,
// not latency, bounds the rate; 8 waves per SIMD, as the pixel pass runs.
#include <hip/hip_runtime.h>

#include "calibrate.h"

namespace fpho {

constexpr int kCalIters = 32768;

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void fma_chains_kernel(float *__restrict__ out, float a, float b) {
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = threadIdx.x * 1e-3f + j;
    for (int i = 0; i < kCalIters; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = fmaf(v[j], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) s += v[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void fma_exp_kernel(float *__restrict__ out, float a, float b) {
    float v[8], w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        v[j] = threadIdx.x * 1e-3f + j;
        w[j] = -1.0f - j;
    }
    for (int i = 0; i < kCalIters / 4; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v[j] = fmaf(v[j], a, b);
            w[j] = -__builtin_amdgcn_exp2f(w[j]);   // exp argument stays in [-1, 0)
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j] + w[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// as fma_chains_kernel, each step an accumulate v += k * m (the 4-byte VOP2
// encoding v_fmac_f32_e32, as most of the pixel pass's FMAs)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void fmac_chains_kernel(float *__restrict__ out, float k, float b) {
    float v[16], m[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        v[j] = threadIdx.x * 1e-3f + j;
        m[j] = b * (j + 1);
    }
    for (int i = 0; i < kCalIters; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = fmaf(k, m[j], v[j]);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) s += v[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16 independent v_exp_f32 chains per lane (the negation folds into the
// next exp's source modifier)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void exp_chains_kernel(float *__restrict__ out) {
    float w[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) w[j] = -1.0f + threadIdx.x * 1e-4f - j * 1e-2f;
    for (int i = 0; i < kCalIters / 4; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) w[j] = -__builtin_amdgcn_exp2f(w[j]);  // stays in [-1, 0)
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) s += w[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16 independent FP32 multiply (kind 4) or add (kind 5) chains per lane
template <bool MUL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void muladd_chains_kernel(float *__restrict__ out, float k) {
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = 1.0f + threadIdx.x * 1e-6f + j * 1e-3f;
    for (int i = 0; i < kCalIters; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = MUL ? v[j] * k : v[j] + k;
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) s += v[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16 independent FMA chains per lane whose operands are all vector
// registers (kind 6): v += m * k with per-lane m, k
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void fmac_vgpr_chains_kernel(float *__restrict__ out, float b) {
    float v[16], m[16], k[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        v[j] = threadIdx.x * 1e-3f + j;
        m[j] = b * (j + 1) + threadIdx.x * 1e-9f;
        k[j] = 1e-7f + threadIdx.x * 1e-12f;
    }
    for (int i = 0; i < kCalIters; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = fmaf(k[j], m[j], v[j]);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) s += v[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipError_t launch_issue_calibration(int kind, int blocks, float *out, hipStream_t stream,
                                    int64_t *instr_per_wave) {
    if (kind == 0) {
        hipLaunchKernelGGL(fma_chains_kernel, dim3(blocks), dim3(256), 0, stream, out,
                           0.999f, 0.001f);
        *instr_per_wave = 16LL * kCalIters;
    } else if (kind == 2) {
        hipLaunchKernelGGL(fmac_chains_kernel, dim3(blocks), dim3(256), 0, stream, out,
                           1e-7f, 0.001f);
        *instr_per_wave = 16LL * kCalIters;
    } else if (kind == 4) {
        hipLaunchKernelGGL(muladd_chains_kernel<true>, dim3(blocks), dim3(256), 0, stream, out,
                           0.9999999f);
        *instr_per_wave = 16LL * kCalIters;
    } else if (kind == 5) {
        hipLaunchKernelGGL(muladd_chains_kernel<false>, dim3(blocks), dim3(256), 0, stream, out,
                           1e-7f);
        *instr_per_wave = 16LL * kCalIters;
    } else if (kind == 6) {
        hipLaunchKernelGGL(fmac_vgpr_chains_kernel, dim3(blocks), dim3(256), 0, stream, out,
                           0.001f);
        *instr_per_wave = 16LL * kCalIters;
    } else if (kind == 3) {
        hipLaunchKernelGGL(exp_chains_kernel, dim3(blocks), dim3(256), 0, stream, out);
        *instr_per_wave = 16LL * (kCalIters / 4);
    } else {
        hipLaunchKernelGGL(fma_exp_kernel, dim3(blocks), dim3(256), 0, stream, out, 0.999f,
                           0.001f);
        *instr_per_wave = 16LL * (kCalIters / 4);
    }
    return hipGetLastError();
}

}  // namespace fpho
