// measure.hip -- libfpho_measure.so: measurement-only kernels for bench.py.
// Not part of the product library or its C ABI (include/fpho.h); nothing in
// forcepho_amd loads it.
//
//   VALU issue calibration.  The pixel pass is bound by VALU issue.  These
//   kernels measure, on the device at hand, how many wave64 VALU
//   instructions a SIMD issues per cycle for the instruction types that
//   dominate it: FP32 FMA (non-packed, as the build compiles with packed FP32
//   disabled) and FMA interleaved with v_exp_f32, FMA in its 4-byte
//   accumulate encoding, v_exp_f32 alone (the exp-issue ceiling of the pixel
//   pass), FP32 multiply and add alone, and FMA with vector-register operands
//   only.  Each lane runs independent dependency chains so that issue, not
//   latency, bounds the rate; 8 waves per SIMD, as the pixel pass runs.
//
//   In-cutoff pair count.  The number of (pixel, image Gaussian) pairs with
//   exp argument <= 36 for a batch and proposals, with the pixel pass's own
//   FP32 test: the reference algorithm's FLOP count for the roofline.
#include <hip/hip_runtime.h>

#include "measure.h"
#include "../forcepho_amd/csrc/kernels.h"

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

static hipError_t launch_issue_calibration(int kind, int blocks, float *out,
                                           hipStream_t stream, int64_t *instr_per_wave) {
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

// Pairs inside the cutoff, one wave per tile (queue order does not matter).
template <int K>
__global__ __launch_bounds__(256) void count_pairs_kernel(
    const TileDesc *__restrict__ tiles, int32_t n_tiles, int32_t waves_per_group,
    const ExpDesc *__restrict__ exps, const ImgGauss *__restrict__ gauss,
    unsigned long long *__restrict__ count) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = blockIdx.x * waves_per_group + w;
    if (tile >= n_tiles) return;
    const int lane = threadIdx.x & (kWave - 1);
    const TileDesc t = tiles[tile];
    const ExpDesc &e = exps[t.exp];
    const ImgGauss *__restrict__ g = gauss + e.gauss_off;
    const int ng = e.n_src * e.n_psf;
    constexpr float kFar = 1.0e6f;   // lanes past the tile end
    float x[K], y[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = k * kWave + lane;
        x[k] = p < t.npix ? e.xpix[t.pix0 + p] - e.x0 : kFar;
        y[k] = p < t.npix ? e.ypix[t.pix0 + p] - e.y0 : kFar;
    }
    unsigned long long n = 0;
    for (int i = 0; i < ng; ++i) {
        const ImgGauss &G = g[i];
        if (G.xhi < t.xlo || G.xlo > t.xhi || G.yhi < t.ylo || G.ylo > t.yhi) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float dx = x[k] - G.xcen;
            const float dy = y[k] - G.ycen;
            const float vx = G.sfxx * dx + G.sfxy * dy;
            const float vy = G.sfyy * dy + G.sfxy * dx;
            const float ea = dx * vx + dy * vy;
            n += __popcll(__builtin_amdgcn_ballot_w64(!(ea < kCutScaled)));
        }
    }
    if (lane == 0) atomicAdd(count, n);
}

template <int K>
static void launch_count_k(const PlanView &v, unsigned groups, unsigned long long *count) {
    hipLaunchKernelGGL(count_pairs_kernel<K>, dim3(groups), dim3(64 * v.eval.waves_per_group),
                       0, nullptr, v.eval.tiles, (int32_t)v.eval.n_tiles,
                       (int32_t)v.eval.waves_per_group, v.eval.exps, v.eval.gauss, count);
}

}  // namespace fpho

using namespace fpho;

namespace {
thread_local char g_msg[256];
int fail(const char *what, hipError_t e) {
    snprintf(g_msg, sizeof(g_msg), "%s: %s", what, hipGetErrorString(e));
    return FPHO_E_HIP;
}
}  // namespace

extern "C" {

const char *fpho_measure_last_error(void) { return g_msg; }

int fpho_measure_issue_calibration(int kind, double *ms, double *wave_instructions,
                                   int *clock_khz) {
    if (!ms || !wave_instructions || !clock_khz) {
        snprintf(g_msg, sizeof(g_msg), "null argument");
        return FPHO_E_ARG;
    }
    if (kind < 0 || kind > 6) {
        snprintf(g_msg, sizeof(g_msg), "kind must be 0 to 6 (got %d)", kind);
        return FPHO_E_ARG;
    }
    int dev = 0, cus = 0, clk = 0;
    hipError_t err;
    if ((err = hipGetDevice(&dev)) != hipSuccess) return fail("hipGetDevice", err);
    if ((err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) !=
        hipSuccess)
        return fail("hipDeviceGetAttribute", err);
    if ((err = hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, dev)) != hipSuccess)
        return fail("hipDeviceGetAttribute", err);
    const int blocks = 8 * cus;   // 8 workgroups of 4 waves: 8 waves per SIMD
    float *out = nullptr;
    if ((err = hipMalloc((void **)&out, sizeof(float) * (size_t)blocks * 256)) != hipSuccess)
        return fail("hipMalloc", err);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int64_t per_wave = 0;
    float t = 0.0f;
    err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    for (int rep = 0; rep < 3 && err == hipSuccess; ++rep) {
        if ((err = hipEventRecord(e0, nullptr)) != hipSuccess) break;
        if ((err = launch_issue_calibration(kind, blocks, out, nullptr, &per_wave)) !=
            hipSuccess)
            break;
        if ((err = hipEventRecord(e1, nullptr)) != hipSuccess) break;
        if ((err = hipEventSynchronize(e1)) != hipSuccess) break;
        err = hipEventElapsedTime(&t, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (err != hipSuccess) return fail("issue calibration", err);
    *ms = t;
    *wave_instructions = (double)blocks * 4.0 * (double)per_wave;
    *clock_khz = clk;
    return FPHO_OK;
}

int fpho_measure_batch_count_pairs(fpho_batch *b, const fpho_source *props, int64_t *n_in) {
    if (!b || !props || !n_in) {
        snprintf(g_msg, sizeof(g_msg), "null argument");
        return FPHO_E_ARG;
    }
    PlanView v;
    int rc = batch_plan_view(b, &v);
    if (rc) {
        snprintf(g_msg, sizeof(g_msg), "%s", fpho_last_error());
        return rc;
    }
    unsigned long long *cnt = nullptr;
    hipError_t err = hipMalloc((void **)&cnt, sizeof(unsigned long long));
    if (err != hipSuccess) return fail("hipMalloc", err);
    err = hipMemset(cnt, 0, sizeof(unsigned long long));
    if (err == hipSuccess)
        err = launch_convert(v.eval.exps, v.n_exp, props, const_cast<ImgGauss *>(v.eval.gauss),
                             const_cast<ImgGauss64 *>(v.eval.gauss64),
                             const_cast<int4 *>(v.eval.gboxes), v.status, nullptr);
    if (err == hipSuccess && v.eval.n_tiles > 0) {
        const int W = v.eval.waves_per_group;
        const unsigned groups = (unsigned)((v.eval.n_tiles + W - 1) / W);
        switch (v.eval.pix_per_lane) {
            case 1: launch_count_k<1>(v, groups, cnt); break;
            case 2: launch_count_k<2>(v, groups, cnt); break;
            case 4: launch_count_k<4>(v, groups, cnt); break;
            default: launch_count_k<8>(v, groups, cnt); break;
        }
        err = hipGetLastError();
    }
    unsigned long long h = 0;
    if (err == hipSuccess) err = hipMemcpy(&h, cnt, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(cnt);
    if (err != hipSuccess) return fail("pair count", err);
    *n_in = (int64_t)h;
    return FPHO_OK;
}

}  // extern "C"
