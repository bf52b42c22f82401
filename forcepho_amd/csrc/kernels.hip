// kernels.hip -- gfx950 kernels of the forcepho likelihood-and-gradient path.
//
//   convert_kernel      proposal -> image-plane Gaussians + Jacobians
//                       (forcepho/src/compute_gaussians_kernel.cu:45-127,
//                        forcepho/src/common_kernel.cc:175-240)
//   eval_kernel<K>      per-pixel mixture, residual, chi2 and the 7 gradient
//                       partials per source (common_kernel.cc:57-143,
//                       compute_gaussians_kernel.cu:262-301); one wave per
//                       tile of 64*K pixels, no LDS, no float atomics.
//   reduce_kernel       fixed-order sum of tile partials (FP64 or FP32)
//                       (replaces the shared-memory Accumulator,
//                        compute_gaussians_kernel.cu:133-198)
//   theta_kernel        parameter vector -> fpho_source structs, Sersic
//                       amplitudes from bicubic B-splines in FP64
//                       (forcepho/sources.py:607-694)
//   stack_kernel        lnlike = -chi2/2 and stacked gradient
//                       (forcepho/model.py:331-332,370-387)
#include <hip/hip_runtime.h>

#include "fpho_internal.h"
#include "kernels.h"

namespace fpho {

// ------------------------------------------------------------------ convert

// Build one image Gaussian in FP64: the source set-up of
// compute_gaussians_kernel.cu:76-124 and the convolution and Jacobian of
// GetGaussianAndJacobian (common_kernel.cc:175-240).  The per-Gaussian cost
// is negligible next to the pixel pass, so it runs in double and is rounded
// once into the FP32 record the pixel pass streams.
__device__ __forceinline__ void make_image_gaussian(const ExpDesc &e,
                                                    const fpho_source &src, int s,
                                                    int p, ImgGauss &g,
                                                    ImgGauss64 &g64) {
    using M = M22<double>;
    const float *Dp = e.D + 4 * s;
    const float *CWp = e.CW + 4 * s;
    const M D{Dp[0], Dp[1], Dp[2], Dp[3]};

    double sn, cs;
    sincos((double)src.pa, &sn, &cs);
    const double q = src.q;
    const M R{cs, -sn, sn, cs};
    const M dR{-sn, -cs, cs, -sn};
    const double iq = 1.0 / q;
    const M S{iq, 0.0, 0.0, q};
    const M dS{-iq * iq, 0.0, 0.0, 1.0};

    const M DR = mul(D, R);
    const M T = mul(DR, S);
    const M dT_dq = mul(DR, dS);
    const M dT_dpa = mul(mul(D, dR), S);

    // source centre in pixels: CW (ra - crval, dec - crval) + crpix
    const double s0 = (double)src.ra - (double)e.crval[2 * s + 0];
    const double s1 = (double)src.dec - (double)e.crval[2 * s + 1];
    const double xs = (double)CWp[0] * s0 + (double)CWp[1] * s1 + (double)e.crpix[2 * s + 0];
    const double ys = (double)CWp[2] * s0 + (double)CWp[3] * s1 + (double)e.crpix[2 * s + 1];

    const fpho_psf_gauss pg = e.psf[p];
    const int bin = pg.sersic_radius_bin;
    const double r2 = e.rad2[bin];

    // covariance = r2 T T^T + C_psf ; F = covariance^-1
    const M Tt = transpose(T);
    const M cov = add(scale(mul(T, Tt), r2), M{pg.Cxx, pg.Cxy, pg.Cxy, pg.Cyy});
    const M F = inverse(cov);
    const double detF = det(F);

    const double flux = src.fluxes[e.band];
    const double am = src.mixture_amplitudes[bin];
    const double tmp = (double)e.G * (double)pg.amp * sqrt(detF) * (0.5 / 3.14159265358979323846);
    const double amp = tmp * flux * am;
    const double xc = xs + (double)pg.xcen;
    const double yc = ys + (double)pg.ycen;

    // dSigma = r2 (T dT^T + dT T^T); dF = -F dSigma F
    const M dSq = scale(add(mul(T, transpose(dT_dq)), mul(dT_dq, Tt)), r2);
    const M dSp = scale(add(mul(T, transpose(dT_dpa)), mul(dT_dpa, Tt)), r2);
    const M dFq = scale(mul(F, mul(dSq, F)), -1.0);
    const M dFp = scale(mul(F, mul(dSp, F)), -1.0);

    // bounding box of the cutoff ellipse d^T F d <= 36: |dx| <= 6 sqrt(cov_xx),
    // widened (0.1% + 0.01 pix) so that no pair the FP32 test accepts is
    // outside it; skipping boxed-out tiles then changes no result bit.
    const double hx = 6.0 * sqrt(cov.a) * 1.001 + 0.01;
    const double hy = 6.0 * sqrt(cov.d) * 1.001 + 0.01;
    g.xlo = order_key((float)(xc - hx));
    g.xhi = order_key((float)(xc + hx));
    g.ylo = order_key((float)(yc - hy));
    g.yhi = order_key((float)(yc + hy));
    g.xcen = (float)xc;
    g.ycen = (float)yc;
    g.fxx = (float)F.a;
    g.fyy = (float)F.d;
    g.fxy = (float)F.c;
    g.amp = (float)amp;
    g.hc = (float)(1.0 - (F.a + F.d) * (1.0 / 24.0));
    g.pad0 = 0.0f;
    g.dA_dQ = (float)(0.5 * amp * trace(mul(cov, dFq)));
    g.dA_dPA = (float)(0.5 * amp * trace(mul(cov, dFp)));
    g.dA_dFlux = (float)(tmp * am);
    g.dA_dSersic = (float)(tmp * flux * src.damplitude_dnsersic[bin]);
    g.dA_drh = (float)(tmp * flux * src.damplitude_drh[bin]);
    g.dFxx_dQ = (float)dFq.a;
    g.dFyy_dQ = (float)dFq.d;
    g.dFxy_dQ = (float)dFq.c;
    g.dFxx_dPA = (float)dFp.a;
    g.dFyy_dPA = (float)dFp.d;
    g.dFxy_dPA = (float)dFp.c;
    g.pad1 = 0.0f;
    g.dx_dAlpha = CWp[0];
    g.dy_dAlpha = CWp[2];
    g.dx_dDelta = CWp[1];
    g.dy_dDelta = CWp[3];
    g.pad2[0] = g.pad2[1] = g.pad2[2] = g.pad2[3] = 0.0f;

    g64.xcen = xc;
    g64.ycen = yc;
    g64.fxx = F.a;
    g64.fyy = F.d;
    g64.fxy = F.c;
    g64.amp = amp;
    g64.pad0 = 0.0;
    g64.pad1 = 0.0;
}

// One workgroup per exposure; lanes stride over its n_src * n_psf Gaussians.
__global__ __launch_bounds__(128) void convert_kernel(
    const ExpDesc *__restrict__ exps, const fpho_source *__restrict__ props,
    ImgGauss *__restrict__ out, ImgGauss64 *__restrict__ out64) {
    const ExpDesc e = exps[blockIdx.x];
    const int n = e.n_src * e.n_psf;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int s = j / e.n_psf;
        const int p = j - s * e.n_psf;
        ImgGauss g;
        ImgGauss64 g64;
        make_image_gaussian(e, props[e.src_off + s], s, p, g, g64);
        out[e.gauss_off + j] = g;
        out64[e.gauss_off + j] = g64;
    }
}

// --------------------------------------------------------------- pixel pass

// exp(-E/2) as one hardware exp2: exp2(E * (-log2(e)/2)).
constexpr float kHalfLog2e = -0.72134752044448170368f;
constexpr float kMaskedCoord = 1.0e6f;  // lanes past the tile end: far away

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// True when the Gaussian's cutoff box (gb = xlo, xhi, ylo, yhi keys) misses
// the tile's pixel box tb; one 16-byte scalar load, integer compares only.
__device__ __forceinline__ bool boxed_out(int cull, const int4 gb, const int4 tb) {
    const int miss = (gb.y < tb.x ? 1 : 0) | (gb.x > tb.y ? 2 : 0) | (gb.w < tb.z ? 4 : 0) |
                     (gb.z > tb.w ? 8 : 0);
    return cull != 0 && miss != 0;
}

// Value of one image Gaussian at one pixel in FP64 (bright terms only).
__device__ __forceinline__ double pair_value64(const ImgGauss64 &g, float xp, float yp) {
    const double dx = (double)xp - g.xcen;
    const double dy = (double)yp - g.ycen;
    const double vx = g.fxx * dx + g.fxy * dy;
    const double vy = g.fyy * dy + g.fxy * dx;
    const double ea = dx * vx + dy * vy;
    const double h = 1.0 + (vx * vx + vy * vy - g.fxx - g.fyy) * (1.0 / 24.0);
    return g.amp * exp(-0.5 * ea) * h;
}

// Evaluate one wave tile.  `tile` must be wave-uniform (an SGPR value) so
// every descriptor and Gaussian load below is a scalar load.
//   MIXED = true : residual accumulated in FP64; a term whose FP32 value
//                  exceeds kBrightSigma * sigma of its pixel is re-evaluated
//                  in FP64 (a wave branch taken only near bright cores).
//   MIXED = false: FP32 residual, as forcepho/src/common_kernel.cc:57-83.
template <int K, bool MIXED>
__device__ __forceinline__ void eval_tile(int tile, const TileDesc *__restrict__ tiles,
                                          const ExpDesc *__restrict__ exps,
                                          const ImgGauss *__restrict__ gauss,
                                          const ImgGauss64 *__restrict__ gauss64,
                                          float *__restrict__ gpart,
                                          double *__restrict__ cpart, int cull) {
    const int lane = threadIdx.x & (kWave - 1);
    const TileDesc t = tiles[tile];
    const ExpDesc &e = exps[t.exp];
    const ImgGauss *__restrict__ g = gauss + e.gauss_off;
    const ImgGauss64 *__restrict__ g64 = gauss64 + e.gauss_off;
    const int n_src = e.n_src;
    const int n_psf = e.n_psf;
    const int ng = n_src * n_psf;

    float x[K], y[K], d[K], ie[K], r[K], thr[K];
    double rd[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = k * kWave + lane;
        if (p < t.npix) {
            const int i = t.pix0 + p;
            x[k] = e.xpix[i];
            y[k] = e.ypix[i];
            d[k] = e.data[i];
            ie[k] = e.ierr[i];
        } else {
            x[k] = kMaskedCoord;
            y[k] = kMaskedCoord;
            d[k] = 0.0f;
            ie[k] = 0.0f;
        }
        r[k] = d[k];
        rd[k] = (double)d[k];
        thr[k] = ie[k] > 0.0f ? kBrightSigma / ie[k] : __builtin_huge_valf();
    }

    // A Gaussian whose cutoff box misses the tile's pixel box contributes
    // nothing to any pixel of the tile (scalar test, no vector work).
    const int4 tbox = make_int4(t.xlo, t.xhi, t.ylo, t.yhi);
#define FPHO_BOXED_OUT(G) boxed_out(cull, *reinterpret_cast<const int4 *>(&(G).xlo), tbox)

    // ---- pass 1: residual = data - sum_g C_g   (common_kernel.cc:57-83)
    for (int i = 0; i < ng; ++i) {
        if (FPHO_BOXED_OUT(g[i])) continue;
        const float xc = g[i].xcen, yc = g[i].ycen;
        const float fxx = g[i].fxx, fyy = g[i].fyy, fxy = g[i].fxy;
        const float amp = g[i].amp, hc = g[i].hc;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float dx = x[k] - xc;
            const float dy = y[k] - yc;
            const float vx = fxx * dx + fxy * dy;
            const float vy = fyy * dy + fxy * dx;
            const float ea = dx * vx + dy * vy;
            if (!(ea > FPHO_MAX_EXP_ARG)) {
                const float gp = __builtin_amdgcn_exp2f(ea * kHalfLog2e);
                const float h = hc + (vx * vx + vy * vy) * (1.0f / 24.0f);
                const float c = amp * gp * h;
                if (MIXED) {
                    if (fabsf(c) > thr[k])
                        rd[k] -= pair_value64(g64[i], x[k], y[k]);
                    else
                        rd[k] -= (double)c;
                } else {
                    r[k] -= c;
                }
            }
        }
    }

    // ---- chi2 = sum (r^2 - d^2) ierr^2 in FP64 (compute_gaussians_kernel.cu:276-287)
    double c2 = 0.0;
    float rr[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = k * kWave + lane;
        if (MIXED) {
            r[k] = (float)rd[k];
            const double rs = rd[k] * (double)ie[k];
            const double ds = (double)d[k] * (double)ie[k];
            c2 += (rs - ds) * (rs + ds);
        } else {
            const float rs = r[k] * ie[k];
            const float ds = d[k] * ie[k];
            double c = (double)rs - (double)ds;
            c *= (double)rs + (double)ds;
            c2 += c;
        }
        if (e.residual != nullptr && p < t.npix) e.residual[t.pix0 + p] = r[k];
        rr[k] = -2.0f * ((r[k] * ie[k]) * ie[k]);  // -2 r ierr^2
    }
    c2 = wave_sum(c2);
    if (lane == 0) cpart[tile] = c2;

    // ---- pass 2: per source, 7 partials   (common_kernel.cc:95-143)
    float *__restrict__ out = gpart + t.part_off;
    for (int s = 0; s < n_src; ++s) {
        float a0 = 0.f, sx = 0.f, sy = 0.f, a3 = 0.f, a4 = 0.f, a5 = 0.f, a6 = 0.f;
        const int i0 = s * n_psf;
        bool touched = false;
        for (int i = i0; i < i0 + n_psf; ++i) {
            const ImgGauss &G = g[i];
            if (FPHO_BOXED_OUT(G)) continue;
            touched = true;
            const float xc = G.xcen, yc = G.ycen;
            const float fxx = G.fxx, fyy = G.fyy, fxy = G.fxy;
            const float amp = G.amp, hc = G.hc;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float dx = x[k] - xc;
                const float dy = y[k] - yc;
                const float vx = fxx * dx + fxy * dy;
                const float vy = fyy * dy + fxy * dx;
                const float ea = dx * vx + dy * vy;
                if (!(ea > FPHO_MAX_EXP_ARG)) {
                    const float gp = __builtin_amdgcn_exp2f(ea * kHalfLog2e);
                    const float h = hc + (vx * vx + vy * vy) * (1.0f / 24.0f);
                    float dC_dA = rr[k] * gp;
                    const float c_h = dC_dA * amp;
                    dC_dA *= h;
                    const float C = c_h * h;
                    const float dC_dx = C * vx - c_h * (fxx * vx + fxy * vy) * (1.0f / 12.0f);
                    const float dC_dy = C * vy - c_h * (fyy * vy + fxy * vx) * (1.0f / 12.0f);
                    const float dC_dfx = -0.5f * C * dx * dx - c_h * (1.0f - 2.0f * dx * vx) * (1.0f / 24.0f);
                    const float dC_dfy = -0.5f * C * dy * dy - c_h * (1.0f - 2.0f * dy * vy) * (1.0f / 24.0f);
                    const float dC_dfxy = -C * dx * dy + c_h * (dy * vx + dx * vy) * (1.0f / 12.0f);
                    a0 += G.dA_dFlux * dC_dA;
                    sx += dC_dx;
                    sy += dC_dy;
                    a3 += G.dA_dQ * dC_dA + dC_dfx * G.dFxx_dQ + dC_dfxy * G.dFxy_dQ + dC_dfy * G.dFyy_dQ;
                    a4 += G.dA_dPA * dC_dA + dC_dfx * G.dFxx_dPA + dC_dfxy * G.dFxy_dPA + dC_dfy * G.dFyy_dPA;
                    a5 += G.dA_dSersic * dC_dA;
                    a6 += G.dA_drh * dC_dA;
                }
            }
        }
        float *o = out + kNParams * s;
        if (!touched) {
            // no Gaussian of this source reaches the tile
            if (lane < kNParams) o[lane] = 0.0f;
            continue;
        }
        a0 = wave_sum(a0);
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        a3 = wave_sum(a3);
        a4 = wave_sum(a4);
        a5 = wave_sum(a5);
        a6 = wave_sum(a6);
        if (lane == 0) {
            // the sky-position derivatives share the source's CW matrix
            const float *CW = e.CW + 4 * s;
            o[0] = a0;
            o[1] = CW[0] * sx + CW[2] * sy;  // dx/dalpha, dy/dalpha
            o[2] = CW[1] * sx + CW[3] * sy;  // dx/ddelta, dy/ddelta
            o[3] = a3;
            o[4] = a4;
            o[5] = a5;
            o[6] = a6;
        }
    }
#undef FPHO_BOXED_OUT
}

// Non-persistent: wave w of workgroup b evaluates queue[b*W + w].
template <int K, bool MIXED>
__global__ __launch_bounds__(256) void eval_kernel(
    const TileDesc *__restrict__ tiles, const int32_t *__restrict__ queue,
    int32_t n_tiles, int32_t waves_per_group, const ExpDesc *__restrict__ exps,
    const ImgGauss *__restrict__ gauss, const ImgGauss64 *__restrict__ gauss64,
    float *__restrict__ gpart, double *__restrict__ cpart, int32_t cull) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = blockIdx.x * waves_per_group + w;
    if (slot >= n_tiles) return;
    const int tile = __builtin_amdgcn_readfirstlane(queue[slot]);
    eval_tile<K, MIXED>(tile, tiles, exps, gauss, gauss64, gpart, cpart, cull);
}

// Persistent: every wave pulls tiles from `counter` until the queue is empty.
// Each wave reaches the exit once the counter passes n_tiles.
template <int K, bool MIXED>
__global__ __launch_bounds__(256) void eval_persistent_kernel(
    const TileDesc *__restrict__ tiles, const int32_t *__restrict__ queue,
    int32_t n_tiles, int32_t *__restrict__ counter, const ExpDesc *__restrict__ exps,
    const ImgGauss *__restrict__ gauss, const ImgGauss64 *__restrict__ gauss64,
    float *__restrict__ gpart, double *__restrict__ cpart, int32_t cull) {
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        int slot = 0;
        if (lane == 0) slot = atomicAdd(counter, 1);
        slot = __builtin_amdgcn_readfirstlane(__shfl(slot, 0, kWave));
        if (slot >= n_tiles) break;
        const int tile = __builtin_amdgcn_readfirstlane(queue[slot]);
        eval_tile<K, MIXED>(tile, tiles, exps, gauss, gauss64, gpart, cpart, cull);
    }
}

__global__ void zero_counter_kernel(int32_t *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// ------------------------------------------------------------------ reduce

// One workgroup per (patch, band).  Column j of the band's [ntiles][n_src*7]
// partial block is summed by one lane in tile order; chi2 partials by a
// fixed-shape tree.  The result does not depend on tile execution order.
template <typename Acc>
__global__ __launch_bounds__(256) void reduce_kernel(
    const OutDesc *__restrict__ outs, const float *__restrict__ gpart,
    const double *__restrict__ cpart, double *__restrict__ chi2,
    double *__restrict__ dchi2) {
    __shared__ Acc sh[256];
    const OutDesc o = outs[blockIdx.x];
    const int ncol = o.n_src * kNParams;
    for (int j = threadIdx.x; j < ncol; j += blockDim.x) {
        const float *col = gpart + o.part0 + j;
        Acc acc = 0;
        int t = 0;
        for (; t + 4 <= o.ntiles; t += 4) {
            const float v0 = col[(int64_t)(t + 0) * ncol];
            const float v1 = col[(int64_t)(t + 1) * ncol];
            const float v2 = col[(int64_t)(t + 2) * ncol];
            const float v3 = col[(int64_t)(t + 3) * ncol];
            acc += (Acc)v0;
            acc += (Acc)v1;
            acc += (Acc)v2;
            acc += (Acc)v3;
        }
        for (; t < o.ntiles; ++t) acc += (Acc)col[(int64_t)t * ncol];
        dchi2[o.dchi2_out + j] = (double)acc;
    }
    Acc c = 0;
    for (int t = threadIdx.x; t < o.ntiles; t += blockDim.x) c += (Acc)cpart[o.tile0 + t];
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) chi2[o.chi2_out] = (double)sh[0];
}

// ------------------------------------------------- proposals from parameters

// Cubic B-spline basis on knot span l (t[l] <= x < t[l+1]): N[0..3] are the
// values of N_{l-3..l,3}(x), dN their derivatives (Cox-de Boor recursion).
__device__ inline void bspline3(const double *t, int l, double x, double N[4],
                                double dN[4]) {
    double left[4], right[4], N2[3];
    N[0] = 1.0;
    for (int j = 1; j <= 3; ++j) {
        left[j] = x - t[l + 1 - j];
        right[j] = t[l + j] - x;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            const double tmp = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        N[j] = saved;
        if (j == 2) {
            N2[0] = N[0];
            N2[1] = N[1];
            N2[2] = N[2];
        }
    }
    // d/dx N_{i,3} = 3 (N_{i,2}/(t[i+3]-t[i]) - N_{i+1,2}/(t[i+4]-t[i+1])),
    // N2[r] = N_{l-2+r,2}; N_{l-3,2} = N_{l+1,2} = 0 on this span.
    for (int r = 0; r < 4; ++r) {
        const int i = l - 3 + r;
        const double a = (r >= 1) ? N2[r - 1] / (t[i + 3] - t[i]) : 0.0;
        const double b = (r <= 2) ? N2[r] / (t[i + 4] - t[i + 1]) : 0.0;
        dN[r] = 3.0 * (a - b);
    }
}

__device__ inline int knot_span(const double *t, int n_t, double x) {
    // largest l in [3, n_t-5] with t[l] <= x  (x already clamped)
    int lo = 3, hi = n_t - 5;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(64) void theta_kernel(
    const SrcDesc *__restrict__ srcs, int32_t n_srcs, const double *__restrict__ theta,
    SplineView sp, fpho_source *__restrict__ props) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_srcs) return;
    const SrcDesc sd = srcs[i];
    const double *th = theta + sd.theta_off;
    fpho_source o;
    for (int b = 0; b < FPHO_MAXBANDS; ++b) o.fluxes[b] = 0.0f;
    for (int m = 0; m < FPHO_MAXRADII; ++m) {
        o.mixture_amplitudes[m] = 0.0f;
        o.damplitude_drh[m] = 0.0f;
        o.damplitude_dnsersic[m] = 0.0f;
    }
    const int nb = sd.n_bands;
    for (int b = 0; b < nb; ++b) o.fluxes[b] = (float)th[b];
    const double ra = th[nb + 0], dec = th[nb + 1], q = th[nb + 2], pa = th[nb + 3];
    const double n = th[nb + 4], rh = th[nb + 5];
    o.ra = (float)ra;
    o.dec = (float)dec;
    o.q = (float)q;
    o.pa = (float)pa;
    o.nsersic = (float)n;
    o.rh = (float)rh;

    if (n < sp.nsersic_min) {
        // one-hot at the radius closest to rh (forcepho/sources.py:662-665)
        int best = 0;
        double bd = fabs(rh - sp.radii[0]);
        for (int m = 1; m < sp.n_amp; ++m) {
            const double dd = fabs(rh - sp.radii[m]);
            if (dd < bd) {
                bd = dd;
                best = m;
            }
        }
        o.mixture_amplitudes[best] = 1.0f;
    } else if (rh < 0.0) {
        o.mixture_amplitudes[0] = 1.0f;  // forcepho/sources.py:666-668
    } else {
        const double xn = fmin(fmax(n, sp.tx[3]), sp.tx[sp.n_tx - 4]);
        const double xr = fmin(fmax(rh, sp.ty[3]), sp.ty[sp.n_ty - 4]);
        const int lx = knot_span(sp.tx, sp.n_tx, xn);
        const int ly = knot_span(sp.ty, sp.n_ty, xr);
        double Nx[4], dNx[4], Ny[4], dNy[4];
        bspline3(sp.tx, lx, xn, Nx, dNx);
        bspline3(sp.ty, ly, xr, Ny, dNy);
        const int ny = sp.n_ty - 4;
        const int nx = sp.n_tx - 4;
        for (int m = 0; m < sp.n_amp; ++m) {
            const double *c = sp.coeffs + (int64_t)m * nx * ny;
            double v = 0.0, vn = 0.0, vr = 0.0;
            for (int a = 0; a < 4; ++a) {
                const double *row = c + (int64_t)(lx - 3 + a) * ny + (ly - 3);
                double rv = 0.0, rd = 0.0;
                for (int b = 0; b < 4; ++b) {
                    rv += row[b] * Ny[b];
                    rd += row[b] * dNy[b];
                }
                v += Nx[a] * rv;
                vn += dNx[a] * rv;
                vr += Nx[a] * rd;
            }
            o.mixture_amplitudes[m] = (float)v;
            o.damplitude_dnsersic[m] = (float)vn;
            o.damplitude_drh[m] = (float)vr;
        }
    }
    props[sd.proposal] = o;
}

// ------------------------------------------------------- gradient stacking

__global__ __launch_bounds__(64) void stack_kernel(
    const SrcDesc *__restrict__ srcs, int32_t n_srcs, const PatchOut *__restrict__ pouts,
    int32_t n_patches, const double *__restrict__ chi2, const double *__restrict__ dchi2,
    double *__restrict__ lnlike, double *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_srcs) {
        const SrcDesc sd = srcs[i];
        const double *dv = dchi2 + sd.dchi2_base;
        double *gv = grad + sd.theta_off;
        const int nb = sd.n_bands;
        for (int b = 0; b < nb; ++b) gv[b] = -0.5 * dv[(int64_t)b * sd.src_stride];
        for (int j = 1; j < kNParams; ++j) {
            double acc = 0.0;
            for (int b = 0; b < nb; ++b) acc += dv[(int64_t)b * sd.src_stride + j];
            gv[nb + j - 1] = -0.5 * acc;
        }
    }
    if (i < n_patches) {
        const PatchOut po = pouts[i];
        double acc = 0.0;
        for (int b = 0; b < po.n_bands; ++b) acc += chi2[po.chi2_off + b];
        lnlike[i] = -0.5 * acc;
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_convert(const ExpDesc *exps, int64_t n_exp, const fpho_source *props,
                          ImgGauss *out, ImgGauss64 *out64, hipStream_t stream) {
    if (n_exp <= 0) return hipSuccess;
    hipLaunchKernelGGL(convert_kernel, dim3((unsigned)n_exp), dim3(128), 0, stream,
                       exps, props, out, out64);
    return hipGetLastError();
}

template <int K, bool MIXED>
static hipError_t launch_eval_k(const EvalArgs &a, hipStream_t stream) {
    const int W = a.waves_per_group;
    if (a.persistent) {
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, a.counter);
        hipLaunchKernelGGL((eval_persistent_kernel<K, MIXED>), dim3((unsigned)a.grid),
                           dim3(64 * W), 0, stream, a.tiles, a.queue, (int32_t)a.n_tiles,
                           a.counter, a.exps, a.gauss, a.gauss64, a.gpart, a.cpart,
                           (int32_t)a.cull);
    } else {
        const int64_t groups = (a.n_tiles + W - 1) / W;
        hipLaunchKernelGGL((eval_kernel<K, MIXED>), dim3((unsigned)groups), dim3(64 * W), 0,
                           stream, a.tiles, a.queue, (int32_t)a.n_tiles, (int32_t)W, a.exps,
                           a.gauss, a.gauss64, a.gpart, a.cpart, (int32_t)a.cull);
    }
    return hipGetLastError();
}

template <bool MIXED>
static hipError_t launch_eval_p(const EvalArgs &a, hipStream_t stream) {
    switch (a.pix_per_lane) {
        case 1: return launch_eval_k<1, MIXED>(a, stream);
        case 2: return launch_eval_k<2, MIXED>(a, stream);
        case 4: return launch_eval_k<4, MIXED>(a, stream);
        case 8: return launch_eval_k<8, MIXED>(a, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_eval(const EvalArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    return a.precision == kPrecMixed ? launch_eval_p<true>(a, stream)
                                     : launch_eval_p<false>(a, stream);
}

hipError_t launch_reduce(const OutDesc *outs, int64_t n_outs, const float *gpart,
                         const double *cpart, double *chi2, double *dchi2, int fp64,
                         hipStream_t stream) {
    if (n_outs <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL(reduce_kernel<double>, dim3((unsigned)n_outs), dim3(256), 0, stream,
                           outs, gpart, cpart, chi2, dchi2);
    else
        hipLaunchKernelGGL(reduce_kernel<float>, dim3((unsigned)n_outs), dim3(256), 0, stream,
                           outs, gpart, cpart, chi2, dchi2);
    return hipGetLastError();
}

hipError_t launch_theta(const SrcDesc *srcs, int64_t n_srcs, const double *theta,
                        const SplineView &sp, fpho_source *props, hipStream_t stream) {
    if (n_srcs <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n_srcs + 63) / 64);
    hipLaunchKernelGGL(theta_kernel, dim3(blocks), dim3(64), 0, stream, srcs, (int32_t)n_srcs,
                       theta, sp, props);
    return hipGetLastError();
}

hipError_t launch_stack(const SrcDesc *srcs, int64_t n_srcs, const PatchOut *pouts,
                        int64_t n_patches, const double *chi2, const double *dchi2,
                        double *lnlike, double *grad, hipStream_t stream) {
    const int64_t n = n_srcs > n_patches ? n_srcs : n_patches;
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 63) / 64);
    hipLaunchKernelGGL(stack_kernel, dim3(blocks), dim3(64), 0, stream, srcs, (int32_t)n_srcs,
                       pouts, (int32_t)n_patches, chi2, dchi2, lnlike, grad);
    return hipGetLastError();
}

}  // namespace fpho
