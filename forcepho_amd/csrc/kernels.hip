// kernels.hip -- gfx950 kernels of the forcepho likelihood-and-gradient path.
//
//   convert_kernel      proposal -> image-plane Gaussians + Jacobians
//                       (forcepho/src/compute_gaussians_kernel.cu:45-127,
//                        forcepho/src/common_kernel.cc:175-240)
//   eval_kernel<K>      per-pixel mixture, residual, chi2 and the 7 gradient
//                       partials per source (common_kernel.cc:57-143,
//                       compute_gaussians_kernel.cu:262-301); one wave per
//                       tile of 64*K pixels, no float atomics; LDS holds only
//                       each wave's own slot masks and activity bits.
//   reduce_kernel       fixed-order sum of tile partials (FP64 or FP32)
//                       (replaces the shared-memory Accumulator,
//                        compute_gaussians_kernel.cu:133-198)
//   theta_kernel        parameter vector -> source records (FP32 fpho_source
//                       and/or FP64 Source64), Sersic amplitudes from
//                       bicubic B-splines in FP64 (forcepho/sources.py:607-694)
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
// Returns the FPHO_STATUS_* bits of what it met (0 for a valid Gaussian).
// Src is fpho_source (the FP32 proposal record) or Source64 (the same fields
// in FP64, from the parameter-vector path).
template <typename Src>
__device__ __forceinline__ uint32_t make_image_gaussian(const ExpDesc &e, const Src &src,
                                                        int s, int p, ImgGauss &g,
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
    const double xc = xs + (double)pg.xcen - (double)e.x0;   // relative to the origin
    const double yc = ys + (double)pg.ycen - (double)e.y0;

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

    // value parameters in the scaled variables of the pixel pass
    const double sc = kScaleS, sc2 = sc * sc;
    g.xcen = (float)xc;
    g.ycen = (float)yc;
    g.sfxx = (float)(sc * F.a);
    g.sfyy = (float)(sc * F.d);
    g.sfxy = (float)(sc * F.c);
    g.hcs = (float)(sc2 * (24.0 - F.a - F.d));
    g.ampA = (float)(amp / (24.0 * sc2));
    // |C| = |ampA| gp h' <= |ampA| (|hcs| + s^2 tr(F) max_E E e^{-E/2}), since
    // |v'|^2 <= s^2 lambda_max(F) E and E e^{-E/2} <= 2/e; widened by 1%
    g.peak = (float)(1.01 * fabs(amp / (24.0 * sc2)) *
                     (fabs(sc2 * (24.0 - F.a - F.d)) + sc2 * (F.a + F.d) * 0.73575888234288467));

    // gradient epilogue constants: Jacobian of (A, F) with respect to the
    // source parameters folded with amp and the powers of s (derivation in
    // eval_tile)
    const double dA_dq = 0.5 * amp * trace(mul(cov, dFq));
    const double dA_dpa = 0.5 * amp * trace(mul(cov, dFp));
    const double c24 = 1.0 / (24.0 * sc2);
    g.cA_flux = (float)(tmp * am * c24);
    g.cA_q = (float)(dA_dq * c24);
    g.cA_pa = (float)(dA_dpa * c24);
    g.cA_n = (float)(tmp * flux * src.damplitude_dnsersic[bin] * c24);
    g.cA_rh = (float)(tmp * flux * src.damplitude_drh[bin] * c24);
    g.k_axx = (float)(-amp / (48.0 * sc2));
    g.k_axy = (float)(-amp / (24.0 * sc2));
    g.k_e = (float)(-amp / 24.0);
    g.k_u = (float)(amp / (12.0 * sc));
    g.k_x1 = (float)(amp / (24.0 * sc2 * sc));
    g.k_x2 = (float)(-amp / (12.0 * sc2));
    g.dFxx_dQ = (float)dFq.a;
    g.dFyy_dQ = (float)dFq.d;
    g.dFxy_dQ = (float)dFq.c;
    g.dFxx_dPA = (float)dFp.a;
    g.dFyy_dPA = (float)dFp.d;
    g.dFxy_dPA = (float)dFp.c;
    g.xcen_lo = (float)(xc - (double)g.xcen);
    g.ycen_lo = (float)(yc - (double)g.ycen);
    g.pad1 = 0.0f;

    g64.xcen = xc;
    g64.ycen = yc;
    g64.sfxx = sc * F.a;
    g64.sfyy = sc * F.d;
    g64.sfxy = sc * F.c;
    g64.hcs = sc2 * (24.0 - F.a - F.d);
    g64.ampA = amp / (24.0 * sc2);
    g64.pad0 = 0.0;

    // what the reference would turn into NaN or Inf silently
    uint32_t flags = 0;
    if (!(isfinite(src.ra) && isfinite(src.dec) && isfinite(src.q) && isfinite(src.pa) &&
          isfinite(flux) && isfinite(am) && isfinite(src.damplitude_dnsersic[bin]) &&
          isfinite(src.damplitude_drh[bin])))
        flags |= FPHO_STATUS_PROPOSAL_NONFINITE;
    if (!(src.q > 0)) flags |= FPHO_STATUS_Q_NONPOSITIVE;
    if (!(det(cov) > 0.0)) flags |= FPHO_STATUS_COVARIANCE;
    if (!(isfinite(amp) && isfinite(xc) && isfinite(yc) && isfinite(F.a) && isfinite(F.c) &&
          isfinite(F.d) && isfinite(dA_dq) && isfinite(dA_dpa)))
        flags |= FPHO_STATUS_GAUSSIAN_NONFINITE;
    return flags;
}

// One workgroup per exposure; lanes stride over its n_src * n_psf Gaussians.
// Invalid inputs set bits of *status (a vector atomic, taken only then).
template <typename Src>
__global__ __launch_bounds__(128) void convert_kernel(
    const ExpDesc *__restrict__ exps, const Src *__restrict__ props,
    ImgGauss *__restrict__ out, ImgGauss64 *__restrict__ out64, int4 *__restrict__ boxes,
    uint32_t *__restrict__ status) {
    const ExpDesc e = exps[blockIdx.x];
    const int n = e.n_src * e.n_psf;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int s = j / e.n_psf;
        const int p = j - s * e.n_psf;
        ImgGauss g;
        ImgGauss64 g64;
        const uint32_t flags = make_image_gaussian(e, props[e.src_off + s], s, p, g, g64);
        if (flags) atomicOr(status, flags);
        out[e.gauss_off + j] = g;
        out64[e.gauss_off + j] = g64;
        boxes[e.gauss_off + j] = make_int4(g.xlo, g.xhi, g.ylo, g.yhi);
    }
}

// --------------------------------------------------------------- pixel pass

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
// the tile's pixel box tb; integer compares only.
__device__ __forceinline__ bool boxed_out(int cull, const int4 gb, const int4 tb) {
    const int miss = (gb.y < tb.x ? 1 : 0) | (gb.x > tb.y ? 2 : 0) | (gb.w < tb.z ? 4 : 0) |
                     (gb.z > tb.w ? 8 : 0);
    return cull != 0 && miss != 0;
}

// Value of one image Gaussian at one pixel in FP64 (bright terms only), in
// the scaled variables of ImgGauss: ampA 2^E' (hcs + v'.v').  2^E' for
// E' in [36 s, 0] is 2^n 2^f with n = rint(E'), |f| <= 1/2, and 2^f =
// (e^y)^1024 with y = f ln2 / 1024, |y| < 3.4e-4, e^y by its cubic Taylor
// polynomial (truncation 6e-16, amplified 1024-fold by the squarings to
// < 1e-12 relative) -- two constants, so the rare FP64 path holds few
// registers.
__device__ __forceinline__ double pair_value64(const ImgGauss64 &g, float xp, float yp) {
    const double dx = (double)xp - g.xcen;
    const double dy = (double)yp - g.ycen;
    const double vx = g.sfxx * dx + g.sfxy * dy;
    const double vy = g.sfyy * dy + g.sfxy * dx;
    const double ea = dx * vx + dy * vy;
    const double n = __builtin_rint(ea);
    const double y = (ea - n) * (0.69314718055994530942 / 1024.0);
    double p = fma(y, fma(y, fma(y, 1.0 / 6.0, 0.5), 1.0), 1.0);
#pragma unroll
    for (int j = 0; j < 10; ++j) p *= p;
    const double h = fma(vx, vx, fma(vy, vy, g.hcs));
    return g.ampA * __builtin_ldexp(p, (int)n) * h;
}

// Evaluate one wave tile.  `tile` must be wave-uniform (an SGPR value) so
// every descriptor and Gaussian record load below is a scalar load (the
// box scan alone reads 64 boxes with one vector load).
//   MIXED = true : residual accumulated in FP64; a term whose FP32 value
//                  exceeds kBrightSigma * sigma of the tile's quietest pixel
//                  is re-evaluated in FP64 (a wave branch taken only near
//                  bright cores; about 3% of slot evaluations on 512 uneven
//                  patches of 32^2-256^2 pixel exposures).
//   MIXED = false: FP32 residual, as forcepho/src/common_kernel.cc:57-83.
//
// Scaled variables (ImgGauss): d = pixel - centre, v' = s F d, E' = d.v' =
// s E, gp = exp(-E/2) = exp2(E'), h' = hcs + v'.v' = 24 s^2 H.
// Gradient (common_kernel.cc:95-143): with e = -2 r ierr^2 gp and
// a = e H = a' / (24 s^2), a' = e h', the reference's per-pair terms are
//   dC/dA   = a
//   dC/dx   = amp (a vx - e (F v)_x / 12)
//   dC/dFxx = amp (-a dx^2 / 2 - e / 24 + e dx vx / 12)
//   dC/dFyy = amp (-a dy^2 / 2 - e / 24 + e dy vy / 12)
//   dC/dFxy = amp (-a dx dy + e (dy vx + dx vy) / 12)
// With v = v'/s, F v = (s F v')/s^2, dx vx = u'/s (u' = dx v'x), they are
// linear in 12 pixel sums of a', a' v'x, e v'x, a' v'y, e v'y,
// a' dx^2, a' dy^2, a' dx dy, e, e u', e w' (w' = dy v'y), e cr'
// (cr' = dy v'x + dx v'y); each sum costs one FMA per pair (e w' is kept as
// e E' = e (u' + w') and e u' subtracted once) and is turned into the 7
// source partials once per Gaussian (constants in ImgGauss).
//
// Pass 1 tests the cutoff boxes of 64 Gaussians at once (one vector load, a
// ballot) against the boxes of the tile's 64-pixel slots and visits, in
// index order, only those that meet a slot, at those slots.  It
// records, per Gaussian, which of the K pixel slots hold a pixel inside the
// cutoff (a K-bit mask, one byte in LDS per Gaussian for the first mask_cap
// Gaussians) and which Gaussians have any (one bit each in LDS, 64 per word,
// for the first 64 act_words Gaussians); pass 2 visits only those Gaussians
// and evaluates only those slots.  Gaussians past either capacity are
// visited with every slot: a pixel beyond the cutoff adds exact zeros.
template <int K, bool MIXED, bool CENTRE>
__device__ __forceinline__ void eval_tile(int tile, const TileDesc *__restrict__ tiles,
                                          const ExpDesc *__restrict__ exps,
                                          const ImgGauss *__restrict__ gauss,
                                          const ImgGauss64 *__restrict__ gauss64,
                                          const int4 *__restrict__ gboxes,
                                          const int4 *__restrict__ slot_boxes,
                                          float *__restrict__ gpart,
                                          double *__restrict__ cpart, int cull,
                                          uint8_t *__restrict__ amask, int mask_cap,
                                          uint64_t *__restrict__ act, int act_words) {
    const int lane = threadIdx.x & (kWave - 1);
    const TileDesc t = tiles[tile];
    const ExpDesc &e = exps[t.exp];
    const ImgGauss *__restrict__ g = gauss + e.gauss_off;
    const ImgGauss64 *__restrict__ g64 = gauss64 + e.gauss_off;
    const int4 *__restrict__ gb = gboxes + e.gauss_off;
    const int n_src = e.n_src;
    const int n_psf = e.n_psf;
    const int ng = n_src * n_psf;

    float x[K], y[K], r[K];
    double rd[K];
    float tmin = __builtin_huge_valf();   // smallest bright threshold in the tile
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = k * kWave + lane;
        float dk = 0.0f;
        if (p < t.npix) {
            const int i = t.pix0 + p;
            x[k] = e.xpix[i] - e.x0;
            y[k] = e.ypix[i] - e.y0;
            dk = e.data[i];
            const float iek = e.ierr[i];
            if (MIXED && iek > 0.0f) tmin = fminf(tmin, kBrightSigma / iek);
        } else {
            x[k] = kMaskedCoord;
            y[k] = kMaskedCoord;
        }
        r[k] = dk;
        rd[k] = (double)dk;
    }

    // A Gaussian whose cutoff box misses the box of a 64-pixel slot adds
    // nothing to that slot's pixels.  Lane k < K holds slot k's box; a slot
    // with no pixel has an empty box.
    int4 sbox = make_int4(INT_MAX, INT_MIN, INT_MAX, INT_MIN);
    if (lane < K && lane * kWave < t.npix) sbox = slot_boxes[t.chunk0 + lane];

    // A term counts as bright above 0.1 sigma of the tile's quietest pixel;
    // a Gaussian whose peak stays below that cannot have a bright term here.
    if (MIXED) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tmin = fminf(tmin, __shfl_xor(tmin, off, kWave));
        // readfirstlane is an integer builtin: pass the float's bits
        tmin = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(tmin)));
    }

    // ---- pass 1: residual = data - sum_g C_g   (common_kernel.cc:57-83)
    // The boxes of 64 Gaussians are tested at once (lane j: Gaussian
    // base + j) against the K slot boxes; only Gaussians whose box meets a
    // slot are visited, in order, and only at those slots.  act[] keeps, per
    // 64 Gaussians, the ones with a pixel inside the cutoff.
    for (int base = 0; base < ng; base += kWave) {
        const int j = base + lane;
        uint32_t sm = 0;   // slots whose box meets Gaussian j's box
        if (j < ng) {
            if (cull) {
                const int4 b = gb[j];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int4 q = make_int4(__builtin_amdgcn_readlane(sbox.x, k),
                                             __builtin_amdgcn_readlane(sbox.y, k),
                                             __builtin_amdgcn_readlane(sbox.z, k),
                                             __builtin_amdgcn_readlane(sbox.w, k));
                    sm |= boxed_out(1, b, q) ? 0u : 1u << k;
                }
            } else {
                sm = (1u << K) - 1;
            }
        }
        uint64_t todo = __builtin_amdgcn_ballot_w64(sm != 0);
        uint64_t active = 0;
        while (todo) {
            const int bit = __builtin_ctzll(todo);
            todo &= todo - 1;
            const int i = base + bit;
            const uint32_t smi = __builtin_amdgcn_readlane(sm, bit);
            uint32_t m = 0;   // slots with a pixel inside the cutoff (wave-uniform)
            const float xc = g[i].xcen, yc = g[i].ycen;
            const float xcl = CENTRE ? g[i].xcen_lo : 0.0f, ycl = CENTRE ? g[i].ycen_lo : 0.0f;
            const float sfxx = g[i].sfxx, sfyy = g[i].sfyy, sfxy = g[i].sfxy;
            const float hcs = g[i].hcs, ampA = g[i].ampA;
            // a term is bright when |c| > tb: a Gaussian whose peak stays
            // below the tile threshold cannot have one
            const float tb = g[i].peak > tmin ? tmin : __builtin_huge_valf();
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (!(smi & (1u << k))) continue;
                float dx = x[k] - xc;
                float dy = y[k] - yc;
                if (CENTRE) {   // the pixel offset from the FP64 centre
                    dx -= xcl;
                    dy -= ycl;
                }
                const float vx = sfxx * dx + sfxy * dy;
                const float vy = sfyy * dy + sfxy * dx;
                const float ea = dx * vx + dy * vy;      // s E
                const bool in = !(ea < kCutScaled);       // E <= 36 (cutoff)
                if (__builtin_amdgcn_ballot_w64(in) == 0) continue;
                m |= 1u << k;
                {
                    // a lane beyond the cutoff adds an exact zero
                    const float gp = __builtin_amdgcn_exp2f(in ? ea : -__builtin_huge_valf());
                    const float h = fmaf(vx, vx, fmaf(vy, vy, hcs));
                    const float c = ampA * gp * h;
                    if (MIXED) {
                        if (fabsf(c) > tb) {
                            // keep the FP64 copies of the coordinates inside
                            // this rare branch (hoisted out of the Gaussian
                            // loop they would hold 4K more registers)
                            float xk = x[k], yk = y[k];
                            asm volatile("" : "+v"(xk), "+v"(yk));
                            rd[k] -= pair_value64(g64[i], xk, yk);
                        } else
                            rd[k] -= (double)c;
                    } else {
                        r[k] -= c;
                    }
                }
            }
            if (m) active |= 1ull << bit;
            if (i < mask_cap && lane == 0) amask[i] = (uint8_t)m;
        }
        if (lane == 0 && base / kWave < act_words) act[base / kWave] = active;
    }

    // pass 2 of this wave reads the masks and activity words of its lane 0
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();

    // ---- chi2 = sum (r^2 - d^2) ierr^2 in FP64 (compute_gaussians_kernel.cu:276-287)
    // (pixel indices recomputed from an opaque copy of the lane id: kept
    // from the load above, they would hold K registers across pass 1)
    int lane2 = lane;
    asm volatile("" : "+v"(lane2));
    double c2 = 0.0;
    float rr[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = k * kWave + lane2;
        float dk = 0.0f, iek = 0.0f;
        if (p < t.npix) {
            dk = e.data[t.pix0 + p];
            iek = e.ierr[t.pix0 + p];
        }
        if (MIXED) {
            r[k] = (float)rd[k];
            const double rs = rd[k] * (double)iek;
            const double ds = (double)dk * (double)iek;
            c2 += (rs - ds) * (rs + ds);
        } else {
            const float rs = r[k] * iek;
            const float ds = dk * iek;
            double c = (double)rs - (double)ds;
            c *= (double)rs + (double)ds;
            c2 += c;
        }
        if (e.residual != nullptr && p < t.npix) e.residual[t.pix0 + p] = r[k];
        rr[k] = -2.0f * ((r[k] * iek) * iek);  // -2 r ierr^2
        // finish this slot here: deferred, its loaded values would spill
        asm volatile("" : "+v"(c2), "+v"(rr[k]));
    }
    c2 = wave_sum(c2);
    if (lane == 0) cpart[tile] = c2;

    // ---- pass 2: per source, 7 partials   (common_kernel.cc:95-143)
    float *__restrict__ out = gpart + t.part_off;
    for (int s = 0; s < n_src; ++s) {
        float p0 = 0.f, sx = 0.f, sy = 0.f, p3 = 0.f, p4 = 0.f, p5 = 0.f, p6 = 0.f;
        const int i0 = s * n_psf;
        bool touched = false;
        for (int w = i0 / kWave; w <= (i0 + n_psf - 1) / kWave; ++w) {
          // the source's Gaussians in this word of the activity bits
          const uint64_t word = w < act_words ? act[w] : ~0ull;
          const int lo = i0 - w * kWave, hi = i0 + n_psf - w * kWave;   // [lo, hi)
          uint64_t sel = word;
          if (lo > 0) sel &= ~0ull << lo;
          if (hi < kWave) sel &= (1ull << hi) - 1;
          sel = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(sel >> 32)) << 32) |
                (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)sel);
          while (sel) {
            const int i = w * kWave + __builtin_ctzll(sel);
            sel &= sel - 1;
            const ImgGauss &G = g[i];
            // slots to evaluate: pass 1's mask, else (past the LDS capacity)
            // every slot
            const uint32_t m = i < mask_cap
                ? (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)amask[i]) : 0xFFu;
            touched = true;
            const float xc = G.xcen, yc = G.ycen;
            const float xcl = CENTRE ? G.xcen_lo : 0.0f, ycl = CENTRE ? G.ycen_lo : 0.0f;
            const float sfxx = G.sfxx, sfyy = G.sfyy, sfxy = G.sfxy, hcs = G.hcs;
            // the 12 per-Gaussian pixel sums (over this lane's K pixels)
            float PA = 0.f, X1 = 0.f, X2 = 0.f, Y1 = 0.f, Y2 = 0.f, Pxx = 0.f;
            float Pyy = 0.f, Pxy = 0.f, Pe = 0.f, Peu = 0.f, Pea = 0.f, Pecr = 0.f;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (!(m & (1u << k))) continue;
                float dx = x[k] - xc;
                float dy = y[k] - yc;
                if (CENTRE) {
                    dx -= xcl;
                    dy -= ycl;
                }
                const float vx = sfxx * dx + sfxy * dy;
                const float vy = sfyy * dy + sfxy * dx;
                const float u = dx * vx;
                const float ea = fmaf(dy, vy, u);
                // a pixel beyond the cutoff gets gp = exp2(-inf) = 0, so all
                // its terms below are exact zeros (every factor is finite)
                const bool in = !(ea < kCutScaled);
                const float gp = __builtin_amdgcn_exp2f(in ? ea : -__builtin_huge_valf());
                const float h = fmaf(vx, vx, fmaf(vy, vy, hcs));
                const float ee = rr[k] * gp;
                const float a = ee * h;
                PA += a;
                X1 = fmaf(a, vx, X1);
                X2 = fmaf(ee, vx, X2);   // sum e v'x (s F applied below)
                Y1 = fmaf(a, vy, Y1);
                Y2 = fmaf(ee, vy, Y2);   // sum e v'y
                const float ax = a * dx;
                const float ay = a * dy;
                Pxx = fmaf(ax, dx, Pxx);
                Pxy = fmaf(ax, dy, Pxy);
                Pyy = fmaf(ay, dy, Pyy);
                Pe += ee;
                Peu = fmaf(ee, u, Peu);
                Pea = fmaf(ee, ea, Pea);
                Pecr = fmaf(ee, fmaf(dy, vx, dx * vy), Pecr);
            }
            // per-Gaussian epilogue: image-space derivative sums -> 7 partials
            const float gfx = G.k_axx * Pxx + G.k_e * Pe + G.k_u * Peu;
            const float gfy = G.k_axx * Pyy + G.k_e * Pe + G.k_u * (Pea - Peu);
            const float gfxy = G.k_axy * Pxy + G.k_u * Pecr;
            p0 = fmaf(G.cA_flux, PA, p0);
            p5 = fmaf(G.cA_n, PA, p5);
            p6 = fmaf(G.cA_rh, PA, p6);
            p3 = fmaf(G.cA_q, PA, p3);
            p3 = fmaf(G.dFxx_dQ, gfx, p3);
            p3 = fmaf(G.dFyy_dQ, gfy, p3);
            p3 = fmaf(G.dFxy_dQ, gfxy, p3);
            p4 = fmaf(G.cA_pa, PA, p4);
            p4 = fmaf(G.dFxx_dPA, gfx, p4);
            p4 = fmaf(G.dFyy_dPA, gfy, p4);
            p4 = fmaf(G.dFxy_dPA, gfxy, p4);
            // sum e (sF v')_x = sfxx sum e v'x + sfxy sum e v'y (and for y)
            const float fX = sfxx * X2 + sfxy * Y2;
            const float fY = sfyy * Y2 + sfxy * X2;
            sx = fmaf(G.k_x1, X1, fmaf(G.k_x2, fX, sx));
            sy = fmaf(G.k_x1, Y1, fmaf(G.k_x2, fY, sy));
          }
        }
        float *o = out + kNParams * s;
        if (!touched) {
            // no Gaussian of this source reaches the tile
            if (lane < kNParams) o[lane] = 0.0f;
            continue;
        }
        p0 = wave_sum(p0);
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        p3 = wave_sum(p3);
        p4 = wave_sum(p4);
        p5 = wave_sum(p5);
        p6 = wave_sum(p6);
        if (lane == 0) {
            // the sky-position derivatives share the source's CW matrix
            const float *CW = e.CW + 4 * s;
            o[0] = p0;
            o[1] = CW[0] * sx + CW[2] * sy;  // dx/dalpha, dy/dalpha
            o[2] = CW[1] * sx + CW[3] * sy;  // dx/ddelta, dy/ddelta
            o[3] = p3;
            o[4] = p4;
            o[5] = p5;
            o[6] = p6;
        }
    }
}

// Workgroup b runs on XCD b % 8 (round-robin dispatch); map it to queue
// group (chunk-row, XCD, offset) so that each XCD takes chunks of C
// consecutive groups (tiles of one exposure share image Gaussians in its L2
// and scalar caches) and the chunks interleave across XCDs (the cost-sorted
// queue stays balanced).  A bijection on [0, n); the tail past the last full
// row of 8*C groups maps to itself.
constexpr int kXcdChunk = 64;   // queue groups per XCD chunk (measured: 16, 64, 256)

template <int C>
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
    constexpr int X = 8, span = X * C;
    const int full = (n / span) * span;
    if (b >= full) return b;
    const int xcd = b % X, local = b / X;
    return (local / C) * span + xcd * C + (local % C);
}

// Eight waves per SIMD (the most a SIMD holds) caps the kernels at 64
// VGPRs (measured at K = 8: 5.7% faster than 6 waves at 76 VGPRs).
// eval_kernel fits without spills; the persistent kernel at K = 8 spills a
// few values once per tile, outside every Gaussian loop.
#define FPHO_EVAL_OCCUPANCY __attribute__((amdgpu_waves_per_eu(8)))

// Non-persistent: wave w of workgroup b evaluates queue[b*W + w].
template <int K, bool MIXED, bool CENTRE>
__global__ __launch_bounds__(256) FPHO_EVAL_OCCUPANCY void eval_kernel(
    const TileDesc *__restrict__ tiles, const int32_t *__restrict__ queue,
    int32_t n_tiles, int32_t waves_per_group, const ExpDesc *__restrict__ exps,
    const ImgGauss *__restrict__ gauss, const ImgGauss64 *__restrict__ gauss64,
    const int4 *__restrict__ gboxes, const int4 *__restrict__ slot_boxes,
    float *__restrict__ gpart, double *__restrict__ cpart,
    int32_t cull, int32_t mask_cap, int32_t act_words) {
    extern __shared__ uint64_t lds_words[];
    uint8_t *lds_mask = reinterpret_cast<uint8_t *>(lds_words);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int slot = xcd_swizzle<kXcdChunk>((int)blockIdx.x, (int)gridDim.x) *
                     waves_per_group + w;
    if (slot >= n_tiles) return;
    const int tile = __builtin_amdgcn_readfirstlane(queue[slot]);
    eval_tile<K, MIXED, CENTRE>(tile, tiles, exps, gauss, gauss64, gboxes, slot_boxes, gpart, cpart,
                        cull, lds_mask + w * mask_cap, mask_cap,
                        reinterpret_cast<uint64_t *>(lds_mask + waves_per_group * mask_cap) +
                            w * act_words, act_words);
}

// Persistent: every wave pulls tiles from `counter` until the queue is empty.
// Each wave reaches the exit once the counter passes n_tiles.
template <int K, bool MIXED, bool CENTRE>
__global__ __launch_bounds__(256) FPHO_EVAL_OCCUPANCY void eval_persistent_kernel(
    const TileDesc *__restrict__ tiles, const int32_t *__restrict__ queue,
    int32_t n_tiles, int32_t *__restrict__ counter, const ExpDesc *__restrict__ exps,
    const ImgGauss *__restrict__ gauss, const ImgGauss64 *__restrict__ gauss64,
    const int4 *__restrict__ gboxes, const int4 *__restrict__ slot_boxes,
    float *__restrict__ gpart, double *__restrict__ cpart,
    int32_t cull, int32_t mask_cap, int32_t act_words) {
    extern __shared__ uint64_t lds_words[];
    uint8_t *lds_mask = reinterpret_cast<uint8_t *>(lds_words);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int W = (int)(blockDim.x >> 6);
    for (;;) {
        int slot = 0;
        if (lane == 0) slot = atomicAdd(counter, 1);
        slot = __builtin_amdgcn_readfirstlane(__shfl(slot, 0, kWave));
        if (slot >= n_tiles) break;
        const int tile = __builtin_amdgcn_readfirstlane(queue[slot]);
        eval_tile<K, MIXED, CENTRE>(tile, tiles, exps, gauss, gauss64, gboxes, slot_boxes, gpart,
                            cpart, cull, lds_mask + w * mask_cap, mask_cap,
                            reinterpret_cast<uint64_t *>(lds_mask + W * mask_cap) +
                                w * act_words, act_words);
    }
}

__global__ void zero_counter_kernel(int32_t *counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

// ------------------------------------------------------------------ reduce

// One workgroup per (patch, band).  Column j of the band's [ntiles][n_src*7]
// partial block is cut into S segments of consecutive tiles, S a function of
// the column count only (S * ncol <= kReduceSlots); each segment is summed in
// tile order by one lane, then each column's S segment sums in segment
// order.  chi2 partials go through a fixed-shape tree.  The result depends
// only on (ntiles, ncol), not on tile execution order or the launch.
constexpr int kReduceSlots = 4096;
constexpr int kReduceThreads = 1024;   // a single patch has few (patch, band) groups

template <typename Acc>
__global__ __launch_bounds__(kReduceThreads) void reduce_kernel(
    const OutDesc *__restrict__ outs, const float *__restrict__ gpart,
    const double *__restrict__ cpart, double *__restrict__ chi2,
    double *__restrict__ dchi2) {
    __shared__ Acc seg[kReduceSlots];
    __shared__ Acc sh[kReduceThreads];
    const OutDesc o = outs[blockIdx.x];
    const int ncol = o.n_src * kNParams;
    const int S = max(1, min(16, kReduceSlots / ncol));
    const int len = (o.ntiles + S - 1) / S;   // tiles per segment
    for (int w = threadIdx.x; w < ncol * S; w += blockDim.x) {
        const int j = w % ncol, r = w / ncol;
        const int t0 = r * len, t1 = min(o.ntiles, t0 + len);
        const float *col = gpart + o.part0 + j;
        Acc acc = 0;
        int t = t0;
        for (; t + 4 <= t1; t += 4) {
            const float v0 = col[(int64_t)(t + 0) * ncol];
            const float v1 = col[(int64_t)(t + 1) * ncol];
            const float v2 = col[(int64_t)(t + 2) * ncol];
            const float v3 = col[(int64_t)(t + 3) * ncol];
            acc += (Acc)v0;
            acc += (Acc)v1;
            acc += (Acc)v2;
            acc += (Acc)v3;
        }
        for (; t < t1; ++t) acc += (Acc)col[(int64_t)t * ncol];
        seg[r * ncol + j] = acc;
    }
    Acc c = 0;
    for (int t = threadIdx.x; t < o.ntiles; t += blockDim.x) c += (Acc)cpart[o.tile0 + t];
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int j = threadIdx.x; j < ncol; j += blockDim.x) {
        Acc acc = 0;
        for (int r = 0; r < S; ++r) acc += seg[r * ncol + j];
        dchi2[o.dchi2_out + j] = (double)acc;
    }
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

// Writes the FP32 record to props and/or the FP64 record to props64 (either
// may be null).
__global__ __launch_bounds__(64) void theta_kernel(
    const SrcDesc *__restrict__ srcs, int32_t n_srcs, const double *__restrict__ theta,
    SplineView sp, fpho_source *__restrict__ props, Source64 *__restrict__ props64,
    uint32_t *__restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_srcs) return;
    const SrcDesc sd = srcs[i];
    const double *th = theta + sd.theta_off;
    {
        bool finite = true;
        for (int j = 0; j < sd.n_bands + 6; ++j) finite = finite && isfinite(th[j]);
        if (!finite) atomicOr(status, FPHO_STATUS_THETA_NONFINITE);
    }
    Source64 o;
    for (int b = 0; b < FPHO_MAXBANDS; ++b) o.fluxes[b] = 0.0;
    for (int m = 0; m < FPHO_MAXRADII; ++m) {
        o.mixture_amplitudes[m] = 0.0;
        o.damplitude_drh[m] = 0.0;
        o.damplitude_dnsersic[m] = 0.0;
    }
    const int nb = sd.n_bands;
    for (int b = 0; b < nb; ++b) o.fluxes[b] = th[b];
    o.ra = th[nb + 0];
    o.dec = th[nb + 1];
    o.q = th[nb + 2];
    o.pa = th[nb + 3];
    o.nsersic = th[nb + 4];
    o.rh = th[nb + 5];
    const double n = o.nsersic, rh = o.rh;

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
        o.mixture_amplitudes[best] = 1.0;
    } else if (rh < 0.0) {
        o.mixture_amplitudes[0] = 1.0;  // forcepho/sources.py:666-668
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
            o.mixture_amplitudes[m] = v;
            o.damplitude_dnsersic[m] = vn;
            o.damplitude_drh[m] = vr;
        }
    }
    if (props64) props64[sd.proposal] = o;
    if (props) {
        fpho_source f;
        f.ra = (float)o.ra;
        f.dec = (float)o.dec;
        f.q = (float)o.q;
        f.pa = (float)o.pa;
        f.nsersic = (float)o.nsersic;
        f.rh = (float)o.rh;
        for (int b = 0; b < FPHO_MAXBANDS; ++b) f.fluxes[b] = (float)o.fluxes[b];
        for (int m = 0; m < FPHO_MAXRADII; ++m) {
            f.mixture_amplitudes[m] = (float)o.mixture_amplitudes[m];
            f.damplitude_drh[m] = (float)o.damplitude_drh[m];
            f.damplitude_dnsersic[m] = (float)o.damplitude_dnsersic[m];
        }
        props[sd.proposal] = f;
    }
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

// ------------------------------------------------------ compat Response out

// dchi2 [n_bands][n_src][7] (double) -> Response[n_bands] (float, stride
// MAXSOURCES * 7, compute_gaussians_kernel.cu:304-310), tail zeroed.
__global__ __launch_bounds__(256) void response_kernel(const double *__restrict__ dchi2,
                                                       int32_t n_src,
                                                       fpho_response *__restrict__ out) {
    const int b = blockIdx.x;
    constexpr int kStride = FPHO_MAXSOURCES * kNParams;
    for (int j = threadIdx.x; j < kStride; j += blockDim.x)
        out[b].dchi2_dparam[j] =
            j < n_src * kNParams ? (float)dchi2[(int64_t)b * n_src * kNParams + j] : 0.0f;
}

// ---------------------------------------------------------------- launchers

template <typename Src>
static hipError_t launch_convert_t(const ExpDesc *exps, int64_t n_exp, const Src *props,
                                   ImgGauss *out, ImgGauss64 *out64, int4 *boxes,
                                   uint32_t *status, hipStream_t stream) {
    if (n_exp <= 0) return hipSuccess;
    hipLaunchKernelGGL(convert_kernel<Src>, dim3((unsigned)n_exp), dim3(128), 0, stream,
                       exps, props, out, out64, boxes, status);
    return hipGetLastError();
}

hipError_t launch_convert(const ExpDesc *exps, int64_t n_exp, const fpho_source *props,
                          ImgGauss *out, ImgGauss64 *out64, int4 *boxes, uint32_t *status,
                          hipStream_t stream) {
    return launch_convert_t(exps, n_exp, props, out, out64, boxes, status, stream);
}

hipError_t launch_convert(const ExpDesc *exps, int64_t n_exp, const Source64 *props,
                          ImgGauss *out, ImgGauss64 *out64, int4 *boxes, uint32_t *status,
                          hipStream_t stream) {
    return launch_convert_t(exps, n_exp, props, out, out64, boxes, status, stream);
}

hipError_t launch_response(const double *dchi2, int32_t n_bands, int32_t n_src,
                           fpho_response *out, hipStream_t stream) {
    if (n_bands <= 0) return hipSuccess;
    hipLaunchKernelGGL(response_kernel, dim3((unsigned)n_bands), dim3(256), 0, stream, dchi2,
                       n_src, out);
    return hipGetLastError();
}

template <int K, bool MIXED, bool CENTRE>
static hipError_t launch_eval_k(const EvalArgs &a, hipStream_t stream) {
    const int W = a.waves_per_group;
    // per wave: the slot mask bytes, then the activity words
    const size_t lds = (size_t)W * ((size_t)a.mask_cap + 8 * (size_t)a.act_words);
    if (a.persistent) {
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, a.counter);
        hipLaunchKernelGGL((eval_persistent_kernel<K, MIXED, CENTRE>), dim3((unsigned)a.grid),
                           dim3(64 * W), lds, stream, a.tiles, a.queue, (int32_t)a.n_tiles,
                           a.counter, a.exps, a.gauss, a.gauss64, a.gboxes, a.slot_boxes, a.gpart, a.cpart,
                           (int32_t)a.cull, (int32_t)a.mask_cap, (int32_t)a.act_words);
    } else {
        const int64_t groups = (a.n_tiles + W - 1) / W;
        hipLaunchKernelGGL((eval_kernel<K, MIXED, CENTRE>), dim3((unsigned)groups), dim3(64 * W), lds,
                           stream, a.tiles, a.queue, (int32_t)a.n_tiles, (int32_t)W, a.exps,
                           a.gauss, a.gauss64, a.gboxes, a.slot_boxes, a.gpart, a.cpart, (int32_t)a.cull,
                           (int32_t)a.mask_cap, (int32_t)a.act_words);
    }
    return hipGetLastError();
}

template <bool MIXED, bool CENTRE>
static hipError_t launch_eval_p(const EvalArgs &a, hipStream_t stream) {
    switch (a.pix_per_lane) {
        case 1: return launch_eval_k<1, MIXED, CENTRE>(a, stream);
        case 2: return launch_eval_k<2, MIXED, CENTRE>(a, stream);
        case 4: return launch_eval_k<4, MIXED, CENTRE>(a, stream);
        case 8: return launch_eval_k<8, MIXED, CENTRE>(a, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_eval(const EvalArgs &a, hipStream_t stream) {
    if (a.n_tiles <= 0) return hipSuccess;
    switch (a.precision) {
        case kPrecMixed: return launch_eval_p<true, false>(a, stream);
        case kPrecMixedCentre: return launch_eval_p<true, true>(a, stream);
        default: return launch_eval_p<false, false>(a, stream);
    }
}

hipError_t launch_reduce(const OutDesc *outs, int64_t n_outs, const float *gpart,
                         const double *cpart, double *chi2, double *dchi2, int fp64,
                         hipStream_t stream) {
    if (n_outs <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL(reduce_kernel<double>, dim3((unsigned)n_outs), dim3(kReduceThreads), 0,
                           stream, outs, gpart, cpart, chi2, dchi2);
    else
        hipLaunchKernelGGL(reduce_kernel<float>, dim3((unsigned)n_outs), dim3(kReduceThreads), 0,
                           stream, outs, gpart, cpart, chi2, dchi2);
    return hipGetLastError();
}

hipError_t launch_theta(const SrcDesc *srcs, int64_t n_srcs, const double *theta,
                        const SplineView &sp, fpho_source *props, Source64 *props64,
                        uint32_t *status, hipStream_t stream) {
    if (n_srcs <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n_srcs + 63) / 64);
    hipLaunchKernelGGL(theta_kernel, dim3(blocks), dim3(64), 0, stream, srcs, (int32_t)n_srcs,
                       theta, sp, props, props64, status);
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
