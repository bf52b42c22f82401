// fpho_internal.h -- device-side data layout of the MI355X likelihood engine.
//
// HBM layout (all structure-of-arrays, per patch unless noted):
//   pixels   data, ierr, xpix, ypix, residual : float[npix]  (superpixel order,
//            exposures concatenated, padding pixels carry ierr = 0)
//   ImgGauss workspace : one 128-byte record per (exposure, source, psf comp)
//            written by the convert kernel and read by the pixel kernel with
//            wave-uniform (scalar) loads; bytes 16-47 are all the residual
//            pass needs.
//   ExpDesc  : one per exposure (batch-global), 128 bytes.
//   TileDesc : one per wave tile of 64*K pixels of one exposure.
//   partials : float[n_src*7] per tile + double chi2 per tile; summed in a
//              fixed order by the reduce kernel (deterministic results).
#pragma once

#include <stdint.h>
#include "../../include/fpho.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FPHO_HD __host__ __device__
#define FPHO_HDI __host__ __device__ inline
#else
#define FPHO_HD
#define FPHO_HDI inline
#endif

namespace fpho {

constexpr int kWave = 64;
constexpr int kNParams = FPHO_NPARAMS;

// One image-plane Gaussian (a source Sersic component convolved with one PSF
// component) in the form the pixel pass consumes.  It carries the same
// information as ImageGaussian in forcepho/src/common_kernel.cc:20-46 (value
// parameters and the Jacobian to the 7 on-sky parameters), re-expressed to
// minimise vector instructions per (pixel, Gaussian) pair:
//
//   with s = -log2(e)/2, the pass uses v' = s F d (d = pixel - centre), so
//   that E' = d.v' = s E is directly the exp2 argument (exp(-E/2) =
//   exp2(E')), and 24 s^2 H = hcs + v'.v' with hcs = s^2 (24 - Fxx - Fyy);
//   the value is C = ampA exp2(E') (hcs + v'.v') with ampA = amp/(24 s^2).
//
//   The gradient pass accumulates, per Gaussian and lane, the 12 pixel sums
//   of the monomials the derivative terms are made of (see eval_tile) and
//   converts them once per Gaussian with the constants below, which fold in
//   amp, the powers of s and the Jacobian (dA/dflux, dA/dq, dF/dq, ...).
//
//   box   : bounding box of the cutoff ellipse (exp argument <= 36, widened by
//           a safety margin), as order-preserving integer keys, so the
//           per-tile skip test runs on the scalar ALU.
struct alignas(16) ImgGauss {
    int32_t xlo, xhi, ylo, yhi;              // box
    float xcen, ycen, sfxx, sfyy;            // value parameters (hot)
    float sfxy, hcs, ampA, peak;             // peak: bound on |C| over all pixels
    // gradient epilogue constants (cold)
    float cA_flux, cA_q, cA_pa, cA_n;        // dA/d(.) / (24 s^2)
    float cA_rh, k_axx, k_axy, k_e;          // -amp/(48s^2), -amp/(24s^2), -amp/24
    float k_u, k_x1, k_x2, dFxx_dQ;          // amp/(12s), amp/(24s^3), -amp/(12s^2)
    float dFyy_dQ, dFxy_dQ, dFxx_dPA, dFyy_dPA;
    float dFxy_dPA, xcen_lo, ycen_lo, pad1;  // centre = xcen + xcen_lo (FP64 to ~1e-14)
};
static_assert(sizeof(ImgGauss) == 128, "ImgGauss must be 128 bytes");

// s = -log2(e)/2 and the cutoff exp argument 36 in scaled units (36 s).
constexpr double kScaleS = -0.72134752044448170368;
constexpr float kCutScaled = (float)(36.0 * kScaleS);

// Order-preserving map of a float to a signed integer (a < b <=> key(a) <
// key(b) for all non-NaN floats), so ranges compare with integer instructions.
FPHO_HDI int32_t order_key(float f) {
    int32_t i;
    __builtin_memcpy(&i, &f, 4);
    return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}

// Everything the kernels need to know about one exposure of one patch.
struct alignas(16) ExpDesc {
    const float *data, *ierr, *xpix, *ypix;  // patch pixel arrays
    float *residual;                         // nullptr: not written
    const float *rad2;                       // [n_radii]
    const float *D, *CW;                     // at (exposure, source 0): [n_src][4]
    const float *crpix, *crval;              // at (exposure, source 0): [n_src][2]
    const fpho_psf_gauss *psf;               // psfgauss + psfgauss_start[e]
    int64_t gauss_off;                       // first ImgGauss of this exposure
    int32_t src_off;                         // first fpho_source of this patch
    int32_t band;                            // band index within the patch
    int32_t n_src, n_psf, n_radii;
    float G;
    // pixel origin of the exposure: the pass works in coordinates relative
    // to it (exact for pixel coordinates), so FP32 centres keep their
    // precision however far the patch sits from the frame's origin
    float x0, y0;
};
static_assert(sizeof(ExpDesc) == 128, "ExpDesc must be 128 bytes");

// One wave tile: pixels [pix0, pix0 + npix) of exposure `exp`, with the
// bounding box of their coordinates as order keys.
struct alignas(8) TileDesc {
    int32_t exp;       // ExpDesc index
    int32_t npix;      // <= 64*K
    int32_t pix0;      // offset into the patch pixel arrays
    int32_t chunk0;    // first of the tile's 64-pixel slot boxes (plan-wide index)
    int64_t part_off;  // first float of this tile's gradient partials
    int32_t xlo, xhi, ylo, yhi;
};
static_assert(sizeof(TileDesc) == 40, "TileDesc must be 40 bytes");

// One reduce output: band `band` of a patch, its tiles [tile0, tile0+ntiles)
// are contiguous and their partials contiguous from part0 (stride n_src*7).
struct alignas(8) OutDesc {
    int64_t tile0;
    int64_t part0;
    int64_t chi2_out;   // index into chi2 output
    int64_t dchi2_out;  // index into dchi2 output ([n_src][7] block)
    int32_t ntiles;
    int32_t n_src;
};
static_assert(sizeof(OutDesc) == 40, "OutDesc must be 40 bytes");

// One active source of the batch (for proposals and gradient stacking).
struct alignas(8) SrcDesc {
    int64_t theta_off;   // first parameter of this source
    int64_t dchi2_base;  // dchi2 index of (band 0, this source, param 0)
    int32_t n_bands;     // flux parameters of this source
    int32_t src_stride;  // n_src * 7 of its patch (band stride in dchi2)
    int32_t proposal;    // global fpho_source index
    int32_t pad;
};
static_assert(sizeof(SrcDesc) == 32, "SrcDesc must be 32 bytes");

// The fields of fpho_source in FP64: what the parameter-vector path hands
// the conversion stage, so that positions and shapes reach it unrounded
// (an FP32 ra/dec offset is quantised at ~3e-6 pixel, which moves the
// gradient of a bright source's position by up to the stated tolerance).
struct alignas(16) Source64 {
    double ra, dec, q, pa, nsersic, rh;
    double fluxes[FPHO_MAXBANDS];
    double mixture_amplitudes[FPHO_MAXRADII];
    double damplitude_drh[FPHO_MAXRADII];
    double damplitude_dnsersic[FPHO_MAXRADII];
};
static_assert(sizeof(Source64) == 528, "Source64 must be 528 bytes");

struct alignas(8) PatchOut {
    int64_t chi2_off;    // first band of this patch in chi2 output
    int32_t n_bands;
    int32_t pad;
};

// FP64 copy of an image Gaussian's value parameters in the scaled variables
// of ImgGauss (sF, hcs = s^2 (24 - tr F), ampA = amp / (24 s^2)), used to
// re-evaluate the few (pixel, Gaussian) terms that are bright relative to
// the pixel noise.
struct alignas(16) ImgGauss64 {
    double xcen, ycen, sfxx, sfyy, sfxy, hcs, ampA, pad0;
};
static_assert(sizeof(ImgGauss64) == 64, "ImgGauss64 must be 64 bytes");

// Precision modes of the pixel pass and reductions.
enum Precision : int32_t {
    kPrecFP32 = 0,        // FP32 residual, FP32 cross-tile sums (ablation)
    kPrecFP64Reduce = 1,  // FP32 residual (as the reference kernel), FP64 sums
    kPrecMixed = 2,       // FP64 residual accumulation, bright terms in FP64,
                          // FP64 sums (default)
    kPrecMixedCentre = 3, // kPrecMixed with pixel offsets from the hi/lo centre
};
// Terms brighter than kBrightSigma * sigma_pixel are evaluated in FP64.
constexpr float kBrightSigma = 0.1f;

// ------------------------------------------------------------ 2x2 algebra
// Row-major [[a, b], [c, d]]; the order of forcepho/src/matrix22.cc.
template <typename T>
struct M22 {
    T a, b, c, d;
};
template <typename T>
FPHO_HD inline M22<T> mul(const M22<T> &x, const M22<T> &y) {
    return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d,
            x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
template <typename T>
FPHO_HD inline M22<T> transpose(const M22<T> &x) { return {x.a, x.c, x.b, x.d}; }
template <typename T>
FPHO_HD inline M22<T> add(const M22<T> &x, const M22<T> &y) {
    return {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d};
}
template <typename T>
FPHO_HD inline M22<T> scale(const M22<T> &x, T s) {
    return {x.a * s, x.b * s, x.c * s, x.d * s};
}
template <typename T>
FPHO_HD inline T det(const M22<T> &x) { return x.a * x.d - x.b * x.c; }
template <typename T>
FPHO_HD inline T trace(const M22<T> &x) { return x.a + x.d; }
template <typename T>
FPHO_HD inline M22<T> inverse(const M22<T> &x) {
    const T id = T(1) / det(x);
    return {x.d * id, -x.b * id, -x.c * id, x.a * id};
}

}  // namespace fpho
