// kernels.h -- host-side launch wrappers of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "fpho_internal.h"

namespace fpho {

struct EvalArgs {
    const TileDesc *tiles;
    const int32_t *queue;  // tile order (cost-sorted)
    int64_t n_tiles;
    int32_t *counter;      // persistent work-queue counter (device)
    const ExpDesc *exps;
    const ImgGauss *gauss;
    const ImgGauss64 *gauss64;
    const int4 *gboxes;    // [n_gauss] cutoff boxes (order keys), as in ImgGauss
    const int4 *slot_boxes;  // per tile (from TileDesc::chunk0), its 64-pixel slot boxes
    float *gpart;
    double *cpart;
    int pix_per_lane;
    int waves_per_group;
    int persistent;
    int64_t grid;          // persistent: number of workgroups
    int precision;         // Precision mode
    int cull;              // skip Gaussians whose cutoff box misses the tile
    int mask_cap;          // Gaussians per exposure with an LDS slot mask (<= kMaskCapMax,
                           // a multiple of 8 so the activity words stay aligned)
    int act_words;         // 64-bit activity words per wave (<= kActWordsMax)
};

// LDS bytes per wave for the pass-1 slot masks, at most (4 waves: 32 KiB).
constexpr int kMaskCapMax = 8192;
constexpr int kActWordsMax = 1024;   // activity bits for 65536 Gaussians per wave

struct SplineView {
    const double *coeffs;  // [n_amp][n_tx-4][n_ty-4]
    const double *tx;
    const double *ty;
    const double *radii;
    int32_t n_tx, n_ty, n_amp, pad;
    double nsersic_min;
};

hipError_t launch_convert(const ExpDesc *exps, int64_t n_exp, const fpho_source *props,
                          ImgGauss *out, ImgGauss64 *out64, int4 *boxes, uint32_t *status,
                          hipStream_t stream);
hipError_t launch_convert(const ExpDesc *exps, int64_t n_exp, const Source64 *props,
                          ImgGauss *out, ImgGauss64 *out64, int4 *boxes, uint32_t *status,
                          hipStream_t stream);
hipError_t launch_eval(const EvalArgs &a, hipStream_t stream);
hipError_t launch_response(const double *dchi2, int32_t n_bands, int32_t n_src,
                           fpho_response *out, hipStream_t stream);
hipError_t launch_reduce(const OutDesc *outs, int64_t n_outs, const float *gpart,
                         const double *cpart, double *chi2, double *dchi2, int fp64,
                         hipStream_t stream);
hipError_t launch_theta(const SrcDesc *srcs, int64_t n_srcs, const double *theta,
                        const SplineView &sp, fpho_source *props, Source64 *props64,
                        uint32_t *status, hipStream_t stream);
hipError_t launch_stack(const SrcDesc *srcs, int64_t n_srcs, const PatchOut *pouts,
                        int64_t n_patches, const double *chi2, const double *dchi2,
                        double *lnlike, double *grad, hipStream_t stream);

// A batch's current plan as the pixel pass sees it, for the measurement
// library (libfpho_measure.so); not part of the C ABI.  Rebuilds a stale
// plan first.
struct PlanView {
    EvalArgs eval;
    int64_t n_exp;
    uint32_t *status;
};
int batch_plan_view(fpho_batch *b, PlanView *out);

}  // namespace fpho
