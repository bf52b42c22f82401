/* fpho.h -- C ABI of libfpho_hip.so, the MI355X (gfx950) likelihood-and-gradient
 * engine for forcepho patches.
 *
 * Plain C: no torch or HIP types cross this boundary.  Every function returns 0
 * on success and a nonzero FPHO_E* code on failure; fpho_last_error() then
 * returns a message (thread-local).  No exception crosses the ABI.
 *
 * Reference interfaces replaced (paths relative to the forcepho repository):
 *   - forcepho/src/compute_gaussians_kernel.cu:212-214
 *       extern "C" __global__ EvaluateProposal(patch*, proposal*, chi2*, dchi2*)
 *     -> fpho_evaluate (synchronous host in/out, as pycuda cuda.In/cuda.Out at
 *        forcepho/proposal.py:133-136) and fpho_evaluate_async (device pointers,
 *        stream ordered, capturable in a hipGraph).
 *   - forcepho/src/compute_gaussians_kernel.cc:227  (pybind11 CPU kernel,
 *       one band per call) -> same results via fpho_evaluate for all bands.
 *   - forcepho/patches/device_patch.py:268-300  send_to_device   -> fpho_patch_create
 *   - forcepho/patches/device_patch.py:302-311  replace_device_meta_ptrs
 *                                                 -> fpho_patch_set_meta
 *   - forcepho/patches/device_patch.py:346-367  swap_on_device   -> fpho_patch_swap_data_residual
 *   - forcepho/patches/device_patch.py:369-375  retrieve_array   -> fpho_patch_read
 *   - forcepho/patches/device_patch.py:377-393  free             -> fpho_patch_free
 *   - forcepho/sources.py:656-694 (scipy RectBivariateSpline amplitudes) and
 *     forcepho/sources.py:607-621 (Galaxy.proposal) -> fpho_spline_create +
 *        fpho_batch_theta_to_proposals_async (FP64 on device)
 *   - forcepho/model.py:331-332,370-387 (lnlike = -chi2/2, stack_grad)
 *        -> fpho_batch_stack_grad_async
 *   There is no reference counterpart for the batch object (many patches in one
 *   persistent launch); it exists for samplers that advance many patches at once.
 */
#ifndef FPHO_H
#define FPHO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compatibility limits, forcepho/src/kernel_limits.h:5-26.  MAXSOURCES only
 * shapes the compat Response layout; the native outputs are sized by the
 * patch's own n_sources (up to FPHO_MAX_NATIVE_SOURCES). */
#define FPHO_MAXBANDS 30
#define FPHO_MAXSOURCES 30
#define FPHO_NPARAMS 7
#define FPHO_MAXRADII 10
#define FPHO_MAX_EXP_ARG 36.0f
#define FPHO_MAX_NATIVE_SOURCES 128

/* error codes */
#define FPHO_OK 0
#define FPHO_E_ARG 1      /* invalid argument / inconsistent layout */
#define FPHO_E_LIMIT 2    /* a limit was exceeded                   */
#define FPHO_E_HIP 3      /* HIP runtime error                      */
#define FPHO_E_STATE 4    /* object in the wrong state              */
#define FPHO_E_VALUE 5    /* non-physical input value (q<=0, NaN)   */

/* Precision modes.  Pair arithmetic (the bulk) is FP32 in all modes; the
 * convert stage (per image Gaussian) is FP64 in all modes.
 *   FPHO_PREC_FP32        FP32 residual, FP32 cross-tile sums (ablation)
 *   FPHO_PREC_FP64_REDUCE FP32 residual as the reference kernel computes it
 *                         (common_kernel.cc:57-83), FP64 cross-tile sums
 *   FPHO_PREC_MIXED       FP64 residual accumulation; terms brighter than
 *                         0.1 sigma of their pixel re-evaluated in FP64;
 *                         FP64 cross-tile sums (default)
 *   FPHO_PREC_MIXED_CENTRE FPHO_PREC_MIXED with every pixel offset taken from
 *                         the FP64 centre (FP32 high part, then the low
 *                         part), so the FP32 terms and the gradient pass
 *                         see the centre the FP64 bright terms see: position
 *                         derivatives that nearly cancel over the pixels
 *                         keep their accuracy; about 8% slower */
#define FPHO_PREC_FP32 0
#define FPHO_PREC_FP64_REDUCE 1
#define FPHO_PREC_MIXED 2
#define FPHO_PREC_MIXED_CENTRE 3

/* == class Source, forcepho/src/proposal.cu:13-27 (264 bytes). */
typedef struct fpho_source {
    float ra, dec;
    float q, pa;
    float nsersic, rh;
    float fluxes[FPHO_MAXBANDS];
    float mixture_amplitudes[FPHO_MAXRADII];
    float damplitude_drh[FPHO_MAXRADII];
    float damplitude_dnsersic[FPHO_MAXRADII];
} fpho_source;

/* == class PSFSourceGaussian, forcepho/src/patch.cc:103-119 (28 bytes). */
typedef struct fpho_psf_gauss {
    float amp;
    float xcen, ycen;
    float Cxx, Cyy, Cxy;
    int32_t sersic_radius_bin;
} fpho_psf_gauss;

/* == class Response, forcepho/src/proposal.cu:33-39 (840 bytes): the
 * gradients of one band, dchi2_dparam[source * 7 + param] with params in the
 * order d/d(flux, ra, dec, q, pa, nsersic, rh).  pycuda callers allocate
 * n_bands of these (response_struct_dtype, forcepho/proposal.py:242-244) and
 * reshape them to [n_bands, MAXSOURCES, 7] (forcepho/proposal.py:139-141). */
typedef struct fpho_response {
    float dchi2_dparam[FPHO_MAXSOURCES * FPHO_NPARAMS];
} fpho_response;

/* Bits of the device status word.  The reference kernel turns such inputs
 * into NaN or Inf silently (forcepho/src/common_kernel.cc:175-240); the
 * stream-ordered entry points cannot return an error for values that only
 * exist on the device, so their kernels record what they met here and
 * fpho_patch_status / fpho_batch_status read it after the work finished.
 * The results of an evaluation that set a bit are undefined (usually NaN). */
#define FPHO_STATUS_PROPOSAL_NONFINITE 0x1u  /* NaN/Inf in ra, dec, q, pa, flux or amplitudes */
#define FPHO_STATUS_Q_NONPOSITIVE 0x2u       /* axis ratio q <= 0                          */
#define FPHO_STATUS_COVARIANCE 0x4u          /* image Gaussian covariance not positive
                                                definite (det <= 0 or NaN)                */
#define FPHO_STATUS_GAUSSIAN_NONFINITE 0x8u  /* non-finite image Gaussian parameter        */
#define FPHO_STATUS_THETA_NONFINITE 0x10u    /* non-finite parameter vector entry           */

/* Host-side description of one patch: the arrays of forcepho/patches/patch.py:
 * 121-197 (PatchBase.build_patch contract) with the struct layout of
 * forcepho/src/patch.cc:23-100.  All pointers are HOST pointers, read only
 * during the call. */
typedef struct fpho_patch_desc {
    /* pixel data, all exposures concatenated (padding included) */
    int64_t npix;
    const float *data, *ierr, *xpix, *ypix;
    int32_t n_bands, n_exp;
    const int16_t *band_start, *band_N;          /* [n_bands]            */
    const int32_t *exposure_start, *exposure_N;  /* [n_exp]              */
    /* scene metadata (forcepho/patches/patch.py:242-396) */
    int32_t n_sources, n_radii;
    const float *rad2;               /* [n_radii]                         */
    const float *D;                  /* [n_exp][n_sources][2][2]          */
    const float *crpix;              /* [n_exp][n_sources][2]             */
    const float *crval;              /* [n_exp][n_sources][2]             */
    const float *CW;                 /* [n_exp][n_sources][2][2]          */
    const float *G;                  /* [n_exp]                           */
    const int32_t *n_psf_per_source; /* [n_bands]                         */
    const fpho_psf_gauss *psfgauss;  /* [n_psfgauss]                      */
    int64_t n_psfgauss;
    const int32_t *psfgauss_start;   /* [n_exp]                           */
} fpho_patch_desc;

typedef struct fpho_patch fpho_patch;
typedef struct fpho_batch fpho_batch;
typedef struct fpho_spline fpho_spline;

/* Launch shape of the pixel kernel.  Filled by the autotuner; defaults are
 * used when never set. */
typedef struct fpho_launch {
    int32_t pix_per_lane;     /* K: pixels per lane, one of 1,2,4,8         */
    int32_t waves_per_group;  /* workgroup = 64*waves_per_group lanes (1..4)*/
    int32_t persistent;       /* 1: persistent waves pull tiles from a queue*/
    int32_t groups_per_cu;    /* persistent grid = groups_per_cu * n_CU     */
    int32_t cull;             /* 1: skip (tile, Gaussian) pairs whose cutoff
                                 box misses the tile (results unchanged)    */
} fpho_launch;

/* ---------------------------------------------------------------- runtime */
const char *fpho_last_error(void);
int fpho_version(void);
/* Select the HIP device used by the calling thread (once per process is
 * enough; objects remember their device). */
int fpho_init(int device);
/* Number of compute units of the current device (for grid sizing). */
int fpho_device_cu_count(int *n_cu);

/* ------------------------------------------------------------------ patch */
/* Copy a patch to HBM.  With return_residual != 0 a residual buffer is
 * allocated and every evaluation writes residual = data - model into it
 * (compute_gaussians_kernel.cu:273-274). */
int fpho_patch_create(const fpho_patch_desc *desc, int return_residual,
                      fpho_patch **out);
/* Replace the scene metadata only (n_sources, n_radii, rad2, D, crpix, crval,
 * CW, G, n_psf_per_source, psfgauss, psfgauss_start); pixel fields of desc are
 * ignored.  Invalidates batches that contain the patch. */
int fpho_patch_set_meta(fpho_patch *p, const fpho_patch_desc *desc);
/* Swap the device data and residual buffers (fixed-source subtraction). */
int fpho_patch_swap_data_residual(fpho_patch *p);
/* Copy a device pixel array to host: which = 0 data, 1 ierr, 2 xpix,
 * 3 ypix, 4 residual.  host_out holds npix floats. */
int fpho_patch_read(fpho_patch *p, int which, float *host_out);
int fpho_patch_info(const fpho_patch *p, int32_t *n_bands, int32_t *n_sources,
                    int64_t *npix, int64_t *n_pairs);
int fpho_patch_set_launch(fpho_patch *p, const fpho_launch *launch);
/* Precision of the pixel pass and reductions (FPHO_PREC_*). */
int fpho_patch_set_precision(fpho_patch *p, int mode);
/* Start (on != 0) or stop writing residual = data - model at every
 * evaluation.  Stopping keeps the residual buffer, so a later
 * fpho_patch_swap_data_residual still works; starting allocates it if the
 * patch was created without one.  (forcepho sets patch.return_residual on
 * the host, forcepho/patches/device_patch.py:288-293.) */
int fpho_patch_set_residual_writes(fpho_patch *p, int on);
/* Read (and with clear != 0, reset) the status word of the patch's stream-
 * ordered evaluations (FPHO_STATUS_* bits); waits for the device first.
 * The word is reset whenever the patch's plan is rebuilt. */
int fpho_patch_status(fpho_patch *p, uint32_t *status, int clear);
int fpho_patch_free(fpho_patch *p);

/* Synchronous drop-in: proposal (host, n_sources entries) in, chi2[n_bands]
 * (double) and dchi2[n_bands][n_sources][7] (float, order d/d(flux, ra, dec,
 * q, pa, nsersic, rh)) out, exactly the quantities of
 * compute_gaussians_kernel.cu:283-310.  chi2 = sum (r^2 - d^2) * ierr^2. */
int fpho_evaluate(fpho_patch *p, const fpho_source *proposal_host,
                  double *chi2_host, float *dchi2_host);
/* Stream-ordered: all pointers are device pointers; dchi2 is double
 * [n_bands][n_sources][7].  No host synchronisation.  It allocates only
 * when the patch changed since its last evaluation (new metadata, a swap,
 * a launch or residual-write change): then its plan is rebuilt first, or,
 * if `stream` is capturing a graph, FPHO_E_STATE is returned instead.
 * Invalid proposal values set bits of the status word (fpho_patch_status). */
int fpho_evaluate_async(fpho_patch *p, const fpho_source *proposal_dev,
                        double *chi2_dev, double *dchi2_dev, void *stream);

/* Zero-allocation variant of fpho_evaluate_async.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during an
 * evaluation (tight serving or sampling loops, harnesses that forbid
 * hipMalloc).  Same arguments, results bit for bit those of
 * fpho_evaluate_async.  The image Gaussians, tile partials and queue counter
 * live in a fixed 64 MiB scratch array that every GPU the library runs on
 * holds until the process ends; a patch whose work buffers do not fit returns
 * FPHO_E_LIMIT.  The tile descriptors are the patch's own: after any change to
 * the patch (new metadata, a swap, a launch or residual-write change) call
 * fpho_evaluate_async once, which rebuilds them; until then this returns
 * FPHO_E_STATE.  All calls on one GPU share its scratch array, so calls that
 * overlap on one GPU (on different streams or threads) race: order them on
 * one stream.  Invalid proposal values set the patch's status word. */
int fpho_evaluate_async_noalloc(fpho_patch *p, const fpho_source *proposal_dev,
                                double *chi2_dev, double *dchi2_dev, void *stream);

/* The reference's output layout, for callers written against
 * EvaluateProposal(patch, proposal, pchi2, pdchi2_dp)
 * (forcepho/src/compute_gaussians_kernel.cu:212-214, 304-310): chi2[n_bands]
 * and Response[n_bands], entries [0, 7 n_sources) of each band written
 * (the reference leaves the rest unwritten; here it is zeroed).  Patches
 * with more than FPHO_MAXSOURCES sources return FPHO_E_LIMIT.
 * fpho_evaluate_response: host pointers, synchronous, as fpho_evaluate.
 * fpho_evaluate_response_async: device pointers, stream ordered, allocation
 * rules of fpho_evaluate_async. */
int fpho_evaluate_response(fpho_patch *p, const fpho_source *proposal_host,
                           double *chi2_host, fpho_response *response_host);
int fpho_evaluate_response_async(fpho_patch *p, const fpho_source *proposal_dev,
                                 double *chi2_dev, fpho_response *response_dev,
                                 void *stream);

/* ------------------------------------------------------------------ batch */
/* Many patches evaluated by one set of launches: convert, one pixel kernel
 * (a grid launch or a persistent queue) over (patch, band, exposure,
 * pixel-block) tiles sorted by cost, one fixed-order FP64 reduction.
 * Patches stay owned by the caller and must outlive the batch.  The batch
 * keeps the layout (n_bands, n_sources of every patch) it was created
 * with: once a member's source count changes (fpho_patch_set_meta), every
 * batch call returns FPHO_E_STATE, since the caller's buffers were sized by
 * fpho_batch_layout.  Other member changes (new metadata of the same
 * shape, a swap, residual writes) rebuild the plan at the next call,
 * which allocates; while the call's stream is capturing a graph they
 * return FPHO_E_STATE instead. */
int fpho_batch_create(fpho_patch *const *patches, int n_patches,
                      fpho_batch **out);
/* Offsets of each patch's slice in the concatenated arrays:
 * src_off[i] (fpho_source index), chi2_off[i] ([band] index),
 * dchi2_off[i] ([band][src][7] index), theta_off[i] (parameter index);
 * each array has n_patches+1 entries (the last is the total). */
int fpho_batch_layout(const fpho_batch *b, int64_t *src_off, int64_t *chi2_off,
                      int64_t *dchi2_off, int64_t *theta_off);
int fpho_batch_info(const fpho_batch *b, int64_t *n_tiles, int64_t *n_pairs,
                    int64_t *npix, int64_t *n_gauss);
int fpho_batch_set_launch(fpho_batch *b, const fpho_launch *launch);
int fpho_batch_set_precision(fpho_batch *b, int mode);
int fpho_batch_evaluate_async(fpho_batch *b, const fpho_source *proposals_dev,
                              double *chi2_dev, double *dchi2_dev, void *stream);
/* Record HIP events around the three stages of every evaluation (on the
 * evaluation's stream); fpho_batch_last_timing waits for the last evaluation
 * and returns milliseconds of convert, pixel pass and reduce.  Timing adds
 * event records to the stream; leave it off when capturing graphs. */
int fpho_batch_set_timing(fpho_batch *b, int on);
int fpho_batch_last_timing(fpho_batch *b, float *ms3);
/* Status word of the batch's stream-ordered calls (FPHO_STATUS_* bits, from
 * the proposal and conversion kernels); waits for the device first, resets
 * the word with clear != 0 (a host call: do not make it inside a capture).
 * The word is reset whenever the batch's plan is rebuilt. */
int fpho_batch_status(fpho_batch *b, uint32_t *status, int clear);
int fpho_batch_free(fpho_batch *b);

/* ------------------------------------------ device proposals and gradients */
/* Bicubic tensor-product B-spline tables for the Sersic mixture amplitudes:
 * coeffs[n_amp][n_tx-4][n_ty-4] (FP64), knots tx[n_tx] (nsersic), ty[n_ty]
 * (rh), radii[n_amp].  nsersic_min is the table's smallest nsersic: below it
 * (and for rh < 0) the amplitudes become one-hot, as forcepho/sources.py:
 * 656-694 does. */
int fpho_spline_create(const double *coeffs, int n_amp, const double *tx,
                       int n_tx, const double *ty, int n_ty,
                       const double *radii, double nsersic_min,
                       fpho_spline **out);
int fpho_spline_free(fpho_spline *s);
/* theta (device, FP64) holds per source [flux_1..flux_NB, ra, dec, q, pa,
 * nsersic, rh] (forcepho/sources.py:559-592 order); writes fpho_source
 * structs with amplitudes and their derivatives evaluated in FP64. */
int fpho_batch_theta_to_proposals_async(fpho_batch *b, const fpho_spline *s,
                                        const double *theta_dev,
                                        fpho_source *proposals_dev,
                                        void *stream);
/* theta -> chi2 and dchi2 of every patch in one stream-ordered call: the
 * source records are built from theta as fpho_batch_theta_to_proposals_async
 * does, but the conversion stage reads them in FP64 (the FP32 record rounds
 * a source's ra/dec offset to ~3e-6 pixel, which moves the position gradient
 * of a bright source by up to the stated tolerance); then the pixel pass and
 * the reduction of fpho_batch_evaluate_async.  proposals_dev, if not NULL,
 * receives the FP32 records too.  Replaces forcepho's per-evaluation
 * Scene.get_proposal + EvaluateProposal round trip
 * (forcepho/sources.py:80-97,607-694; compute_gaussians_kernel.cu:212-214)
 * for callers that hold the parameter vector on the device. */
int fpho_batch_evaluate_theta_async(fpho_batch *b, const fpho_spline *s,
                                    const double *theta_dev, fpho_source *proposals_dev,
                                    double *chi2_dev, double *dchi2_dev, void *stream);
/* lnlike[i] = -1/2 sum_b chi2 of patch i, grad_theta = -1/2 * stacked dchi2
 * (forcepho/model.py:331-332, 370-387). */
int fpho_batch_stack_grad_async(fpho_batch *b, const double *chi2_dev,
                                const double *dchi2_dev, double *lnlike_dev,
                                double *grad_theta_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FPHO_H */
