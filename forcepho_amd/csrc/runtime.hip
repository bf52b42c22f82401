// runtime.hip -- host side of libfpho_hip.so: device memory for patches, tile
// plans, batches, the synchronous drop-in and the stream-ordered entry points.
//
// Ownership: a patch owns its HBM copies of the pixel and metadata arrays
// (forcepho/patches/device_patch.py:268-393 holds the same arrays as pycuda
// DeviceAllocations).  A plan (single patch or batch) owns its descriptors,
// the image-Gaussian workspace and the tile partials; nothing is allocated
// during an evaluation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fpho.h"
#include "fpho_internal.h"
#include "kernels.h"

using namespace fpho;

// ------------------------------------------------------------------ errors
namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return fail(FPHO_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));    \
    } while (0)

// Device buffer with RAII release.
template <typename T>
struct DevVec {
    T *ptr = nullptr;
    int64_t n = 0;
    DevVec() = default;
    DevVec(const DevVec &) = delete;
    DevVec &operator=(const DevVec &) = delete;
    ~DevVec() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    hipError_t alloc(int64_t count) {
        release();
        if (count <= 0) return hipSuccess;
        hipError_t e = hipMalloc((void **)&ptr, sizeof(T) * (size_t)count);
        if (e == hipSuccess) n = count;
        return e;
    }
    hipError_t upload(const T *host, int64_t count) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count <= 0) return e;
        return hipMemcpy(ptr, host, sizeof(T) * (size_t)count, hipMemcpyHostToDevice);
    }
    void swap(DevVec &o) {
        std::swap(ptr, o.ptr);
        std::swap(n, o.n);
    }
};

// Page-locked host buffer (hipHostMalloc): copies from and to it run
// asynchronously, without a staging copy inside the runtime.
template <typename T>
struct PinnedVec {
    T *ptr = nullptr;
    int64_t n = 0;
    PinnedVec() = default;
    PinnedVec(const PinnedVec &) = delete;
    PinnedVec &operator=(const PinnedVec &) = delete;
    ~PinnedVec() { release(); }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    hipError_t alloc(int64_t count) {
        release();
        if (count <= 0) return hipSuccess;
        hipError_t e = hipHostMalloc((void **)&ptr, sizeof(T) * (size_t)count,
                                     hipHostMallocDefault);
        if (e == hipSuccess) n = count;
        return e;
    }
};

bool finite_all(const float *p, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

fpho_launch default_launch(bool batch) {
    fpho_launch l;
    l.pix_per_lane = 4;
    l.waves_per_group = 4;
    l.persistent = batch ? 1 : 0;
    l.groups_per_cu = 8;
    l.cull = 1;
    return l;
}

int check_launch(const fpho_launch &l) {
    if (l.pix_per_lane != 1 && l.pix_per_lane != 2 && l.pix_per_lane != 4 &&
        l.pix_per_lane != 8)
        return fail(FPHO_E_ARG, "pix_per_lane must be 1, 2, 4 or 8 (got %d)", l.pix_per_lane);
    if (l.waves_per_group < 1 || l.waves_per_group > 4)
        return fail(FPHO_E_ARG, "waves_per_group must be in 1..4 (got %d)", l.waves_per_group);
    if (l.persistent != 0 && l.persistent != 1)
        return fail(FPHO_E_ARG, "persistent must be 0 or 1");
    if (l.groups_per_cu < 1 || l.groups_per_cu > 64)
        return fail(FPHO_E_ARG, "groups_per_cu must be in 1..64");
    if (l.cull != 0 && l.cull != 1) return fail(FPHO_E_ARG, "cull must be 0 or 1");
    return FPHO_OK;
}

int cu_count_cached() {
    static int n = 0;
    if (n == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
    }
    return n;
}

}  // namespace

// ------------------------------------------------------------------- objects

// Tile plan over one or more patches (single-patch drop-in or batch).
struct Plan {
    DevVec<ExpDesc> exps;
    DevVec<TileDesc> tiles;
    DevVec<int32_t> queue;
    DevVec<OutDesc> outs;
    DevVec<ImgGauss> gauss;
    DevVec<ImgGauss64> gauss64;
    DevVec<int4> gboxes;
    DevVec<int4> slot_boxes;
    DevVec<float> gpart;
    DevVec<double> cpart;
    DevVec<int32_t> counter;
    DevVec<uint32_t> status;        // FPHO_STATUS_* bits of the plan's kernels
    DevVec<double> dchi2_tmp;       // single patch: dchi2 before the compat Response layout
    DevVec<SrcDesc> srcs;
    DevVec<PatchOut> pouts;
    DevVec<Source64> src64;         // batch: FP64 source records of the theta path
    int64_t n_exp = 0, n_tiles = 0, n_outs = 0, n_gauss = 0, n_srcs = 0, n_patches = 0;
    int64_t n_pairs = 0, npix = 0, n_part = 0;
    int32_t max_ng = 0;  // most image Gaussians of one exposure
    std::vector<int64_t> src_off, chi2_off, dchi2_off, theta_off;  // per patch, +1
    fpho_launch launch;
    int precision = kPrecMixed;
    bool valid = false;
    // optional per-stage timing: events before convert, eval, reduce and after
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Plan() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

struct fpho_patch {
    int device = 0;
    int64_t npix = 0;
    int32_t n_bands = 0, n_exp = 0, n_sources = 0, n_radii = 0;
    int64_t n_psfgauss = 0;
    std::vector<int16_t> band_start, band_N;
    std::vector<int32_t> exposure_start, exposure_N, n_psf_per_source, psfgauss_start;
    std::vector<float> G_host;
    // coordinate box of every 64-pixel chunk of every exposure (chunks start
    // at the exposure start): xlo, xhi, ylo, yhi as order keys
    std::vector<int64_t> chunk_start;  // [n_exp+1] first chunk of exposure e
    std::vector<int32_t> chunk_box;    // [n_chunks][4], coordinates relative to (x0, y0)
    std::vector<float> x0, y0;         // [n_exp] pixel origin of each exposure
    DevVec<float> data, ierr, xpix, ypix, residual;
    DevVec<float> rad2, D, crpix, crval, CW;
    DevVec<fpho_psf_gauss> psf;
    DevVec<uint32_t> status;      // FPHO_STATUS_* bits of its evaluations
    bool write_residual = false;  // evaluations write residual = data - model
    uint64_t generation = 1;  // bumps on metadata change, a data/residual swap or
                              // a residual-write change
    fpho_launch launch = default_launch(false);
    int precision = kPrecMixed;
    Plan plan;
    uint64_t plan_generation = 0;
    // staging for the synchronous drop-in: the proposal, and chi2 [n_bands]
    // followed by dchi2 [n_bands][n_sources][7] in one buffer, so a call
    // makes one copy each way from page-locked memory and one wait
    DevVec<fpho_source> prop_dev;
    DevVec<double> out_dev;
    PinnedVec<fpho_source> prop_host;
    PinnedVec<double> out_host;
};

struct fpho_batch {
    std::vector<fpho_patch *> patches;
    std::vector<uint64_t> generations;
    std::vector<int32_t> n_sources0;  // each patch's source count at creation
    fpho_launch launch = default_launch(true);
    int precision = kPrecMixed;
    Plan plan;
};

struct fpho_spline {
    DevVec<double> coeffs, tx, ty, radii;
    SplineView view;
};

// ------------------------------------------------------------ validation

namespace {

int validate_meta(const fpho_patch_desc *d, int32_t n_bands, int32_t n_exp) {
    if (d->n_sources < 1 || d->n_sources > FPHO_MAX_NATIVE_SOURCES)
        return fail(FPHO_E_LIMIT, "n_sources=%d outside 1..%d", d->n_sources,
                    FPHO_MAX_NATIVE_SOURCES);
    if (d->n_radii < 1 || d->n_radii > FPHO_MAXRADII)
        return fail(FPHO_E_LIMIT, "n_radii=%d outside 1..%d", d->n_radii, FPHO_MAXRADII);
    if (!d->rad2 || !d->D || !d->crpix || !d->crval || !d->CW || !d->G ||
        !d->n_psf_per_source || !d->psfgauss || !d->psfgauss_start)
        return fail(FPHO_E_ARG, "null metadata pointer");
    const int64_t nes = (int64_t)n_exp * d->n_sources;
    if (!finite_all(d->rad2, d->n_radii) || !finite_all(d->D, 4 * nes) ||
        !finite_all(d->CW, 4 * nes) || !finite_all(d->crpix, 2 * nes) ||
        !finite_all(d->crval, 2 * nes) || !finite_all(d->G, n_exp))
        return fail(FPHO_E_VALUE, "non-finite astrometry or calibration value");
    for (int b = 0; b < n_bands; ++b) {
        const int np = d->n_psf_per_source[b];
        if (np < 1 || np > 4096) return fail(FPHO_E_LIMIT, "n_psf_per_source[%d]=%d", b, np);
        if ((int64_t)np * d->n_sources > (1 << 20))
            return fail(FPHO_E_LIMIT, "too many image Gaussians in band %d", b);
    }
    for (int b = 0; b < n_bands; ++b) {
        for (int e = d->band_start[b]; e < d->band_start[b] + d->band_N[b]; ++e) {
            const int64_t s0 = d->psfgauss_start[e];
            const int np = d->n_psf_per_source[b];
            if (s0 < 0 || s0 + np > d->n_psfgauss)
                return fail(FPHO_E_ARG, "psfgauss_start[%d]=%lld out of range", e,
                            (long long)s0);
            for (int p = 0; p < np; ++p) {
                const fpho_psf_gauss &g = d->psfgauss[s0 + p];
                if (g.sersic_radius_bin < 0 || g.sersic_radius_bin >= d->n_radii)
                    return fail(FPHO_E_ARG, "sersic_radius_bin %d out of range",
                                g.sersic_radius_bin);
                const float v[6] = {g.amp, g.xcen, g.ycen, g.Cxx, g.Cyy, g.Cxy};
                if (!finite_all(v, 6)) return fail(FPHO_E_VALUE, "non-finite PSF Gaussian");
                if (!(g.Cxx > 0.0f && g.Cyy > 0.0f && g.Cxx * g.Cyy - g.Cxy * g.Cxy > 0.0f))
                    return fail(FPHO_E_VALUE,
                                "PSF covariance not positive definite (exposure %d)", e);
            }
        }
    }
    return FPHO_OK;
}

int validate_pixels(const fpho_patch_desc *d) {
    if (d->npix < 1 || d->npix > ((int64_t)1 << 31) - 1)
        return fail(FPHO_E_LIMIT, "npix=%lld outside 1..2^31-1", (long long)d->npix);
    if (d->n_bands < 1 || d->n_bands > FPHO_MAXBANDS)
        return fail(FPHO_E_LIMIT, "n_bands=%d outside 1..%d", d->n_bands, FPHO_MAXBANDS);
    if (d->n_exp < 1) return fail(FPHO_E_ARG, "n_exp=%d", d->n_exp);
    if (!d->data || !d->ierr || !d->xpix || !d->ypix || !d->band_start || !d->band_N ||
        !d->exposure_start || !d->exposure_N)
        return fail(FPHO_E_ARG, "null pixel or index pointer");
    for (int b = 0; b < d->n_bands; ++b) {
        if (d->band_start[b] < 0 || d->band_N[b] < 0 ||
            d->band_start[b] + d->band_N[b] > d->n_exp)
            return fail(FPHO_E_ARG, "band %d exposures [%d, +%d) outside 0..%d", b,
                        d->band_start[b], d->band_N[b], d->n_exp);
    }
    for (int e = 0; e < d->n_exp; ++e) {
        if (d->exposure_start[e] < 0 || d->exposure_N[e] < 0 ||
            (int64_t)d->exposure_start[e] + d->exposure_N[e] > d->npix)
            return fail(FPHO_E_ARG, "exposure %d pixels [%d, +%d) outside 0..%lld", e,
                        d->exposure_start[e], d->exposure_N[e], (long long)d->npix);
    }
    if (!finite_all(d->data, d->npix) || !finite_all(d->ierr, d->npix) ||
        !finite_all(d->xpix, d->npix) || !finite_all(d->ypix, d->npix))
        return fail(FPHO_E_VALUE, "non-finite pixel data, ierr or coordinate");
    for (int64_t i = 0; i < d->npix; ++i)
        if (d->ierr[i] < 0.0f) return fail(FPHO_E_VALUE, "negative ierr at pixel %lld",
                                           (long long)i);
    return FPHO_OK;
}

int upload_meta(fpho_patch *p, const fpho_patch_desc *d) {
    const int64_t nes = (int64_t)p->n_exp * d->n_sources;
    p->n_sources = d->n_sources;
    p->n_radii = d->n_radii;
    p->n_psfgauss = d->n_psfgauss;
    p->n_psf_per_source.assign(d->n_psf_per_source, d->n_psf_per_source + p->n_bands);
    p->psfgauss_start.assign(d->psfgauss_start, d->psfgauss_start + p->n_exp);
    p->G_host.assign(d->G, d->G + p->n_exp);
    HIP_TRY(p->rad2.upload(d->rad2, d->n_radii));
    HIP_TRY(p->D.upload(d->D, 4 * nes));
    HIP_TRY(p->CW.upload(d->CW, 4 * nes));
    HIP_TRY(p->crpix.upload(d->crpix, 2 * nes));
    HIP_TRY(p->crval.upload(d->crval, 2 * nes));
    HIP_TRY(p->psf.upload(d->psfgauss, d->n_psfgauss));
    ++p->generation;
    return FPHO_OK;
}

// ------------------------------------------------------------- planning

struct HostPlan {
    std::vector<ExpDesc> exps;
    std::vector<TileDesc> tiles;
    std::vector<int4> slot_boxes;   // per tile, the boxes of its 64-pixel slots
    std::vector<int64_t> cost;
    std::vector<OutDesc> outs;
    std::vector<SrcDesc> srcs;
    std::vector<PatchOut> pouts;
    int64_t n_gauss = 0, n_part = 0, n_pairs = 0, npix = 0;
    int32_t max_ng = 0;   // most image Gaussians of one exposure
};

// Append one patch to a host plan.  Tiles: each exposure is cut into tiles of
// 64*K consecutive pixels (superpixel order keeps them compact on the sky).
void plan_patch(const fpho_patch *p, int K, int64_t src_base, int64_t chi2_base,
                int64_t dchi2_base, int64_t theta_base, int32_t patch_index, HostPlan &hp) {
    const int tile_pix = kWave * K;
    const int ns = p->n_sources;
    for (int b = 0; b < p->n_bands; ++b) {
        OutDesc od;
        od.tile0 = (int64_t)hp.tiles.size();
        od.part0 = hp.n_part;
        od.chi2_out = chi2_base + b;
        od.dchi2_out = dchi2_base + (int64_t)b * ns * kNParams;
        od.n_src = ns;
        const int np = p->n_psf_per_source[b];
        for (int e = p->band_start[b]; e < p->band_start[b] + p->band_N[b]; ++e) {
            ExpDesc ed;
            ed.data = p->data.ptr;
            ed.ierr = p->ierr.ptr;
            ed.xpix = p->xpix.ptr;
            ed.ypix = p->ypix.ptr;
            ed.residual = p->write_residual ? p->residual.ptr : nullptr;
            ed.rad2 = p->rad2.ptr;
            ed.D = p->D.ptr + (int64_t)4 * ns * e;
            ed.CW = p->CW.ptr + (int64_t)4 * ns * e;
            ed.crpix = p->crpix.ptr + (int64_t)2 * ns * e;
            ed.crval = p->crval.ptr + (int64_t)2 * ns * e;
            ed.psf = p->psf.ptr + p->psfgauss_start[e];
            ed.gauss_off = hp.n_gauss;
            ed.src_off = (int32_t)src_base;
            ed.band = b;
            ed.n_src = ns;
            ed.n_psf = np;
            ed.n_radii = p->n_radii;
            ed.G = p->G_host[e];
            ed.x0 = p->x0[e];
            ed.y0 = p->y0[e];
            const int32_t ei = (int32_t)hp.exps.size();
            hp.exps.push_back(ed);
            hp.n_gauss += (int64_t)ns * np;
            hp.max_ng = std::max(hp.max_ng, ns * np);
            const int32_t n_e = p->exposure_N[e];
            hp.npix += n_e;
            hp.n_pairs += (int64_t)n_e * ns * np;
            for (int32_t off = 0; off < n_e; off += tile_pix) {
                TileDesc t;
                t.exp = ei;
                t.npix = std::min(tile_pix, n_e - off);
                t.pix0 = p->exposure_start[e] + off;
                t.chunk0 = (int32_t)hp.slot_boxes.size();
                t.part_off = hp.n_part;
                // union of the boxes of the tile's 64-pixel chunks
                const int64_t c0 = p->chunk_start[e] + off / kWave;
                const int64_t c1 = p->chunk_start[e] + (off + t.npix + kWave - 1) / kWave;
                t.xlo = p->chunk_box[4 * c0 + 0];
                t.xhi = p->chunk_box[4 * c0 + 1];
                t.ylo = p->chunk_box[4 * c0 + 2];
                t.yhi = p->chunk_box[4 * c0 + 3];
                for (int64_t c = c0; c < c1; ++c)
                    hp.slot_boxes.push_back(make_int4(p->chunk_box[4 * c + 0],
                                                      p->chunk_box[4 * c + 1],
                                                      p->chunk_box[4 * c + 2],
                                                      p->chunk_box[4 * c + 3]));
                for (int64_t c = c0 + 1; c < c1; ++c) {
                    t.xlo = std::min(t.xlo, p->chunk_box[4 * c + 0]);
                    t.xhi = std::max(t.xhi, p->chunk_box[4 * c + 1]);
                    t.ylo = std::min(t.ylo, p->chunk_box[4 * c + 2]);
                    t.yhi = std::max(t.yhi, p->chunk_box[4 * c + 3]);
                }
                hp.n_part += (int64_t)ns * kNParams;
                hp.tiles.push_back(t);
                // cost model: both passes visit every (pixel, Gaussian) pair
                hp.cost.push_back((int64_t)ns * np * t.npix);
            }
        }
        od.ntiles = (int32_t)((int64_t)hp.tiles.size() - od.tile0);
        hp.outs.push_back(od);
    }
    int64_t theta = theta_base;
    for (int s = 0; s < ns; ++s) {
        SrcDesc sd;
        sd.theta_off = theta;
        sd.dchi2_base = dchi2_base + (int64_t)s * kNParams;
        sd.n_bands = p->n_bands;
        sd.src_stride = ns * kNParams;
        sd.proposal = (int32_t)(src_base + s);
        sd.pad = 0;
        hp.srcs.push_back(sd);
        theta += p->n_bands + 6;
    }
    PatchOut po;
    po.chi2_off = chi2_base;
    po.n_bands = p->n_bands;
    po.pad = patch_index;
    hp.pouts.push_back(po);
}

// Tile order of the queue: most expensive first (longest-processing-time
// first); ties keep canonical order, so tiles of one exposure stay adjacent.
std::vector<int32_t> cost_order(const HostPlan &hp) {
    std::vector<int32_t> queue(hp.tiles.size());
    std::iota(queue.begin(), queue.end(), 0);
    std::stable_sort(queue.begin(), queue.end(),
                     [&](int32_t a, int32_t b) { return hp.cost[a] > hp.cost[b]; });
    return queue;
}

// On failure the plan stays invalid (its buffers may be partly replaced).
int build_plan(const std::vector<fpho_patch *> &patches, const fpho_launch &launch,
               Plan &plan) {
    plan.valid = false;
    HostPlan hp;
    plan.src_off.assign(1, 0);
    plan.chi2_off.assign(1, 0);
    plan.dchi2_off.assign(1, 0);
    plan.theta_off.assign(1, 0);
    for (size_t i = 0; i < patches.size(); ++i) {
        const fpho_patch *p = patches[i];
        plan_patch(p, launch.pix_per_lane, plan.src_off.back(), plan.chi2_off.back(),
                   plan.dchi2_off.back(), plan.theta_off.back(), (int32_t)i, hp);
        plan.src_off.push_back(plan.src_off.back() + p->n_sources);
        plan.chi2_off.push_back(plan.chi2_off.back() + p->n_bands);
        plan.dchi2_off.push_back(plan.dchi2_off.back() +
                                 (int64_t)p->n_bands * p->n_sources * kNParams);
        plan.theta_off.push_back(plan.theta_off.back() +
                                 (int64_t)p->n_sources * (p->n_bands + 6));
    }
    if ((int64_t)hp.tiles.size() > INT32_MAX || (int64_t)hp.slot_boxes.size() > INT32_MAX)
        return fail(FPHO_E_LIMIT, "too many tiles (%zu)", hp.tiles.size());
    // Queue order: most expensive tiles first (longest-processing-time first);
    // ties keep canonical order, so tiles of one exposure stay adjacent.
    const std::vector<int32_t> queue = cost_order(hp);

    HIP_TRY(plan.exps.upload(hp.exps.data(), (int64_t)hp.exps.size()));
    HIP_TRY(plan.tiles.upload(hp.tiles.data(), (int64_t)hp.tiles.size()));
    HIP_TRY(plan.slot_boxes.upload(hp.slot_boxes.data(), (int64_t)hp.slot_boxes.size()));
    HIP_TRY(plan.queue.upload(queue.data(), (int64_t)queue.size()));
    HIP_TRY(plan.outs.upload(hp.outs.data(), (int64_t)hp.outs.size()));
    HIP_TRY(plan.srcs.upload(hp.srcs.data(), (int64_t)hp.srcs.size()));
    HIP_TRY(plan.pouts.upload(hp.pouts.data(), (int64_t)hp.pouts.size()));
    HIP_TRY(plan.gauss.alloc(hp.n_gauss));
    HIP_TRY(plan.gauss64.alloc(hp.n_gauss));
    HIP_TRY(plan.gboxes.alloc(hp.n_gauss));
    HIP_TRY(plan.gpart.alloc(std::max<int64_t>(hp.n_part, 1)));
    HIP_TRY(plan.cpart.alloc(std::max<int64_t>((int64_t)hp.tiles.size(), 1)));
    HIP_TRY(plan.counter.alloc(1));
    HIP_TRY(hipMemset(plan.counter.ptr, 0, sizeof(int32_t)));
    HIP_TRY(plan.status.alloc(1));
    HIP_TRY(hipMemset(plan.status.ptr, 0, sizeof(uint32_t)));
    if (patches.size() == 1) {
        const fpho_patch *p = patches[0];
        HIP_TRY(plan.dchi2_tmp.alloc((int64_t)p->n_bands * p->n_sources * kNParams));
    } else {
        plan.dchi2_tmp.release();
    }
    HIP_TRY(plan.src64.alloc(std::max<int64_t>((int64_t)hp.srcs.size(), 1)));
    plan.n_exp = (int64_t)hp.exps.size();
    plan.n_tiles = (int64_t)hp.tiles.size();
    plan.n_outs = (int64_t)hp.outs.size();
    plan.n_gauss = hp.n_gauss;
    plan.n_part = hp.n_part;
    plan.max_ng = hp.max_ng;
    plan.n_srcs = (int64_t)hp.srcs.size();
    plan.n_patches = (int64_t)patches.size();
    plan.n_pairs = hp.n_pairs;
    plan.npix = hp.npix;
    plan.launch = launch;
    plan.valid = true;
    return FPHO_OK;
}

// Device pointers and counts of a plan, wherever its buffers live (the
// plan's own allocations, or the zero-allocation scratch area).
struct PlanPtrs {
    const ExpDesc *exps = nullptr;
    const TileDesc *tiles = nullptr;
    const int32_t *queue = nullptr;
    const int4 *slot_boxes = nullptr;
    const OutDesc *outs = nullptr;
    ImgGauss *gauss = nullptr;
    ImgGauss64 *gauss64 = nullptr;
    int4 *gboxes = nullptr;
    float *gpart = nullptr;
    double *cpart = nullptr;
    int32_t *counter = nullptr;
    uint32_t *status = nullptr;
    int64_t n_exp = 0, n_tiles = 0, n_outs = 0;
    int32_t max_ng = 0;
    fpho_launch launch;
    int precision = kPrecMixed;
    bool timing = false;
    hipEvent_t *ev = nullptr;
};

PlanPtrs plan_ptrs(Plan &plan, uint32_t *status) {
    PlanPtrs q;
    q.exps = plan.exps.ptr;
    q.tiles = plan.tiles.ptr;
    q.queue = plan.queue.ptr;
    q.slot_boxes = plan.slot_boxes.ptr;
    q.outs = plan.outs.ptr;
    q.gauss = plan.gauss.ptr;
    q.gauss64 = plan.gauss64.ptr;
    q.gboxes = plan.gboxes.ptr;
    q.gpart = plan.gpart.ptr;
    q.cpart = plan.cpart.ptr;
    q.counter = plan.counter.ptr;
    q.status = status;
    q.n_exp = plan.n_exp;
    q.n_tiles = plan.n_tiles;
    q.n_outs = plan.n_outs;
    q.max_ng = plan.max_ng;
    q.launch = plan.launch;
    q.precision = plan.precision;
    q.timing = plan.timing;
    q.ev = plan.ev;
    return q;
}

EvalArgs eval_args(const PlanPtrs &q) {
    EvalArgs a;
    a.tiles = q.tiles;
    a.queue = q.queue;
    a.n_tiles = q.n_tiles;
    a.counter = q.counter;
    a.exps = q.exps;
    a.gauss = q.gauss;
    a.gauss64 = q.gauss64;
    a.gboxes = q.gboxes;
    a.slot_boxes = q.slot_boxes;
    a.precision = q.precision;
    a.gpart = q.gpart;
    a.cpart = q.cpart;
    a.pix_per_lane = q.launch.pix_per_lane;
    a.cull = q.launch.cull;
    a.mask_cap = std::min(kMaskCapMax, (q.max_ng + 7) & ~7);
    a.act_words = std::min(kActWordsMax, (q.max_ng + 63) / 64);
    a.waves_per_group = q.launch.waves_per_group;
    a.persistent = q.launch.persistent;
    const int64_t groups_needed = (q.n_tiles + a.waves_per_group - 1) / a.waves_per_group;
    a.grid = std::min<int64_t>(groups_needed,
                               (int64_t)q.launch.groups_per_cu * cu_count_cached());
    if (a.grid < 1) a.grid = 1;
    return a;
}

EvalArgs eval_args(Plan &plan) { return eval_args(plan_ptrs(plan, plan.status.ptr)); }

// convert -> pixel pass -> reduce on `stream`
template <typename Src>
int run_ptrs(const PlanPtrs &q, const Src *props, double *chi2, double *dchi2,
             hipStream_t stream) {
    if (q.timing) HIP_TRY(hipEventRecord(q.ev[0], stream));
    HIP_TRY(launch_convert(q.exps, q.n_exp, props, q.gauss, q.gauss64, q.gboxes, q.status,
                           stream));
    const EvalArgs a = eval_args(q);
    if (q.timing) HIP_TRY(hipEventRecord(q.ev[1], stream));
    HIP_TRY(launch_eval(a, stream));
    if (q.timing) HIP_TRY(hipEventRecord(q.ev[2], stream));
    HIP_TRY(launch_reduce(q.outs, q.n_outs, q.gpart, q.cpart, chi2, dchi2,
                          q.precision != kPrecFP32, stream));
    if (q.timing) HIP_TRY(hipEventRecord(q.ev[3], stream));
    return FPHO_OK;
}

// A single patch's plan reports into the patch's status word, a batch's
// into the plan's.
template <typename Src>
int run_plan(Plan &plan, uint32_t *status, const Src *props, double *chi2, double *dchi2,
             hipStream_t stream) {
    return run_ptrs(plan_ptrs(plan, status), props, chi2, dchi2, stream);
}

}  // namespace

// ------------------------------------------- zero-allocation scratch area
// The work buffers of fpho_evaluate_async_noalloc -- image Gaussians, their
// FP64 copies and boxes, the tile partials and the work-queue counter --
// live in this fixed array.  The HIP runtime places one copy on every GPU
// the library is loaded on and keeps it until the process ends.  In host
// code the array's name only stands in for the symbol: its address on a
// device comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is ever passed to kernels.
constexpr int64_t kScratchBytes = (int64_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kScratchDevices] = {};

// Address of the scratch array on `device`, which must be current; looked up
// on the first call for that device and remembered.
int scratch_address(int device, unsigned char **out) {
    if (device < 0 || device >= kScratchDevices)
        return fail(FPHO_E_LIMIT, "device %d outside 0..%d", device, kScratchDevices - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(FPHO_E_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s",
                        device, hipGetErrorString(e));
        g_scratch_dev[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_dev[device];
    return FPHO_OK;
}

int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

// True when `stream` is capturing a graph (plans are not rebuilt then: a
// rebuild allocates and copies synchronously).
bool capturing(hipStream_t stream) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &st) != hipSuccess) return false;
    return st != hipStreamCaptureStatusNone;
}

int ensure_patch_plan(fpho_patch *p, hipStream_t stream) {
    if (p->plan.valid && p->plan_generation == p->generation) return FPHO_OK;
    if (capturing(stream))
        return fail(FPHO_E_STATE,
                    "the patch changed since its last evaluation and its plan must be "
                    "rebuilt (allocates); evaluate it once outside the graph capture first");
    std::vector<fpho_patch *> one{p};
    const int rc = build_plan(one, p->launch, p->plan);
    if (rc) return rc;
    HIP_TRY(hipMemset(p->status.ptr, 0, sizeof(uint32_t)));
    p->plan.precision = p->precision;
    p->plan_generation = p->generation;
    return FPHO_OK;
}

int ensure_batch_plan(fpho_batch *b, hipStream_t stream) {
    for (size_t i = 0; i < b->patches.size(); ++i)
        if (b->patches[i]->n_sources != b->n_sources0[i])
            return fail(FPHO_E_STATE,
                        "patch %zu of the batch now has %d sources (%d when the batch was "
                        "created); the batch layout no longer holds: create a new batch",
                        i, b->patches[i]->n_sources, b->n_sources0[i]);
    bool stale = !b->plan.valid;
    for (size_t i = 0; i < b->patches.size() && !stale; ++i)
        stale = b->generations[i] != b->patches[i]->generation;
    if (!stale) return FPHO_OK;
    if (capturing(stream))
        return fail(FPHO_E_STATE,
                    "a patch of the batch changed since the last evaluation and the plan must "
                    "be rebuilt (allocates); evaluate once outside the graph capture first");
    const int rc = build_plan(b->patches, b->launch, b->plan);
    if (rc) return rc;
    b->plan.precision = b->precision;
    for (size_t i = 0; i < b->patches.size(); ++i) b->generations[i] = b->patches[i]->generation;
    return FPHO_OK;
}

// Wait for the device, then read (and with `clear`, reset) a status word.
int read_status(uint32_t *word_dev, uint32_t *status, int clear) {
    *status = 0;
    if (!word_dev) return FPHO_OK;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(status, word_dev, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (clear) HIP_TRY(hipMemset(word_dev, 0, sizeof(uint32_t)));
    return FPHO_OK;
}

int check_proposal_host(const fpho_patch *p, const fpho_source *prop) {
    for (int s = 0; s < p->n_sources; ++s) {
        const fpho_source &src = prop[s];
        if (!(src.q > 0.0f) || !std::isfinite(src.q) || !std::isfinite(src.pa) ||
            !std::isfinite(src.ra) || !std::isfinite(src.dec))
            return fail(FPHO_E_VALUE, "source %d: q must be > 0 and positions finite", s);
        for (int b = 0; b < p->n_bands; ++b)
            if (!std::isfinite(src.fluxes[b]))
                return fail(FPHO_E_VALUE, "source %d: non-finite flux in band %d", s, b);
        for (int m = 0; m < p->n_radii; ++m)
            if (!std::isfinite(src.mixture_amplitudes[m]) ||
                !std::isfinite(src.damplitude_drh[m]) ||
                !std::isfinite(src.damplitude_dnsersic[m]))
                return fail(FPHO_E_VALUE, "source %d: non-finite mixture amplitude", s);
    }
    return FPHO_OK;
}

// Synchronous evaluation through the patch's page-locked staging: host
// proposal in, chi2 [n_bands] and dchi2 [n_bands][n_sources][7] (double) out
// in out_host.
int evaluate_sync(fpho_patch *p, const fpho_source *prop) {
    int rc = check_proposal_host(p, prop);
    if (rc) return rc;
    rc = ensure_patch_plan(p, nullptr);
    if (rc) return rc;
    const int64_t nd = (int64_t)p->n_bands * p->n_sources * kNParams;
    const int64_t nout = p->n_bands + nd;
    if (p->prop_dev.n < p->n_sources) HIP_TRY(p->prop_dev.alloc(p->n_sources));
    if (p->out_dev.n < nout) HIP_TRY(p->out_dev.alloc(nout));
    if (p->prop_host.n < p->n_sources) HIP_TRY(p->prop_host.alloc(p->n_sources));
    if (p->out_host.n < nout) HIP_TRY(p->out_host.alloc(nout));
    std::memcpy(p->prop_host.ptr, prop, sizeof(fpho_source) * p->n_sources);
    HIP_TRY(hipMemcpyAsync(p->prop_dev.ptr, p->prop_host.ptr,
                           sizeof(fpho_source) * p->n_sources, hipMemcpyHostToDevice, nullptr));
    rc = run_plan(p->plan, p->status.ptr, p->prop_dev.ptr, p->out_dev.ptr,
                  p->out_dev.ptr + p->n_bands,
                  nullptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(p->out_host.ptr, p->out_dev.ptr, sizeof(double) * nout,
                           hipMemcpyDeviceToHost, nullptr));
    HIP_TRY(hipStreamSynchronize(nullptr));
    return FPHO_OK;
}

}  // namespace

// =================================================================== C ABI
extern "C" {

const char *fpho_last_error(void) { return g_err.c_str(); }

int fpho_version(void) { return 1; }

int fpho_init(int device) {
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(FPHO_E_ARG, "device %d outside 0..%d", device, n - 1);
    HIP_TRY(hipSetDevice(device));
    return FPHO_OK;
}

int fpho_device_cu_count(int *n_cu) {
    if (!n_cu) return fail(FPHO_E_ARG, "null n_cu");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    return FPHO_OK;
}

int fpho_patch_create(const fpho_patch_desc *d, int return_residual, fpho_patch **out) {
    if (!d || !out) return fail(FPHO_E_ARG, "null argument");
    *out = nullptr;
    int rc = validate_pixels(d);
    if (rc) return rc;
    rc = validate_meta(d, d->n_bands, d->n_exp);
    if (rc) return rc;
    fpho_patch *p = new (std::nothrow) fpho_patch();
    if (!p) return fail(FPHO_E_LIMIT, "out of host memory");
    (void)hipGetDevice(&p->device);
    p->npix = d->npix;
    p->n_bands = d->n_bands;
    p->n_exp = d->n_exp;
    p->band_start.assign(d->band_start, d->band_start + d->n_bands);
    p->band_N.assign(d->band_N, d->band_N + d->n_bands);
    p->exposure_start.assign(d->exposure_start, d->exposure_start + d->n_exp);
    p->exposure_N.assign(d->exposure_N, d->exposure_N + d->n_exp);
    p->write_residual = return_residual != 0;
    p->chunk_start.assign(1, 0);
    for (int ex = 0; ex < d->n_exp; ++ex) {
        const int64_t s0 = d->exposure_start[ex], n = d->exposure_N[ex];
        // origin: the exposure's first pixel (a real pixel: padding comes last)
        const float ox = n > 0 ? std::floor(d->xpix[s0]) : 0.0f;
        const float oy = n > 0 ? std::floor(d->ypix[s0]) : 0.0f;
        p->x0.push_back(ox);
        p->y0.push_back(oy);
        for (int64_t c = 0; c < n; c += kWave) {
            float xlo = d->xpix[s0 + c] - ox, xhi = xlo, ylo = d->ypix[s0 + c] - oy, yhi = ylo;
            for (int64_t i = s0 + c; i < s0 + std::min<int64_t>(n, c + kWave); ++i) {
                xlo = std::min(xlo, d->xpix[i] - ox);
                xhi = std::max(xhi, d->xpix[i] - ox);
                ylo = std::min(ylo, d->ypix[i] - oy);
                yhi = std::max(yhi, d->ypix[i] - oy);
            }
            p->chunk_box.push_back(order_key(xlo));
            p->chunk_box.push_back(order_key(xhi));
            p->chunk_box.push_back(order_key(ylo));
            p->chunk_box.push_back(order_key(yhi));
        }
        p->chunk_start.push_back((int64_t)p->chunk_box.size() / 4);
    }
    hipError_t e = hipSuccess;
    if ((e = p->data.upload(d->data, d->npix)) != hipSuccess ||
        (e = p->ierr.upload(d->ierr, d->npix)) != hipSuccess ||
        (e = p->xpix.upload(d->xpix, d->npix)) != hipSuccess ||
        (e = p->ypix.upload(d->ypix, d->npix)) != hipSuccess) {
        delete p;
        return fail(FPHO_E_HIP, "pixel upload failed: %s", hipGetErrorString(e));
    }
    if (p->write_residual) {
        if ((e = p->residual.alloc(d->npix)) != hipSuccess ||
            (e = hipMemset(p->residual.ptr, 0, sizeof(float) * d->npix)) != hipSuccess) {
            delete p;
            return fail(FPHO_E_HIP, "residual allocation failed: %s", hipGetErrorString(e));
        }
    }
    if ((e = p->status.alloc(1)) != hipSuccess ||
        (e = hipMemset(p->status.ptr, 0, sizeof(uint32_t))) != hipSuccess) {
        delete p;
        return fail(FPHO_E_HIP, "status word allocation failed: %s", hipGetErrorString(e));
    }
    rc = upload_meta(p, d);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return FPHO_OK;
}

int fpho_patch_set_meta(fpho_patch *p, const fpho_patch_desc *d) {
    if (!p || !d) return fail(FPHO_E_ARG, "null argument");
    fpho_patch_desc m = *d;
    m.n_bands = p->n_bands;
    m.n_exp = p->n_exp;
    m.band_start = p->band_start.data();
    m.band_N = p->band_N.data();
    const int rc = validate_meta(&m, p->n_bands, p->n_exp);
    if (rc) return rc;
    return upload_meta(p, d);
}

int fpho_patch_swap_data_residual(fpho_patch *p) {
    if (!p) return fail(FPHO_E_ARG, "null patch");
    if (!p->residual.ptr)
        return fail(FPHO_E_STATE, "patch has no residual buffer (create it with "
                                  "return_residual or call fpho_patch_set_residual_writes)");
    p->data.swap(p->residual);
    ++p->generation;
    return FPHO_OK;
}

int fpho_patch_read(fpho_patch *p, int which, float *host_out) {
    if (!p || !host_out) return fail(FPHO_E_ARG, "null argument");
    const DevVec<float> *src = nullptr;
    switch (which) {
        case 0: src = &p->data; break;
        case 1: src = &p->ierr; break;
        case 2: src = &p->xpix; break;
        case 3: src = &p->ypix; break;
        case 4: src = &p->residual; break;
        default: return fail(FPHO_E_ARG, "unknown array %d", which);
    }
    if (!src->ptr) return fail(FPHO_E_STATE, "array %d is not allocated", which);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host_out, src->ptr, sizeof(float) * p->npix, hipMemcpyDeviceToHost));
    return FPHO_OK;
}

int fpho_patch_info(const fpho_patch *p, int32_t *n_bands, int32_t *n_sources, int64_t *npix,
                    int64_t *n_pairs) {
    if (!p) return fail(FPHO_E_ARG, "null patch");
    if (n_bands) *n_bands = p->n_bands;
    if (n_sources) *n_sources = p->n_sources;
    if (npix) *npix = p->npix;
    if (n_pairs) {
        int64_t acc = 0;
        for (int b = 0; b < p->n_bands; ++b)
            for (int e = p->band_start[b]; e < p->band_start[b] + p->band_N[b]; ++e)
                acc += (int64_t)p->exposure_N[e] * p->n_sources * p->n_psf_per_source[b];
        *n_pairs = acc;
    }
    return FPHO_OK;
}

int fpho_patch_set_launch(fpho_patch *p, const fpho_launch *l) {
    if (!p || !l) return fail(FPHO_E_ARG, "null argument");
    const int rc = check_launch(*l);
    if (rc) return rc;
    p->launch = *l;
    p->plan.valid = false;
    return FPHO_OK;
}

int fpho_patch_set_precision(fpho_patch *p, int mode) {
    if (!p) return fail(FPHO_E_ARG, "null patch");
    if (mode < kPrecFP32 || mode > kPrecMixedCentre) return fail(FPHO_E_ARG, "precision mode %d", mode);
    p->precision = mode;
    p->plan.precision = mode;
    return FPHO_OK;
}

int fpho_patch_set_residual_writes(fpho_patch *p, int on) {
    if (!p) return fail(FPHO_E_ARG, "null patch");
    const bool w = on != 0;
    if (w && !p->residual.ptr) {
        HIP_TRY(p->residual.alloc(p->npix));
        HIP_TRY(hipMemset(p->residual.ptr, 0, sizeof(float) * p->npix));
    }
    if (w != p->write_residual) {
        p->write_residual = w;
        ++p->generation;
    }
    return FPHO_OK;
}

int fpho_patch_status(fpho_patch *p, uint32_t *status, int clear) {
    if (!p || !status) return fail(FPHO_E_ARG, "null argument");
    return read_status(p->status.ptr, status, clear);
}

int fpho_patch_free(fpho_patch *p) {
    if (!p) return FPHO_OK;
    (void)hipDeviceSynchronize();
    delete p;
    return FPHO_OK;
}

int fpho_evaluate(fpho_patch *p, const fpho_source *prop, double *chi2, float *dchi2) {
    if (!p || !prop || !chi2 || !dchi2) return fail(FPHO_E_ARG, "null argument");
    const int rc = evaluate_sync(p, prop);
    if (rc) return rc;
    const int64_t nd = (int64_t)p->n_bands * p->n_sources * kNParams;
    std::memcpy(chi2, p->out_host.ptr, sizeof(double) * p->n_bands);
    for (int64_t i = 0; i < nd; ++i) dchi2[i] = (float)p->out_host.ptr[p->n_bands + i];
    return FPHO_OK;
}

int fpho_evaluate_async(fpho_patch *p, const fpho_source *prop_dev, double *chi2_dev,
                        double *dchi2_dev, void *stream) {
    if (!p || !prop_dev || !chi2_dev || !dchi2_dev) return fail(FPHO_E_ARG, "null argument");
    int rc = ensure_patch_plan(p, (hipStream_t)stream);
    if (rc) return rc;
    return run_plan(p->plan, p->status.ptr, prop_dev, chi2_dev, dchi2_dev,
                    (hipStream_t)stream);
}

int fpho_evaluate_async_noalloc(fpho_patch *p, const fpho_source *prop_dev, double *chi2_dev,
                                double *dchi2_dev, void *stream) {
    if (!p || !prop_dev || !chi2_dev || !dchi2_dev) return fail(FPHO_E_ARG, "null argument");
    int device = 0;
    HIP_TRY(hipGetDevice(&device));
    if (device != p->device)
        return fail(FPHO_E_STATE, "the patch lives on device %d, device %d is current",
                    p->device, device);
    Plan &plan = p->plan;
    if (!plan.valid || p->plan_generation != p->generation)
        return fail(FPHO_E_STATE,
                    "the patch changed since its descriptors were built; evaluate it once "
                    "with fpho_evaluate_async (which allocates) before evaluating it here");
    // work buffers in the scratch array, each 256-byte aligned
    int64_t off = 0;
    auto place = [&](int64_t bytes) {
        const int64_t at = off;
        off = align256(off + bytes);
        return at;
    };
    const int64_t o_gauss = place((int64_t)sizeof(ImgGauss) * plan.n_gauss);
    const int64_t o_gauss64 = place((int64_t)sizeof(ImgGauss64) * plan.n_gauss);
    const int64_t o_gboxes = place((int64_t)sizeof(int4) * plan.n_gauss);
    const int64_t o_gpart = place((int64_t)sizeof(float) * std::max<int64_t>(plan.n_part, 1));
    const int64_t o_cpart = place((int64_t)sizeof(double) * std::max<int64_t>(plan.n_tiles, 1));
    const int64_t o_counter = place((int64_t)sizeof(int32_t));
    if (off > kScratchBytes)
        return fail(FPHO_E_LIMIT,
                    "the patch needs %lld bytes of work buffers (%lld image Gaussians, %lld "
                    "tiles), more than the %lld-byte scratch area of the zero-allocation "
                    "entry point; use fpho_evaluate_async",
                    (long long)off, (long long)plan.n_gauss, (long long)plan.n_tiles,
                    (long long)kScratchBytes);
    unsigned char *base = nullptr;
    const int rc = scratch_address(device, &base);
    if (rc) return rc;
    PlanPtrs q = plan_ptrs(plan, p->status.ptr);
    q.gauss = reinterpret_cast<ImgGauss *>(base + o_gauss);
    q.gauss64 = reinterpret_cast<ImgGauss64 *>(base + o_gauss64);
    q.gboxes = reinterpret_cast<int4 *>(base + o_gboxes);
    q.gpart = reinterpret_cast<float *>(base + o_gpart);
    q.cpart = reinterpret_cast<double *>(base + o_cpart);
    q.counter = reinterpret_cast<int32_t *>(base + o_counter);
    q.timing = false;
    return run_ptrs(q, prop_dev, chi2_dev, dchi2_dev, (hipStream_t)stream);
}

int fpho_evaluate_response(fpho_patch *p, const fpho_source *prop, double *chi2,
                           fpho_response *resp) {
    if (!p || !prop || !chi2 || !resp) return fail(FPHO_E_ARG, "null argument");
    if (p->n_sources > FPHO_MAXSOURCES)
        return fail(FPHO_E_LIMIT, "%d sources do not fit the Response layout (MAXSOURCES %d)",
                    p->n_sources, FPHO_MAXSOURCES);
    const int rc = evaluate_sync(p, prop);
    if (rc) return rc;
    std::memcpy(chi2, p->out_host.ptr, sizeof(double) * p->n_bands);
    const int nw = p->n_sources * kNParams;
    for (int b = 0; b < p->n_bands; ++b) {
        const double *src = p->out_host.ptr + p->n_bands + (int64_t)b * nw;
        for (int j = 0; j < FPHO_MAXSOURCES * kNParams; ++j)
            resp[b].dchi2_dparam[j] = j < nw ? (float)src[j] : 0.0f;
    }
    return FPHO_OK;
}

int fpho_evaluate_response_async(fpho_patch *p, const fpho_source *prop_dev, double *chi2_dev,
                                 fpho_response *resp_dev, void *stream) {
    if (!p || !prop_dev || !chi2_dev || !resp_dev) return fail(FPHO_E_ARG, "null argument");
    if (p->n_sources > FPHO_MAXSOURCES)
        return fail(FPHO_E_LIMIT, "%d sources do not fit the Response layout (MAXSOURCES %d)",
                    p->n_sources, FPHO_MAXSOURCES);
    int rc = ensure_patch_plan(p, (hipStream_t)stream);
    if (rc) return rc;
    rc = run_plan(p->plan, p->status.ptr, prop_dev, chi2_dev, p->plan.dchi2_tmp.ptr,
                  (hipStream_t)stream);
    if (rc) return rc;
    HIP_TRY(launch_response(p->plan.dchi2_tmp.ptr, p->n_bands, p->n_sources, resp_dev,
                            (hipStream_t)stream));
    return FPHO_OK;
}

int fpho_batch_create(fpho_patch *const *patches, int n, fpho_batch **out) {
    if (!patches || !out || n < 1) return fail(FPHO_E_ARG, "need at least one patch");
    *out = nullptr;
    fpho_batch *b = new (std::nothrow) fpho_batch();
    if (!b) return fail(FPHO_E_LIMIT, "out of host memory");
    for (int i = 0; i < n; ++i) {
        if (!patches[i]) {
            delete b;
            return fail(FPHO_E_ARG, "patch %d is null", i);
        }
        b->patches.push_back(patches[i]);
        b->generations.push_back(0);
        b->n_sources0.push_back(patches[i]->n_sources);
    }
    const int rc = ensure_batch_plan(b, nullptr);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return FPHO_OK;
}

int fpho_batch_layout(const fpho_batch *b, int64_t *src_off, int64_t *chi2_off,
                      int64_t *dchi2_off, int64_t *theta_off) {
    if (!b) return fail(FPHO_E_ARG, "null batch");
    const size_t n = b->plan.src_off.size();
    for (size_t i = 0; i < n; ++i) {
        if (src_off) src_off[i] = b->plan.src_off[i];
        if (chi2_off) chi2_off[i] = b->plan.chi2_off[i];
        if (dchi2_off) dchi2_off[i] = b->plan.dchi2_off[i];
        if (theta_off) theta_off[i] = b->plan.theta_off[i];
    }
    return FPHO_OK;
}

int fpho_batch_info(const fpho_batch *b, int64_t *n_tiles, int64_t *n_pairs, int64_t *npix,
                    int64_t *n_gauss) {
    if (!b) return fail(FPHO_E_ARG, "null batch");
    if (n_tiles) *n_tiles = b->plan.n_tiles;
    if (n_pairs) *n_pairs = b->plan.n_pairs;
    if (npix) *npix = b->plan.npix;
    if (n_gauss) *n_gauss = b->plan.n_gauss;
    return FPHO_OK;
}

int fpho_batch_set_launch(fpho_batch *b, const fpho_launch *l) {
    if (!b || !l) return fail(FPHO_E_ARG, "null argument");
    const int rc = check_launch(*l);
    if (rc) return rc;
    b->launch = *l;
    b->plan.valid = false;    // rebuilt now; on failure, again at the next call
    return ensure_batch_plan(b, nullptr);
}

int fpho_batch_set_precision(fpho_batch *b, int mode) {
    if (!b) return fail(FPHO_E_ARG, "null batch");
    if (mode < kPrecFP32 || mode > kPrecMixedCentre) return fail(FPHO_E_ARG, "precision mode %d", mode);
    b->precision = mode;
    b->plan.precision = mode;
    return FPHO_OK;
}

int fpho_batch_evaluate_async(fpho_batch *b, const fpho_source *props, double *chi2,
                              double *dchi2, void *stream) {
    if (!b || !props || !chi2 || !dchi2) return fail(FPHO_E_ARG, "null argument");
    const int rc = ensure_batch_plan(b, (hipStream_t)stream);
    if (rc) return rc;
    return run_plan(b->plan, b->plan.status.ptr, props, chi2, dchi2, (hipStream_t)stream);
}

int fpho_batch_set_timing(fpho_batch *b, int on) {
    if (!b) return fail(FPHO_E_ARG, "null batch");
    if (on) {
        for (auto &e : b->plan.ev)
            if (!e) HIP_TRY(hipEventCreate(&e));
    }
    b->plan.timing = on != 0;
    return FPHO_OK;
}

int fpho_batch_last_timing(fpho_batch *b, float *ms3) {
    if (!b || !ms3) return fail(FPHO_E_ARG, "null argument");
    if (!b->plan.timing) return fail(FPHO_E_STATE, "timing is not enabled");
    HIP_TRY(hipEventSynchronize(b->plan.ev[3]));
    for (int i = 0; i < 3; ++i)
        HIP_TRY(hipEventElapsedTime(&ms3[i], b->plan.ev[i], b->plan.ev[i + 1]));
    return FPHO_OK;
}

int fpho_batch_status(fpho_batch *b, uint32_t *status, int clear) {
    if (!b || !status) return fail(FPHO_E_ARG, "null argument");
    return read_status(b->plan.valid ? b->plan.status.ptr : nullptr, status, clear);
}

int fpho_batch_free(fpho_batch *b) {
    if (!b) return FPHO_OK;
    (void)hipDeviceSynchronize();
    delete b;
    return FPHO_OK;
}

int fpho_spline_create(const double *coeffs, int n_amp, const double *tx, int n_tx,
                       const double *ty, int n_ty, const double *radii, double nsersic_min,
                       fpho_spline **out) {
    if (!coeffs || !tx || !ty || !radii || !out) return fail(FPHO_E_ARG, "null argument");
    *out = nullptr;
    if (n_amp < 1 || n_amp > FPHO_MAXRADII) return fail(FPHO_E_LIMIT, "n_amp=%d", n_amp);
    if (n_tx < 8 || n_ty < 8) return fail(FPHO_E_ARG, "cubic splines need >= 8 knots");
    for (int i = 1; i < n_tx; ++i)
        if (!(tx[i] >= tx[i - 1])) return fail(FPHO_E_ARG, "tx not sorted");
    for (int i = 1; i < n_ty; ++i)
        if (!(ty[i] >= ty[i - 1])) return fail(FPHO_E_ARG, "ty not sorted");
    fpho_spline *s = new (std::nothrow) fpho_spline();
    if (!s) return fail(FPHO_E_LIMIT, "out of host memory");
    const int64_t nc = (int64_t)n_amp * (n_tx - 4) * (n_ty - 4);
    hipError_t e;
    if ((e = s->coeffs.upload(coeffs, nc)) != hipSuccess ||
        (e = s->tx.upload(tx, n_tx)) != hipSuccess ||
        (e = s->ty.upload(ty, n_ty)) != hipSuccess ||
        (e = s->radii.upload(radii, n_amp)) != hipSuccess) {
        delete s;
        return fail(FPHO_E_HIP, "spline upload failed: %s", hipGetErrorString(e));
    }
    s->view.coeffs = s->coeffs.ptr;
    s->view.tx = s->tx.ptr;
    s->view.ty = s->ty.ptr;
    s->view.radii = s->radii.ptr;
    s->view.n_tx = n_tx;
    s->view.n_ty = n_ty;
    s->view.n_amp = n_amp;
    s->view.pad = 0;
    s->view.nsersic_min = nsersic_min;
    *out = s;
    return FPHO_OK;
}

int fpho_spline_free(fpho_spline *s) {
    if (!s) return FPHO_OK;
    (void)hipDeviceSynchronize();
    delete s;
    return FPHO_OK;
}

int fpho_batch_theta_to_proposals_async(fpho_batch *b, const fpho_spline *s,
                                        const double *theta, fpho_source *props,
                                        void *stream) {
    if (!b || !s || !theta || !props) return fail(FPHO_E_ARG, "null argument");
    const int rc = ensure_batch_plan(b, (hipStream_t)stream);
    if (rc) return rc;
    HIP_TRY(launch_theta(b->plan.srcs.ptr, b->plan.n_srcs, theta, s->view, props, nullptr,
                         b->plan.status.ptr, (hipStream_t)stream));
    return FPHO_OK;
}

int fpho_batch_evaluate_theta_async(fpho_batch *b, const fpho_spline *s, const double *theta,
                                    fpho_source *props, double *chi2, double *dchi2,
                                    void *stream) {
    if (!b || !s || !theta || !chi2 || !dchi2) return fail(FPHO_E_ARG, "null argument");
    const int rc = ensure_batch_plan(b, (hipStream_t)stream);
    if (rc) return rc;
    HIP_TRY(launch_theta(b->plan.srcs.ptr, b->plan.n_srcs, theta, s->view, props,
                         b->plan.src64.ptr, b->plan.status.ptr, (hipStream_t)stream));
    return run_plan(b->plan, b->plan.status.ptr, (const Source64 *)b->plan.src64.ptr, chi2,
                    dchi2,
                    (hipStream_t)stream);
}

int fpho_batch_stack_grad_async(fpho_batch *b, const double *chi2, const double *dchi2,
                                double *lnlike, double *grad, void *stream) {
    if (!b || !chi2 || !dchi2 || !lnlike || !grad) return fail(FPHO_E_ARG, "null argument");
    const int rc = ensure_batch_plan(b, (hipStream_t)stream);
    if (rc) return rc;
    HIP_TRY(launch_stack(b->plan.srcs.ptr, b->plan.n_srcs, b->plan.pouts.ptr,
                         b->plan.n_patches, chi2, dchi2, lnlike, grad, (hipStream_t)stream));
    return FPHO_OK;
}

}  // extern "C"

// Internal hook of the measurement library (declared in kernels.h, not in the
// C ABI).
namespace fpho {
int batch_plan_view(fpho_batch *b, PlanView *out) {
    if (!b || !out) return fail(FPHO_E_ARG, "null argument");
    const int rc = ensure_batch_plan(b, nullptr);
    if (rc) return rc;
    out->eval = eval_args(b->plan);
    out->n_exp = b->plan.n_exp;
    out->status = b->plan.status.ptr;
    return FPHO_OK;
}
}  // namespace fpho
