/* A plain C host of libfpho_hip.so (test code): no Python, no torch.  It
 * reads a packed patch and proposals from raw little-endian files in the
 * directory argv[1] (written by tests/test_gpu_c_host.py), creates the
 * device patch with fpho_patch_create, evaluates every proposal with the
 * synchronous drop-in fpho_evaluate and, when the sources fit the
 * reference's MAXSOURCES, with its compat form fpho_evaluate_response, and
 * writes chi2 (double), dchi2 (float) and the Response blocks back to that
 * directory.  This is the call sequence a C, Go or
 * Julia binding of forcepho's GPUProposer.evaluate_proposal makes
 * (forcepho/proposal.py:96-152). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fpho.h"

static const char *g_dir;

static void *read_file(const char *name, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", g_dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    void *buf = malloc(bytes > 0 ? bytes : 1);
    if (!buf || fread(buf, 1, bytes, f) != bytes) {
        fprintf(stderr, "short read of %s (%zu bytes)\n", path, bytes);
        exit(2);
    }
    fclose(f);
    return buf;
}

static void write_file(const char *name, const void *buf, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", g_dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(buf, 1, bytes, f) != bytes) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(2);
    }
    fclose(f);
}

#define CHECK(call)                                                              \
    do {                                                                         \
        int rc_ = (call);                                                        \
        if (rc_ != FPHO_OK) {                                                    \
            fprintf(stderr, "%s -> %d: %s\n", #call, rc_, fpho_last_error());    \
            return 1;                                                            \
        }                                                                        \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s DIR\n", argv[0]);
        return 2;
    }
    g_dir = argv[1];
    long long npix, n_bands, n_exp, n_sources, n_radii, n_psfgauss, n_props;
    {
        char path[4096];
        snprintf(path, sizeof(path), "%s/meta.txt", g_dir);
        FILE *f = fopen(path, "r");
        if (!f || fscanf(f, "%lld %lld %lld %lld %lld %lld %lld", &npix, &n_bands, &n_exp,
                         &n_sources, &n_radii, &n_psfgauss, &n_props) != 7) {
            fprintf(stderr, "bad %s\n", path);
            return 2;
        }
        fclose(f);
    }
    fpho_patch_desc d;
    memset(&d, 0, sizeof(d));
    d.npix = npix;
    d.data = read_file("data.f32", sizeof(float) * npix);
    d.ierr = read_file("ierr.f32", sizeof(float) * npix);
    d.xpix = read_file("xpix.f32", sizeof(float) * npix);
    d.ypix = read_file("ypix.f32", sizeof(float) * npix);
    d.n_bands = (int32_t)n_bands;
    d.n_exp = (int32_t)n_exp;
    d.band_start = read_file("band_start.i16", sizeof(int16_t) * n_bands);
    d.band_N = read_file("band_N.i16", sizeof(int16_t) * n_bands);
    d.exposure_start = read_file("exposure_start.i32", sizeof(int32_t) * n_exp);
    d.exposure_N = read_file("exposure_N.i32", sizeof(int32_t) * n_exp);
    d.n_sources = (int32_t)n_sources;
    d.n_radii = (int32_t)n_radii;
    d.rad2 = read_file("rad2.f32", sizeof(float) * n_radii);
    d.D = read_file("D.f32", sizeof(float) * n_exp * n_sources * 4);
    d.crpix = read_file("crpix.f32", sizeof(float) * n_exp * n_sources * 2);
    d.crval = read_file("crval.f32", sizeof(float) * n_exp * n_sources * 2);
    d.CW = read_file("CW.f32", sizeof(float) * n_exp * n_sources * 4);
    d.G = read_file("G.f32", sizeof(float) * n_exp);
    d.n_psf_per_source = read_file("n_psf_per_source.i32", sizeof(int32_t) * n_bands);
    d.psfgauss = read_file("psfgauss.bin", sizeof(fpho_psf_gauss) * n_psfgauss);
    d.n_psfgauss = n_psfgauss;
    d.psfgauss_start = read_file("psfgauss_start.i32", sizeof(int32_t) * n_exp);
    const fpho_source *props =
        read_file("proposals.bin", sizeof(fpho_source) * n_sources * n_props);

    CHECK(fpho_init(0));
    fpho_patch *p = NULL;
    CHECK(fpho_patch_create(&d, 0, &p));
    double *chi2 = malloc(sizeof(double) * n_bands * n_props);
    float *dchi2 = malloc(sizeof(float) * n_bands * n_sources * 7 * n_props);
    for (long long k = 0; k < n_props; ++k)
        CHECK(fpho_evaluate(p, props + k * n_sources, chi2 + k * n_bands,
                            dchi2 + k * n_bands * n_sources * 7));
    /* the same proposals through the compat output: Response[n_bands] per
     * proposal (forcepho/src/proposal.cu:33-39), for patches whose sources
     * fit its MAXSOURCES rows */
    long long n_resp = n_sources <= FPHO_MAXSOURCES ? n_props : 0;
    double *chi2_r = malloc(sizeof(double) * n_bands * (n_resp > 0 ? n_resp : 1));
    fpho_response *resp = malloc(sizeof(fpho_response) * n_bands * (n_resp > 0 ? n_resp : 1));
    for (long long k = 0; k < n_resp; ++k)
        CHECK(fpho_evaluate_response(p, props + k * n_sources, chi2_r + k * n_bands,
                                     resp + k * n_bands));
    CHECK(fpho_patch_free(p));
    write_file("chi2.f64", chi2, sizeof(double) * n_bands * n_props);
    write_file("dchi2.f32", dchi2, sizeof(float) * n_bands * n_sources * 7 * n_props);
    write_file("chi2_response.f64", chi2_r, sizeof(double) * n_bands * n_resp);
    write_file("response.bin", resp, sizeof(fpho_response) * n_bands * n_resp);
    free(chi2_r);
    free(resp);
    printf("evaluated %lld proposals of a %lld-pixel patch, %lld bands, %lld sources\n",
           n_props, npix, n_bands, n_sources);
    const void *owned[] = {d.data, d.ierr, d.xpix, d.ypix, d.band_start, d.band_N,
                           d.exposure_start, d.exposure_N, d.rad2, d.D, d.crpix, d.crval,
                           d.CW, d.G, d.n_psf_per_source, d.psfgauss, d.psfgauss_start, props};
    for (size_t i = 0; i < sizeof(owned) / sizeof(owned[0]); ++i) free((void *)owned[i]);
    free(chi2);
    free(dchi2);
    return 0;
}
