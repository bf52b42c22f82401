/* fpho_oracle.c -- FP64 CPU restatement of forcepho's likelihood and gradient.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never part of the product.
 *
 * Follows, in double precision:
 *   - source set-up and convolution: forcepho/src/compute_gaussians_kernel.cc:
 *     45-114 (T = D R(pa) S(q), centre = CW (sky - crval) + crpix, PSF block at
 *     psfgauss_start[e] for every source) with the Jacobian written as in
 *     forcepho/slow/gaussmodel.py:273-332 (dSigma = T s dT^T + dT s T^T,
 *     dF = -F dSigma F, dA/dq = K/2 tr(Sigma dF/dq)); divisions by amplitude
 *     or flux of the slow path are written as the equivalent products.
 *   - pixel value and image-parameter derivatives: forcepho/slow/gaussmodel.py:
 *     351-476 (second-order pixel-integral term H = 1 + (vx^2+vy^2-fxx-fyy)/24).
 *   - per-pair cutoff exp argument > 36 skipped (common_kernel.cc:73,108;
 *     kernel_limits.h:17), switchable.
 *   - chi2 convention sum (r^2 - d^2) ierr^2 (compute_gaussians_kernel.cu:
 *     277-286), gradient dchi2/dtheta = sum -2 r ierr^2 dm/dtheta.
 *
 * The patch description has the field order of fpho_patch_desc; it is declared
 * here independently so the oracle does not include product headers.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t npix;
    const float *data, *ierr, *xpix, *ypix;
    int32_t n_bands, n_exp;
    const int16_t *band_start, *band_N;
    const int32_t *exposure_start, *exposure_N;
    int32_t n_sources, n_radii;
    const float *rad2, *D, *crpix, *crval, *CW, *G;
    const int32_t *n_psf_per_source;
    const void *psfgauss;
    int64_t n_psfgauss;
    const int32_t *psfgauss_start;
} oracle_patch;

/* fpho_source / Source record: 66 floats */
enum { SRC_RA = 0, SRC_DEC = 1, SRC_Q = 2, SRC_PA = 3, SRC_FLUX = 6, SRC_AMP = 36,
       SRC_DADR = 46, SRC_DADN = 56, SRC_STRIDE = 66 };

typedef struct {
    float amp, xcen, ycen, Cxx, Cyy, Cxy;
    int32_t bin;
} psf_rec;

typedef struct {
    double amp, x, y, fxx, fyy, fxy;
    double jac[7][6]; /* rows: flux, ra, dec, q, pa, n, rh; cols A,x,y,Fxx,Fyy,Fxy */
} gauss64;

static void mat_mul(const double a[4], const double b[4], double o[4]) {
    double t[4];
    t[0] = a[0] * b[0] + a[1] * b[2];
    t[1] = a[0] * b[1] + a[1] * b[3];
    t[2] = a[2] * b[0] + a[3] * b[2];
    t[3] = a[2] * b[1] + a[3] * b[3];
    memcpy(o, t, sizeof(t));
}
static void mat_T(const double a[4], double o[4]) {
    double t[4] = {a[0], a[2], a[1], a[3]};
    memcpy(o, t, sizeof(t));
}

/* Image Gaussians of source s in exposure e of band b. */
static void convert_source(const oracle_patch *p, const double *src, int e, int b, int s,
                           gauss64 *out) {
    const int ns = p->n_sources;
    const int np = p->n_psf_per_source[b];
    const float *Df = p->D + 4 * ((int64_t)ns * e + s);
    const float *CWf = p->CW + 4 * ((int64_t)ns * e + s);
    const float *crpix = p->crpix + 2 * ((int64_t)ns * e + s);
    const float *crval = p->crval + 2 * ((int64_t)ns * e + s);
    double D[4], CW[4];
    for (int i = 0; i < 4; ++i) {
        D[i] = Df[i];
        CW[i] = CWf[i];
    }
    const double q = src[SRC_Q], pa = src[SRC_PA];
    const double c = cos(pa), sn = sin(pa);
    const double R[4] = {c, -sn, sn, c};
    const double dR[4] = {-sn, -c, c, -sn};
    const double S[4] = {1.0 / q, 0.0, 0.0, q};
    const double dS[4] = {-1.0 / (q * q), 0.0, 0.0, 1.0};
    double T[4], dTq[4], dTp[4], tmp[4];
    mat_mul(D, R, tmp);
    mat_mul(tmp, S, T);
    mat_mul(tmp, dS, dTq);
    mat_mul(D, dR, tmp);
    mat_mul(tmp, S, dTp);
    const double s0 = (double)src[SRC_RA] - crval[0];
    const double s1 = (double)src[SRC_DEC] - crval[1];
    const double xs = CW[0] * s0 + CW[1] * s1 + crpix[0];
    const double ys = CW[2] * s0 + CW[3] * s1 + crpix[1];
    const double flux = src[SRC_FLUX + b];
    const double G = p->G[e];
    const psf_rec *psf = (const psf_rec *)p->psfgauss + p->psfgauss_start[e];
    double Tt[4], TTt[4], dTqt[4], dTpt[4];
    mat_T(T, Tt);
    mat_T(dTq, dTqt);
    mat_T(dTp, dTpt);
    mat_mul(T, Tt, TTt);
    for (int k = 0; k < np; ++k) {
        const psf_rec pg = psf[k];
        const int bin = pg.bin;
        const double r2 = p->rad2[bin];
        const double am = src[SRC_AMP + bin];
        const double dadn = src[SRC_DADN + bin];
        const double dadr = src[SRC_DADR + bin];
        double Sig[4] = {r2 * TTt[0] + pg.Cxx, r2 * TTt[1] + pg.Cxy, r2 * TTt[2] + pg.Cxy,
                         r2 * TTt[3] + pg.Cyy};
        const double det = Sig[0] * Sig[3] - Sig[1] * Sig[2];
        double F[4] = {Sig[3] / det, -Sig[1] / det, -Sig[2] / det, Sig[0] / det};
        const double detF = F[0] * F[3] - F[1] * F[2];
        const double norm = G * pg.amp * sqrt(detF) / (2.0 * M_PI);
        gauss64 *g = out + k;
        g->amp = norm * flux * am;
        g->x = xs + pg.xcen;
        g->y = ys + pg.ycen;
        g->fxx = F[0];
        g->fyy = F[3];
        g->fxy = F[2];
        /* dSigma/dq = r2 (T dT^T + dT T^T) */
        double a1[4], a2[4], dSq[4], dSp[4], dFq[4], dFp[4], t1[4];
        mat_mul(T, dTqt, a1);
        mat_mul(dTq, Tt, a2);
        for (int i = 0; i < 4; ++i) dSq[i] = r2 * (a1[i] + a2[i]);
        mat_mul(T, dTpt, a1);
        mat_mul(dTp, Tt, a2);
        for (int i = 0; i < 4; ++i) dSp[i] = r2 * (a1[i] + a2[i]);
        mat_mul(F, dSq, t1);
        mat_mul(t1, F, dFq);
        mat_mul(F, dSp, t1);
        mat_mul(t1, F, dFp);
        for (int i = 0; i < 4; ++i) {
            dFq[i] = -dFq[i];
            dFp[i] = -dFp[i];
        }
        double SdF[4];
        mat_mul(Sig, dFq, SdF);
        const double dA_dq = 0.5 * g->amp * (SdF[0] + SdF[3]);
        mat_mul(Sig, dFp, SdF);
        const double dA_dpa = 0.5 * g->amp * (SdF[0] + SdF[3]);
        memset(g->jac, 0, sizeof(g->jac));
        g->jac[0][0] = norm * am;              /* dA/dflux   */
        g->jac[1][1] = CW[0];                  /* dx/dra     */
        g->jac[1][2] = CW[2];                  /* dy/dra     */
        g->jac[2][1] = CW[1];                  /* dx/ddec    */
        g->jac[2][2] = CW[3];                  /* dy/ddec    */
        g->jac[3][0] = dA_dq;
        g->jac[3][3] = dFq[0];
        g->jac[3][4] = dFq[3];
        g->jac[3][5] = dFq[2];
        g->jac[4][0] = dA_dpa;
        g->jac[4][3] = dFp[0];
        g->jac[4][4] = dFp[3];
        g->jac[4][5] = dFp[2];
        g->jac[5][0] = norm * flux * dadn;     /* dA/dnsersic */
        g->jac[6][0] = norm * flux * dadr;     /* dA/drh      */
    }
}

/* Counts of one Gaussian in one pixel and d(counts)/d(A, x, y, Fxx, Fyy, Fxy).
 * Returns 0 (and leaves outputs untouched) when the pair is cut off. */
static int pixel_gauss(const gauss64 *g, double xp, double yp, int cutoff, double *C,
                       double dC[6]) {
    const double dx = xp - g->x, dy = yp - g->y;
    const double vx = g->fxx * dx + g->fxy * dy;
    const double vy = g->fyy * dy + g->fxy * dx;
    const double ea = dx * vx + dy * vy;
    if (cutoff && ea > 36.0) return 0;
    const double Gp = exp(-0.5 * ea);
    const double H = 1.0 + (vx * vx + vy * vy - g->fxx - g->fyy) / 24.0;
    const double cA = Gp * H;  /* counts per unit amplitude */
    *C = g->amp * cA;
    if (dC) {
        const double ch = g->amp * Gp; /* C / H */
        dC[0] = cA;
        dC[1] = *C * vx - ch * (g->fxx * vx + g->fxy * vy) / 12.0;
        dC[2] = *C * vy - ch * (g->fyy * vy + g->fxy * vx) / 12.0;
        dC[3] = -0.5 * *C * dx * dx - ch * (1.0 - 2.0 * dx * vx) / 24.0;
        dC[4] = -0.5 * *C * dy * dy - ch * (1.0 - 2.0 * dy * vy) / 24.0;
        dC[5] = -1.0 * *C * dx * dy + ch * (dy * vx + dx * vy) / 12.0;
    }
    return 1;
}

/* Pixels per work unit: exposures are cut into chunks of this many pixels,
 * evaluated in parallel, and each band's sums are taken over its units in
 * a fixed order (exposure, then chunk), so the result does not depend on
 * the number of threads. */
#define ORACLE_CHUNK 4096

/* chi2[n_bands], dchi2[n_bands][n_sources][7] (or NULL: no gradient, e.g.
 * to render a model); residual/model [npix] or NULL, for source records in
 * FP64 (SRC_STRIDE doubles each, the fields of the Source struct).  Returns
 * 0 on success. */
int oracle_evaluate64(const oracle_patch *p, const double *proposal, int cutoff, double *chi2,
                      double *dchi2, double *residual, double *model) {
    const int ns = p->n_sources;
    const int ne = p->n_exp;
    int status = 0;
    int *band_of = (int *)malloc(sizeof(int) * (size_t)(ne > 0 ? ne : 1));
    int64_t *goff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ne + 1));
    int64_t *uoff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ne + 1));
    if (!band_of || !goff || !uoff) {
        free(band_of);
        free(goff);
        free(uoff);
        return 1;
    }
    for (int e = 0; e < ne; ++e) band_of[e] = -1;
    for (int b = 0; b < p->n_bands; ++b)
        for (int e = p->band_start[b]; e < p->band_start[b] + p->band_N[b]; ++e) band_of[e] = b;
    goff[0] = 0;
    uoff[0] = 0;
    for (int e = 0; e < ne; ++e) {
        const int b = band_of[e];
        const int64_t ng = b < 0 ? 0 : (int64_t)ns * p->n_psf_per_source[b];
        const int64_t nu = b < 0 ? 0 : (p->exposure_N[e] + ORACLE_CHUNK - 1) / ORACLE_CHUNK;
        goff[e + 1] = goff[e] + ng;
        uoff[e + 1] = uoff[e] + nu;
    }
    const int64_t n_units = uoff[ne];
    const int64_t nd = (int64_t)ns * 7;
    gauss64 *gs = (gauss64 *)malloc(sizeof(gauss64) * (size_t)(goff[ne] > 0 ? goff[ne] : 1));
    int *unit_exp = (int *)malloc(sizeof(int) * (size_t)(n_units > 0 ? n_units : 1));
    double *uc2 = (double *)calloc((size_t)(n_units > 0 ? n_units : 1), sizeof(double));
    double *ud = dchi2 ? (double *)calloc((size_t)(n_units > 0 ? n_units : 1) * (size_t)nd,
                                          sizeof(double))
                       : NULL;
    if (!gs || !unit_exp || !uc2 || (dchi2 && !ud)) {
        free(band_of); free(goff); free(uoff); free(gs); free(unit_exp); free(uc2); free(ud);
        return 1;
    }
    for (int e = 0; e < ne; ++e)
        for (int64_t u = uoff[e]; u < uoff[e + 1]; ++u) unit_exp[u] = e;

    /* image Gaussians of every exposure */
#pragma omp parallel for schedule(dynamic, 1)
    for (int e = 0; e < ne; ++e) {
        const int b = band_of[e];
        if (b < 0) continue;
        const int np = p->n_psf_per_source[b];
        for (int s = 0; s < ns; ++s)
            convert_source(p, proposal + (int64_t)SRC_STRIDE * s, e, b, s, gs + goff[e] + s * np);
    }

    /* pixel chunks */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t u = 0; u < n_units; ++u) {
        const int e = unit_exp[u];
        const int b = band_of[e];
        const int np = p->n_psf_per_source[b];
        const int ng = ns * np;
        const gauss64 *g = gs + goff[e];
        const int64_t i0 = p->exposure_start[e] + (u - uoff[e]) * ORACLE_CHUNK;
        int64_t i1 = i0 + ORACLE_CHUNK;
        if (i1 > (int64_t)p->exposure_start[e] + p->exposure_N[e])
            i1 = (int64_t)p->exposure_start[e] + p->exposure_N[e];
        double c2 = 0.0;
        double *du = ud ? ud + u * nd : NULL;
        for (int64_t i = i0; i < i1; ++i) {
            const double xp = p->xpix[i], yp = p->ypix[i];
            const double d = p->data[i], ie = p->ierr[i];
            double m = 0.0, C;
            for (int k = 0; k < ng; ++k)
                if (pixel_gauss(g + k, xp, yp, cutoff, &C, NULL)) m += C;
            const double r = d - m;
            if (residual) residual[i] = r;
            if (model) model[i] = m;
            c2 += (r * ie - d * ie) * (r * ie + d * ie);
            const double w = -2.0 * r * ie * ie;
            if (!du || w == 0.0) continue;
            for (int s = 0; s < ns; ++s) {
                double *ds = du + 7 * s;
                for (int k = s * np; k < (s + 1) * np; ++k) {
                    double dC[6];
                    if (!pixel_gauss(g + k, xp, yp, cutoff, &C, dC)) continue;
                    for (int j = 0; j < 7; ++j) {
                        double acc = 0.0;
                        for (int l = 0; l < 6; ++l) acc += g[k].jac[j][l] * dC[l];
                        ds[j] += w * acc;
                    }
                }
            }
        }
        uc2[u] = c2;
    }

    /* per band, in unit order */
    for (int b = 0; b < p->n_bands; ++b) {
        double c2 = 0.0;
        double *db = dchi2 ? dchi2 + (int64_t)b * nd : NULL;
        if (db) memset(db, 0, sizeof(double) * (size_t)nd);
        for (int e = p->band_start[b]; e < p->band_start[b] + p->band_N[b]; ++e)
            for (int64_t u = uoff[e]; u < uoff[e + 1]; ++u) {
                c2 += uc2[u];
                if (db)
                    for (int64_t j = 0; j < nd; ++j) db[j] += ud[u * nd + j];
            }
        chi2[b] = c2;
    }
    free(band_of); free(goff); free(uoff); free(gs); free(unit_exp); free(uc2); free(ud);
    return status;
}

/* The same for float32 Source records (the struct the kernels read): each
 * field is widened exactly to double. */
int oracle_evaluate(const oracle_patch *p, const float *proposal, int cutoff, double *chi2,
                    double *dchi2, double *residual, double *model) {
    const int64_t n = (int64_t)SRC_STRIDE * p->n_sources;
    double *src = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!src) return 1;
    for (int64_t i = 0; i < n; ++i) src[i] = proposal[i];
    const int rc = oracle_evaluate64(p, src, cutoff, chi2, dchi2, residual, model);
    free(src);
    return rc;
}

/* The image Gaussians (amp, x, y, fxx, fyy, fxy then the 7x6 Jacobian,
 * 48 doubles each) of exposure e for all sources; out has n_sources *
 * n_psf_per_source[band] * 48 doubles. */
int oracle_image_gaussians(const oracle_patch *p, const float *proposal, int e, int b,
                           double *out) {
    const int ns = p->n_sources;
    const int np = p->n_psf_per_source[b];
    gauss64 *gs = (gauss64 *)malloc(sizeof(gauss64) * (size_t)ns * np);
    if (!gs) return 1;
    double src[SRC_STRIDE];
    for (int s = 0; s < ns; ++s) {
        for (int i = 0; i < SRC_STRIDE; ++i) src[i] = proposal[(int64_t)SRC_STRIDE * s + i];
        convert_source(p, src, e, b, s, gs + s * np);
    }
    for (int k = 0; k < ns * np; ++k) {
        double *o = out + (int64_t)48 * k;
        o[0] = gs[k].amp;
        o[1] = gs[k].x;
        o[2] = gs[k].y;
        o[3] = gs[k].fxx;
        o[4] = gs[k].fyy;
        o[5] = gs[k].fxy;
        memcpy(o + 6, gs[k].jac, sizeof(gs[k].jac));
    }
    free(gs);
    return 0;
}
