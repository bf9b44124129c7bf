/*
 * swd_oracle_decon.c -- CPU restatement of the reference's wavenumber-domain
 * PSF deconvolution (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp, DC),
 * single MPI rank.  TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of
 * the HIP path, never by the product.
 *
 * What follows the reference exactly: window geometry (DC:84-129), edge
 * replication (DC:300-340), the raised-cosine output taper with PI = 3.1415926
 * (DC:355-382), window placement and bounds checks (DC:543-615), the float
 * sums of magnitudes with their 1e-15 / 1e-10 guards, the PSF lattice clamp,
 * fft_max as a float max times epsilon, the float tmpfloat, and the overlap-add
 * in window order with its float / double mixing (DC:653-673).
 *
 * What is restated: the reference transforms each 2D window with FFTW in
 * double (DC:479-493); FFTW is not in this image.  Here the unnormalised DFTs
 * are evaluated directly, in the fixed order below, which the HIP kernels
 * follow operation for operation:
 *  - twiddles tw[t] = (cos(2 pi t / n), sin(2 pi t / n)) from libm;
 *  - forward, real input x[r][c] supported on rows offx.. and columns offz..:
 *    along z first, half spectrum k1 = 0..n/2 (y = sum_c x e^{-i..}, fma
 *    chains in c order), then along x over pairs of wavenumbers (p, n - p),
 *    whose twiddles are conjugate: P = sum_r cos(.) y, Q = sum_r sin(.) y
 *    (fma chains in r order), M[p] = P - iQ, M[n-p] = P + iQ; the k1 > n/2
 *    half is the conjugate of the computed half, so the fft_max maximum runs
 *    over the computed half;
 *  - the image normalisation by pp_max is folded into the Wiener divisor:
 *    d = (M S*) / (pp_max * tmpfloat), pp_max * tmpfloat being exact in double
 *    (a product of two floats), instead of dividing the window before the
 *    transform;
 *  - inverse only at the output points, real part, along x first over all k0,
 *    then along z over the half spectrum with weight 2 on k1 = 1..(n-1)/2.
 * The FFT algorithm's own rounding is not reproduced; agreement with it is
 * pinned by the DFT definition (numpy FFT, to ~1e-13).
 *
 * The FFTW 2D plan views the flat window buffer as [fullwz][fullwx]
 * (DC:490-493) while the data are laid out [fullwx][fullwz]; only square
 * windows (fullwx == fullwz) are meaningful, and only those are accepted.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "swd_oracle.h"

/* Derived sizes (DC:84-129); returns -1 when the reference would stop or the
 * window is not square. */
int or_decon_geometry(const or_decon_params *p, or_decon_geom *g)
{
    if (p->nx < 1 || p->nz < 1 || p->psfwinx < 1 || p->psfwinz < 1 || p->snapwinx < 2 || p->snapwinz < 2) return -1;
    if (p->psfwinx % 2 == 0 || p->psfwinz % 2 == 0 || p->filterwinx % 2 == 0 || p->filterwinz % 2 == 0) return -1;
    g->halfpsfwx = p->psfwinx / 2;
    g->halfpsfwz = p->psfwinz / 2;
    g->halfltwx = p->filterwinx / 2;
    g->halfltwz = p->filterwinz / 2;
    g->nzwpsf = (int)ceil(1.0 * p->nz / p->psfwinz);
    g->nxwpsf = (int)ceil(1.0 * p->nx / p->psfwinx);
    g->fullwx = p->filterwinx + p->psfwinx - 1;
    g->fullwz = p->filterwinz + p->psfwinz - 1;
    g->midpx = g->fullwx / 2;
    g->midpz = g->fullwz / 2;
    g->outwinx = 2 * (p->psfwinx / 10) + 2 * p->snapwinx + 1;
    g->outwinz = 2 * (p->psfwinz / 10) + 2 * p->snapwinz + 1;
    g->halfoutwx = g->outwinx / 2;
    g->halfoutwz = g->outwinz / 2;
    g->stepwinx = g->outwinx - p->snapwinx;
    g->stepwinz = g->outwinz - p->snapwinz;
    g->stepnumx = (int)ceil(1.0 * (p->nx + g->fullwx) / g->stepwinx) + 1;
    g->stepnumz = (int)ceil(1.0 * (p->nz + g->fullwz) / g->stepwinz) + 1;
    g->totalnum = g->stepnumx * g->stepnumz;
    g->nxp = g->stepnumx * g->stepwinx;
    g->nzp = g->stepnumz * g->stepwinz;
    if (p->filterwinx < p->psfwinx || p->filterwinz < p->psfwinz) return -1;
    if (p->filterwinx < g->outwinx || p->filterwinz < g->outwinz) return -1;
    if (g->fullwx != g->fullwz) return -1;
    if (g->nxwpsf < 7 || g->nzwpsf < 7) return -1; /* the PSF lattice clamp needs cells 3 .. nw-4 */
    return 0;
}

/* Raised-cosine output window coefr[outwinx][outwinz] (DC:355-382). */
void or_decon_coefr(const or_decon_params *p, const or_decon_geom *g, float *coefr)
{
    const double pi = 3.1415926;
    for (int ix = 0; ix < g->outwinx; ix++)
        for (int iz = 0; iz < g->outwinz; iz++) {
            float cx, cz;
            if (ix <= p->snapwinx - 1) cx = (float)(0.5000 - 0.5 * cos(pi * ix / (p->snapwinx - 1)));
            else if (ix >= g->outwinx - p->snapwinx) cx = (float)(0.5000 + 0.5 * cos(pi * (ix - g->outwinx + p->snapwinx) / (p->snapwinx - 1)));
            else cx = 1.0f;
            if (iz <= p->snapwinz - 1) cz = (float)(0.5000 - 0.5 * cos(pi * iz / (p->snapwinz - 1)));
            else if (iz >= g->outwinz - p->snapwinz) cz = (float)(0.5000 + 0.5 * cos(pi * (iz - g->outwinz + p->snapwinz) / (p->snapwinz - 1)));
            else cz = 1.0f;
            coefr[ix * g->outwinz + iz] = cx * cz;
        }
}

/* [nx][nz] -> [nxp][nzp]: the last column, then the last row, replicated (DC:303-339). */
void or_decon_expand(const float *a, int nx, int nz, float *af, int nxp, int nzp)
{
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) af[ix * nzp + iz] = a[ix * nz + iz];
    for (int ix = nx; ix < nxp; ix++)
        for (int iz = 0; iz < nz; iz++) af[ix * nzp + iz] = af[(nx - 1) * nzp + iz];
    for (int iz = nz; iz < nzp; iz++)
        for (int ix = 0; ix < nxp; ix++) af[ix * nzp + iz] = af[ix * nzp + nz - 1];
}

void or_twiddles(int n, double *c, double *s)
{
    for (int t = 0; t < n; t++) {
        c[t] = cos(2.0 * M_PI * t / n);
        s[t] = sin(2.0 * M_PI * t / n);
    }
}

/* Half spectrum re/im[n][n/2+1] of the n x n real buffer that holds x[mx][mz]
 * at rows offx.., columns offz.. and zeros elsewhere (forward, unnormalised). */
void or_decon_spectrum(const double *x, int mx, int mz, int offx, int offz, int n, double *re, double *im)
{
    const int nh = n / 2 + 1;
    double *c = malloc(sizeof(double) * n), *s = malloc(sizeof(double) * n);
    double *yr = malloc(sizeof(double) * mx * nh), *yi = malloc(sizeof(double) * mx * nh);
    or_twiddles(n, c, s);
    for (int r = 0; r < mx; r++)
        for (int k1 = 0; k1 < nh; k1++) {
            double ar = 0.0, ai = 0.0;
            for (int j = 0; j < mz; j++) {
                const int t = (int)(((long)k1 * (j + offz)) % n);
                const double v = x[r * mz + j];
                ar = fma(v, c[t], ar);
                ai = fma(-v, s[t], ai);
            }
            yr[r * nh + k1] = ar;
            yi[r * nh + k1] = ai;
        }
    /* pairs (p, n - p): P = sum c y, Q = sum s y, M[p] = P - iQ, M[n-p] = P + iQ */
    for (int p = 0; p < nh; p++)
        for (int k1 = 0; k1 < nh; k1++) {
            double pr = 0.0, pi = 0.0, qr = 0.0, qi = 0.0;
            for (int r = 0; r < mx; r++) {
                const int t = (int)(((long)p * (r + offx)) % n);
                const double a = yr[r * nh + k1], b = yi[r * nh + k1];
                pr = fma(a, c[t], pr);
                pi = fma(b, c[t], pi);
                qr = fma(a, s[t], qr);
                qi = fma(b, s[t], qi);
            }
            re[p * nh + k1] = pr + qi;
            im[p * nh + k1] = pi - qr;
            if (p > 0 && 2 * p != n) {
                re[(n - p) * nh + k1] = pr - qi;
                im[(n - p) * nh + k1] = pi + qr;
            }
        }
    free(c);
    free(s);
    free(yr);
    free(yi);
}

/* Real part of the unnormalised inverse DFT of the Hermitian spectrum whose
 * half dr/di[n][n/2+1] is given, at rows r0..r0+nr-1, columns c0..c0+nc-1. */
void or_decon_inverse(const double *dr, const double *di, int n, int r0, int nr, int c0, int nc, double *out)
{
    const int nh = n / 2 + 1;
    double *c = malloc(sizeof(double) * n), *s = malloc(sizeof(double) * n);
    double *qr = malloc(sizeof(double) * nr * nh), *qi = malloc(sizeof(double) * nr * nh);
    or_twiddles(n, c, s);
    for (int i = 0; i < nr; i++)
        for (int k1 = 0; k1 < nh; k1++) {
            double ar = 0.0, ai = 0.0;
            for (int k0 = 0; k0 < n; k0++) {
                const int t = (int)(((long)k0 * (r0 + i)) % n);
                const double a = dr[k0 * nh + k1], b = di[k0 * nh + k1];
                ar = fma(a, c[t], ar);
                ar = fma(-b, s[t], ar);
                ai = fma(b, c[t], ai);
                ai = fma(a, s[t], ai);
            }
            qr[i * nh + k1] = ar;
            qi[i * nh + k1] = ai;
        }
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < nc; j++) {
            double o = 0.0;
            for (int k1 = 0; k1 < nh; k1++) {
                const double w = (k1 == 0 || 2 * k1 == n) ? 1.0 : 2.0;
                const int t = (int)(((long)k1 * (c0 + j)) % n);
                o = fma(w * qr[i * nh + k1], c[t], o);
                o = fma(-(w * qi[i * nh + k1]), s[t], o);
            }
            out[i * nc + j] = o;
        }
    free(c);
    free(s);
    free(qr);
    free(qi);
}

/* PSF lattice cell of a window origin (DC:583-596). */
void or_decon_cell(const or_decon_params *p, const or_decon_geom *g, int ix, int iz, int *cx, int *cz)
{
    int ixf = (ix - g->halfpsfwx) / p->psfwinx, izf = (iz - g->halfpsfwz) / p->psfwinz;
    if (ixf <= 2) ixf = 3;
    if (ixf >= g->nxwpsf - 3) ixf = g->nxwpsf - 4;
    if (izf <= 2) izf = 3;
    if (izf >= g->nzwpsf - 3) izf = g->nzwpsf - 4;
    *cx = ixf;
    *cz = izf;
}

/* Normalised PSF half spectrum of lattice cell (cx, cz) and the Wiener
 * denominators tden[n][n/2+1] (DC:597-648): the PSF patch is divided by its
 * float sum of magnitudes before the transform, as in the reference. */
void or_decon_psf(const or_decon_params *p, const or_decon_geom *g, const float *psff, int cx, int cz, double *sr,
                  double *si, float *tden)
{
    const int n = g->fullwx, nh = n / 2 + 1, mx = p->psfwinx, mz = p->psfwinz;
    const int ixp = g->halfpsfwx + cx * p->psfwinx, izp = g->halfpsfwz + cz * p->psfwinz;
    float psf_max = 0.0f;
    for (int i = 0; i < mx; i++)
        for (int j = 0; j < mz; j++)
            psf_max = (float)((double)psf_max + fabs((double)psff[(size_t)(ixp + i - g->halfpsfwx) * g->nzp + (izp + j - g->halfpsfwz)]));
    double *x = malloc(sizeof(double) * mx * mz);
    for (int i = 0; i < mx; i++)
        for (int j = 0; j < mz; j++)
            x[i * mz + j] = (double)psff[(size_t)(ixp + i - g->halfpsfwx) * g->nzp + (izp + j - g->halfpsfwz)] / (double)psf_max;
    or_decon_spectrum(x, mx, mz, 0, 0, n, sr, si);
    float fft_max = 0.0f;
    for (int k = 0; k < n * nh; k++) {
        const float q = (float)(sr[k] * sr[k] + si[k] * si[k]);
        if (fft_max < q) fft_max = q;
    }
    fft_max *= p->epsilon;
    if (fft_max == 0.0f) fft_max = 1.0e-15f;
    for (int k = 0; k < n * nh; k++) tden[k] = (float)((sr[k] * sr[k] + si[k] * si[k]) + (double)fft_max);
    free(x);
}

/* The deconvolved real part of window `step` at its outwinx x outwinz output
 * points (DC:543-674). */
void or_decon_window(const or_decon_params *p, const or_decon_geom *g, const float *imgf, const float *psff, int step,
                     double *out)
{
    const int n = g->fullwx, nh = n / 2 + 1, nb = n * nh, mx = p->filterwinx, mz = p->filterwinz;
    const int ix = g->stepwinx * (step / g->stepnumz), iz = g->stepwinz * (step % g->stepnumz);
    double *x = calloc((size_t)mx * mz, sizeof(double));
    double *mr = malloc(sizeof(double) * nb), *mi = malloc(sizeof(double) * nb);
    double *sr = malloc(sizeof(double) * nb), *si = malloc(sizeof(double) * nb);
    float *tden = malloc(sizeof(float) * nb);
    /* image window and its float sum of magnitudes (DC:553-576) */
    float pp_max = 0.0f;
    for (int i = 0; i < mx; i++)
        for (int j = 0; j < mz; j++) {
            const int iix = ix - g->halfltwx + i, iiz = iz - g->halfltwz + j;
            if (iix >= 0 && iix < g->nxp && iiz >= 0 && iiz < g->nzp) {
                x[i * mz + j] = imgf[(size_t)iix * g->nzp + iiz];
                pp_max = (float)((double)pp_max + fabs(x[i * mz + j]));
            }
        }
    if (pp_max <= 1.0e-15) pp_max = 1.0e-10f;
    or_decon_spectrum(x, mx, mz, g->midpx - g->halfltwx, g->midpz - g->halfltwz, n, mr, mi);
    int cx, cz;
    or_decon_cell(p, g, ix, iz, &cx, &cz);
    or_decon_psf(p, g, psff, cx, cz, sr, si, tden);
    for (int k = 0; k < nb; k++) {
        const double den = (double)pp_max * (double)tden[k];
        const double dr = (mr[k] * sr[k] + mi[k] * si[k]) / den;
        const double di = (-mr[k] * si[k] + mi[k] * sr[k]) / den;
        mr[k] = dr;
        mi[k] = di;
    }
    or_decon_inverse(mr, mi, n, g->halfltwx - g->halfoutwx, g->outwinx, g->halfltwz - g->halfoutwz, g->outwinz, out);
    free(x);
    free(mr);
    free(mi);
    free(sr);
    free(si);
    free(tden);
}

/* Whole deconvolution, windows accumulated in step order (one rank):
 * dec, coef are [nx][nz] (DC:543-731). */
int or_decon(const or_decon_params *p, const float *image, const float *psf, float *dec, float *coef)
{
    or_decon_geom g;
    if (or_decon_geometry(p, &g) != 0) return -1;
    const size_t npf = (size_t)g.nxp * g.nzp;
    float *imgf = malloc(sizeof(float) * npf), *psff = malloc(sizeof(float) * npf);
    float *acc = calloc(npf, sizeof(float)), *accc = calloc(npf, sizeof(float));
    float *coefr = malloc(sizeof(float) * g.outwinx * g.outwinz);
    or_decon_expand(image, p->nx, p->nz, imgf, g.nxp, g.nzp);
    or_decon_expand(psf, p->nx, p->nz, psff, g.nxp, g.nzp);
    or_decon_coefr(p, &g, coefr);
    const int nout = g.outwinx * g.outwinz;
    double *win = malloc(sizeof(double) * (size_t)g.totalnum * nout);
#pragma omp parallel for schedule(dynamic)
    for (int step = 0; step < g.totalnum; step++) or_decon_window(p, &g, imgf, psff, step, win + (size_t)step * nout);
    for (int step = 0; step < g.totalnum; step++) {
        const int ix = g.stepwinx * (step / g.stepnumz), iz = g.stepwinz * (step % g.stepnumz);
        for (int i = 0; i < g.outwinx; i++)
            for (int j = 0; j < g.outwinz; j++) {
                const int iix = ix - g.halfoutwx + i, iiz = iz - g.halfoutwz + j;
                if (iix >= 0 && iix < g.nxp && iiz >= 0 && iiz < g.nzp) {
                    const size_t ipp = (size_t)iix * g.nzp + iiz;
                    acc[ipp] = (float)((double)acc[ipp] + win[(size_t)step * nout + i * g.outwinz + j] * (double)coefr[i * g.outwinz + j]);
                    accc[ipp] = accc[ipp] + coefr[i * g.outwinz + j];
                }
            }
    }
    for (int ix = 0; ix < p->nx; ix++)
        for (int iz = 0; iz < p->nz; iz++) {
            dec[ix * p->nz + iz] = acc[(size_t)ix * g.nzp + iz];
            if (coef) coef[ix * p->nz + iz] = accc[(size_t)ix * g.nzp + iz];
        }
    free(imgf);
    free(psff);
    free(acc);
    free(accc);
    free(coefr);
    free(win);
    return 0;
}
