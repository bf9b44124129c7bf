/*
 * swd_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see swd_oracle.h).
 *
 * Setup routines restate the reference's host code with its exact mixed
 * float/double arithmetic.  The reference host files are C++ compiled with
 * <math.h>, so exp()/log() of a float argument resolve to the float overloads
 * (expf/logf) while pow(float,int) promotes to double; this file spells those
 * choices out.  Build with -ffp-contract=off: the reference host build has no
 * FMA contraction (x86-64 baseline ISA), and the propagation loops use fma()
 * only where written.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "swd_oracle.h"

#define REF_PI 3.1415926 /* RC:7, forward/headrtm.h:5 */

void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_frame_size(const or_grid *g)
{
    return 2 * g->Lc * (g->nx + g->nz - 2 * g->Lc);
}

/* RC:1858-1878: staggered-grid coefficients, float accumulators with double
 * intermediate products. */
void or_cal_xishu(int Lx, float *rx)
{
    int m, i;
    float s1, s2;
    for (m = 1; m <= Lx; m++) {
        s1 = 1.0f;
        s2 = 1.0f;
        for (i = 1; i < m; i++) {
            s1 = (float)(s1 * (2.0 * i - 1) * (2.0 * i - 1));
            s2 = (float)(s2 * ((2.0 * m - 1) * (2.0 * m - 1) - (2.0 * i - 1) * (2.0 * i - 1)));
        }
        for (i = m + 1; i <= Lx; i++) {
            s1 = (float)(s1 * (2.0 * i - 1) * (2.0 * i - 1));
            s2 = (float)(s2 * ((2.0 * m - 1) * (2.0 * m - 1) - (2.0 * i - 1) * (2.0 * i - 1)));
        }
        s2 = fabsf(s2);
        rx[m - 1] = (float)(pow(-1.0, m + 1) * s1 / (s2 * (2.0 * m - 1)));
    }
}

/* RC:1284-1299 (ricker_wave flag 2): float phase, expf, double product. */
void or_ricker(float *rick, int nt, float f0, float t0, float dt)
{
    const float pi = 3.1415927f;
    int it;
    for (it = 0; it < nt; it++) {
        float temp = pi * f0 * ((float)it * dt - t0);
        temp = temp * temp;
        rick[it] = (float)((1.0 - 2.0 * (double)temp) * (double)expf(-temp));
    }
}

/* RC:891-1129 get_absorbing_parameters; the driver passes pml+Lc as "pml"
 * (RC:467).  N=2, Rc=1e-5, Vpmax=5500 (hard-coded, RC:923), alpha_max=15*pi. */
void or_cpml(int ntx, int ntz, int pml, float dx, float dz, float dt,
             float *a_x, float *a_x_half, float *b_x, float *b_x_half,
             float *a_z, float *a_z_half, float *b_z, float *b_z_half)
{
    const int N = 2;
    const float Rc = 1.0e-5f, pi = 3.1415927f;
    const float alpha_max = pi * 15;
    const float Vpmax = 5500;
    float thickness = (float)pml * dz;
    float d0 = (float)((float)(-(N + 1)) * Vpmax * logf(Rc) / (2.0 * thickness));
    float *tmp = (float *)calloc((size_t)(ntx > ntz ? ntx : ntz) * 4, sizeof(float));
    int pass;
    for (pass = 0; pass < 2; pass++) {
        int n = pass == 0 ? ntx : ntz;
        float h = pass == 0 ? dx : dz;
        float *a = pass == 0 ? a_x : a_z, *ah = pass == 0 ? a_x_half : a_z_half;
        float *b = pass == 0 ? b_x : b_z, *bh = pass == 0 ? b_x_half : b_z_half;
        float *d = tmp, *dh = tmp + n, *al = tmp + 2 * n, *alh = tmp + 3 * n;
        float start = (float)pml * h, end = (float)(n - pml - 1) * h;
        int i;
        for (i = 0; i < 4 * n; i++) tmp[i] = 0.0f;
        for (i = 0; i < n; i++) { a[i] = 0.0f; ah[i] = 0.0f; b[i] = 0.0f; bh[i] = 0.0f; }
        for (i = 0; i < n; i++) {
            float cur = (float)i * h, delta;
            if (cur <= start) {
                delta = start - cur;
                d[i] = (float)(d0 * pow((double)(delta / thickness), 2));
                al[i] = (float)(alpha_max * (1.0 - (double)(delta / thickness)) + 0.1 * alpha_max);
            }
            if (cur >= end) {
                delta = cur - end;
                d[i] = (float)(d0 * pow((double)(delta / thickness), 2));
                al[i] = (float)(alpha_max * (1.0 - (double)(delta / thickness)) + 0.1 * alpha_max);
            }
        }
        for (i = 0; i < n; i++) {
            float cur = (float)((i + 0.5) * h), delta;
            if (cur <= start) {
                delta = start - cur;
                dh[i] = (float)(d0 * pow((double)(delta / thickness), 2));
                alh[i] = (float)(alpha_max * (1.0 - (double)(delta / thickness)) + 0.1 * alpha_max);
            }
            if (cur >= end) {
                delta = cur - end;
                dh[i] = (float)(d0 * pow((double)(delta / thickness), 2));
                alh[i] = (float)(alpha_max * (1.0 - (double)(delta / thickness)) + 0.1 * alpha_max);
            }
        }
        for (i = 0; i < n; i++) {
            const float k = 1.0f;
            if (al[i] < 0.0f) al[i] = 0.0f;
            if (alh[i] < 0.0f) alh[i] = 0.0f;
            b[i] = expf(-(d[i] / k + al[i]) * dt);
            if (d[i] > 1.0e-6) a[i] = (float)(d[i] / (k * (d[i] + k * al[i])) * (b[i] - 1.0));
            bh[i] = expf(-(dh[i] / k + alh[i]) * dt);
            if (dh[i] > 1.0e-6) ah[i] = (float)(dh[i] / (k * (dh[i] + k * alh[i])) * (bh[i] - 1.0));
        }
    }
    free(tmp);
}

/* RC:1428-1468 model_in_pml: replicate the interior edge into the CPML. */
void or_model_in_pml(float *vp, int ntz, int ntx, int pml)
{
    int iz, ix;
    for (iz = pml; iz <= ntz - pml - 1; iz++) {
        for (ix = 0; ix <= pml - 1; ix++) vp[ix * ntz + iz] = vp[pml * ntz + iz];
        for (ix = ntx - pml; ix < ntx; ix++) vp[ix * ntz + iz] = vp[(ntx - pml - 1) * ntz + iz];
    }
    for (ix = 0; ix < ntx; ix++) {
        for (iz = 0; iz <= pml - 1; iz++) vp[ix * ntz + iz] = vp[ix * ntz + pml];
        for (iz = ntz - pml; iz < ntz; iz++) vp[ix * ntz + iz] = vp[ix * ntz + (ntz - pml - 1)];
    }
}

/* RC:1473-1569 ini_model_mine: normalised Gaussian (sigma = window/2) over a
 * (2w+1)^2 window with clamped indices, double accumulation in the order
 * izw outer / ixw inner; the first 8 interior rows keep the input. */
void or_smooth(float *vp, int ntz, int ntx, int pml, int window)
{
    int windxz = 2 * window + 1;
    double sigma = window / 2.0, sum;
    double *coef = (double *)malloc(sizeof(double) * windxz * windxz);
    double *old = (double *)malloc(sizeof(double) * (size_t)ntz * ntx);
    int ix, iz, ip;
    sum = 0.0;
    for (ix = 0; ix < windxz; ix++)
        for (iz = 0; iz < windxz; iz++) {
            ip = ix * windxz + iz;
            coef[ip] = 1.0 / (2.0 * REF_PI * sigma * sigma) *
                       exp(-((ix - window) * (ix - window) + (iz - window) * (iz - window)) / (2.0 * sigma * sigma));
            sum += coef[ip];
        }
    for (ip = 0; ip < windxz * windxz; ip++) coef[ip] /= sum;
    for (ip = 0; ip < ntz * ntx; ip++) old[ip] = vp[ip];
#pragma omp parallel for private(ix)
    for (iz = pml; iz <= ntz - pml - 1; iz++) {
        for (ix = pml; ix <= ntx - pml - 1; ix++) {
            double s = 0.0;
            int izw, ixw;
            for (izw = iz - window; izw <= iz + window; izw++)
                for (ixw = ix - window; ixw <= ix + window; ixw++) {
                    int ipp = (izw - iz + window) * windxz + (ixw - ix + window);
                    int iz1 = izw < 0 ? 0 : (izw > ntz - 1 ? ntz - 1 : izw);
                    int ix1 = ixw < 0 ? 0 : (ixw > ntx - 1 ? ntx - 1 : ixw);
                    s += old[ix1 * ntz + iz1] * coef[ipp];
                }
            vp[ix * ntz + iz] = (float)s;
            if (iz < pml + 8) vp[ix * ntz + iz] = (float)old[ix * ntz + iz];
        }
    }
    or_model_in_pml(vp, ntz, ntx, pml);
    free(coef);
    free(old);
}

/* FK:595-596 / FC:299-300 / FK:1263-1264: 2^ceil(log2 n) as the reference computes it. */
int or_nfft(int n)
{
    int K = (int)ceil(log(1.0 * n) / log(2.0));
    return (int)pow(2.0, K);
}

/* FC:354-377: raised-cosine ramps of width layerz on both ends of kz. */
int or_kz_taper(float *coef, int nfftz, float dz)
{
    float dkz = (float)(1.0 / (float)(nfftz * dz));
    float kcut = (float)(1.0 / 15.0 / dz);
    int layerz = (int)(kcut / dkz), upz, downz, iz;
    if (layerz > nfftz / 6 || layerz < 1) return -1; /* FC:360-362 */
    upz = layerz;
    downz = nfftz - layerz;
    for (iz = 0; iz < nfftz; iz++) {
        if (iz <= upz) coef[iz] = (float)(0.5000 - 0.5 * cos(REF_PI * iz / layerz));
        if (iz > upz && iz < downz) coef[iz] = 1.0f;
        if (iz >= downz) coef[iz] = (float)(0.5000 + 0.5 * cos(REF_PI * (iz - downz) / layerz));
    }
    return layerz;
}

/* FK:1257-1324: wavelet Hilbert transform through an NZ-point DFT
 * (NZ = 2^ceil(log2 nt)); bins 1..NZ/2-1 times -i, NZ/2..NZ-1 times +i, DC 0,
 * unnormalised inverse divided by NZ.  The reference runs this in FP32 cuFFT;
 * here it is a direct O(NZ^2) DFT in double. */
void or_hilbert1d(const float *rick, float *hrick, int nt)
{
    int NZ = or_nfft(nt), k, j;
    double *cs = (double *)malloc(sizeof(double) * NZ * 2), *sn = cs + NZ;
    double *Xr = (double *)calloc((size_t)NZ * 2, sizeof(double)), *Xi = Xr + NZ;
    for (k = 0; k < NZ; k++) {
        cs[k] = cos(2.0 * 3.14159265358979323846 * k / NZ);
        sn[k] = sin(2.0 * 3.14159265358979323846 * k / NZ);
    }
    for (k = 0; k < NZ; k++) {
        double sr = 0, si = 0;
        for (j = 0; j < nt; j++) {
            int m = (int)(((long long)j * k) % NZ);
            sr += rick[j] * cs[m];
            si -= rick[j] * sn[m];
        }
        /* multiply by -i below NZ/2, +i from NZ/2 */
        if (k == 0) { Xr[k] = 0; Xi[k] = 0; }
        else if (k < NZ / 2) { Xr[k] = si; Xi[k] = -sr; }
        else { Xr[k] = -si; Xi[k] = sr; }
    }
    for (j = 0; j < nt; j++) {
        double sr = 0;
        for (k = 0; k < NZ; k++) {
            int m = (int)(((long long)j * k) % NZ);
            sr += Xr[k] * cs[m] - Xi[k] * sn[m];
        }
        hrick[j] = (float)(sr / NZ);
    }
    free(cs);
    free(Xr);
}

/* RC:150-200: source node and the two-borehole receiver rows of shot `is`.
 * Returns r_n (2 x rows within offsetz of the source depth), fills r_iz with
 * the left-column rows followed by the same rows for the right column. */
int or_geometry(int nx, int nz, int pml, int Lc, float dx, float dz,
                float sx0, float shotdx, float shotdep, float offsetz, int is,
                int *s_ix, int *s_iz, int *r_iz, int max_rn)
{
    int pmlc = pml + Lc, iz, i = 0, rn;
    (void)nx;
    *s_ix = pmlc + (int)(sx0 / dx) + (int)(shotdx / dx) * is;
    *s_iz = pmlc + (int)(shotdep / dz);
    for (iz = 0; iz < nz; iz++)
        if (fabs((double)(*s_iz - iz - pmlc)) * dz <= offsetz) i++;
    rn = 2 * i;
    if (r_iz) {
        if (rn > max_rn) return -1;
        i = 0;
        for (iz = 0; iz < nz; iz++)
            if (fabs((double)(*s_iz - iz - pmlc)) * dz <= offsetz) {
                r_iz[i] = pmlc + iz;
                r_iz[i + rn / 2] = pmlc + iz;
                i++;
            }
    }
    return rn;
}

/* RC:603-637 with t0n = 0: two running float integrations, negated. */
void or_adjoint_source(const float *d, float *rms, int ntr, int nt, float dt)
{
    float *f = (float *)malloc(sizeof(float) * (size_t)ntr * nt);
    int tr, it;
    memcpy(f, d, sizeof(float) * (size_t)ntr * nt);
    for (tr = 0; tr < ntr; tr++) {
        float t = 0.0f;
        for (it = 0; it < nt; it++) {
            size_t ip = (size_t)tr * nt + it;
            float inc = f[ip] * dt;
            t = t + inc;
            f[ip] = t;
        }
    }
    for (tr = 0; tr < ntr; tr++) {
        float t = 0.0f;
        for (it = 0; it < nt; it++) {
            size_t ip = (size_t)tr * nt + it;
            float inc = f[ip] * dt;
            t = t + inc;
            rms[ip] = -t;
        }
    }
    free(f);
}

/* FC:697-705: density point scatterers rho += 1e-3*vp on the swin lattice. */
void or_scatterers(float *rho, const float *vp, int ntx, int ntz, int pmlc, int swin)
{
    int ix, iz;
    for (ix = pmlc + swin / 2; ix < ntx - pmlc - swin / 2; ix++)
        for (iz = pmlc + swin / 2; iz < ntz - pmlc - 5; iz++)
            if ((ix - (pmlc + swin / 2)) % swin == 0 && (iz - (pmlc + swin / 2)) % swin == 0)
                rho[ix * ntz + iz] = (float)(rho[ix * ntz + iz] + vp[ix * ntz + iz] * 1.0e-3);
}

/* RC:663-679: image /= illum + 1e-3 * max|illum| (double quotient). */
void or_image_normalize(float *image, const float *illum, int ntp)
{
    float remax = 0;
    int ip;
    for (ip = 0; ip < ntp; ip++)
        if (remax < fabsf(illum[ip])) remax = fabsf(illum[ip]);
    for (ip = 0; ip < ntp; ip++)
        image[ip] = (float)((double)image[ip] / ((double)illum[ip] + 1.0e-3 * (double)remax));
}

/* RC:722-735: out = -(d2/dx2 + d2/dz2)(image) * vp^2 on the interior (double
 * intermediates), out is [nx][nz]; acc += out. */
void or_image_laplacian(const float *I, const float *vp, float *out, float *acc,
                        int ntx, int ntz, int pmlc, float dx, float dz)
{
    int ix, iz, nz = ntz - 2 * pmlc;
    for (iz = pmlc; iz <= ntz - pmlc - 1; iz++)
        for (ix = pmlc; ix <= ntx - pmlc - 1; ix++) {
            int ip = ix * ntz + iz, ipp = (ix - pmlc) * nz + iz - pmlc;
            double v = -(((double)I[ip + ntz] - 2.0 * I[ip] + I[ip - ntz]) / (double)(dx * dx) +
                         ((double)I[ip + 1] - 2.0 * I[ip] + I[ip - 1]) / (double)(dz * dz)) *
                       vp[ip] * vp[ip];
            out[ipp] = (float)v;
            if (acc) acc[ipp] = acc[ipp] + out[ipp];
        }
}

/* RC:754-768 */
void or_edge_zero(float *mig, int nx, int nz, int Lc)
{
    int ix, iz;
    for (ix = 0; ix < nx; ix++)
        for (iz = 0; iz < nz; iz++)
            if (ix <= 3 || ix >= nx - Lc || iz < 7 || iz >= nz - Lc) mig[ix * nz + iz] = 0.0f;
}

/* ---- propagation twins ---- */
#define REAL float
#define SFX f32
#define FMA fmaf
#define CEILR ceilf
#include "swd_oracle_prop.inc"
#undef REAL
#undef SFX
#undef FMA
#undef CEILR

#define REAL double
#define SFX f64
#define FMA fma
#define CEILR ceil
#include "swd_oracle_prop.inc"
#undef REAL
#undef SFX
#undef FMA
#undef CEILR
