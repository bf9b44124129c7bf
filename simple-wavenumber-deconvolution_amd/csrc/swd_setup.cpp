// swd_setup.cpp -- host-side setup of the FDTD/RTM path with the reference's
// semantics: parameter.txt parsing, geometry, FD and CPML coefficients, model
// I/O and smoothing, kz taper, wavelet Hilbert transform, adjoint-source
// integration and image post-processing.
//
// The reference computes these on the host in C++ with a mix of float and
// double; each routine here keeps the same precision at each step (float
// arguments to std::exp/std::log pick the float overloads exactly as the
// reference's <math.h> calls do; pow(float,int) promotes to double).  This
// file is compiled with -ffp-contract=off so no multiply-add is fused.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <complex>
#include <vector>

#include "swd.h"
#include "swd_internal.h"

namespace {
constexpr double kRefPi = 3.1415926;   // RC:7, forward/headrtm.h:5
constexpr float kPiF = 3.1415927f;     // RC:1251, RC:907

template <class F>
bool read_pair(FILE *fp, const char *fmt, F *v)
{
    char line[256];
    if (!std::fgets(line, sizeof line, fp)) return false;   // label line
    if (std::fscanf(fp, "\n") < 0) return false;
    if (std::fscanf(fp, fmt, v) != 1) return false;
    std::fscanf(fp, "\n");
    return true;
}
}  // namespace

extern "C" int swd_read_parameters(const char *path, swd_params *o)
{
    if (!path || !o) return swd_fail(SWD_E_ARG, "swd_read_parameters: null argument");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return swd_fail(SWD_E_IO, "Cannot open the parameters file %s", path);
    bool ok = read_pair(fp, "%d", &o->nx) && read_pair(fp, "%d", &o->nz) && read_pair(fp, "%d", &o->pml) &&
              read_pair(fp, "%d", &o->Lc) && read_pair(fp, "%f", &o->dx) && read_pair(fp, "%f", &o->dz) &&
              read_pair(fp, "%d", &o->itmax) && read_pair(fp, "%f", &o->dt) && read_pair(fp, "%f", &o->f0) &&
              read_pair(fp, "%d", &o->ns) && read_pair(fp, "%f", &o->sx0) && read_pair(fp, "%f", &o->shotdx) &&
              read_pair(fp, "%f", &o->shotdep) && read_pair(fp, "%f", &o->recdix) &&
              read_pair(fp, "%f", &o->offsetz) && read_pair(fp, "%d", &o->scatter_flag) &&
              read_pair(fp, "%d", &o->swin) && read_pair(fp, "%d", &o->window);
    std::fclose(fp);
    if (!ok) return swd_fail(SWD_E_IO, "parameter file %s is short or malformed", path);
    return SWD_OK;
}

extern "C" int swd_decon_read_parameters(const char *path, swd_decon_params *o, swd_decon_files *f)
{
    if (!path || !o) return swd_fail(SWD_E_ARG, "swd_decon_read_parameters: null argument");
    FILE *fp = std::fopen(path, "r");
    if (!fp) return swd_fail(SWD_E_IO, "Cannot open the parameters file %s", path);
    swd_decon_files tmp;
    std::memset(&tmp, 0, sizeof tmp);
    bool ok = read_pair(fp, "%d", &o->nx) && read_pair(fp, "%d", &o->nz) && read_pair(fp, "%d", &o->psfwinx) &&
              read_pair(fp, "%d", &o->psfwinz) && read_pair(fp, "%d", &o->filterwinx) &&
              read_pair(fp, "%d", &o->filterwinz) && read_pair(fp, "%d", &o->snapwinx) &&
              read_pair(fp, "%d", &o->snapwinz) && read_pair(fp, "%f", &o->epsilon);
    if (ok)   // file names and flags; a file that stops after epsilon leaves them empty
        read_pair(fp, "%255s", tmp.vp) && read_pair(fp, "%255s", tmp.vp0) && read_pair(fp, "%255s", tmp.rho) &&
            read_pair(fp, "%255s", tmp.image) && read_pair(fp, "%255s", tmp.psf) &&
            read_pair(fp, "%255s", tmp.decon) && read_pair(fp, "%255s", tmp.inv_psf) &&
            read_pair(fp, "%d", &tmp.shift_flag) && read_pair(fp, "%f", &tmp.reg_eps);
    std::fclose(fp);
    if (!ok) return swd_fail(SWD_E_IO, "parameter file %s is short or malformed", path);
    if (f) *f = tmp;
    return SWD_OK;
}

extern "C" int swd_cal_xishu(int Lx, float *rx)
{
    if (Lx < 1 || Lx > SWD_MAX_LC || !rx) return swd_fail(SWD_E_ARG, "swd_cal_xishu: Lc must be 1..%d", SWD_MAX_LC);
    for (int m = 1; m <= Lx; m++) {
        float s1 = 1.0f, s2 = 1.0f;
        const double om = 2.0 * m - 1;
        for (int i = 1; i <= Lx; i++) {
            if (i == m) continue;
            const double oi = 2.0 * i - 1;
            s1 = static_cast<float>(s1 * oi * oi);
            s2 = static_cast<float>(s2 * (om * om - oi * oi));
        }
        s2 = std::fabs(s2);
        rx[m - 1] = static_cast<float>(std::pow(-1.0, m + 1) * s1 / (s2 * om));
    }
    return SWD_OK;
}

extern "C" int swd_cfl_vmax(const swd_grid *g, const float *rc, double *vmax)
{
    if (!g || !vmax || g->Lc < 1 || g->Lc > SWD_MAX_LC || !(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0))
        return swd_fail(SWD_E_ARG, "swd_cfl_vmax: bad grid");
    float own[SWD_MAX_LC];
    if (!rc) {
        swd_cal_xishu(g->Lc, own);
        rc = own;
    }
    double sum = 0.0;   // tmprc of RC:110-114
    for (int i = 0; i < g->Lc; i++) sum += std::fabs(static_cast<double>(rc[i]));
    const double h = std::sqrt(1.0 / (static_cast<double>(g->dx) * g->dx) + 1.0 / (static_cast<double>(g->dz) * g->dz));
    *vmax = 1.0 / (sum * g->dt * h);
    return SWD_OK;
}

extern "C" int swd_check_model(const swd_grid *g, const float *rc, const float *vp, const float *rho)
{
    if (!g || !vp || !rho) return swd_fail(SWD_E_ARG, "swd_check_model: null argument");
    double lim = 0.0;
    const int e = swd_cfl_vmax(g, rc, &lim);
    if (e != SWD_OK) return e;
    const long long pmlc = static_cast<long long>(g->pml) + g->Lc;
    const long long ntz = g->nz + 2 * pmlc, ntp = (g->nx + 2 * pmlc) * ntz;
    float vmax = 0.f;
    long long imax = 0;
    for (long long ip = 0; ip < ntp; ip++) {
        if (!(vp[ip] > 0.f) || !(rho[ip] > 0.f) || !std::isfinite(vp[ip]) || !std::isfinite(rho[ip]))
            return swd_fail(SWD_E_ARG, "model must be positive and finite (vp=%g rho=%g at ix=%lld iz=%lld)", vp[ip],
                            rho[ip], ip / ntz, ip % ntz);
        if (vp[ip] > vmax) {
            vmax = vp[ip];
            imax = ip;
        }
    }
    if (vmax > lim)
        return swd_fail(SWD_E_RANGE,
                        "vp=%g m/s at ix=%lld iz=%lld exceeds the stability limit %.3f m/s of dx=%g dz=%g dt=%g Lc=%d",
                        vmax, imax / ntz, imax % ntz, lim, g->dx, g->dz, g->dt, g->Lc);
    return SWD_OK;
}

extern "C" int swd_ricker(float *rick, int nt, float f0, float t0, float dt)
{
    if (!rick || nt < 1) return swd_fail(SWD_E_ARG, "swd_ricker: bad arguments");
    for (int it = 0; it < nt; it++) {
        float ph = kPiF * f0 * (static_cast<float>(it) * dt - t0);
        ph = ph * ph;
        rick[it] = static_cast<float>((1.0 - 2.0 * ph) * std::exp(-ph));
    }
    return SWD_OK;
}

namespace {
// One axis of get_absorbing_parameters (RC:959-1120).
void cpml_axis(int n, int pml, float h, float thickness, float d0, float dt,
               float *a, float *ah, float *b, float *bh)
{
    const float alpha_max = kPiF * 15;
    const float start = static_cast<float>(pml) * h;
    const float end = static_cast<float>(n - pml - 1) * h;
    std::vector<float> d(n, 0.f), dh(n, 0.f), al(n, 0.f), alh(n, 0.f);
    auto profile = [&](float delta, float &dd, float &aa) {
        const float q = delta / thickness;
        dd = static_cast<float>(d0 * std::pow(static_cast<double>(q), 2));
        aa = static_cast<float>(alpha_max * (1.0 - q) + 0.1 * alpha_max);
    };
    for (int i = 0; i < n; i++) {
        const float xi = static_cast<float>(i) * h;
        if (xi <= start) profile(start - xi, d[i], al[i]);
        if (xi >= end) profile(xi - end, d[i], al[i]);
        const float xh = static_cast<float>((i + 0.5) * h);
        if (xh <= start) profile(start - xh, dh[i], alh[i]);
        if (xh >= end) profile(xh - end, dh[i], alh[i]);
    }
    for (int i = 0; i < n; i++) {
        const float k = 1.0f;
        if (al[i] < 0.0f) al[i] = 0.0f;
        if (alh[i] < 0.0f) alh[i] = 0.0f;
        b[i] = std::exp(-(d[i] / k + al[i]) * dt);
        a[i] = d[i] > 1.0e-6 ? static_cast<float>(d[i] / (k * (d[i] + k * al[i])) * (b[i] - 1.0)) : 0.0f;
        bh[i] = std::exp(-(dh[i] / k + alh[i]) * dt);
        ah[i] = dh[i] > 1.0e-6 ? static_cast<float>(dh[i] / (k * (dh[i] + k * alh[i])) * (bh[i] - 1.0)) : 0.0f;
    }
}
}  // namespace

extern "C" int swd_cpml(int ntx, int ntz, int pml, float dx, float dz, float dt,
                        float *a_x, float *a_x_half, float *b_x, float *b_x_half,
                        float *a_z, float *a_z_half, float *b_z, float *b_z_half)
{
    if (ntx < 2 * pml + 1 || ntz < 2 * pml + 1 || pml < 1 || !a_x || !a_x_half || !b_x || !b_x_half ||
        !a_z || !a_z_half || !b_z || !b_z_half)
        return swd_fail(SWD_E_ARG, "swd_cpml: bad arguments");
    const int N = 2;
    const float Rc = 1.0e-5f, Vpmax = 5500;   // Vpmax fixed (RC:923)
    const float thickness = static_cast<float>(pml) * dz;
    const float d0 = static_cast<float>(static_cast<float>(-(N + 1)) * Vpmax * std::log(Rc) / (2.0 * thickness));
    cpml_axis(ntx, pml, dx, thickness, d0, dt, a_x, a_x_half, b_x, b_x_half);
    cpml_axis(ntz, pml, dz, thickness, d0, dt, a_z, a_z_half, b_z, b_z_half);
    return SWD_OK;
}

extern "C" int swd_model_in_pml(float *vp, int ntz, int ntx, int pml)
{
    if (!vp || pml < 0 || ntx <= 2 * pml || ntz <= 2 * pml) return swd_fail(SWD_E_ARG, "swd_model_in_pml: bad arguments");
    for (int iz = pml; iz <= ntz - pml - 1; iz++) {
        for (int ix = 0; ix < pml; ix++) vp[ix * ntz + iz] = vp[pml * ntz + iz];
        for (int ix = ntx - pml; ix < ntx; ix++) vp[ix * ntz + iz] = vp[(ntx - pml - 1) * ntz + iz];
    }
    for (int ix = 0; ix < ntx; ix++) {
        float *col = vp + static_cast<size_t>(ix) * ntz;
        for (int iz = 0; iz < pml; iz++) col[iz] = col[pml];
        for (int iz = ntz - pml; iz < ntz; iz++) col[iz] = col[ntz - pml - 1];
    }
    return SWD_OK;
}

extern "C" int swd_read_model(const char *path, float *vp, float *rho, int ntx, int ntz, int pml)
{
    if (!path || !vp || !rho) return swd_fail(SWD_E_ARG, "swd_read_model: null argument");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return swd_fail(SWD_E_IO, "Cannot open the model file %s", path);
    const int nz = ntz - 2 * pml;
    bool ok = true;
    for (int ix = pml; ix < ntx - pml && ok; ix++)
        ok = std::fread(vp + static_cast<size_t>(ix) * ntz + pml, sizeof(float), nz, fp) == static_cast<size_t>(nz);
    std::fclose(fp);
    if (!ok) return swd_fail(SWD_E_IO, "model file %s is shorter than %d x %d floats", path, ntx - 2 * pml, nz);
    int rc = swd_model_in_pml(vp, ntz, ntx, pml);
    if (rc) return rc;
    for (size_t ip = 0; ip < static_cast<size_t>(ntx) * ntz; ip++) rho[ip] = 1000.0f;   // RC:1198-1201
    return SWD_OK;
}

extern "C" int swd_smooth_model(float *vp, int ntz, int ntx, int pml, int window)
{
    if (!vp || window < 0 || ntx <= 2 * pml || ntz <= 2 * pml) return swd_fail(SWD_E_ARG, "swd_smooth_model: bad arguments");
    const int w = window, wn = 2 * w + 1;
    const double sigma = w / 2.0;
    std::vector<double> coef(wn * wn);
    double sum = 0.0;
    for (int ix = 0; ix < wn; ix++)
        for (int iz = 0; iz < wn; iz++) {
            const int r2 = (ix - w) * (ix - w) + (iz - w) * (iz - w);
            coef[ix * wn + iz] = 1.0 / (2.0 * kRefPi * sigma * sigma) * std::exp(-r2 / (2.0 * sigma * sigma));
            sum += coef[ix * wn + iz];
        }
    for (double &c : coef) c /= sum;
    std::vector<double> old(vp, vp + static_cast<size_t>(ntx) * ntz);
#pragma omp parallel for schedule(static)
    for (int iz = pml; iz <= ntz - pml - 1; iz++) {
        for (int ix = pml; ix <= ntx - pml - 1; ix++) {
            const size_t ip = static_cast<size_t>(ix) * ntz + iz;
            if (iz < pml + 8) {   // the top rows keep the input (RC:1551-1554)
                vp[ip] = static_cast<float>(old[ip]);
                continue;
            }
            double s = 0.0;
            for (int izw = iz - w; izw <= iz + w; izw++) {
                const int z1 = izw < 0 ? 0 : (izw > ntz - 1 ? ntz - 1 : izw);
                for (int ixw = ix - w; ixw <= ix + w; ixw++) {
                    const int x1 = ixw < 0 ? 0 : (ixw > ntx - 1 ? ntx - 1 : ixw);
                    s += old[static_cast<size_t>(x1) * ntz + z1] * coef[(izw - iz + w) * wn + (ixw - ix + w)];
                }
            }
            vp[ip] = static_cast<float>(s);
        }
    }
    return swd_model_in_pml(vp, ntz, ntx, pml);
}

extern "C" int swd_nfft(int n)
{
    if (n < 1) return swd_fail(SWD_E_ARG, "swd_nfft: n < 1");
    const int K = static_cast<int>(std::ceil(std::log(1.0 * n) / std::log(2.0)));
    return static_cast<int>(std::pow(2.0, K));
}

extern "C" int swd_kz_taper(float *coef, int nfftz, float dz)
{
    if (!coef || nfftz < 2 || (nfftz & (nfftz - 1))) return swd_fail(SWD_E_ARG, "swd_kz_taper: nfftz must be a power of two");
    const float dkz = static_cast<float>(1.0 / static_cast<float>(nfftz * dz));
    const float kcut = static_cast<float>(1.0 / 15.0 / dz);
    const int layerz = static_cast<int>(kcut / dkz);
    if (layerz > nfftz / 6 || layerz < 1)
        return swd_fail(SWD_E_RANGE, "Damping wavenumber is too long (layerz=%d, nfftz=%d)", layerz, nfftz);
    const int upz = layerz, downz = nfftz - layerz;
    for (int iz = 0; iz < nfftz; iz++) {
        if (iz <= upz) coef[iz] = static_cast<float>(0.5000 - 0.5 * std::cos(kRefPi * iz / layerz));
        else if (iz < downz) coef[iz] = 1.0f;
        else coef[iz] = static_cast<float>(0.5000 + 0.5 * std::cos(kRefPi * (iz - downz) / layerz));
    }
    return layerz;
}

namespace {
// Iterative radix-2 FFT in double (sign -1 forward, +1 inverse, unnormalised).
void fft64(std::vector<std::complex<double>> &a, int sign)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / static_cast<double>(len);
        for (size_t k = 0; k < len / 2; k++) {
            const std::complex<double> w(std::cos(ang * k), std::sin(ang * k));
            for (size_t i = k; i < n; i += len) {
                const std::complex<double> t = a[i + len / 2] * w;
                a[i + len / 2] = a[i] - t;
                a[i] += t;
            }
        }
    }
}
}  // namespace

extern "C" int swd_hilbert1d(const float *rick, float *hrick, int nt)
{
    if (!rick || !hrick || nt < 2) return swd_fail(SWD_E_ARG, "swd_hilbert1d: bad arguments");
    const int NZ = swd_nfft(nt);
    std::vector<std::complex<double>> a(NZ);
    for (int i = 0; i < nt; i++) a[i] = rick[i];
    fft64(a, -1);
    const std::complex<double> mi(0, -1), pi(0, 1);
    a[0] = 0;
    for (int k = 1; k < NZ / 2; k++) a[k] *= mi;   // FK:1294-1298
    for (int k = NZ / 2; k < NZ; k++) a[k] *= pi;  // FK:1299-1303 (Nyquist gets +i)
    fft64(a, +1);
    for (int i = 0; i < nt; i++) hrick[i] = static_cast<float>(a[i].real() / NZ);
    return SWD_OK;
}

extern "C" int swd_geometry(int nx, int nz, int pml, int Lc, float dx, float dz, float sx0, float shotdx,
                            float shotdep, float offsetz, int is, int *s_ix, int *s_iz, int *r_iz, int max_rn)
{
    (void)nx;
    if (!s_ix || !s_iz) return swd_fail(SWD_E_ARG, "swd_geometry: null argument");
    const int pmlc = pml + Lc;
    *s_ix = pmlc + static_cast<int>(sx0 / dx) + static_cast<int>(shotdx / dx) * is;
    *s_iz = pmlc + static_cast<int>(shotdep / dz);
    std::vector<int> rows;
    for (int iz = 0; iz < nz; iz++)
        if (std::fabs(static_cast<double>(*s_iz - iz - pmlc)) * dz <= offsetz) rows.push_back(pmlc + iz);
    const int rn = 2 * static_cast<int>(rows.size());
    if (r_iz) {
        if (rn > max_rn) return swd_fail(SWD_E_RANGE, "The trace number of %d th source is out of range", is + 1);
        for (size_t i = 0; i < rows.size(); i++) {
            r_iz[i] = rows[i];
            r_iz[i + rows.size()] = rows[i];
        }
    }
    return rn;
}

extern "C" int swd_adjoint_source(const float *d, float *rms, int ntr, int nt, float dt)
{
    if (!d || !rms || ntr < 0 || nt < 1) return swd_fail(SWD_E_ARG, "swd_adjoint_source: bad arguments");
    std::vector<float> f(nt);
    for (int tr = 0; tr < ntr; tr++) {
        const float *di = d + static_cast<size_t>(tr) * nt;
        float *ro = rms + static_cast<size_t>(tr) * nt;
        float acc = 0.0f;
        for (int it = 0; it < nt; it++) {
            const float inc = di[it] * dt;
            acc = acc + inc;
            f[it] = acc;
        }
        acc = 0.0f;
        for (int it = 0; it < nt; it++) {
            const float inc = f[it] * dt;
            acc = acc + inc;
            ro[it] = -acc;
        }
    }
    return SWD_OK;
}

extern "C" int swd_scatterers(float *rho, const float *vp, int ntx, int ntz, int pmlc, int swin)
{
    if (!rho || !vp || swin < 1) return swd_fail(SWD_E_ARG, "swd_scatterers: bad arguments");
    const int o = pmlc + swin / 2;
    for (int ix = o; ix < ntx - pmlc - swin / 2; ix += swin)
        for (int iz = o; iz < ntz - pmlc - 5; iz += swin) {
            const size_t ip = static_cast<size_t>(ix) * ntz + iz;
            rho[ip] = static_cast<float>(rho[ip] + vp[ip] * 1.0e-3);
        }
    return SWD_OK;
}

extern "C" int swd_image_normalize(float *image, const float *illum, int ntp)
{
    if (!image || !illum || ntp < 0) return swd_fail(SWD_E_ARG, "swd_image_normalize: bad arguments");
    float remax = 0;
    for (int ip = 0; ip < ntp; ip++) remax = std::fabs(illum[ip]) > remax ? std::fabs(illum[ip]) : remax;
    for (int ip = 0; ip < ntp; ip++) image[ip] = static_cast<float>(image[ip] / (illum[ip] + 1.0e-3 * remax));
    return SWD_OK;
}

extern "C" int swd_image_laplacian(const float *I, const float *vp, float *out, float *acc,
                                   int ntx, int ntz, int pmlc, float dx, float dz)
{
    if (!I || !vp || !out || pmlc < 1) return swd_fail(SWD_E_ARG, "swd_image_laplacian: bad arguments");
    const int nz = ntz - 2 * pmlc;
    const double hx = dx * dx, hz = dz * dz;   // float products, as in RC:727-728
    for (int ix = pmlc; ix <= ntx - pmlc - 1; ix++)
        for (int iz = pmlc; iz <= ntz - pmlc - 1; iz++) {
            const size_t ip = static_cast<size_t>(ix) * ntz + iz;
            const size_t o = static_cast<size_t>(ix - pmlc) * nz + iz - pmlc;
            const double lx = (I[ip + ntz] - 2.0 * I[ip] + I[ip - ntz]) / hx;
            const double lz = (I[ip + 1] - 2.0 * I[ip] + I[ip - 1]) / hz;
            out[o] = static_cast<float>(-(lx + lz) * vp[ip] * vp[ip]);
            if (acc) acc[o] = acc[o] + out[o];
        }
    return SWD_OK;
}

extern "C" int swd_edge_zero(float *mig, int nx, int nz, int Lc)
{
    if (!mig) return swd_fail(SWD_E_ARG, "swd_edge_zero: null argument");
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++)
            if (ix <= 3 || ix >= nx - Lc || iz < 7 || iz >= nz - Lc) mig[static_cast<size_t>(ix) * nz + iz] = 0.0f;
    return SWD_OK;
}
