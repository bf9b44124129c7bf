// swd_decon.hip -- wavenumber-domain PSF deconvolution on gfx950: the
// reference's windowed Wiener deconvolution of a migrated image by the nearest
// PSF of a point-scatterer lattice (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp,
// DC:543-731), one process, every window on one device.
//
// Per window of side n (fullw = filterwin + psfwin - 1) the reference runs two
// forward 2D FFTs, a Wiener division and one inverse 2D FFT in double, then
// overlap-adds a tapered outwin x outwin block in window order.  Here:
//   k_decon_sums     one thread per window: the float sum of magnitudes of the
//                    image window (pp_max, DC:555-574), a serial chain kept in
//                    the reference's order;
//   k_decon_cells    one workgroup per PSF lattice cell (windows share cells,
//                    2484 windows use 374 cells at the reference size): the
//                    normalised PSF's half spectrum and the Wiener denominators
//                    tmpfloat = float(|S|^2 + eps * max|S|^2) (DC:583-648);
//   k_decon_windows  one workgroup per window: half spectrum of the image
//                    window, Wiener division, inverse only at the output block
//                    (DC:653-673);
//   k_decon_gather   one thread per output cell: the tapered overlap-add of
//                    the windows covering it, in window order (DC:653-673).
// The transforms are direct DFTs as small complex matrix products: n = 101 is
// prime, the operands are LDS-resident, and the window-independent operand
// (twiddles, constant over a launch) is wave-uniform and read through the
// scalar cache, so each lane owns one z wavenumber k1 and the fp64 FMAs take
// one vector operand from LDS and one scalar operand.  All arithmetic is in the
// fixed order of the host oracle (fma chains in index order, the Hermitian half
// spectrum along z, the image normalisation folded into the exact divisor
// pp_max * tmpfloat), so results agree with it bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include <hip/hip_runtime.h>

#include "swd.h"
#include "swd_internal.h"

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return swd_fail(e_ == hipErrorOutOfMemory ? SWD_E_NOMEM : SWD_E_DEVICE, "%s failed: %s (%s:%d)", \
                            #expr, hipGetErrorString(e_), __FILE__, __LINE__);                            \
    } while (0)

namespace {

constexpr int kMaxN = 127;        // window side: the half spectrum (n/2+1) fits one wavefront

struct DGeo {
    int nx, nz, n, nh;
    int mxi, mzi, offx, offz;     // image window support and its place in the n x n buffer
    int mxp, mzp;                 // PSF patch support (placed at the origin)
    int outx, outz, r0x, c0z;     // output block of the inverse
    int halfltwx, halfltwz, halfpsfwx, halfpsfwz, psfwinx, psfwinz, nxwpsf, nzwpsf;
    int halfoutwx, halfoutwz, stepwinx, stepwinz, stepnumx, stepnumz, totalnum, nxp, nzp;
    int ncx, ncz;                 // PSF lattice cells 3..nxwpsf-4 x 3..nzwpsf-4
    float epsilon;
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The edge-replicated [nxp][nzp] image (DC:300-340) at (x, z), zero outside it
// (the window bounds checks of DC:565 and DC:664).
__device__ __forceinline__ float expanded(const float *__restrict__ a, const DGeo &G, int x, int z)
{
    if (x < 0 || x >= G.nxp || z < 0 || z >= G.nzp) return 0.0f;
    return a[(size_t)min(x, G.nx - 1) * G.nz + min(z, G.nz - 1)];
}

// PSF lattice cell of window s (DC:547-548, 583-596), as a cell-table index.
__device__ __forceinline__ int cell_of(const DGeo &G, int s)
{
    const int ix = G.stepwinx * (s / G.stepnumz), iz = G.stepwinz * (s % G.stepnumz);
    int cx = (ix - G.halfpsfwx) / G.psfwinx, cz = (iz - G.halfpsfwz) / G.psfwinz;
    if (cx <= 2) cx = 3;
    if (cx >= G.nxwpsf - 3) cx = G.nxwpsf - 4;
    if (cz <= 2) cz = 3;
    if (cz >= G.nzwpsf - 3) cz = G.nzwpsf - 4;
    return (cx - 3) * G.ncz + (cz - 3);
}

// ---------------------------------------------------------------- stages
// Rows pass: Y[r][k1] = sum_c X[r][c] e^{-2 pi i k1 (c + offz) / n} for the
// half spectrum k1 < nh; lane = k1, wave w owns rows w + NW rr (rr < R, X is
// zero-padded to NW R rows); X is read as an LDS broadcast, the twiddle per
// lane from the LDS table.
template <int NW, int R, class XT>
__device__ __forceinline__ void rows_pass(const XT *X, int mx, int mz, int offz, int n, int nh, const double2 *tw,
                                          double2 *Y)
{
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int k1 = min(lane, nh - 1);
    double yr[R], yi[R];
#pragma unroll
    for (int rr = 0; rr < R; rr++) yr[rr] = yi[rr] = 0.0;
    int t = (k1 * offz) % n;
    const XT *Xw = X + w * mz;
    for (int c = 0; c < mz; c++) {
        const double2 e = tw[t];
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const double x = static_cast<double>(Xw[rr * NW * mz + c]);
            yr[rr] = fma(x, e.x, yr[rr]);
            yi[rr] = fma(-x, e.y, yi[rr]);
        }
        t += k1;
        if (t >= n) t -= n;
    }
    if (lane < nh) {
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const int r = w + NW * rr;
            if (r < mx) Y[r * nh + lane] = make_double2(yr[rr], yi[rr]);
        }
    }
}

// Columns pass: M[k0][k1] = sum_r e^{-2 pi i k0 (r + offx) / n} Y[r][k1] for
// the P wavenumbers k0 = w P + kk of wave w.  A[r][k0] (rows of NW P entries,
// zero-padded) is the launch-constant twiddle table, read wave-uniformly
// through the scalar cache; the four waves a SIMD holds cover each other's
// scalar-load latency.
template <int NW, int P>
__device__ __forceinline__ void cols_pass(const double2 *Y, int mx, int nh, const double2 *__restrict__ A,
                                          double (&mr)[P], double (&mi)[P])
{
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int k1 = min(lane, nh - 1);
    constexpr int LD = NW * P;
    const double2 *Aw = A + w * P;
#pragma unroll
    for (int kk = 0; kk < P; kk++) mr[kk] = mi[kk] = 0.0;
    for (int r = 0; r < mx; r++) {
        const double2 y = Y[r * nh + k1];
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
            const double2 a = Aw[r * LD + kk];
            mr[kk] = fma(y.x, a.x, mr[kk]);
            mr[kk] = fma(y.y, a.y, mr[kk]);
            mi[kk] = fma(y.y, a.x, mi[kk]);
            mi[kk] = fma(-y.x, a.y, mi[kk]);
        }
    }
}

// Inverse along x at the output rows: Q[i][k1] = sum_k0 e^{+2 pi i k0 (r0x + i) / n} D[k0][k1]
// for the PQ rows i = w PQ + ii of wave w, U wavenumbers k0 a step; Tq[k0][i]
// (rows of NW PQ entries) is read through the scalar cache.  D and Tq are
// zero past row n, up to a whole number of steps.
template <int NW, int PQ>
__device__ __forceinline__ void inv_rows_pass(const double2 *D, int n, int nh, const double2 *__restrict__ Tq,
                                              double (&qr)[PQ], double (&qi)[PQ])
{
    constexpr int U = PQ >= 8 ? 1 : 8 / PQ;
    constexpr int LD = NW * PQ;
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int k1 = min(lane, nh - 1);
    const double2 *Tw = Tq + w * PQ;
#pragma unroll
    for (int ii = 0; ii < PQ; ii++) qr[ii] = qi[ii] = 0.0;
    for (int k0 = 0; k0 < n; k0 += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {   // rows k0 >= n of D and Tq are zero: they add nothing
            const double2 d = D[(k0 + u) * nh + k1];
#pragma unroll
            for (int ii = 0; ii < PQ; ii++) {
                const double2 e = Tw[(k0 + u) * LD + ii];
                qr[ii] = fma(d.x, e.x, qr[ii]);
                qr[ii] = fma(-d.y, e.y, qr[ii]);
                qi[ii] = fma(d.y, e.x, qi[ii]);
                qi[ii] = fma(d.x, e.y, qi[ii]);
            }
        }
    }
}

// Rows of D and of the Tq table: n rounded up to whole steps of inv_rows_pass.
__host__ __device__ constexpr int inv_rows_pad(int n, int PQ)
{
    return (n + (PQ >= 8 ? 1 : 8 / PQ) - 1) / (PQ >= 8 ? 1 : 8 / PQ) * (PQ >= 8 ? 1 : 8 / PQ);
}

// ---------------------------------------------------------------- kernels
__global__ __launch_bounds__(256) void k_decon_sums(DGeo G, const float *__restrict__ img, float *__restrict__ pp)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= G.totalnum) return;
    const int x0 = G.stepwinx * (s / G.stepnumz) - G.halfltwx, z0 = G.stepwinz * (s % G.stepnumz) - G.halfltwz;
    // float(double(acc) + |v|) of DC:568 is the float sum acc + |v| (binary64
    // holds more than twice binary32's precision, so the double rounding is exact)
    float acc = 0.0f;
    for (int i = 0; i < G.mxi; i++) {
        const int x = x0 + i;
        if (x < 0 || x >= G.nxp) continue;
        const float *row = img + (size_t)min(x, G.nx - 1) * G.nz;
        for (int j = 0; j < G.mzi; j++) {
            const int z = z0 + j;
            if (z < 0 || z >= G.nzp) continue;
            acc = acc + fabsf(row[min(z, G.nz - 1)]);
        }
    }
    if (static_cast<double>(acc) <= 1.0e-15) acc = 1.0e-10f;
    pp[s] = acc;
}

template <int NW, int R, int P>
__global__ __launch_bounds__(NW * 64) void k_decon_cells(DGeo G, const float *__restrict__ psf,
                                                         const double2 *__restrict__ tw_g,
                                                         const double2 *__restrict__ Apsf, double2 *__restrict__ S,
                                                         float *__restrict__ T)
{
    constexpr int NT = NW * 64;
    extern __shared__ __align__(16) unsigned char lds[];
    const int n = G.n, nh = G.nh, mx = G.mxp, mz = G.mzp;
    double2 *tw = reinterpret_cast<double2 *>(lds);
    double2 *Y = tw + n;                                                   // mx (+1) rows
    double *X = reinterpret_cast<double *>(Y + (mx + 1) * nh);             // NW R rows
    float *scal = reinterpret_cast<float *>(X + NW * R * mz);              // [0] psf_max, [1] max |S|^2 (bits)
    const int cell = blockIdx.x;
    const int cx = 3 + cell / G.ncz, cz = 3 + cell % G.ncz;
    const int x0 = cx * G.psfwinx, z0 = cz * G.psfwinz;   // ixp - halfpsfwx, izp - halfpsfwz (DC:597-612)
    for (int t = threadIdx.x; t < n; t += NT) tw[t] = tw_g[t];
    for (int e = threadIdx.x; e < (mx + 1) * nh; e += NT) Y[e] = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) {
        float acc = 0.0f;   // psf_max (DC:617-619), float sum in the reference's order
        for (int i = 0; i < mx; i++)
            for (int j = 0; j < mz; j++) acc = acc + fabsf(expanded(psf, G, x0 + i, z0 + j));
        scal[0] = acc;
        reinterpret_cast<unsigned *>(scal)[1] = 0u;
    }
    __syncthreads();
    const double pm = static_cast<double>(scal[0]);
    for (int e = threadIdx.x; e < NW * R * mz; e += NT) {
        const int i = e / mz;
        X[e] = i < mx ? static_cast<double>(expanded(psf, G, x0 + i, z0 + e % mz)) / pm : 0.0;
    }
    __syncthreads();
    rows_pass<NW, R>(X, mx, mz, 0, n, nh, tw, Y);
    __syncthreads();
    double mr[P], mi[P];
    cols_pass<NW, P>(Y, mx, nh, Apsf, mr, mi);
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    // fft_max: max over the bins of float(|S|^2) (DC:625-633); non-negative
    // floats order as their bit patterns
    float qmax = 0.0f;
#pragma unroll
    for (int kk = 0; kk < P; kk++)
        if (w * P + kk < n && lane < nh) qmax = fmaxf(qmax, static_cast<float>(mr[kk] * mr[kk] + mi[kk] * mi[kk]));
    atomicMax(reinterpret_cast<unsigned *>(scal) + 1, __float_as_uint(qmax));
    __syncthreads();
    float fft_max = __uint_as_float(reinterpret_cast<unsigned *>(scal)[1]) * G.epsilon;
    if (fft_max == 0.0f) fft_max = 1.0e-15f;
    if (lane < nh) {
        double2 *Sc = S + (size_t)cell * n * nh;
        float *Tc = T + (size_t)cell * n * nh;
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
            const int k0 = w * P + kk;
            if (k0 < n) {
                Sc[k0 * nh + lane] = make_double2(mr[kk], mi[kk]);
                Tc[k0 * nh + lane] = static_cast<float>((mr[kk] * mr[kk] + mi[kk] * mi[kk]) + static_cast<double>(fft_max));
            }
        }
    }
}

template <int NW, int R, int P, int PQ>
__global__ __launch_bounds__(NW * 64) void k_decon_windows(DGeo G, const float *__restrict__ img,
                                                           const float *__restrict__ pp,
                                                           const double2 *__restrict__ tw_g,
                                                           const double2 *__restrict__ Aimg,
                                                           const double2 *__restrict__ Tinv,
                                                           const double2 *__restrict__ S,
                                                           const float *__restrict__ T, double *__restrict__ wout)
{
    constexpr int NT = NW * 64;
    extern __shared__ __align__(16) unsigned char lds[];
    const int n = G.n, nh = G.nh, mx = G.mxi, mz = G.mzi;
    double2 *tw = reinterpret_cast<double2 *>(lds);
    double2 *Rg = tw + n;                                            // Y, then D, then Q
    float *X = reinterpret_cast<float *>(Rg + (mx + 1) * nh);        // image window beside Y, NW R rows
    const int s = blockIdx.x;
    const int x0 = G.stepwinx * (s / G.stepnumz) - G.halfltwx, z0 = G.stepwinz * (s % G.stepnumz) - G.halfltwz;
    for (int t = threadIdx.x; t < n; t += NT) tw[t] = tw_g[t];
    for (int e = threadIdx.x; e < NW * R * mz; e += NT) {
        const int i = e / mz;
        X[e] = i < mx ? expanded(img, G, x0 + i, z0 + e % mz) : 0.0f;
    }
    for (int e = mx * nh + threadIdx.x; e < (mx + 1) * nh; e += NT) Rg[e] = make_double2(0.0, 0.0);
    __syncthreads();
    rows_pass<NW, R>(X, mx, mz, G.offz, n, nh, tw, Rg);
    __syncthreads();
    double mr[P], mi[P];
    cols_pass<NW, P>(Rg, mx, nh, Aimg, mr, mi);
    // Wiener division (DC:642-651) with the image normalisation in the divisor
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int k1 = min(lane, nh - 1);
    const int cell = uniform(cell_of(G, s));
    const double ppd = static_cast<double>(pp[s]);
    const double2 *Sc = S + (size_t)cell * n * nh;
    const float *Tc = T + (size_t)cell * n * nh;
#pragma unroll
    for (int kk = 0; kk < P; kk++) {
        const int k0 = w * P + kk;
        if (k0 < n) {
            const double2 sv = Sc[k0 * nh + k1];
            const double den = ppd * static_cast<double>(Tc[k0 * nh + k1]);
            const double dr = (mr[kk] * sv.x + mi[kk] * sv.y) / den;
            const double di = (-mr[kk] * sv.y + mi[kk] * sv.x) / den;
            mr[kk] = dr;
            mi[kk] = di;
        }
    }
    __syncthreads();   // every wave is past its reads of Y
    if (lane < nh) {
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
            const int k0 = w * P + kk;
            if (k0 < n) Rg[k0 * nh + lane] = make_double2(mr[kk], mi[kk]);
        }
    }
    for (int e = n * nh + threadIdx.x; e < inv_rows_pad(n, PQ) * nh; e += NT) Rg[e] = make_double2(0.0, 0.0);
    __syncthreads();
    const int outx = G.outx, outz = G.outz;
    double qr[PQ], qi[PQ];
    inv_rows_pass<NW, PQ>(Rg, n, nh, Tinv, qr, qi);
    __syncthreads();   // every wave is past its reads of D
    if (lane < nh) {
#pragma unroll
        for (int ii = 0; ii < PQ; ii++) {
            const int i = w * PQ + ii;
            if (i < outx) Rg[i * nh + lane] = make_double2(qr[ii], qi[ii]);
        }
    }
    __syncthreads();
    // inverse along z at the output columns, real part of the Hermitian sum
    double *ws = wout + (size_t)s * outx * outz;
    for (int o = threadIdx.x; o < outx * outz; o += NT) {
        const int i = o / outz, j = o % outz;
        const int step = (G.c0z + j) % n;
        int t = 0;
        double acc = 0.0;
        for (int kk = 0; kk < nh; kk++) {
            const double wgt = (kk == 0 || 2 * kk == n) ? 1.0 : 2.0;
            const double2 q = Rg[i * nh + kk];
            const double2 e = tw[t];
            acc = fma(wgt * q.x, e.x, acc);
            acc = fma(-(wgt * q.y), e.y, acc);
            t += step;
            if (t >= n) t -= n;
        }
        ws[o] = acc;
    }
}

__global__ __launch_bounds__(256) void k_decon_gather(DGeo G, const double *__restrict__ wout,
                                                      const float *__restrict__ coefr, float *__restrict__ dec,
                                                      float *__restrict__ coef)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= G.nx * G.nz) return;
    const int x = idx / G.nz, z = idx % G.nz;
    // windows sx with 0 <= x - (stepwinx sx - halfoutwx) < outx, ascending
    const int ax = x + G.halfoutwx - G.outx + 1, az = z + G.halfoutwz - G.outz + 1;
    const int sx0 = ax <= 0 ? 0 : (ax + G.stepwinx - 1) / G.stepwinx;
    const int sz0 = az <= 0 ? 0 : (az + G.stepwinz - 1) / G.stepwinz;
    const int sx1 = min((x + G.halfoutwx) / G.stepwinx, G.stepnumx - 1);
    const int sz1 = min((z + G.halfoutwz) / G.stepwinz, G.stepnumz - 1);
    float acc = 0.0f, cf = 0.0f;
    for (int sx = sx0; sx <= sx1; sx++) {
        const int i = x - (G.stepwinx * sx - G.halfoutwx);
        for (int sz = sz0; sz <= sz1; sz++) {
            const int j = z - (G.stepwinz * sz - G.halfoutwz);
            const size_t s = (size_t)sx * G.stepnumz + sz;
            const float c = coefr[i * G.outz + j];
            acc = static_cast<float>(static_cast<double>(acc) + wout[(s * G.outx + i) * G.outz + j] * static_cast<double>(c));
            cf = cf + c;
        }
    }
    dec[idx] = acc;
    coef[idx] = cf;
}

// Launch configurations: rows / wavenumbers per wavefront of each stage, with
// kNW wavefronts per workgroup.  The first fits the reference's windows
// (filterwin 71, psfwin 31, n 101, outwin 29), the second any window up to kMaxN.
struct KCfg {
    int R, P, PQ, Rc, Pc;
};
constexpr int kNW = 16;
constexpr KCfg kCfgs[2] = {{5, 7, 2, 2, 7}, {8, 8, 8, 8, 8}};

// ---------------------------------------------------------------- host
int decon_geometry(const swd_decon_params &p, DGeo &G)
{
    if (p.nx < 1 || p.nz < 1 || p.psfwinx < 1 || p.psfwinz < 1 || p.filterwinx < 1 || p.filterwinz < 1)
        return swd_fail(SWD_E_ARG, "deconvolution: sizes must be positive");
    if (p.psfwinx % 2 == 0 || p.psfwinz % 2 == 0 || p.filterwinx % 2 == 0 || p.filterwinz % 2 == 0)
        return swd_fail(SWD_E_RANGE, "psfwin and filterwin must be odd (DC:89-101)");
    if (p.snapwinx < 2 || p.snapwinz < 2) return swd_fail(SWD_E_RANGE, "snapwin must be at least 2");
    if (!(p.epsilon >= 0.0f)) return swd_fail(SWD_E_ARG, "epsilon must be non-negative");
    G = DGeo{};
    G.nx = p.nx;
    G.nz = p.nz;
    G.halfpsfwx = p.psfwinx / 2;
    G.halfpsfwz = p.psfwinz / 2;
    G.halfltwx = p.filterwinx / 2;
    G.halfltwz = p.filterwinz / 2;
    G.psfwinx = p.psfwinx;
    G.psfwinz = p.psfwinz;
    G.nzwpsf = static_cast<int>(std::ceil(1.0 * p.nz / p.psfwinz));
    G.nxwpsf = static_cast<int>(std::ceil(1.0 * p.nx / p.psfwinx));
    const int fullwx = p.filterwinx + p.psfwinx - 1, fullwz = p.filterwinz + p.psfwinz - 1;
    const int midpx = fullwx / 2, midpz = fullwz / 2;
    G.outx = 2 * (p.psfwinx / 10) + 2 * p.snapwinx + 1;
    G.outz = 2 * (p.psfwinz / 10) + 2 * p.snapwinz + 1;
    G.halfoutwx = G.outx / 2;
    G.halfoutwz = G.outz / 2;
    G.stepwinx = G.outx - p.snapwinx;
    G.stepwinz = G.outz - p.snapwinz;
    G.stepnumx = static_cast<int>(std::ceil(1.0 * (p.nx + fullwx) / G.stepwinx)) + 1;
    G.stepnumz = static_cast<int>(std::ceil(1.0 * (p.nz + fullwz) / G.stepwinz)) + 1;
    G.totalnum = G.stepnumx * G.stepnumz;
    G.nxp = G.stepnumx * G.stepwinx;
    G.nzp = G.stepnumz * G.stepwinz;
    if (p.filterwinx < p.psfwinx || p.filterwinz < p.psfwinz || p.filterwinx < G.outx || p.filterwinz < G.outz)
        return swd_fail(SWD_E_RANGE, "filterwin must cover psfwin and the output window (DC:96-129)");
    if (fullwx != fullwz)
        return swd_fail(SWD_E_RANGE, "filterwin + psfwin must be equal along x and z (square 2D transform, DC:490-493)");
    if (fullwx > kMaxN) return swd_fail(SWD_E_RANGE, "window side %d exceeds %d", fullwx, kMaxN);
    if (G.nxwpsf < 7 || G.nzwpsf < 7)
        return swd_fail(SWD_E_RANGE, "the PSF lattice needs at least 7 cells along x and z (clamp to 3..n-4, DC:589-596)");
    G.n = fullwx;
    G.nh = fullwx / 2 + 1;
    G.mxi = p.filterwinx;
    G.mzi = p.filterwinz;
    G.offx = midpx - G.halfltwx;
    G.offz = midpz - G.halfltwz;
    G.mxp = p.psfwinx;
    G.mzp = p.psfwinz;
    G.r0x = G.halfltwx - G.halfoutwx;
    G.c0z = G.halfltwz - G.halfoutwz;
    G.ncx = G.nxwpsf - 6;
    G.ncz = G.nzwpsf - 6;
    G.epsilon = p.epsilon;
    return SWD_OK;
}

struct DeconPlan {
    swd_decon_params p{};
    DGeo G{};
    int device = 0;
    int cfg = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {};
    // launch-constant tables
    double2 *tw = nullptr, *Aimg = nullptr, *Apsf = nullptr, *Tinv = nullptr;
    float *coefr = nullptr;
    // workspace
    float *pp = nullptr, *T = nullptr;
    double2 *S = nullptr;
    double *wout = nullptr;
    // host-call staging
    float *io = nullptr;
    size_t lds_cells = 0, lds_windows = 0;
    long long ws_bytes = 0;
    swd_decon_stats st{};
};

template <class T>
int dev_alloc(DeconPlan &P, T **ptr, size_t count)
{
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(ptr), sizeof(T) * count));
    P.ws_bytes += static_cast<long long>(sizeof(T) * count);
    return SWD_OK;
}

void free_plan(DeconPlan *P)
{
    (void)hipSetDevice(P->device);
    for (void *q : {(void *)P->tw, (void *)P->Aimg, (void *)P->Apsf, (void *)P->Tinv, (void *)P->coefr, (void *)P->pp,
                    (void *)P->T, (void *)P->S, (void *)P->wout, (void *)P->io})
        if (q) (void)hipFree(q);
    for (hipEvent_t &e : P->ev)
        if (e) (void)hipEventDestroy(e);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
}

bool cfg_fits(const DGeo &G, const KCfg &c)
{
    return G.mxi <= kNW * c.R && G.n <= kNW * c.P && G.outx <= kNW * c.PQ && G.mxp <= kNW * c.Rc && G.n <= kNW * c.Pc;
}

template <int C>
int set_lds_attrs(const DeconPlan &P)
{
    constexpr KCfg c = kCfgs[C];
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_decon_cells<kNW, c.Rc, c.Pc>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(P.lds_cells)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_decon_windows<kNW, c.R, c.P, c.PQ>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(P.lds_windows)));
    return SWD_OK;
}

int set_lds_limits(DeconPlan &P)
{
    const DGeo &G = P.G;
    const KCfg &c = kCfgs[P.cfg];
    P.lds_cells = sizeof(double2) * (G.n + (G.mxp + 1) * G.nh) + sizeof(double) * kNW * c.Rc * G.mzp + 2 * sizeof(float);
    const size_t region = std::max({sizeof(double2) * (G.mxi + 1) * G.nh + sizeof(float) * kNW * c.R * G.mzi,
                                    sizeof(double2) * inv_rows_pad(G.n, c.PQ) * G.nh, sizeof(double2) * G.outx * G.nh});
    P.lds_windows = sizeof(double2) * G.n + region;
    int max_lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, P.device));
    if (P.lds_windows > static_cast<size_t>(max_lds) || P.lds_cells > static_cast<size_t>(max_lds))
        return swd_fail(SWD_E_RANGE, "deconvolution window needs %zu bytes of LDS, the device has %d", P.lds_windows,
                        max_lds);
    return P.cfg == 0 ? set_lds_attrs<0>(P) : set_lds_attrs<1>(P);
}

// Twiddle tables, zero-padded to the launch configuration: Aimg [mxi+1][kNW P],
// Apsf [mxp+1][kNW Pc], Tinv [inv_rows_pad][kNW PQ].
int upload_tables(DeconPlan &P)
{
    const DGeo &G = P.G;
    const KCfg &c = kCfgs[P.cfg];
    const int n = G.n;
    std::vector<double2> tw(n);
    for (int t = 0; t < n; t++) tw[t] = make_double2(std::cos(2.0 * M_PI * t / n), std::sin(2.0 * M_PI * t / n));
    auto table = [&](int rows, int used_rows, int ld, int cols, auto index) {
        std::vector<double2> A((size_t)rows * ld, make_double2(0.0, 0.0));
        for (int r = 0; r < used_rows; r++)
            for (int k = 0; k < cols; k++) A[(size_t)r * ld + k] = tw[index(r, k)];
        return A;
    };
    const auto aimg = table(G.mxi + 1, G.mxi, kNW * c.P, n,
                            [&](int r, int k0) { return (static_cast<long>(k0) * (r + G.offx)) % n; });
    const auto apsf = table(G.mxp + 1, G.mxp, kNW * c.Pc, n, [&](int r, int k0) { return (static_cast<long>(k0) * r) % n; });
    const int qrows = inv_rows_pad(n, c.PQ);
    const auto tinv = table(qrows, n, kNW * c.PQ, G.outx,
                            [&](int k0, int i) { return (static_cast<long>(k0) * (G.r0x + i)) % n; });
    std::vector<float> coefr((size_t)G.outx * G.outz);
    swd_decon_taper(&P.p, coefr.data());
    if (int r = dev_alloc(P, &P.tw, n)) return r;
    if (int r = dev_alloc(P, &P.Aimg, aimg.size())) return r;
    if (int r = dev_alloc(P, &P.Apsf, apsf.size())) return r;
    if (int r = dev_alloc(P, &P.Tinv, tinv.size())) return r;
    if (int r = dev_alloc(P, &P.coefr, coefr.size())) return r;
    HIP_TRY(hipMemcpy(P.tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P.Aimg, aimg.data(), sizeof(double2) * aimg.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P.Apsf, apsf.data(), sizeof(double2) * apsf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P.Tinv, tinv.data(), sizeof(double2) * tinv.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P.coefr, coefr.data(), sizeof(float) * coefr.size(), hipMemcpyHostToDevice));
    return SWD_OK;
}

template <int C>
void launch_transforms(DeconPlan &P, const float *img, const float *psf)
{
    constexpr KCfg c = kCfgs[C];
    const DGeo &G = P.G;
    k_decon_cells<kNW, c.Rc, c.Pc><<<G.ncx * G.ncz, kNW * 64, P.lds_cells, P.stream>>>(G, psf, P.tw, P.Apsf, P.S, P.T);
    (void)hipEventRecord(P.ev[2], P.stream);
    k_decon_windows<kNW, c.R, c.P, c.PQ><<<G.totalnum, kNW * 64, P.lds_windows, P.stream>>>(
        G, img, P.pp, P.tw, P.Aimg, P.Tinv, P.S, P.T, P.wout);
}

int run_dev(DeconPlan &P, const float *img, const float *psf, float *dec, float *coef)
{
    const DGeo &G = P.G;
    HIP_TRY(hipSetDevice(P.device));
    HIP_TRY(hipEventRecord(P.ev[0], P.stream));
    k_decon_sums<<<(G.totalnum + 255) / 256, 256, 0, P.stream>>>(G, img, P.pp);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(P.ev[1], P.stream));
    if (P.cfg == 0) launch_transforms<0>(P, img, psf);
    else launch_transforms<1>(P, img, psf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(P.ev[3], P.stream));
    k_decon_gather<<<(G.nx * G.nz + 255) / 256, 256, 0, P.stream>>>(G, P.wout, P.coefr, dec, coef);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(P.ev[4], P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    float ms[4];
    for (int k = 0; k < 4; k++) HIP_TRY(hipEventElapsedTime(&ms[k], P.ev[k], P.ev[k + 1]));
    P.st.sums_ms = ms[0];
    P.st.cells_ms = ms[1];
    P.st.windows_ms = ms[2];
    P.st.gather_ms = ms[3];
    P.st.total_ms = static_cast<double>(ms[0]) + ms[1] + ms[2] + ms[3];
    return SWD_OK;
}

}  // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" int swd_decon_taper(const swd_decon_params *p, float *coefr)
{
    if (!p || !coefr) return swd_fail(SWD_E_ARG, "swd_decon_taper: null argument");
    DGeo G;
    if (int r = decon_geometry(*p, G)) return r;
    const double pi = 3.1415926;   // headpsf.h PI
    for (int ix = 0; ix < G.outx; ix++)
        for (int iz = 0; iz < G.outz; iz++) {
            float cx, cz;
            if (ix <= p->snapwinx - 1) cx = static_cast<float>(0.5000 - 0.5 * std::cos(pi * ix / (p->snapwinx - 1)));
            else if (ix >= G.outx - p->snapwinx)
                cx = static_cast<float>(0.5000 + 0.5 * std::cos(pi * (ix - G.outx + p->snapwinx) / (p->snapwinx - 1)));
            else cx = 1.0f;
            if (iz <= p->snapwinz - 1) cz = static_cast<float>(0.5000 - 0.5 * std::cos(pi * iz / (p->snapwinz - 1)));
            else if (iz >= G.outz - p->snapwinz)
                cz = static_cast<float>(0.5000 + 0.5 * std::cos(pi * (iz - G.outz + p->snapwinz) / (p->snapwinz - 1)));
            else cz = 1.0f;
            coefr[ix * G.outz + iz] = cx * cz;
        }
    return SWD_OK;
}

extern "C" int swd_decon_geometry(const swd_decon_params *p, swd_decon_geom *g)
{
    if (!p || !g) return swd_fail(SWD_E_ARG, "swd_decon_geometry: null argument");
    DGeo G;
    if (int r = decon_geometry(*p, G)) return r;
    g->n = G.n;
    g->outwinx = G.outx;
    g->outwinz = G.outz;
    g->stepwinx = G.stepwinx;
    g->stepwinz = G.stepwinz;
    g->stepnumx = G.stepnumx;
    g->stepnumz = G.stepnumz;
    g->windows = G.totalnum;
    g->nxp = G.nxp;
    g->nzp = G.nzp;
    g->cells = G.ncx * G.ncz;
    return SWD_OK;
}

extern "C" int swd_decon_plan_create(const swd_decon_params *p, int device, void **plan)
{
    if (!p || !plan) return swd_fail(SWD_E_ARG, "swd_decon_plan_create: null argument");
    *plan = nullptr;
    DGeo G;
    if (int r = decon_geometry(*p, G)) return r;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return swd_fail(SWD_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return swd_fail(SWD_E_ARG, "device %d outside 0..%d", device, ndev - 1);
    DeconPlan *P = new DeconPlan;
    P->p = *p;
    P->G = G;
    P->device = device;
    P->cfg = cfg_fits(G, kCfgs[0]) ? 0 : 1;
    int r = SWD_OK;
    auto fail = [&](int code) {
        free_plan(P);
        return code;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(swd_fail(SWD_E_DEVICE, "hipSetDevice(%d) failed", device));
    if (hipStreamCreate(&P->stream) != hipSuccess) return fail(swd_fail(SWD_E_DEVICE, "hipStreamCreate failed"));
    for (hipEvent_t &e : P->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(swd_fail(SWD_E_DEVICE, "hipEventCreate failed"));
    if ((r = set_lds_limits(*P)) || (r = upload_tables(*P))) return fail(r);
    const size_t bins = (size_t)G.n * G.nh, cells = (size_t)G.ncx * G.ncz;
    if ((r = dev_alloc(*P, &P->pp, G.totalnum)) || (r = dev_alloc(*P, &P->S, cells * bins)) ||
        (r = dev_alloc(*P, &P->T, cells * bins)) ||
        (r = dev_alloc(*P, &P->wout, (size_t)G.totalnum * G.outx * G.outz)))
        return fail(r);
    P->st.windows = G.totalnum;
    P->st.cells = static_cast<int>(cells);
    P->st.n = G.n;
    *plan = P;
    return SWD_OK;
}

extern "C" int swd_decon_plan_destroy(void *plan)
{
    if (!plan) return swd_fail(SWD_E_ARG, "swd_decon_plan_destroy: null plan");
    free_plan(static_cast<DeconPlan *>(plan));
    return SWD_OK;
}

extern "C" int swd_decon_dev(void *plan, const float *image, const float *psf, float *dec, float *coef)
{
    if (!plan || !image || !psf || !dec || !coef) return swd_fail(SWD_E_ARG, "swd_decon_dev: null argument");
    return run_dev(*static_cast<DeconPlan *>(plan), image, psf, dec, coef);
}

extern "C" int swd_decon(void *plan, const float *image, const float *psf, float *dec, float *coef)
{
    if (!plan || !image || !psf || !dec) return swd_fail(SWD_E_ARG, "swd_decon: null argument");
    DeconPlan &P = *static_cast<DeconPlan *>(plan);
    HIP_TRY(hipSetDevice(P.device));
    const size_t np = (size_t)P.G.nx * P.G.nz;
    if (!P.io) {
        if (int r = dev_alloc(P, &P.io, 4 * np)) return r;
    }
    HIP_TRY(hipMemcpyAsync(P.io, image, sizeof(float) * np, hipMemcpyHostToDevice, P.stream));
    HIP_TRY(hipMemcpyAsync(P.io + np, psf, sizeof(float) * np, hipMemcpyHostToDevice, P.stream));
    if (int r = run_dev(P, P.io, P.io + np, P.io + 2 * np, P.io + 3 * np)) return r;
    HIP_TRY(hipMemcpyAsync(dec, P.io + 2 * np, sizeof(float) * np, hipMemcpyDeviceToHost, P.stream));
    if (coef) HIP_TRY(hipMemcpyAsync(coef, P.io + 3 * np, sizeof(float) * np, hipMemcpyDeviceToHost, P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    return SWD_OK;
}

extern "C" int swd_decon_get_stats(void *plan, swd_decon_stats *st)
{
    if (!plan || !st) return swd_fail(SWD_E_ARG, "swd_decon_get_stats: null argument");
    const DeconPlan &P = *static_cast<DeconPlan *>(plan);
    *st = P.st;
    st->ws_bytes = P.ws_bytes;
    return SWD_OK;
}
