// swd_decon.hip -- wavenumber-domain PSF deconvolution on gfx950: the
// reference's windowed Wiener deconvolution of a migrated image by the nearest
// PSF of a point-scatterer lattice (kkdecon_acoustic/kk_decon_interp_psf_minewindow.cpp,
// DC:543-731), one process, every window on one device.
//
// Per window of side n (fullw = filterwin + psfwin - 1) the reference runs two
// forward 2D FFTs, a Wiener division and one inverse 2D FFT in double, then
// overlap-adds a tapered outwin x outwin block in window order.  Here:
//   k_decon_cells    one workgroup per PSF lattice cell (windows share cells,
//                    2484 windows use 374 cells at the reference size): the
//                    normalised PSF's half spectrum and the Wiener denominators
//                    tmpfloat = float(|S|^2 + eps * max|S|^2) (DC:583-648);
//   k_decon_windows  one workgroup per window: the float sum of magnitudes of
//                    the window (pp_max, DC:555-574, a serial chain in the
//                    reference's order on one wavefront), half spectrum of the
//                    image window on the others, Wiener division, inverse only
//                    at the output block (DC:653-673);
//   k_decon_gather   one thread per output cell: the tapered overlap-add of
//                    the windows covering it, in window order (DC:653-673).
// The transforms are direct DFTs as small complex matrix products: n = 101 is
// prime, the operands are LDS-resident, and the window-independent operand
// (twiddles, constant over a launch) is wave-uniform and read through the
// scalar cache, so each lane owns one z wavenumber k1 and the fp64 FMAs take
// one vector operand from LDS and one scalar operand.  All arithmetic is in the
// fixed order of the host oracle (fma chains in index order, the Hermitian half
// spectrum along z, conjugate wavenumber pairs along x, the image
// normalisation folded into the exact divisor pp_max * tmpfloat), so results
// agree with it bit for bit.
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
// half spectrum k1 < nh; lane = k1, wave w owns rows w + NW rr (rr < R).  X
// must be readable for NW R rows; rows past mx are computed from whatever the
// LDS holds there and never stored.  X is read as an LDS broadcast, the
// twiddle per lane from the LDS table.
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

// Columns pass over pairs of x wavenumbers (p, n - p), whose twiddles are
// conjugate: with (c, s) = (cos, sin)(2 pi p (r + offx) / n),
//   P = sum_r c Y[r][k1],  Q = sum_r s Y[r][k1],
//   M[p] = P - i Q,  M[n - p] = P + i Q,
// half the FMAs of the two transforms.  Wave w owns the PP pairs p = w PP + j;
// slot 2j holds k0 = p, slot 2j + 1 holds k0 = n - p (see slot_k0).
// A[r][p] (rows of NW PP entries, zero-padded) is the launch-constant twiddle
// table, read wave-uniformly through the scalar cache; the waves a SIMD holds
// cover each other's scalar-load latency.
__device__ __forceinline__ int slot_k0(int n, int nh, int p, int b)
{
    if (p >= nh) return -1;
    if (b == 0) return p;
    return (p == 0 || 2 * p == n) ? -1 : n - p;
}

template <int NW, int PP>
__device__ __forceinline__ void cols_pass(const double2 *Y, int mx, int nh, const double2 *__restrict__ A,
                                          double (&mr)[2 * PP], double (&mi)[2 * PP])
{
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int k1 = min(lane, nh - 1);
    constexpr int LD = NW * PP;
    const double2 *Aw = A + w * PP;
    double pr[PP], pi[PP], qr[PP], qi[PP];
#pragma unroll
    for (int j = 0; j < PP; j++) pr[j] = pi[j] = qr[j] = qi[j] = 0.0;
    for (int r = 0; r < mx; r++) {
        const double2 y = Y[r * nh + k1];
#pragma unroll
        for (int j = 0; j < PP; j++) {
            const double2 a = Aw[r * LD + j];
            pr[j] = fma(y.x, a.x, pr[j]);
            pi[j] = fma(y.y, a.x, pi[j]);
            qr[j] = fma(y.x, a.y, qr[j]);
            qi[j] = fma(y.y, a.y, qi[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < PP; j++) {
        mr[2 * j] = pr[j] + qi[j];
        mi[2 * j] = pi[j] - qr[j];
        mr[2 * j + 1] = pr[j] - qi[j];
        mi[2 * j + 1] = pi[j] + qr[j];
    }
}

// Inverse along x at the output rows: Q[i][k1] += sum_k0 e^{+2 pi i k0 (r0x + i) / n} D[k0][k1]
// over the rows k0 = k0base + kl (kl < rows) of one half of D, for the PQ
// output rows i = w PQ + ii of wave w, U wavenumbers a step; Tq[k0][i] (rows
// of NW PQ entries) is read through the scalar cache.  rows is a multiple of
// U; D and Tq are zero past row n.
template <int NW, int PQ>
__device__ __forceinline__ void inv_rows_pass(const double2 *D, int rows, int k0base, int nh,
                                              const double2 *__restrict__ Tq, double (&qr)[PQ], double (&qi)[PQ])
{
    constexpr int U = PQ >= 8 ? 1 : 8 / PQ;
    constexpr int LD = NW * PQ;
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int k1 = min(lane, nh - 1);
    const double2 *Tw = Tq + (size_t)k0base * LD + w * PQ;
    for (int kl = 0; kl < rows; kl += U) {
#pragma unroll
        for (int u = 0; u < U; u++) {   // rows past n of D and Tq are zero: they add nothing
            const double2 d = D[(kl + u) * nh + k1];
#pragma unroll
            for (int ii = 0; ii < PQ; ii++) {
                const double2 e = Tw[(kl + u) * LD + ii];
                qr[ii] = fma(d.x, e.x, qr[ii]);
                qr[ii] = fma(-d.y, e.y, qr[ii]);
                qi[ii] = fma(d.y, e.x, qi[ii]);
                qi[ii] = fma(d.x, e.y, qi[ii]);
            }
        }
    }
}

// D is held in LDS in two halves of k0: the first ceil(n/2) wavenumbers, then
// the rest, each padded with zero rows to a whole number of inv_rows_pass steps.
__host__ __device__ constexpr int inv_step(int PQ) { return PQ >= 8 ? 1 : 8 / PQ; }
__host__ __device__ constexpr int half_rows(int n) { return (n + 1) / 2; }
__host__ __device__ constexpr int half_rows_pad(int n, int PQ)
{
    return (half_rows(n) + inv_step(PQ) - 1) / inv_step(PQ) * inv_step(PQ);
}
// Rows of the Tq table: the second half's steps start at k0 = half_rows(n).
__host__ __device__ constexpr int tq_rows(int n, int PQ) { return half_rows(n) + half_rows_pad(n, PQ); }

// One lane sums |a[0..count)| (float, in index order) from LDS; a is padded
// with zeros to a multiple of 32 entries (adding +0 changes no sum).  The
// reads of the next 32 entries are issued before the adds of the current ones.
__host__ __device__ constexpr int sum_pad(int count) { return (count + 31) / 32 * 32; }

__device__ __forceinline__ float serial_abs_sum(const float *a, int count)
{
    const float4 *a4 = reinterpret_cast<const float4 *>(a);
    const int nb = sum_pad(count) / 32;
    float acc = 0.0f;
    float4 cur[8];
#pragma unroll
    for (int j = 0; j < 8; j++) cur[j] = a4[j];
    for (int b = 0; b < nb; b++) {
        const int bn = min(b + 1, nb - 1);
        float4 nxt[8];
#pragma unroll
        for (int j = 0; j < 8; j++) nxt[j] = a4[bn * 8 + j];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            acc = acc + fabsf(cur[j].x);
            acc = acc + fabsf(cur[j].y);
            acc = acc + fabsf(cur[j].z);
            acc = acc + fabsf(cur[j].w);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) cur[j] = nxt[j];
    }
    return acc;
}

// ---------------------------------------------------------------- kernels
template <int NW, int R, int PP>
__global__ __launch_bounds__(NW * 64) void k_decon_cells(DGeo G, const float *__restrict__ psf,
                                                         const double2 *__restrict__ tw_g,
                                                         const double2 *__restrict__ Apsf, double2 *__restrict__ S,
                                                         float *__restrict__ T)
{
    constexpr int NT = NW * 64;
    extern __shared__ __align__(16) unsigned char lds[];
    const int n = G.n, nh = G.nh, mx = G.mxp, mz = G.mzp;
    double2 *tw = reinterpret_cast<double2 *>(lds);
    double2 *Y = tw + n;                                                   // mx rows
    double *X = reinterpret_cast<double *>(Y + mx * nh);                   // NW R rows
    float *raw = reinterpret_cast<float *>(X + NW * R * mz);               // the patch, padded to 32
    float *scal = raw + sum_pad(mx * mz);                                  // [0] psf_max, [1] max |S|^2 (bits)
    const int cell = blockIdx.x;
    const int cx = 3 + cell / G.ncz, cz = 3 + cell % G.ncz;
    const int x0 = cx * G.psfwinx, z0 = cz * G.psfwinz;   // ixp - halfpsfwx, izp - halfpsfwz (DC:597-612)
    for (int t = threadIdx.x; t < n; t += NT) tw[t] = tw_g[t];
    for (int e = threadIdx.x; e < sum_pad(mx * mz); e += NT)
        raw[e] = e < mx * mz ? expanded(psf, G, x0 + e / mz, z0 + e % mz) : 0.0f;
    __syncthreads();
    if (threadIdx.x == 0) {
        scal[0] = serial_abs_sum(raw, mx * mz);   // psf_max (DC:617-619), in the reference's order
        reinterpret_cast<unsigned *>(scal)[1] = 0u;
    }
    __syncthreads();
    const double pm = static_cast<double>(scal[0]);
    for (int e = threadIdx.x; e < NW * R * mz; e += NT)
        X[e] = e < mx * mz ? static_cast<double>(raw[e]) / pm : 0.0;
    __syncthreads();
    rows_pass<NW, R>(X, mx, mz, 0, n, nh, tw, Y);
    __syncthreads();
    double mr[2 * PP], mi[2 * PP];
    cols_pass<NW, PP>(Y, mx, nh, Apsf, mr, mi);
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    // fft_max: max over the bins of float(|S|^2) (DC:625-633); non-negative
    // floats order as their bit patterns
    float qmax = 0.0f;
#pragma unroll
    for (int q = 0; q < 2 * PP; q++)
        if (slot_k0(n, nh, w * PP + q / 2, q & 1) >= 0 && lane < nh)
            qmax = fmaxf(qmax, static_cast<float>(mr[q] * mr[q] + mi[q] * mi[q]));
    atomicMax(reinterpret_cast<unsigned *>(scal) + 1, __float_as_uint(qmax));
    __syncthreads();
    float fft_max = __uint_as_float(reinterpret_cast<unsigned *>(scal)[1]) * G.epsilon;
    if (fft_max == 0.0f) fft_max = 1.0e-15f;
    if (lane < nh) {
        double2 *Sc = S + (size_t)cell * n * nh;
        float *Tc = T + (size_t)cell * n * nh;
#pragma unroll
        for (int q = 0; q < 2 * PP; q++) {
            const int k0 = slot_k0(n, nh, w * PP + q / 2, q & 1);
            if (k0 >= 0) {
                Sc[k0 * nh + lane] = make_double2(mr[q], mi[q]);
                Tc[k0 * nh + lane] = static_cast<float>((mr[q] * mr[q] + mi[q] * mi[q]) + static_cast<double>(fft_max));
            }
        }
    }
}

// LDS of a window workgroup: the twiddle table and pp_max, then one region
// that holds the image window X (floats) followed by Y, later one half of D,
// later Q.  Rows of X past mx are read (into Y's space) but their results are
// dropped.
__host__ __device__ constexpr size_t win_x_bytes(int mx, int mz) { return ((size_t)mx * mz * 4 + 15) / 16 * 16; }
__host__ __device__ constexpr size_t win_head_bytes(int n) { return sizeof(double2) * n + 16; }

// Adds |X[i][j]| for rows i0 <= i < i1, in row-major order, to acc (one lane).
__device__ __forceinline__ float window_abs_sum(const float *X, int i0, int i1, int mz, float acc)
{
    for (int i = i0; i < i1; i++) {
        const float *r = X + i * mz;
        int j = 0;
        for (; j + 16 <= mz; j += 16) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = r[j + u];
#pragma unroll
            for (int u = 0; u < 16; u++) acc = acc + fabsf(v[u]);
        }
        for (; j < mz; j++) acc = acc + fabsf(r[j]);
    }
    return acc;
}

// NW wavefronts: NW - 1 run the transforms; the last one forms pp_max, the
// float sum of |window| (DC:555-574) as a serial chain in the reference's
// order (rows outer, columns inner), in two halves that run beside the rows
// and the columns passes.  float(double(acc) + |v|) of DC:568 is the float sum
// acc + |v| (binary64 holds more than twice binary32's precision, so the
// double rounding is exact).
template <int NW, int R, int PP, int PQ, int WPE>
__global__ __launch_bounds__(NW * 64, WPE) void k_decon_windows(DGeo G, const float *__restrict__ img,
                                                              const double2 *__restrict__ tw_g,
                                                              const double2 *__restrict__ Aimg,
                                                              const double2 *__restrict__ Tinv,
                                                              const double2 *__restrict__ S,
                                                              const float *__restrict__ T, double *__restrict__ wout)
{
    constexpr int NT = NW * 64, NWC = NW - 1;
    extern __shared__ __align__(16) unsigned char lds[];
    const int n = G.n, nh = G.nh, mx = G.mxi, mz = G.mzi;
    double2 *tw = reinterpret_cast<double2 *>(lds);
    float *ppl = reinterpret_cast<float *>(tw + n);
    unsigned char *reg = lds + win_head_bytes(n);
    float *X = reinterpret_cast<float *>(reg);
    double2 *Y = reinterpret_cast<double2 *>(reg + win_x_bytes(mx, mz));
    double2 *Rg = reinterpret_cast<double2 *>(reg);                  // D halves, then Q
    const int s = blockIdx.x;
    const int lane = threadIdx.x & 63, w = uniform(threadIdx.x >> 6);
    const int x0 = G.stepwinx * (s / G.stepnumz) - G.halfltwx, z0 = G.stepwinz * (s % G.stepnumz) - G.halfltwz;
    for (int t = threadIdx.x; t < n; t += NT) tw[t] = tw_g[t];
    for (int e = threadIdx.x; e < mx * mz; e += NT) X[e] = expanded(img, G, x0 + e / mz, z0 + e % mz);
    __syncthreads();
    // The chain issues one add at a time: at the highest wave priority its
    // instructions go first on its SIMD instead of waiting their turn behind
    // the transform waves.
    float acc = 0.0f;   // pp_max, on lane 0 of wave NWC
    if (w < NWC) rows_pass<NWC, R>(X, mx, mz, G.offz, n, nh, tw, Y);
    else if (lane == 0) {
        __builtin_amdgcn_s_setprio(3);
        acc = window_abs_sum(X, 0, mx / 2, mz, acc);
    }
    __syncthreads();
    double mr[2 * PP], mi[2 * PP];
    if (w < NWC) cols_pass<NWC, PP>(Y, mx, nh, Aimg, mr, mi);
    else if (lane == 0) {
        acc = window_abs_sum(X, mx / 2, mx, mz, acc);
        if (static_cast<double>(acc) <= 1.0e-15) acc = 1.0e-10f;
        *ppl = acc;
        __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    // Wiener division (DC:642-651) with the image normalisation in the divisor
    const int k1 = min(lane, nh - 1);
    if (w < NWC) {
        const int cell = uniform(cell_of(G, s));
        const double ppd = static_cast<double>(*ppl);
        const double2 *Sc = S + (size_t)cell * n * nh;
        const float *Tc = T + (size_t)cell * n * nh;
#pragma unroll
        for (int q = 0; q < 2 * PP; q++) {
            const int k0 = slot_k0(n, nh, w * PP + q / 2, q & 1);
            if (k0 >= 0) {
                const double2 sv = Sc[k0 * nh + k1];
                const double den = ppd * static_cast<double>(Tc[k0 * nh + k1]);
                const double dr = (mr[q] * sv.x + mi[q] * sv.y) / den;
                const double di = (-mr[q] * sv.y + mi[q] * sv.x) / den;
                mr[q] = dr;
                mi[q] = di;
            }
        }
    }
    // inverse along x at the output rows, over the two halves of D in k0 order
    const int outx = G.outx, outz = G.outz;
    const int H = half_rows(n), HP = half_rows_pad(n, PQ);
    double qr[PQ], qi[PQ];
#pragma unroll
    for (int ii = 0; ii < PQ; ii++) qr[ii] = qi[ii] = 0.0;
    for (int h = 0; h < 2; h++) {
        const int lo = h * H, cnt = h ? n - H : H;
        __syncthreads();   // every wave is past its reads of Y / of the previous half
        if (w < NWC && lane < nh) {
#pragma unroll
            for (int q = 0; q < 2 * PP; q++) {
                const int k0 = slot_k0(n, nh, w * PP + q / 2, q & 1);
                if (k0 >= lo && k0 < lo + cnt) Rg[(k0 - lo) * nh + lane] = make_double2(mr[q], mi[q]);
            }
        }
        for (int e = cnt * nh + threadIdx.x; e < HP * nh; e += NT) Rg[e] = make_double2(0.0, 0.0);
        __syncthreads();
        if (w < NWC) inv_rows_pass<NWC, PQ>(Rg, HP, lo, nh, Tinv, qr, qi);
    }
    __syncthreads();   // every wave is past its reads of D
    if (w < NWC && lane < nh) {
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
        double acc2 = 0.0;
        for (int kk = 0; kk < nh; kk++) {
            const double wgt = (kk == 0 || 2 * kk == n) ? 1.0 : 2.0;
            const double2 q = Rg[i * nh + kk];
            const double2 e = tw[t];
            acc2 = fma(wgt * q.x, e.x, acc2);
            acc2 = fma(-(wgt * q.y), e.y, acc2);
            t += step;
            if (t >= n) t -= n;
        }
        ws[o] = acc2;
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
// The window kernel's transforms run on kNW - 1 wavefronts (the last forms
// pp_max); the cell kernel's on all kNW.
struct KCfg {
    int R, PP, PQ;   // window kernel: image rows, wavenumber pairs, output rows per wavefront
    int Rc;          // cell kernel: PSF rows per wavefront
    int wpe;         // wavefronts per SIMD the window kernel is compiled for (2 workgroups a CU when 8)
};
constexpr int kNW = 16;
constexpr int kNWC = kNW - 1;
constexpr int kPPc = 4;   // cell kernel: wavenumber pairs per wavefront, kNW kPPc >= n/2 + 1 for n <= kMaxN
constexpr KCfg kCfgs[2] = {{5, 4, 2, 2, 8}, {9, 5, 9, 8, 4}};

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
    hipEvent_t ev[4] = {};
    // launch-constant tables
    double2 *tw = nullptr, *Aimg = nullptr, *Apsf = nullptr, *Tinv = nullptr;
    float *coefr = nullptr;
    // workspace
    float *T = nullptr;
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
    for (void *q : {(void *)P->tw, (void *)P->Aimg, (void *)P->Apsf, (void *)P->Tinv, (void *)P->coefr,
                    (void *)P->T, (void *)P->S, (void *)P->wout, (void *)P->io})
        if (q) (void)hipFree(q);
    for (hipEvent_t e : P->ev)
        if (e) (void)hipEventDestroy(e);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
}

bool cfg_fits(const DGeo &G, const KCfg &c)
{
    return G.mxi <= kNWC * c.R && G.outx <= kNWC * c.PQ && G.nh <= kNWC * c.PP && G.mxp <= kNW * c.Rc &&
           G.nh <= kNW * kPPc;
}

template <int C>
int set_lds_attrs(const DeconPlan &P)
{
    constexpr KCfg c = kCfgs[C];
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_decon_cells<kNW, c.Rc, kPPc>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(P.lds_cells)));
    HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_decon_windows<kNW, c.R, c.PP, c.PQ, c.wpe>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(P.lds_windows)));
    return SWD_OK;
}

int set_lds_limits(DeconPlan &P)
{
    const DGeo &G = P.G;
    const KCfg &c = kCfgs[P.cfg];
    P.lds_cells = sizeof(double2) * (G.n + G.mxp * G.nh) + sizeof(double) * kNW * c.Rc * G.mzp +
                  sizeof(float) * (sum_pad(G.mxp * G.mzp) + 2);
    const size_t region = std::max({win_x_bytes(G.mxi, G.mzi) + sizeof(double2) * G.mxi * G.nh,
                                    sizeof(float) * kNWC * c.R * G.mzi,
                                    sizeof(double2) * half_rows_pad(G.n, c.PQ) * G.nh,
                                    sizeof(double2) * G.outx * G.nh});
    P.lds_windows = win_head_bytes(G.n) + region;
    int max_lds = 0;
    HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, P.device));
    if (P.lds_windows > static_cast<size_t>(max_lds) || P.lds_cells > static_cast<size_t>(max_lds))
        return swd_fail(SWD_E_RANGE, "deconvolution window needs %zu bytes of LDS, the device has %d", P.lds_windows,
                        max_lds);
    return P.cfg == 0 ? set_lds_attrs<0>(P) : set_lds_attrs<1>(P);
}

// Twiddle tables, zero-padded to the launch configuration: Aimg [mxi][kNWC PP]
// and Apsf [mxp][kNW kPPc] over wavenumber pairs, Tinv [tq_rows][kNWC PQ].
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
    const auto aimg = table(G.mxi, G.mxi, kNWC * c.PP, G.nh,
                            [&](int r, int p) { return (static_cast<long>(p) * (r + G.offx)) % n; });
    const auto apsf = table(G.mxp, G.mxp, kNW * kPPc, G.nh, [&](int r, int p) { return (static_cast<long>(p) * r) % n; });
    const int qrows = tq_rows(n, c.PQ);
    const auto tinv = table(qrows, n, kNWC * c.PQ, G.outx,
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
void launch_cells(DeconPlan &P, const float *psf)
{
    constexpr KCfg c = kCfgs[C];
    k_decon_cells<kNW, c.Rc, kPPc><<<P.G.ncx * P.G.ncz, kNW * 64, P.lds_cells, P.stream>>>(P.G, psf, P.tw, P.Apsf, P.S,
                                                                                        P.T);
}

template <int C>
void launch_windows(DeconPlan &P, const float *img)
{
    constexpr KCfg c = kCfgs[C];
    k_decon_windows<kNW, c.R, c.PP, c.PQ, c.wpe><<<P.G.totalnum, kNW * 64, P.lds_windows, P.stream>>>(
        P.G, img, P.tw, P.Aimg, P.Tinv, P.S, P.T, P.wout);
}

int run_dev(DeconPlan &P, const float *img, const float *psf, float *dec, float *coef)
{
    const DGeo &G = P.G;
    HIP_TRY(hipSetDevice(P.device));
    HIP_TRY(hipEventRecord(P.ev[0], P.stream));
    if (P.cfg == 0) launch_cells<0>(P, psf);
    else launch_cells<1>(P, psf);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(P.ev[1], P.stream));
    if (P.cfg == 0) launch_windows<0>(P, img);
    else launch_windows<1>(P, img);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(P.ev[2], P.stream));
    k_decon_gather<<<(G.nx * G.nz + 255) / 256, 256, 0, P.stream>>>(G, P.wout, P.coefr, dec, coef);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(P.ev[3], P.stream));
    HIP_TRY(hipStreamSynchronize(P.stream));
    float ms[3];
    for (int k = 0; k < 3; k++) HIP_TRY(hipEventElapsedTime(&ms[k], P.ev[k], P.ev[k + 1]));
    P.st.cells_ms = ms[0];
    P.st.windows_ms = ms[1];
    P.st.gather_ms = ms[2];
    P.st.total_ms = static_cast<double>(ms[0]) + ms[1] + ms[2];
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
    if ((r = dev_alloc(*P, &P->S, cells * bins)) ||
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
