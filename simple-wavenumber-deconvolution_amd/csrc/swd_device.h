// swd_device.h -- device helpers shared by the kernel translation units:
// frame (boundary strip) indexing, float4 global accesses through
// address_space(1) pointers, the XCD-aware block order and z windows of a
// column.  Included only by .hip files.
#pragma once
#include <hip/hip_runtime.h>

#include "swd_kernels.h"

namespace {

__device__ __forceinline__ int frame_index(const KGrid &g, int ix, int iz)
{
    const int jx = ix - g.pmlc, jz = iz - g.pmlc, Lc = g.Lc, nx = g.nx, nz = g.nz;
    if (jx < 0 || jx >= nx || jz < 0 || jz >= nz) return -1;
    if (jx < Lc) return jx * nz + jz;
    if (jx >= nx - Lc) return (jx - (nx - 2 * Lc)) * nz + jz;
    if (jz < Lc) return 2 * Lc * nz + (jx - Lc) * 2 * Lc + jz;
    if (jz >= nz - Lc) return 2 * Lc * nz + (jx - Lc) * 2 * Lc + Lc + (jz - (nz - Lc));
    return -1;
}

__device__ __forceinline__ bool interior(const KGrid &g, int ix, int iz)
{
    return ix >= g.pmlc && ix < g.ntx - g.pmlc && iz >= g.pmlc && iz < g.ntz - g.pmlc;
}

// Forward difference from the integer grid to the half grid: the derivative
// the velocity update uses (RK:55-58, RK:104-107).
template <int LC>
__device__ __forceinline__ float dplus(const float *__restrict__ f, size_t ip, size_t s, const float *rc, float inv)
{
    float d = 0.f;
#pragma unroll
    for (int ic = 0; ic < LC; ic++) d = __builtin_fmaf(rc[ic] * (f[ip + (ic + 1) * s] - f[ip - ic * s]), inv, d);
    return d;
}

// Backward difference from the half grid to the integer grid: the derivative
// the pressure update uses (RK:162-166).
template <int LC>
__device__ __forceinline__ float dminus(const float *__restrict__ f, size_t ip, size_t s, const float *rc, float inv)
{
    float d = 0.f;
#pragma unroll
    for (int ic = 0; ic < LC; ic++) d = __builtin_fmaf(rc[ic] * (f[ip + ic * s] - f[ip - (ic + 1) * s]), inv, d);
    return d;
}

// Angle-gather pieces of one cell (the reference's sum_image_GPU_kernel_imageCIG,
// RK:859-901).  cig_products: F = S.R and the squared norms a = |S|^2,
// b = |R|^2 of the source vector S = (-p_s vx_s, -p_s vz_s) and the receiver
// vector R = (p_r vx_r, p_r vz_r) (+p v: the receiver field is stepped with
// (-dx, -dz)).
__device__ __forceinline__ void cig_products(float ps, float vxs, float vzs, float pr, float vxr, float vzr, float &F,
                                             float &a, float &b)
{
    const float fx = -ps * vxs, fz = -ps * vzs;
    const float fxr = pr * vxr, fzr = pr * vzr;
    F = __builtin_fmaf(fx, fxr, fz * fzr);
    a = __builtin_fmaf(fx, fx, fz * fz);
    b = __builtin_fmaf(fxr, fxr, fzr * fzr);
}

// G = |S||R| with sqrt in double rounded to float.
__device__ __forceinline__ float cig_norm(float a, float b)
{
    return static_cast<float>(sqrt(static_cast<double>(a))) * static_cast<float>(sqrt(static_cast<double>(b)));
}

// A cheap test that the angle of F/G cannot reach a bin: thr is a lower
// bound, with margin, of the cosine the last bin needs (-2 when every angle
// can bin).  Rejected cells bin nothing either way, so results do not change.
__device__ __forceinline__ bool cig_may_bin(float F, float a, float b, float thr)
{
    const float t = F / (__builtin_sqrtf(a) * __builtin_sqrtf(b));
    return t >= thr;
}

// Binning of summed products: the half-angle in degrees
// 0.5 acos(pre/pst) 180 / 3.1415926 (acos in double rounded to float), and
// the bins ceil(angle + ii - 0.5), ii = -2..2, inside [2, thetan), each taking
// fma(pp, w_|ii|, bin) with w = 1, w1, w2 (exp(-ii^2/18) rounded to float).
// An undefined angle adds nothing.
__device__ __forceinline__ void cig_bins(float pre, float pst, float pp, int thetan, float w1, float w2, float thr,
                                         float *bins)
{
    const float tht = pre / pst;
    if (!(tht >= thr)) return;   // no bin in reach (or undefined)
    const float ac = static_cast<float>(acos(static_cast<double>(tht)));
    const float ftheta = static_cast<float>(0.5 * static_cast<double>(ac) * 180.0 / 3.1415926);
    if (ftheta != ftheta) return;
#pragma unroll
    for (int ii = -2; ii <= 2; ii++) {
        const int itheta =
            static_cast<int>(ceilf(static_cast<float>(static_cast<double>(ftheta + static_cast<float>(ii)) - 0.5)));
        const float w = ii == 0 ? 1.f : (ii == 1 || ii == -1) ? w1 : w2;
        if (itheta >= 2 && itheta < thetan) bins[itheta] = __builtin_fmaf(pp, w, bins[itheta]);
    }
}

typedef __attribute__((address_space(1))) float gfl;
typedef __attribute__((address_space(1))) const float gcfl;

__device__ __forceinline__ int xcd_block()
{
    const int b = blockIdx.x, per = gridDim.x >> 3;
    return (b & 7) * per + (b >> 3);
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const v4f gcv4;
typedef __attribute__((address_space(1))) v4f gv4;
// 32-bit element offsets (arrays below 4 GiB): the byte offset is formed in
// 32 bits so the access can use a scalar base with a 32-bit vector offset.
typedef __attribute__((address_space(1))) const char gcch;
typedef __attribute__((address_space(1))) char gch;
__device__ __forceinline__ float4 ld4(gcfl *f, unsigned o)
{
    const v4f v = *(gcv4 *)((gcch *)f + (o << 2));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(gfl *f, unsigned o, float4 v)
{
    v4f w = {v.x, v.y, v.z, v.w};
    *(gv4 *)((gch *)f + (o << 2)) = w;
}
__device__ __forceinline__ float ldg(const float *f, unsigned i) { return *(gcfl *)((gcch *)f + (i << 2)); }
__device__ __forceinline__ int ldgi(const int *f, unsigned i)
{
    return *(__attribute__((address_space(1))) const int *)((gcch *)f + (i << 2));
}
__device__ __forceinline__ float4 ld4(gcfl *f, size_t o)
{
    const v4f v = *(gcv4 *)(f + o);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4(gfl *f, size_t o, float4 v)
{
    v4f w = {v.x, v.y, v.z, v.w};
    *(gv4 *)(f + o) = w;
}
__device__ __forceinline__ float &el(float4 &v, int k) { return reinterpret_cast<float *>(&v)[k]; }
__device__ __forceinline__ float elc(const float4 &v, int k) { return reinterpret_cast<const float *>(&v)[k]; }
// v[k] += w for a run-time k in 0..3 (a select per element keeps v in registers)
__device__ __forceinline__ void add_at(float4 &v, int k, float w)
{
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j == k) el(v, j) += w;
}

// rows z-4R .. z+4R+3 of one column (R quads of halo on each side), zero outside
template <int R>
__device__ __forceinline__ void column_window(gcfl *f, size_t o, int z, int ldz, float (&a)[4 * (2 * R + 1)])
{
#pragma unroll
    for (int q = -R; q <= R; q++) {
        const int zq = z + 4 * q;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (zq >= 0 && zq < ldz) v = ld4(f, o + 4 * q);
#pragma unroll
        for (int k = 0; k < 4; k++) a[4 * (q + R) + k] = elc(v, k);
    }
}

}  // namespace

#define DISPATCH_LC(LC, CALL)                          \
    switch (LC) {                                      \
    case 1: { constexpr int L = 1; CALL; } break;      \
    case 2: { constexpr int L = 2; CALL; } break;      \
    case 3: { constexpr int L = 3; CALL; } break;      \
    case 4: { constexpr int L = 4; CALL; } break;      \
    case 5: { constexpr int L = 5; CALL; } break;      \
    case 6: { constexpr int L = 6; CALL; } break;      \
    case 7: { constexpr int L = 7; CALL; } break;      \
    case 8: { constexpr int L = 8; CALL; } break;      \
    default: return hipErrorInvalidValue;              \
    }

