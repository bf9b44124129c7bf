// swd_adjoint.hip -- the exact adjoint step of the Born operator in two
// launches instead of four.
//
// The four per-cell kernels of swd_kernels.hip (k_adj_p1, k_adj_v1, k_adj_v2,
// k_adj_p2) pass their intermediates ex, ez through device memory.  Here a
// workgroup owns a tile of 64 rows x 16 columns and recomputes the pointwise
// halves on the tile and its stencil halo into LDS:
//   k_adj_pv: A_p^T -- ex, ez of the p domain (with the CPML memories of the
//             pressure update), then vx -= D+ ex, vz -= D+ ez on the tile;
//   k_adj_vp: A_v^T -- ex, ez of the velocity domains (with the CPML memories of
//             the velocity update), then p -= D- e, the data injection and the
//             gradient sum.
// Per cell the arithmetic and its order are those of the four kernels, so the
// results are the same bit for bit.  A halo cell's CPML memory is read by its
// neighbours while its owner updates it, so each memory is double-buffered by
// step parity: step it reads slot (it+1)&1 and writes slot it&1 (slot 0 the
// instance's pvx/pvz/psx/psz, slot 1 its ex/ez/e2/e3 arrays).  A cell whose
// damping coefficient is zero never reads or writes its memory, and every
// other cell of the domain writes it every step, so no stale value is read.
#include "swd_device.h"
#include "swd_kernels.h"
#include "swd_launch.h"

namespace {

constexpr int kTX = 16, kNZ = 64;

// The workgroup's tile: instances fastest, then z tiles, then x tiles, each
// XCD taking a contiguous eighth of that order (workgroups are dealt to the
// eight XCDs in turn), so a tile's halo neighbours and the model tile the
// instances share are mostly read through the same L2.  False for the padding
// workgroups of the last eighth.
__device__ __forceinline__ bool adj_tile(const KGrid &g, int ninst, int &inst, int &X0, int &Z0)
{
    const int nzb = (g.ntz + kNZ - 1) / kNZ, nxb = (g.ntx + kTX - 1) / kTX;
    const int L = xcd_block();
    if (L >= ninst * nzb * nxb) return false;
    inst = L % ninst;
    Z0 = (L / ninst) % nzb * kNZ;
    X0 = L / (ninst * nzb) * kTX;
    return true;
}

template <int LC>
__global__ __launch_bounds__(256) void k_adj_pv(KGrid g, const KInst *__restrict__ insts, int ninst, int it)
{
    constexpr int XL = LC - 1, NXE = kTX + 2 * LC - 1;   // ex: columns X0-XL .. X0+kTX+LC-1
    constexpr int ZL = LC - 1, NZE = kNZ + 2 * LC - 1;   // ez: rows Z0-ZL .. Z0+kNZ+LC-1
    __shared__ float sEX[NXE * kNZ];
    __shared__ float sEZ[kTX * NZE];
    int inst, X0, Z0;
    if (!adj_tile(g, ninst, inst, X0, Z0)) return;
    const KInst &I = insts[inst];
    const int ntx = g.ntx, ntz = g.ntz;
    const size_t ldz = static_cast<size_t>(g.ldz);
    const bool odd = it & 1;
    const float *pxo = odd ? I.pvx : I.ex, *pzo = odd ? I.pvz : I.ez;
    float *pxn = odd ? I.ex : I.pvx, *pzn = odd ? I.ez : I.pvz;
    // ex, ez = K p (plus the transposed CPML terms) on [LC, ntx-LC] x [LC, ntz-LC], else 0
    for (int q = threadIdx.x; q < NXE * kNZ; q += 256) {
        const int c = q / kNZ, r = q - c * kNZ;
        const int ix = X0 - XL + c, iz = Z0 + r;
        float ex = 0.f;
        if (ix >= LC && ix <= ntx - LC && iz >= LC && iz <= ntz - LC) {
            const size_t ip = static_cast<size_t>(ix) * ldz + iz;
            const float t = g.K[ip] * I.p[ip];
            ex = t;
            const float ax = g.ax[ix];
            if (ax != 0.f) {
                const float u = t + pxo[ip];
                ex = __builtin_fmaf(ax, u, t);
                if (c >= XL && c < XL + kTX) pxn[ip] = g.bx[ix] * u;
            }
        }
        sEX[q] = ex;
    }
    for (int q = threadIdx.x; q < kTX * NZE; q += 256) {
        const int c = q / NZE, r = q - c * NZE;
        const int ix = X0 + c, iz = Z0 - ZL + r;
        float ez = 0.f;
        if (ix >= LC && ix <= ntx - LC && iz >= LC && iz <= ntz - LC) {
            const size_t ip = static_cast<size_t>(ix) * ldz + iz;
            const float t = g.K[ip] * I.p[ip];
            ez = t;
            const float az = g.az[iz];
            if (az != 0.f) {
                const float u = t + pzo[ip];
                ez = __builtin_fmaf(az, u, t);
                if (r >= ZL && r < ZL + kNZ) pzn[ip] = g.bz[iz] * u;
            }
        }
        sEZ[q] = ez;
    }
    __syncthreads();
    // vhat -= D+ e (transpose of the backward difference)
    for (int q = threadIdx.x; q < kTX * kNZ; q += 256) {
        const int c = q / kNZ, r = q - c * kNZ;
        const int ix = X0 + c, iz = Z0 + r;
        if (ix >= ntx || iz >= ntz) continue;
        float gx = 0.f, gz = 0.f;
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float e1 = sEX[(c + XL + ic + 1) * kNZ + r], e0 = sEX[(c + XL - ic) * kNZ + r];
            const float f1 = sEZ[c * NZE + r + ZL + ic + 1], f0 = sEZ[c * NZE + r + ZL - ic];
            gx = __builtin_fmaf(g.rc[ic] * (e1 - e0), g.idx, gx);
            gz = __builtin_fmaf(g.rc[ic] * (f1 - f0), g.idz, gz);
        }
        const size_t ip = static_cast<size_t>(ix) * ldz + iz;
        I.vx[ip] -= gx;
        I.vz[ip] -= gz;
    }
}

template <int LC>
__global__ __launch_bounds__(256) void k_adj_vp(KGrid g, const KInst *__restrict__ insts, int ninst, int it, int first)
{
    constexpr int XL = LC, NXE = kTX + 2 * LC - 1;   // ex: columns X0-LC .. X0+kTX+LC-2
    constexpr int ZL = LC, NZE = kNZ + 2 * LC - 1;   // ez: rows Z0-LC .. Z0+kNZ+LC-2
    __shared__ float sEX[NXE * kNZ];
    __shared__ float sEZ[kTX * NZE];
    int inst, X0, Z0;
    if (!adj_tile(g, ninst, inst, X0, Z0)) return;
    const KInst &I = insts[inst];
    const int ntx = g.ntx, ntz = g.ntz;
    const size_t ldz = static_cast<size_t>(g.ldz);
    if (!first) {
        const bool odd = it & 1;
        const float *sxo = odd ? I.psx : I.e2, *szo = odd ? I.psz : I.e3;
        float *sxn = odd ? I.e2 : I.psx, *szn = odd ? I.e3 : I.psz;
        // adjoint of dp/dx on the vx domain, of dp/dz on the vz domain
        for (int q = threadIdx.x; q < NXE * kNZ; q += 256) {
            const int c = q / kNZ, r = q - c * kNZ;
            const int ix = X0 - XL + c, iz = Z0 + r;
            float ex = 0.f;
            if (ix >= LC - 1 && ix <= ntx - LC - 1 && iz >= LC && iz <= ntz - LC) {
                const size_t ip = static_cast<size_t>(ix) * ldz + iz;
                const float bv = g.Bx[ip] * I.vx[ip];
                ex = bv;
                const float a = g.axh[ix];
                if (a != 0.f) {
                    const float u = bv + sxo[ip];
                    ex = __builtin_fmaf(a, u, bv);
                    if (c >= XL && c < XL + kTX) sxn[ip] = g.bxh[ix] * u;
                }
            }
            sEX[q] = ex;
        }
        for (int q = threadIdx.x; q < kTX * NZE; q += 256) {
            const int c = q / NZE, r = q - c * NZE;
            const int ix = X0 + c, iz = Z0 - ZL + r;
            float ez = 0.f;
            if (ix >= LC && ix <= ntx - LC && iz >= LC - 1 && iz <= ntz - LC - 1) {
                const size_t ip = static_cast<size_t>(ix) * ldz + iz;
                const float bv = g.Bz[ip] * I.vz[ip];
                ez = bv;
                const float a = g.azh[iz];
                if (a != 0.f) {
                    const float u = bv + szo[ip];
                    ez = __builtin_fmaf(a, u, bv);
                    if (r >= ZL && r < ZL + kNZ) szn[ip] = g.bzh[iz] * u;
                }
            }
            sEZ[q] = ez;
        }
        __syncthreads();
    }
    // phat -= D- e; the data injection R^T d; grad += (K*S0)*phat on the interior
    for (int q = threadIdx.x; q < kTX * kNZ; q += 256) {
        const int c = q / kNZ, r = q - c * kNZ;
        const int ix = X0 + c, iz = Z0 + r;
        if (ix >= ntx || iz >= ntz) continue;
        const size_t ip = static_cast<size_t>(ix) * ldz + iz;
        float ph = I.p[ip];
        if (!first) {
            float gx = 0.f, gz = 0.f;
#pragma unroll
            for (int ic = 0; ic < LC; ic++) {
                const float e1 = sEX[(c + XL + ic) * kNZ + r], e0 = sEX[(c + XL - ic - 1) * kNZ + r];
                const float f1 = sEZ[c * NZE + r + ZL + ic], f0 = sEZ[c * NZE + r + ZL - ic - 1];
                gx = __builtin_fmaf(g.rc[ic] * (e1 - e0), g.idx, gx);
                gz = __builtin_fmaf(g.rc[ic] * (f1 - f0), g.idz, gz);
            }
            ph -= gx + gz;
        }
        if (I.colmap) {
            const int cm = I.colmap[ix];
            if (cm >= 0) {
                const int tr = I.rowmap[cm + iz];
                if (tr >= 0) ph += I.dsum[static_cast<size_t>(tr) * g.nt + it];
            }
        }
        I.p[ip] = ph;
        if (interior(g, ix, iz)) {
            const int im = (ix - g.pmlc) * g.nz + (iz - g.pmlc);
            const float s0 = I.s0[static_cast<size_t>(it - I.s0_t0) * g.nx * g.nz + im];
            I.grad[im] = __builtin_fmaf(g.K[ip] * s0, ph, I.grad[im]);
        }
    }
}

}  // namespace

hipError_t launch_adjoint_step_tiled(const KGrid &g, const KInst *insts, int ninst, int it, bool first, hipStream_t s)
{
    const long long nblk = static_cast<long long>(ninst) * ((g.ntz + kNZ - 1) / kNZ) * ((g.ntx + kTX - 1) / kTX);
    const dim3 gr(static_cast<unsigned>((nblk + 7) / 8 * 8));
    if (!first) {
        DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_adj_pv<L>), gr, dim3(256), 0, s, g, insts, ninst, it));
    }
    DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_adj_vp<L>), gr, dim3(256), 0, s, g, insts, ninst, it, first ? 1 : 0));
    return hipGetLastError();
}
