// swd_kernels.hip -- gfx950 kernels of the 2D acoustic staggered-grid CPML
// FDTD / RTM path: one kernel per sub-step ("plain" kernels), multi-instance.
//
// Layout: every field is [ntx][ldz] floats, z contiguous (the reference's
// ip = ix*ntz+iz, RK:45, with the column stride padded to a multiple of 32
// floats so each column starts on a 128-byte line).  A workgroup is 64 lanes
// along z (one wavefront reads one 256-byte column segment) times 4 columns;
// blockIdx.z selects the instance (shot / wavefield) of a batch.
//
// Arithmetic follows the reference per cell (RK:24-221, 434-817) in the
// canonical order d = fma(rc*(f+ - f-), 1/dx, d), psi = fma(b, psi, a*d),
// v = fma(B, d+psi, v), p = fma(K, ((dvx+psi)+dvz)+psi, p).  The library is
// built with -ffp-contract=off, so only the fmas written here are fused.
// CPML memory variables are read and written only where the damping
// coefficient a is non-zero: elsewhere they stay exactly zero (psi starts at 0
// and psi' = b*0 + 0*d = 0), so skipping them leaves every result unchanged.
#include <algorithm>

#include "swd_device.h"
#include "swd_kernels.h"
#include "swd_launch.h"

namespace {

// ---- quad kernels: one thread updates 4 consecutive z cells of one column ----
// A workgroup covers 64 rows x 16 columns (a wavefront = 4 columns x 64 rows,
// each column read as one 256-byte segment with float4 accesses).  Workgroups
// are ordered instance-fastest and handed to the XCDs in contiguous ranges, so
// the shots of one tile share the model and the halo columns in one L2.
#define QUAD_INDEX                                                         \
    const int L = xcd_block();                                             \
    if (L >= ninst * nzb * nxb) return;                                    \
    const int inst = L % ninst;                                            \
    const int zb = (L / ninst) % nzb, xb = L / (ninst * nzb);              \
    const int x = xb * 16 + (threadIdx.x >> 4);                            \
    const int z = zb * 64 + (threadIdx.x & 15) * 4;                        \
    if (x >= g.ntx || z >= g.ntz) return;                                  \
    const KInst &I = insts[inst];                                          \
    const size_t ldz = g.ldz;                                              \
    const size_t o = static_cast<size_t>(x) * ldz + z;


// Velocity update with CPML (RK:24-118).  NEG: receiver back-propagation,
// called by the reference with (-dx, -dz, +dt) and widened domains
// (RK:623-715, 1111-1127).  COMP: 1 = vx, 2 = vz, 3 = both.
template <int LC, bool NEG, int COMP>
__global__ __launch_bounds__(256) void k_v(KGrid g, const KInst *__restrict__ insts, int ninst, int nzb, int nxb)
{
    constexpr int R = (LC + 3) / 4;
    QUAD_INDEX
    gcfl *P = (gcfl *)I.p;
    if (COMP & 1) {
        if (x >= LC - 1 && x <= g.ntx - LC - 1) {
            const float idx = NEG ? -g.idx : g.idx;
            float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int ic = 0; ic < LC; ic++) {
                const float4 a = ld4(P, o + (ic + 1) * ldz), b = ld4(P, o - ic * ldz);
#pragma unroll
                for (int k = 0; k < 4; k++) el(d, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), idx, el(d, k));
            }
            const float ah = g.axh[x], bh = g.bxh[x];
            const float4 vo = ld4((gcfl *)I.vx, o), B = ld4((gcfl *)g.Bx, o);
            float4 vn = vo;
            float4 pso = make_float4(0.f, 0.f, 0.f, 0.f), psn = pso;
            if (ah != 0.f) psn = pso = ld4((gcfl *)I.psx, o);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                const bool in = NEG ? zk < g.ntz : (zk >= LC && zk <= g.ntz - LC);
                if (in) {
                    float ps = 0.f;
                    if (ah != 0.f) {
                        ps = __builtin_fmaf(bh, elc(pso, k), ah * elc(d, k));
                        el(psn, k) = ps;
                    }
                    el(vn, k) = __builtin_fmaf(elc(B, k), elc(d, k) + ps, elc(vo, k));
                }
            }
            st4((gfl *)I.vx, o, vn);
            if (ah != 0.f) st4((gfl *)I.psx, o, psn);
        }
    }
    if (COMP & 2) {
        if (NEG || (x >= LC && x <= g.ntx - LC)) {
            const float idz = NEG ? -g.idz : g.idz;
            float a[4 * (2 * R + 1)];
            column_window<R>(P, o, z, g.ldz, a);
            const float4 vo = ld4((gcfl *)I.vz, o), B = ld4((gcfl *)g.Bz, o);
            float4 vn = vo;
            bool anyps = false;
            float ahz[4], bhz[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zc = z + k < g.ntz ? z + k : g.ntz - 1;
                ahz[k] = g.azh[zc];
                bhz[k] = g.bzh[zc];
                anyps |= ahz[k] != 0.f;
            }
            float4 pso = make_float4(0.f, 0.f, 0.f, 0.f);
            if (anyps) pso = ld4((gcfl *)I.psz, o);
            float4 psn = pso;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                if (zk >= LC - 1 && zk <= g.ntz - LC - 1) {
                    float d = 0.f;
#pragma unroll
                    for (int ic = 0; ic < LC; ic++)
                        d = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic + 1] - a[4 * R + k - ic]), idz, d);
                    float ps = 0.f;
                    if (ahz[k] != 0.f) {
                        ps = __builtin_fmaf(bhz[k], elc(pso, k), ahz[k] * d);
                        el(psn, k) = ps;
                    }
                    el(vn, k) = __builtin_fmaf(elc(B, k), d + ps, elc(vo, k));
                }
            }
            st4((gfl *)I.vz, o, vn);
            if (anyps) st4((gfl *)I.psz, o, psn);
        }
    }
}

// Pressure update with CPML (RK:121-221; receiver side RK:717-782 when NEG).
// Forward: additive source, receiver recording, boundary-frame saving, AFor
// column capture, Born scattering source, background-sum saving.  Receiver
// side: hard replacement by the adjoint source, then the cross-correlation
// imaging condition with the reconstruction's dp/dt (RK:785-817).
template <int LC, bool NEG>
__global__ __launch_bounds__(256) void k_p(KGrid g, const KInst *__restrict__ insts, int ninst, int nzb, int nxb, int it)
{
    constexpr int R = (LC + 3) / 4;
    QUAD_INDEX
    const float4 pold = ld4((gcfl *)I.p, o);
    float4 pn = pold;
    float4 Sv = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool colin = x >= LC && x <= g.ntx - LC;
    const bool colint = x >= g.pmlc && x < g.ntx - g.pmlc;
    if (colin) {
        const float idx = NEG ? -g.idx : g.idx, idz = NEG ? -g.idz : g.idz;
        gcfl *VX = (gcfl *)I.vx;
        float4 dvx = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float4 a = ld4(VX, o + ic * ldz), b = ld4(VX, o - (ic + 1) * ldz);
#pragma unroll
            for (int k = 0; k < 4; k++) el(dvx, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), idx, el(dvx, k));
        }
        float a[4 * (2 * R + 1)];
        column_window<R>((gcfl *)I.vz, o, z, g.ldz, a);
        const float ax = g.ax[x], bx = g.bx[x];
        float azr[4], bzr[4];
        bool anyz = false;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int zc = z + k < g.ntz ? z + k : g.ntz - 1;
            azr[k] = g.az[zc];
            bzr[k] = g.bz[zc];
            anyz |= azr[k] != 0.f;
        }
        float4 qxo = make_float4(0.f, 0.f, 0.f, 0.f), qzo = qxo;
        if (ax != 0.f) qxo = ld4((gcfl *)I.pvx, o);
        if (anyz) qzo = ld4((gcfl *)I.pvz, o);
        float4 qxn = qxo, qzn = qzo;
        const float4 K = ld4((gcfl *)g.K, o);
        float4 bm = make_float4(0.f, 0.f, 0.f, 0.f), s0 = bm;
        const bool born = !NEG && I.born_m && colint;
        if (born) {
            // background sum of the same step; psi is zero on the interior
            gcfl *BX = (gcfl *)I.bvx;
            float4 bdx = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int ic = 0; ic < LC; ic++) {
                const float4 u = ld4(BX, o + ic * ldz), v = ld4(BX, o - (ic + 1) * ldz);
#pragma unroll
                for (int k = 0; k < 4; k++) el(bdx, k) = __builtin_fmaf(g.rc[ic] * (elc(u, k) - elc(v, k)), g.idx, el(bdx, k));
            }
            float bz_[4 * (2 * R + 1)];
            column_window<R>((gcfl *)I.bvz, o, z, g.ldz, bz_);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float d = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++)
                    d = __builtin_fmaf(g.rc[ic] * (bz_[4 * R + k + ic] - bz_[4 * R + k - ic - 1]), g.idz, d);
                el(s0, k) = elc(bdx, k) + d;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int zk = z + k;
            if (zk >= LC && zk <= g.ntz - LC) {
                float dvz = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++)
                    dvz = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic] - a[4 * R + k - ic - 1]), idz, dvz);
                float qx = 0.f, qz = 0.f;
                if (ax != 0.f) {
                    qx = __builtin_fmaf(bx, elc(qxo, k), ax * elc(dvx, k));
                    el(qxn, k) = qx;
                }
                if (azr[k] != 0.f) {
                    qz = __builtin_fmaf(bzr[k], elc(qzo, k), azr[k] * dvz);
                    el(qzn, k) = qz;
                }
                float S = ((elc(dvx, k) + qx) + dvz) + qz;
                el(Sv, k) = S;
                const bool inter = colint && zk >= g.pmlc && zk < g.ntz - g.pmlc;
                if (born && inter) {
                    const int im = (x - g.pmlc) * g.nz + (zk - g.pmlc);
                    S = __builtin_fmaf(I.born_m[im], elc(s0, k), S);
                }
                el(pn, k) = __builtin_fmaf(elc(K, k), S, elc(pold, k));
            }
        }
        if (ax != 0.f) st4((gfl *)I.pvx, o, qxn);
        if (anyz) st4((gfl *)I.pvz, o, qzn);
        if (!NEG && I.s0_out && colint) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                if (zk >= g.pmlc && zk < g.ntz - g.pmlc)
                    I.s0_out[static_cast<size_t>(it - I.s0_t0) * g.nx * g.nz + (x - g.pmlc) * g.nz + (zk - g.pmlc)] =
                        elc(Sv, k);
            }
        }
    }
    if (!NEG && I.wavelet && x == I.s_ix && I.s_iz >= z && I.s_iz < z + 4) add_at(pn, I.s_iz - z, I.wavelet[it]);   // RK:175-178
    if (I.colmap) {
        const int cm = I.colmap[x];
        if (cm >= 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (z + k >= g.ntz) break;
                const int tr = I.rowmap[cm + z + k];
                if (tr >= 0) {
                    const size_t go = static_cast<size_t>(tr) * g.nt + it;
                    if (NEG) el(pn, k) = I.adj[go];                 // hard replacement (RK:765-780)
                    else if (I.gather) I.gather[go] = elc(pn, k);   // RK:182-197
                }
            }
        }
    }
    st4((gfl *)I.p, o, pn);
    if (!NEG) {
        if (I.strips) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int f = frame_index(g, x, z + k);
                if (f >= 0) I.strips[static_cast<size_t>(it) * g.nframe + f] = elc(pn, k);   // RK:200-219
            }
        }
        if (I.capture) {
            const int c = I.colcap[x];
            if (c >= 0) {
                float *dst = reinterpret_cast<float *>(I.capture + (static_cast<size_t>(it) * I.ncap + c) * I.nfftz + z);
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (z + k < g.ntz) dst[2 * k + I.capcomp] = elc(pn, k);   // FK:324-327, 529-532
            }
        }
    } else if (I.image && colint) {
        const float4 dp = ld4((gcfl *)I.dpdt, o);
        float4 im = ld4((gcfl *)I.image, o), is = ld4((gcfl *)I.illum_s, o), ir = ld4((gcfl *)I.illum_r, o);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int zk = z + k;
            if (zk >= g.pmlc && zk < g.ntz - g.pmlc) {
                el(im, k) = __builtin_fmaf(elc(dp, k), elc(pn, k), elc(im, k));
                el(is, k) = __builtin_fmaf(elc(dp, k), elc(dp, k), elc(is, k));
                el(ir, k) = __builtin_fmaf(elc(pn, k), elc(pn, k), elc(ir, k));
            }
        }
        st4((gfl *)I.image, o, im);
        st4((gfl *)I.illum_s, o, is);
        st4((gfl *)I.illum_r, o, ir);
    }
}

// Source-wavefield reconstruction, pressure (RK:434-485 called with -dt), the
// wavelet subtraction at it+1 (RK:478-481) and the frame restore (RK:575-620).
template <int LC>
__global__ __launch_bounds__(256) void k_src_p(KGrid g, const KInst *__restrict__ insts, int ninst, int nzb, int nxb,
                                               int it)
{
    constexpr int R = (LC + 3) / 4;
    QUAD_INDEX
    float4 pn = ld4((gcfl *)I.p, o);
    const bool colint = x >= g.pmlc && x < g.ntx - g.pmlc;
    if (colint) {
        gcfl *VX = (gcfl *)I.vx;
        float4 dvx = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float4 a = ld4(VX, o + ic * ldz), b = ld4(VX, o - (ic + 1) * ldz);
#pragma unroll
            for (int k = 0; k < 4; k++) el(dvx, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), g.idx, el(dvx, k));
        }
        float a[4 * (2 * R + 1)];
        column_window<R>((gcfl *)I.vz, o, z, g.ldz, a);
        const float4 K = ld4((gcfl *)g.K, o), K0 = ld4((gcfl *)g.K0, o);
        float4 dp = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int zk = z + k;
            if (zk >= g.pmlc && zk < g.ntz - g.pmlc) {
                float dvz = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++)
                    dvz = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic] - a[4 * R + k - ic - 1]), g.idz, dvz);
                const float S = elc(dvx, k) + dvz;
                el(pn, k) = __builtin_fmaf(-elc(K, k), S, elc(pn, k));
                el(dp, k) = elc(K0, k) * S;
            }
        }
        st4((gfl *)I.dpdt, o, dp);
    }
    if (x == I.s_ix && I.s_iz >= z && I.s_iz < z + 4) add_at(pn, I.s_iz - z, -I.wavelet[it + 1]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int f = frame_index(g, x, z + k);
        if (f >= 0) el(pn, k) = I.strips[static_cast<size_t>(it) * g.nframe + f];
    }
    st4((gfl *)I.p, o, pn);
}

// Source-wavefield reconstruction, velocities (RK:488-573 called with -dt).
template <int LC, int COMP>
__global__ __launch_bounds__(256) void k_src_v(KGrid g, const KInst *__restrict__ insts, int ninst, int nzb, int nxb)
{
    constexpr int R = (LC + 3) / 4;
    QUAD_INDEX
    if (x < g.pmlc || x >= g.ntx - g.pmlc) return;
    gcfl *P = (gcfl *)I.p;
    if (COMP & 1) {
        float4 d = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float4 a = ld4(P, o + (ic + 1) * ldz), b = ld4(P, o - ic * ldz);
#pragma unroll
            for (int k = 0; k < 4; k++) el(d, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), g.idx, el(d, k));
        }
        float4 v = ld4((gcfl *)I.vx, o);
        const float4 B = ld4((gcfl *)g.Bx, o);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (z + k >= g.pmlc && z + k < g.ntz - g.pmlc) el(v, k) = __builtin_fmaf(-elc(B, k), elc(d, k), elc(v, k));
        st4((gfl *)I.vx, o, v);
    }
    if (COMP & 2) {
        float a[4 * (2 * R + 1)];
        column_window<R>(P, o, z, g.ldz, a);
        float4 v = ld4((gcfl *)I.vz, o);
        const float4 B = ld4((gcfl *)g.Bz, o);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (z + k >= g.pmlc && z + k < g.ntz - g.pmlc) {
                float d = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++)
                    d = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic + 1] - a[4 * R + k - ic]), g.idz, d);
                el(v, k) = __builtin_fmaf(-elc(B, k), d, elc(v, k));
            }
        }
        st4((gfl *)I.vz, o, v);
    }
}

#define CELL_INDEX                                            \
    const int iz = blockIdx.x * 64 + threadIdx.x;             \
    const int ix = blockIdx.y * 4 + threadIdx.y;              \
    if (ix >= g.ntx || iz >= g.ntz) return;                   \
    const KInst &I = insts[blockIdx.z];                       \
    const size_t ip = static_cast<size_t>(ix) * g.ldz + iz;   \
    const size_t sx = g.ldz;

// ---- exact adjoint of one forward step (Born adjoint) ----
// A_p^T, first half: ex/ez = adjoint of dvx/dvz on the p domain, psi-hat update.
template <int LC>
__global__ __launch_bounds__(256) void k_adj_p1(KGrid g, const KInst *__restrict__ insts)
{
    CELL_INDEX
    float ex = 0.f, ez = 0.f;
    if (ix >= LC && ix <= g.ntx - LC && iz >= LC && iz <= g.ntz - LC) {
        const float t = g.K[ip] * I.p[ip];
        const float ax = g.ax[ix], az = g.az[iz];
        ex = t;
        ez = t;
        if (ax != 0.f) {
            const float u = t + I.pvx[ip];
            ex = __builtin_fmaf(ax, u, t);
            I.pvx[ip] = g.bx[ix] * u;
        }
        if (az != 0.f) {
            const float u = t + I.pvz[ip];
            ez = __builtin_fmaf(az, u, t);
            I.pvz[ip] = g.bz[iz] * u;
        }
    }
    I.ex[ip] = ex;
    I.ez[ip] = ez;
}

// A_p^T, second half: vhat -= D+ e (transpose of the backward difference).
template <int LC>
__global__ __launch_bounds__(256) void k_adj_v1(KGrid g, const KInst *__restrict__ insts)
{
    CELL_INDEX
    float gx = 0.f, gz = 0.f;
#pragma unroll
    for (int ic = 0; ic < LC; ic++) {
        const float e1 = ix + ic + 1 < g.ntx ? I.ex[ip + (ic + 1) * sx] : 0.f;
        const float e0 = ix - ic >= 0 ? I.ex[ip - ic * sx] : 0.f;
        const float f1 = iz + ic + 1 < g.ntz ? I.ez[ip + ic + 1] : 0.f;
        const float f0 = iz - ic >= 0 ? I.ez[ip - ic] : 0.f;
        gx = __builtin_fmaf(g.rc[ic] * (e1 - e0), g.idx, gx);
        gz = __builtin_fmaf(g.rc[ic] * (f1 - f0), g.idz, gz);
    }
    I.vx[ip] -= gx;
    I.vz[ip] -= gz;
}

// A_v^T, first half: adjoint of dp/dx, dp/dz on the velocity domains.
template <int LC>
__global__ __launch_bounds__(256) void k_adj_v2(KGrid g, const KInst *__restrict__ insts)
{
    CELL_INDEX
    float ex = 0.f, ez = 0.f;
    if (ix >= LC - 1 && ix <= g.ntx - LC - 1 && iz >= LC && iz <= g.ntz - LC) {
        const float bv = g.Bx[ip] * I.vx[ip];
        ex = bv;
        const float a = g.axh[ix];
        if (a != 0.f) {
            const float u = bv + I.psx[ip];
            ex = __builtin_fmaf(a, u, bv);
            I.psx[ip] = g.bxh[ix] * u;
        }
    }
    if (ix >= LC && ix <= g.ntx - LC && iz >= LC - 1 && iz <= g.ntz - LC - 1) {
        const float bv = g.Bz[ip] * I.vz[ip];
        ez = bv;
        const float a = g.azh[iz];
        if (a != 0.f) {
            const float u = bv + I.psz[ip];
            ez = __builtin_fmaf(a, u, bv);
            I.psz[ip] = g.bzh[iz] * u;
        }
    }
    I.ex[ip] = ex;
    I.ez[ip] = ez;
}

// A_v^T, second half: phat -= D- e; then the data injection R^T d and the
// gradient sum grad += (K*S0)*phat on the interior.
template <int LC>
__global__ __launch_bounds__(256) void k_adj_p2(KGrid g, const KInst *__restrict__ insts, int it, int first)
{
    CELL_INDEX
    float ph = I.p[ip];
    if (!first) {
        float gx = 0.f, gz = 0.f;
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float e1 = ix + ic < g.ntx ? I.ex[ip + ic * sx] : 0.f;
            const float e0 = ix - ic - 1 >= 0 ? I.ex[ip - (ic + 1) * sx] : 0.f;
            const float f1 = iz + ic < g.ntz ? I.ez[ip + ic] : 0.f;
            const float f0 = iz - ic - 1 >= 0 ? I.ez[ip - (ic + 1)] : 0.f;
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

// Copy gather rows of traces that share a node with a later trace.
__global__ void k_copy_rows(float *__restrict__ gather, const int2 *__restrict__ pairs, int npairs, int nt)
{
    const int k = blockIdx.y;
    if (k >= npairs) return;
    const int2 pr = pairs[k];   // (dst, src)
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < nt; it += gridDim.x * blockDim.x)
        gather[static_cast<size_t>(pr.x) * nt + it] = gather[static_cast<size_t>(pr.y) * nt + it];
}

// Sum rows of traces sharing a node into the row of the node's last trace.
__global__ void k_sum_rows(float *__restrict__ dst, const float *__restrict__ src, const int2 *__restrict__ pairs,
                           int npairs, int nt)
{
    // pairs are (dup, primary), grouped by primary, applied sequentially per primary
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < nt; it += gridDim.x * blockDim.x)
        for (int k = 0; k < npairs; k++) {
            const int2 pr = pairs[k];
            dst[static_cast<size_t>(pr.y) * nt + it] += src[static_cast<size_t>(pr.x) * nt + it];
        }
}

// Sum of per-shot arrays in shot order: dst[i] = src[0][i] + src[1][i] + ...
// With accumulate, dst already holds the sum of earlier shots and the shots
// of src are added to it in order.
__global__ void k_sum_shots(float *__restrict__ dst, const float *__restrict__ src, int nshots, size_t n, int accumulate)
{
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        float acc = accumulate ? dst[i] + src[i] : src[i];
        for (int k = 1; k < nshots; k++) acc += src[static_cast<size_t>(k) * n + i];
        dst[i] = acc;
    }
}

// ---- kz up/down separation on receiver columns (FK:705, 762-777, 52-138) ----
// One workgroup per (time step, receiver column): the captured column
// p + i*htp, zero padded to nfftz, is transformed along z in LDS (radix-2
// decimation in time, twiddles from an FP64-precomputed table), multiplied by
// the taper and split at nfftz/2 (Nyquist goes up), inverse transformed
// (unnormalised, as cuFFT INVERSE), and the real parts at the receiver rows
// are written to the up/down gathers.
__device__ __forceinline__ void lds_fft(float2 *a, float2 *b, int n, int logn, const float2 *__restrict__ tw, int twn,
                                        bool inverse, bool two)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int j = __brev(static_cast<unsigned>(i)) >> (32 - logn);
        if (i < j) {
            float2 t = a[i]; a[i] = a[j]; a[j] = t;
            if (two) { t = b[i]; b[i] = b[j]; b[j] = t; }
        }
    }
    __syncthreads();
    for (int len = 2, stride = twn / 2; len <= n; len <<= 1, stride >>= 1) {
        const int half = len >> 1;
        for (int q = threadIdx.x; q < n / 2; q += blockDim.x) {
            const int k = q & (half - 1);
            const int i0 = (q / half) * len + k, i1 = i0 + half;
            float2 w = tw[k * stride];
            if (inverse) w.y = -w.y;
            {
                const float2 x1 = a[i1], x0 = a[i0];
                const float tr = x1.x * w.x - x1.y * w.y;
                const float ti = x1.x * w.y + x1.y * w.x;
                a[i1] = make_float2(x0.x - tr, x0.y - ti);
                a[i0] = make_float2(x0.x + tr, x0.y + ti);
            }
            if (two) {
                const float2 x1 = b[i1], x0 = b[i0];
                const float tr = x1.x * w.x - x1.y * w.y;
                const float ti = x1.x * w.y + x1.y * w.x;
                b[i1] = make_float2(x0.x - tr, x0.y - ti);
                b[i0] = make_float2(x0.x + tr, x0.y + ti);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_kzsep(const float2 *__restrict__ capture, int ncap, int nfftz, int logn,
                                               const float2 *__restrict__ tw, const float *__restrict__ coef,
                                               const KCapCol *__restrict__ cols, const int2 *__restrict__ traces,
                                               float *__restrict__ up, float *__restrict__ down, int nt)
{
    extern __shared__ float2 lds[];
    float2 *a = lds, *b = lds + nfftz;
    const int c = blockIdx.x, it = blockIdx.y;
    const float2 *src = capture + (static_cast<size_t>(it) * ncap + c) * nfftz;
    for (int i = threadIdx.x; i < nfftz; i += blockDim.x) a[i] = src[i];
    __syncthreads();
    lds_fft(a, b, nfftz, logn, tw, nfftz, false, false);
    const int halfz = nfftz / 2;
    for (int i = threadIdx.x; i < nfftz; i += blockDim.x) {
        const float cf = coef[i];
        const float2 t = a[i];
        const float2 s = make_float2(t.x * cf, t.y * cf);
        if (i <= halfz) { a[i] = s; b[i] = make_float2(0.f, 0.f); }
        else { b[i] = s; a[i] = make_float2(0.f, 0.f); }
    }
    __syncthreads();
    lds_fft(a, b, nfftz, logn, tw, nfftz, true, true);
    const KCapCol cc = cols[c];
    for (int k = threadIdx.x; k < cc.tcount; k += blockDim.x) {
        const int2 t = traces[cc.tstart + k];   // (gather row, iz)
        up[static_cast<size_t>(t.x) * nt + it] = a[t.y].x;
        down[static_cast<size_t>(t.x) * nt + it] = b[t.y].x;
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers
static inline dim3 cell_grid(const KGrid &g, int ninst) { return dim3((g.ntz + 63) / 64, (g.ntx + 3) / 4, ninst); }
static const dim3 kCellBlock(64, 4, 1);

struct QGeom {
    int nzb, nxb, nblocks;
};
static QGeom qgeom(const KGrid &g, int ninst)
{
    QGeom q;
    q.nzb = (g.ntz + 63) / 64;
    q.nxb = (g.ntx + 15) / 16;
    q.nblocks = (ninst * q.nzb * q.nxb + 7) / 8 * 8;
    return q;
}

hipError_t launch_v(const KGrid &g, const KInst *insts, int ninst, bool neg, int comp, hipStream_t s)
{
    const QGeom q = qgeom(g, ninst);
#define LV(NEGV, C) hipLaunchKernelGGL((k_v<L, NEGV, C>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb)
    if (!neg) {
        if (comp == 1) { DISPATCH_LC(g.Lc, LV(false, 1)); }
        else if (comp == 2) { DISPATCH_LC(g.Lc, LV(false, 2)); }
        else { DISPATCH_LC(g.Lc, LV(false, 3)); }
    } else {
        if (comp == 1) { DISPATCH_LC(g.Lc, LV(true, 1)); }
        else if (comp == 2) { DISPATCH_LC(g.Lc, LV(true, 2)); }
        else { DISPATCH_LC(g.Lc, LV(true, 3)); }
    }
#undef LV
    return hipGetLastError();
}

hipError_t launch_p(const KGrid &g, const KInst *insts, int ninst, bool neg, int it, hipStream_t s)
{
    const QGeom q = qgeom(g, ninst);
    if (!neg) { DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_p<L, false>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb, it)); }
    else { DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_p<L, true>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb, it)); }
    return hipGetLastError();
}

hipError_t launch_src_p(const KGrid &g, const KInst *insts, int ninst, int it, hipStream_t s)
{
    const QGeom q = qgeom(g, ninst);
    DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_src_p<L>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb, it));
    return hipGetLastError();
}

hipError_t launch_src_v(const KGrid &g, const KInst *insts, int ninst, int comp, hipStream_t s)
{
    const QGeom q = qgeom(g, ninst);
    if (comp == 1) { DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_src_v<L, 1>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb)); }
    else if (comp == 2) { DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_src_v<L, 2>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb)); }
    else { DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_src_v<L, 3>), dim3(q.nblocks), dim3(256), 0, s, g, insts, ninst, q.nzb, q.nxb)); }
    return hipGetLastError();
}

hipError_t launch_adjoint_step(const KGrid &g, const KInst *insts, int ninst, int it, bool first, hipStream_t s)
{
    const dim3 gr = cell_grid(g, ninst);
    if (!first) {
        DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_adj_p1<L>), gr, kCellBlock, 0, s, g, insts));
        DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_adj_v1<L>), gr, kCellBlock, 0, s, g, insts));
        DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_adj_v2<L>), gr, kCellBlock, 0, s, g, insts));
    }
    DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_adj_p2<L>), gr, kCellBlock, 0, s, g, insts, it, first ? 1 : 0));
    return hipGetLastError();
}

hipError_t launch_copy_rows(float *gather, const int2 *pairs, int npairs, int nt, hipStream_t s)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_rows, dim3((nt + 255) / 256, npairs), dim3(256), 0, s, gather, pairs, npairs, nt);
    return hipGetLastError();
}

hipError_t launch_sum_rows(float *dst, const float *src, const int2 *pairs, int npairs, int nt, hipStream_t s)
{
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sum_rows, dim3((nt + 255) / 256), dim3(256), 0, s, dst, src, pairs, npairs, nt);
    return hipGetLastError();
}

hipError_t launch_sum_shots(float *dst, const float *src, int nshots, size_t n, bool accumulate, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_sum_shots, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, dst, src, nshots, n,
                       accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_kzsep(const float2 *capture, int ncap, int nfftz, const float2 *tw, const float *coef,
                        const KCapCol *cols, const int2 *traces, float *up, float *down, int nt, hipStream_t s)
{
    int logn = 0;
    while ((1 << logn) < nfftz) logn++;
    const size_t lds = sizeof(float2) * 2 * nfftz;
    hipLaunchKernelGGL(k_kzsep, dim3(ncap, nt), dim3(256), lds, s, capture, ncap, nfftz, logn, tw, coef, cols, traces,
                       up, down, nt);
    return hipGetLastError();
}

