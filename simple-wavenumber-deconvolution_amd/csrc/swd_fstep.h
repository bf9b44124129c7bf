// swd_fstep.h -- fused leapfrog step (kernel templates, included by the
// swd_fstep_*.hip translation units, one per kernel variant): the velocity and the pressure half of
// one time step of every wavefield of a batch in a single launch.
//
// A workgroup owns a tile of 4 TQ rows (TQ float4 quads per column) x TX
// columns of one wavefield.  It first
// updates the velocities its pressure stencil needs -- vx on the tile's rows
// over columns X0-Lc .. X1+Lc-2 and vz on the tile's columns over quads
// -R .. 15+R (R = ceil(Lc/4)) -- into LDS, storing only its own cells, then
// updates its own pressure cells from LDS.  Each workgroup recomputes its halo
// velocities from the previous slot instead of waiting for a neighbour; the
// double-buffered state (KStep) keeps that slot intact for the whole launch.
// Per cell and step the HBM traffic is one read and one write of p, vx and vz,
// the model (shared by the batch's shots through L2) and, in the absorbing
// frame, the CPML memories.
//
// Arithmetic is the canonical per-cell order of the reference's equations
// (see swd_kernels.hip); results are bit-identical to the sub-step kernels.
//
// Modes (per instance, uniform per workgroup):
//  FWD  forward CPML step (RK:24-221): additive source, receiver recording,
//       boundary-frame saving, AFor column capture.
//  SRC  source-wavefield reconstruction (RK:434-620), interior only, run
//       velocity-then-pressure: the launch of step it first takes the
//       velocities from it+2 to it+1 with the pressure of it+1 (RK:488-573),
//       then the pressure from it+1 to it (RK:434-485) with dp/dt, the
//       wavelet subtraction and the frame restore.  The first backward launch
//       (it = nt-2) keeps the forward's final velocities.  Cells outside the
//       interior are never written and must hold the forward's final state in
//       both slots.
//  RCV  receiver back-propagation with (-dx, -dz) and widened domains
//       (RK:623-782) and hard replacement by the adjoint source.
// A backward launch pairs the SRC and RCV wavefields of a shot in one
// workgroup per tile: the reconstruction's dp/dt stays in registers and the
// imaging condition of step it (RK:785-817) is applied right after the
// receiver pressure of step it, so dp/dt never goes to memory.  The CIG
// variant also bins the angle gathers of step it+1 (RK:859-901): the
// reconstruction's p and velocities at that level are in LDS during the
// source half (kept in registers), and the receiver field's are its previous
// slot.
#pragma once
#include <algorithm>

#include "swd_device.h"
#include "swd_kernels.h"
#include "swd_launch.h"

namespace {

__device__ __forceinline__ float4 zero4() { return make_float4(0.f, 0.f, 0.f, 0.f); }

// Quads s[0 .. 2R] of an LDS column as 4(2R+1) consecutive rows.  The first
// row is not used by the stencils; the empty asm keeps each quad a whole
// 16-byte read (a narrowed read starts 4 bytes off and splits into
// bank-conflicting dword pairs).
template <int R>
__device__ __forceinline__ void lds_window(const float4 *s, float (&a)[4 * (2 * R + 1)])
{
    v4f t[2 * R + 1];
#pragma unroll
    for (int j = 0; j <= 2 * R; j++) t[j] = *reinterpret_cast<const v4f *>(s + j);
#pragma unroll
    for (int j = 0; j <= 2 * R; j++) asm volatile("" ::"v"(t[j]));
#pragma unroll
    for (int j = 0; j <= 2 * R; j++) {
        a[4 * j] = t[j].x;
        a[4 * j + 1] = t[j].y;
        a[4 * j + 2] = t[j].z;
        a[4 * j + 3] = t[j].w;
    }
}

// Direct-to-LDS load of one float4 a lane: lane l of the wavefront writes
// dst[l], dst being the same for the whole wavefront.  The data arrive with
// the wavefront's vector-memory counter; the reader waits for it and a barrier.
__device__ __forceinline__ void ld4_lds(gcfl *f, unsigned o, float4 *dst)
{
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)((gcch *)f + (o << 2)),
                                     (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}

// Velocity-z item i of a tile: the TQ own quads of each column first (TQ
// consecutive lanes per column, so LDS accesses of 8-lane groups stay in one
// column), then the 2R halo quads of each column.
template <int TX, int R, int TQ>
__device__ __forceinline__ void vz_item(int i, int &c, int &qq)
{
    if (i < TX * TQ) {
        c = i / TQ;
        qq = i % TQ;
    } else {
        const int j = i - TX * TQ, h = j % (2 * R);
        c = j / (2 * R);
        qq = h < R ? h - R : TQ + h - R;
    }
}

// Threads per workgroup: one own quad per thread, 128 to 512 threads.
template <int TX, int TQ>
constexpr int fstep_threads() { return TX * TQ < 128 ? 128 : TX * TQ > 512 ? 512 : TX * TQ; }

// Tile geometry: 4 TQ rows (TQ quads) x TX columns, Lc-deep stencils.
template <int LC, int TX, int TQ>
struct FGeo {
    static constexpr int R = (LC + 3) / 4;       // z halo (quads) of a velocity stencil
    static constexpr int RP = 2 * R;             // z halo (quads) of the staged pressure
    static constexpr int OWN = 2 * LC - 1;       // staged column of X0
    static constexpr int NPC = TX + 4 * LC - 2;  // staged pressure columns X0-2Lc+1 .. X1+2Lc-2
    static constexpr int NPQ = TQ + 2 * RP;      // staged pressure quads -RP .. TQ-1+RP
    static constexpr int NVX = TX + 2 * LC - 1;  // vx columns X0-Lc .. X1+Lc-2
    static constexpr int VZQ = TQ + 2 * R;       // vz quads -R .. TQ-1+R
    static constexpr int NT = fstep_threads<TX, TQ>();   // threads per workgroup
    static constexpr int NOWN = (TX * TQ + NT - 1) / NT;  // own quads per thread
    static constexpr int N0 = NPC * NPQ, N1 = NVX * TQ, N2 = TX * VZQ, N3 = TX * TQ;   // staged p, vx, vz, own items
    static constexpr int I0 = (N0 + NT - 1) / NT, I1 = (N1 + NT - 1) / NT, I2 = (N2 + NT - 1) / NT,
                         I3 = (N3 + NT - 1) / NT;   // items per thread
};

// What the reconstruction half of a fast paired tile prepares for the receiver
// half: the model (the same cells), read once for both, and the receiver's
// image accumulators, read during the reconstruction's pressure update.  The
// receiver's previous p, vx and vz come in by direct-to-LDS loads at that time.
template <int LC, int TX, int TQ>
struct FPre {
    using G = FGeo<LC, TX, TQ>;
    float4 bx[G::I1], bz[G::I2], k[G::I3], im[G::I3], is[G::I3], ir[G::I3];
};

// One tile of one wavefield.  FULL: the tile and every halo lie inside the
// interior box free of CPML damping, so no cell needs a domain test, a CPML
// memory, a frame index or an interior test; the per-cell arithmetic is the
// same as in the general case (psi = 0 there; only the sign of an exact zero
// may differ, through the omitted "+ 0").
// Born pair (background first, scattered field second, in one workgroup and
// on the same own cells): the background keeps the divergence sum S of each
// interior cell in acc; the scattered field adds m S_background to its own sum
// (the Born source of the batched k_p, in the same operation order).
__device__ __forceinline__ void born_sum(const KGrid &g, const KStep &I, int x, int zk, float4 &acc, int k, float &S)
{
    if (I.born_m)
        S = __builtin_fmaf(ldg(I.born_m, static_cast<unsigned>((x - g.pmlc) * g.nz + (zk - g.pmlc))), elc(acc, k), S);
    else
        el(acc, k) = S;
}

// PF = 1: a fast reconstruction tile that also prefetches the receiver tile
// (rcv) into pf and the second LDS set (sP2, sVXo, sVZo).  PF = 2: that
// receiver tile, run from the prefetch: its staged p is sP (the caller passes
// sP2), its previous velocities sVXo and sVZo.
template <int LC, int TX, int TQ, bool FULL, bool CIG, bool BORN = false, int PF = 0>
__device__ __forceinline__ void fstep_tile(const KGrid &g, const KStep &I, int X0, int Z0, int it, int first,
                                           float4 *__restrict__ sP, float4 *__restrict__ sVX, float4 *__restrict__ sVZ,
                                           float4 (&dpa)[FGeo<LC, TX, TQ>::NOWN], float4 (&cgk)[3][FGeo<LC, TX, TQ>::NOWN],
                                           FPre<LC, TX, TQ> *pf = nullptr, const KStep *rcv = nullptr,
                                           float4 *sP2 = nullptr, float4 *sVXo = nullptr, float4 *sVZo = nullptr)
{
    using G = FGeo<LC, TX, TQ>;
    constexpr int R = G::R, RP = G::RP, OWN = G::OWN, NPQ = G::NPQ, VZQ = G::VZQ;
    constexpr int N0 = G::N0, N1 = G::N1, N2 = G::N2, N3 = G::N3;
    constexpr int NT = G::NT;
    constexpr int I0 = G::I0, I1 = G::I1, I2 = G::I2, I3 = G::I3;
    static_assert(PF == 0 || (FULL && !CIG && !BORN), "the prefetch runs on fast plain pairs");
    const int tid = threadIdx.x;
    const int mode = I.mode;
    const int pmlc = g.pmlc, ntx = g.ntx, ntz = g.ntz;
    const unsigned ldz = static_cast<unsigned>(g.ldz);
    const bool srcm = mode == KSTEP_SRC, neg = mode == KSTEP_RCV;
    const int po = (it + 1) & 1, pn = it & 1;
    const int vo = srcm ? pn : po, vn = srcm ? po : pn;
    const bool vupd = !(srcm && first);   // the first reconstruction launch keeps the forward's velocities
    const int vsrc = vupd ? vo : vn;
    gcfl *P = (gcfl *)I.p[po];
    gcfl *VXs = (gcfl *)I.vx[vsrc];
    gcfl *VZs = (gcfl *)I.vz[vsrc];
    const float idx = neg ? -g.idx : g.idx, idz = neg ? -g.idz : g.idz;

    // ---- issue the loads that do not depend on LDS: velocities, model, image accumulators
    float4 vxo[I1], bxv[I1], vzo[I2], bzv[I2], kv[I3];
#pragma unroll
    for (int r = 0; r < I1; r++) {
        const int i = tid + NT * r, q = i % TQ, cx = i / TQ;
        const int x = X0 - LC + cx, z = Z0 + 4 * q;
        vxo[r] = bxv[r] = zero4();
        if constexpr (PF == 2) bxv[r] = pf->bx[r];
        else if (i < N1 && (FULL || (x >= 0 && x < ntx && z < ntz))) {
            const unsigned o = static_cast<unsigned>(x) * ldz + z;
            vxo[r] = ld4(VXs, o);
            bxv[r] = ld4((gcfl *)g.Bx, o);
        }
    }
#pragma unroll
    for (int r = 0; r < I2; r++) {
        const int i = tid + NT * r;
        int c, qq;
        vz_item<TX, R, TQ>(i, c, qq);
        const int x = X0 + c, z = Z0 + 4 * qq;
        vzo[r] = bzv[r] = zero4();
        if constexpr (PF == 2) bzv[r] = pf->bz[r];
        else if (i < N2 && (FULL || (x < ntx && z >= 0 && z < ntz))) {
            const unsigned o = static_cast<unsigned>(x) * ldz + z;
            vzo[r] = ld4(VZs, o);
            bzv[r] = ld4((gcfl *)g.Bz, o);
        }
    }
    float4 imv[I3], isv[I3], irv[I3];   // receiver field: image accumulators; reconstruction: imv holds K0
    int cmv[I3];   // receiver map of each own column
#pragma unroll
    for (int r = 0; r < I3; r++) {
        const int i = tid + NT * r, q = i % TQ, c = i / TQ;
        const int x = X0 + c, z = Z0 + 4 * q;
        kv[r] = imv[r] = isv[r] = irv[r] = zero4();
        cmv[r] = -1;
        if constexpr (PF == 2) {
            kv[r] = pf->k[r];
            imv[r] = pf->im[r];
            isv[r] = pf->is[r];
            irv[r] = pf->ir[r];
            if (i < N3 && I.colmap) cmv[r] = ldgi(I.colmap, x);
        } else if (i < N3 && (FULL || (x < ntx && z < ntz))) {
            if (I.colmap) cmv[r] = ldgi(I.colmap, x);
            const unsigned o = static_cast<unsigned>(x) * ldz + z;
            kv[r] = ld4((gcfl *)g.K, o);
            if (srcm) imv[r] = ld4((gcfl *)g.K0, o);
            if (neg && I.image && (FULL || (x >= pmlc && x < ntx - pmlc))) {
                imv[r] = ld4((gcfl *)I.image, o);
                isv[r] = ld4((gcfl *)I.illum_s, o);
                irv[r] = ld4((gcfl *)I.illum_r, o);
            }
        }
    }

    // ---- stage the previous pressure: own columns with z halo, x halo columns on the tile's rows
    if constexpr (PF != 2) {
        float4 st[I0];
#pragma unroll
        for (int r = 0; r < I0; r++) {
            const int i = tid + NT * r, col = i / NPQ, qq = i - col * NPQ - RP;
            const int x = X0 - OWN + col, z = Z0 + 4 * qq;
            const bool need = i < N0 && ((qq >= 0 && qq < TQ) || (col >= OWN && col < OWN + TX));
            st[r] = zero4();
            if (need && (FULL || (x >= 0 && x < ntx && z >= 0 && z < ntz))) st[r] = ld4(P, static_cast<unsigned>(x) * ldz + z);
        }
#pragma unroll
        for (int r = 0; r < I0; r++) {
            const int i = tid + NT * r;
            if (i < N0) sP[i] = st[r];
        }
        __syncthreads();
    }

    // ---- velocity x on columns X0-Lc .. X1+Lc-2 of the tile's rows
#pragma unroll
    for (int r = 0; r < I1; r++) {
        const int i = tid + NT * r;
        if (i >= N1) break;
        const int q = i % TQ, cx = i / TQ;
        const int x = X0 - LC + cx, z = Z0 + 4 * q;
        float4 v = vxo[r];
        if constexpr (PF == 2) v = sVXo[i];
        if (FULL || (x >= 0 && x < ntx && z < ntz)) {
            const unsigned o = static_cast<unsigned>(x) * ldz + z;
            const bool own = cx >= LC && cx < LC + TX;
            const int sa = (cx + LC) * NPQ + q + RP, sb = (cx + LC - 1) * NPQ + q + RP;   // p at x+1, x
            const float4 B = bxv[r];
            if (srcm) {
                const bool upd = vupd && (FULL || (x >= pmlc && x < ntx - pmlc && z + 3 >= pmlc && z < ntz - pmlc));
                if (upd) {
                    float4 d = zero4();
#pragma unroll
                    for (int ic = 0; ic < LC; ic++) {
                        const float4 a = sP[sa + ic * NPQ], b = sP[sb - ic * NPQ];
#pragma unroll
                        for (int k = 0; k < 4; k++) el(d, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), g.idx, el(d, k));
                    }
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const float nv = __builtin_fmaf(-elc(B, k), elc(d, k), elc(v, k));
                        el(v, k) = (FULL || (z + k >= pmlc && z + k < ntz - pmlc)) ? nv : elc(v, k);
                    }
                    if (own) st4((gfl *)I.vx[vn], o, v);
                }
            } else {
                if (FULL || (x >= LC - 1 && x <= ntx - LC - 1)) {
                    float4 d = zero4();
#pragma unroll
                    for (int ic = 0; ic < LC; ic++) {
                        const float4 a = sP[sa + ic * NPQ], b = sP[sb - ic * NPQ];
#pragma unroll
                        for (int k = 0; k < 4; k++) el(d, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), idx, el(d, k));
                    }
                    if (FULL) {
#pragma unroll
                        for (int k = 0; k < 4; k++) el(v, k) = __builtin_fmaf(elc(B, k), elc(d, k), elc(v, k));
                    } else {
                        // the memory is read with the coefficients (ah is zero inside the damping-free box)
                        const float ah = ldg(g.axh, x), bh = ldg(g.bxh, x);
                        float4 pso = zero4();
                        if (x < g.fx0 || x > g.fx1) pso = ld4((gcfl *)I.psx[vo], o);
                        float4 psn = pso;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int zk = z + k;
                            const bool in = neg ? zk < ntz : (zk >= LC && zk <= ntz - LC);
                            const float ps = ah != 0.f ? __builtin_fmaf(bh, elc(pso, k), ah * elc(d, k)) : 0.f;
                            el(psn, k) = in ? ps : elc(pso, k);
                            const float nv = __builtin_fmaf(elc(B, k), elc(d, k) + ps, elc(v, k));
                            el(v, k) = in ? nv : elc(v, k);
                        }
                        if (own && ah != 0.f) st4((gfl *)I.psx[vn], o, psn);
                    }
                }
                if (own) st4((gfl *)I.vx[vn], o, v);
            }
        }
        sVX[i] = v;
    }

    // ---- velocity z on quads -R .. 15+R of the tile's columns
#pragma unroll
    for (int r = 0; r < I2; r++) {
        const int i = tid + NT * r;
        if (i >= N2) break;
        int c, qq;
        vz_item<TX, R, TQ>(i, c, qq);
        const int x = X0 + c, z = Z0 + 4 * qq;
        float4 v = vzo[r];
        if constexpr (PF == 2) v = sVZo[i];
        if (FULL || (x < ntx && z >= 0 && z < ntz)) {
            const unsigned o = static_cast<unsigned>(x) * ldz + z;
            const bool own = qq >= 0 && qq < TQ;
            const float4 B = bzv[r];
            float a[4 * (2 * R + 1)];   // p rows z-4R .. z+4R+3
            lds_window<R>(&sP[(c + OWN) * NPQ + qq + RP - R], a);
            if (srcm) {
                const bool upd = vupd && (FULL || (x >= pmlc && x < ntx - pmlc && z + 3 >= pmlc && z < ntz - pmlc));
                if (upd) {
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        float d = 0.f;
#pragma unroll
                        for (int ic = 0; ic < LC; ic++)
                            d = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic + 1] - a[4 * R + k - ic]), g.idz, d);
                        const float nv = __builtin_fmaf(-elc(B, k), d, elc(v, k));
                        el(v, k) = (FULL || (z + k >= pmlc && z + k < ntz - pmlc)) ? nv : elc(v, k);
                    }
                    if (own) st4((gfl *)I.vz[vn], o, v);
                }
            } else {
                if (FULL || neg || (x >= LC && x <= ntx - LC)) {
                    if (FULL) {
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            float d = 0.f;
#pragma unroll
                            for (int ic = 0; ic < LC; ic++)
                                d = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic + 1] - a[4 * R + k - ic]), idz, d);
                            el(v, k) = __builtin_fmaf(elc(B, k), d, elc(v, k));
                        }
                    } else {
                        float ahz[4], bhz[4];
                        bool anyps = false;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int zc = z + k < ntz ? z + k : ntz - 1;
                            ahz[k] = ldg(g.azh, zc);
                            bhz[k] = ldg(g.bzh, zc);
                            anyps |= ahz[k] != 0.f;
                        }
                        float4 pso = zero4();
                        if (z < g.fz0 || z + 3 > g.fz1) pso = ld4((gcfl *)I.psz[vo], o);
                        float4 psn = pso;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int zk = z + k;
                            const bool in = zk >= LC - 1 && zk <= ntz - LC - 1;
                            float d = 0.f;
#pragma unroll
                            for (int ic = 0; ic < LC; ic++)
                                d = __builtin_fmaf(g.rc[ic] * (a[4 * R + k + ic + 1] - a[4 * R + k - ic]), idz, d);
                            const float ps = ahz[k] != 0.f ? __builtin_fmaf(bhz[k], elc(pso, k), ahz[k] * d) : 0.f;
                            el(psn, k) = (in && ahz[k] != 0.f) ? ps : elc(pso, k);
                            const float nv = __builtin_fmaf(elc(B, k), d + ps, elc(v, k));
                            el(v, k) = in ? nv : elc(v, k);
                        }
                        if (own && anyps) st4((gfl *)I.psz[vn], o, psn);
                    }
                }
                if (own) st4((gfl *)I.vz[vn], o, v);
            }
        }
        sVZ[c * VZQ + qq + R] = v;
    }
    __syncthreads();

    if constexpr (PF == 1) {
        // the receiver tile's previous p (staged as above), vx and vz, straight into the second LDS set,
        // and its image accumulators; the model stays in registers for it
        gcfl *RPp = (gcfl *)rcv->p[po], *RVX = (gcfl *)rcv->vx[po], *RVZ = (gcfl *)rcv->vz[po];
        const int wv = tid & ~63;
#pragma unroll
        for (int r = 0; r < I0; r++) {
            const int i = tid + NT * r, col = i / NPQ, qq = i - col * NPQ - RP;
            const bool need = i < N0 && ((qq >= 0 && qq < TQ) || (col >= OWN && col < OWN + TX));
            if (need) ld4_lds(RPp, static_cast<unsigned>(X0 - OWN + col) * ldz + (Z0 + 4 * qq), sP2 + NT * r + wv);
        }
#pragma unroll
        for (int r = 0; r < I1; r++) {
            const int i = tid + NT * r, q = i % TQ, cx = i / TQ;
            if (i < N1) ld4_lds(RVX, static_cast<unsigned>(X0 - LC + cx) * ldz + (Z0 + 4 * q), sVXo + NT * r + wv);
            pf->bx[r] = bxv[r];
        }
#pragma unroll
        for (int r = 0; r < I2; r++) {
            const int i = tid + NT * r;
            int c, qq;
            vz_item<TX, R, TQ>(i, c, qq);
            if (i < N2) ld4_lds(RVZ, static_cast<unsigned>(X0 + c) * ldz + (Z0 + 4 * qq), sVZo + NT * r + wv);
            pf->bz[r] = bzv[r];
        }
#pragma unroll
        for (int r = 0; r < I3; r++) {
            const int i = tid + NT * r, q = i % TQ, c = i / TQ;
            pf->k[r] = kv[r];
            pf->im[r] = pf->is[r] = pf->ir[r] = zero4();
            if (i < N3 && rcv->image) {
                const unsigned o = static_cast<unsigned>(X0 + c) * ldz + (Z0 + 4 * q);
                pf->im[r] = ld4((gcfl *)rcv->image, o);
                pf->is[r] = ld4((gcfl *)rcv->illum_s, o);
                pf->ir[r] = ld4((gcfl *)rcv->illum_r, o);
            }
        }
    }

    // ---- pressure of the tile's own cells from the velocities in LDS
#pragma unroll
    for (int r = 0; r < I3; r++) {
        const int i = tid + NT * r;
        if (i >= N3) break;
        const int q = i % TQ, c = i / TQ;
        const int x = X0 + c, z = Z0 + 4 * q;
        if (!FULL && (x >= ntx || z >= ntz)) continue;
        const unsigned o = static_cast<unsigned>(x) * ldz + z;
        const float4 pold = sP[(c + OWN) * NPQ + q + RP];
        const float4 K = kv[r];
        float4 dvx = zero4();
        float w[4 * (2 * R + 1)];   // vz rows z-4R .. z+4R+3
        const float sx = srcm ? g.idx : idx;
#define FSTEP_DERIVS                                                                                                    \
    {                                                                                                                   \
        _Pragma("unroll") for (int ic = 0; ic < LC; ic++)                                                               \
        {                                                                                                               \
            const float4 a = sVX[(c + LC + ic) * TQ + q], b = sVX[(c + LC - ic - 1) * TQ + q];                          \
            _Pragma("unroll") for (int k = 0; k < 4; k++) el(dvx, k) =                                                  \
                __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), sx, el(dvx, k));                                     \
        }                                                                                                               \
        _Pragma("unroll") for (int j = 0; j <= 2 * R; j++)                                                              \
        {                                                                                                               \
            const float4 t = sVZ[c * VZQ + q + j];                                                                      \
            _Pragma("unroll") for (int k = 0; k < 4; k++) w[4 * j + k] = elc(t, k);                                     \
        }                                                                                                               \
    }
        const bool colint = FULL || (x >= pmlc && x < ntx - pmlc);
        if (srcm) {
            const bool rowint = FULL || (z + 3 >= pmlc && z < ntz - pmlc);
            const bool hassrc = x == I.s_ix && I.s_iz >= z && I.s_iz < z + 4;
            if (!(colint && rowint) && !hassrc) continue;
            float rst[4];   // frame values of step it, read before the derivatives (RK:575-620)
            int fr[4];
            if (!FULL) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    fr[k] = frame_index(g, x, z + k);
                    rst[k] = fr[k] >= 0 ? ((gcfl *)I.strips)[static_cast<size_t>(it) * g.nframe + fr[k]] : 0.f;
                }
            }
            float4 pv = pold;
            if (colint && rowint) {
                FSTEP_DERIVS
                const float4 K0 = imv[r];
                float4 dp = zero4();
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int zk = z + k;
                    const bool in = FULL || (zk >= pmlc && zk < ntz - pmlc);
                    float dvz = 0.f;
#pragma unroll
                    for (int ic = 0; ic < LC; ic++)
                        dvz = __builtin_fmaf(g.rc[ic] * (w[4 * R + k + ic] - w[4 * R + k - ic - 1]), g.idz, dvz);
                    const float S = elc(dvx, k) + dvz;
                    const float np = __builtin_fmaf(-elc(K, k), S, elc(pv, k));
                    el(pv, k) = in ? np : elc(pv, k);
                    el(dp, k) = in ? elc(K0, k) * S : 0.f;
                }
                dpa[r] = dp;
                if (I.dpdt) st4((gfl *)I.dpdt, o, dp);   // for the up/down imaging that follows the step
                if constexpr (CIG) {
                    // the reconstruction at step it+1 for the receiver half's angle gathers:
                    // p as staged, the velocities just updated to that level
                    cgk[0][r] = pold;
                    cgk[1][r] = sVX[(c + LC) * TQ + q];
                    cgk[2][r] = sVZ[c * VZQ + q + R];
                }
            }
            if (hassrc) add_at(pv, I.s_iz - z, -ldg(I.wavelet, it + 1));   // RK:478-481
            if (!FULL) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (fr[k] >= 0) el(pv, k) = rst[k];   // RK:575-620
            }
            st4((gfl *)I.p[pn], o, pv);
            continue;
        }
        float4 pv = pold;
        if (FULL || (x >= LC && x <= ntx - LC)) {
            FSTEP_DERIVS
            if (FULL) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float dvz = 0.f;
#pragma unroll
                    for (int ic = 0; ic < LC; ic++)
                        dvz = __builtin_fmaf(g.rc[ic] * (w[4 * R + k + ic] - w[4 * R + k - ic - 1]), idz, dvz);
                    float S = elc(dvx, k) + dvz;
                    if constexpr (BORN) born_sum(g, I, x, z + k, dpa[r], k, S);   // the tile is interior
                    el(pv, k) = __builtin_fmaf(elc(K, k), S, elc(pold, k));
                }
            } else {
                const float ax = ldg(g.ax, x), bx = ldg(g.bx, x);
                float azr[4], bzr[4];
                bool anyz = false;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int zc = z + k < ntz ? z + k : ntz - 1;
                    azr[k] = ldg(g.az, zc);
                    bzr[k] = ldg(g.bz, zc);
                    anyz |= azr[k] != 0.f;
                }
                float4 qxo = zero4(), qzo = zero4();
                if (x < g.fx0 || x > g.fx1) qxo = ld4((gcfl *)I.pvx, o);
                if (z < g.fz0 || z + 3 > g.fz1) qzo = ld4((gcfl *)I.pvz, o);
                float4 qxn = qxo, qzn = qzo;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int zk = z + k;
                    const bool in = zk >= LC && zk <= ntz - LC;
                    float dvz = 0.f;
#pragma unroll
                    for (int ic = 0; ic < LC; ic++)
                        dvz = __builtin_fmaf(g.rc[ic] * (w[4 * R + k + ic] - w[4 * R + k - ic - 1]), idz, dvz);
                    const float qx = ax != 0.f ? __builtin_fmaf(bx, elc(qxo, k), ax * elc(dvx, k)) : 0.f;
                    const float qz = azr[k] != 0.f ? __builtin_fmaf(bzr[k], elc(qzo, k), azr[k] * dvz) : 0.f;
                    el(qxn, k) = (in && ax != 0.f) ? qx : elc(qxo, k);
                    el(qzn, k) = (in && azr[k] != 0.f) ? qz : elc(qzo, k);
                    float S = ((elc(dvx, k) + qx) + dvz) + qz;
                    if constexpr (BORN)
                        if (colint && zk >= pmlc && zk < ntz - pmlc) born_sum(g, I, x, zk, dpa[r], k, S);
                    const float np = __builtin_fmaf(elc(K, k), S, elc(pold, k));
                    el(pv, k) = in ? np : elc(pold, k);
                }
                if (ax != 0.f) st4((gfl *)I.pvx, o, qxn);
                if (anyz) st4((gfl *)I.pvz, o, qzn);
            }
        }
#undef FSTEP_DERIVS
        if (!neg && I.wavelet && x == I.s_ix && I.s_iz >= z && I.s_iz < z + 4) add_at(pv, I.s_iz - z, ldg(I.wavelet, it));   // RK:175-178
        {
            const int cm = cmv[r];
            if (cm >= 0) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (!FULL && z + k >= ntz) break;
                    const int tr = ldgi(I.rowmap, cm + z + k);
                    if (tr >= 0) {
                        const size_t go = static_cast<size_t>(tr) * g.nt + it;
                        if (neg) el(pv, k) = ((gcfl *)I.adj)[go];                  // hard replacement (RK:765-780)
                        else if (I.gather) ((gfl *)I.gather)[go] = elc(pv, k);     // RK:182-197
                    }
                }
            }
        }
        st4((gfl *)I.p[pn], o, pv);
        if (!neg) {
            if (!FULL && I.strips) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int f = frame_index(g, x, z + k);
                    if (f >= 0) ((gfl *)I.strips)[static_cast<size_t>(it) * g.nframe + f] = elc(pv, k);   // RK:200-219
                }
            }
            if (I.capture) {
                const int cc = ldgi(I.colcap, x);
                if (cc >= 0) {
                    gfl *dst = (gfl *)reinterpret_cast<float *>(I.capture + (static_cast<size_t>(it) * I.ncap + cc) * I.nfftz + z);
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (FULL || z + k < ntz) dst[2 * k + I.capcomp] = elc(pv, k);   // FK:324-327, 529-532
                }
            }
        } else if (I.image && colint) {
            // imaging condition of step it: the reconstruction's dp/dt and the receiver pressure (RK:785-817)
            const float4 dp = dpa[r];
            float4 im = imv[r], is = isv[r], ir = irv[r];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                const bool in = FULL || (zk >= pmlc && zk < ntz - pmlc);
                const float a = __builtin_fmaf(elc(dp, k), elc(pv, k), elc(im, k));
                const float b = __builtin_fmaf(elc(dp, k), elc(dp, k), elc(is, k));
                const float c2 = __builtin_fmaf(elc(pv, k), elc(pv, k), elc(ir, k));
                el(im, k) = in ? a : elc(im, k);
                el(is, k) = in ? b : elc(is, k);
                el(ir, k) = in ? c2 : elc(ir, k);
            }
            st4((gfl *)I.image, o, im);
            st4((gfl *)I.illum_s, o, is);
            st4((gfl *)I.illum_r, o, ir);
        }
        if constexpr (CIG) {
            // angle gathers of step it+1 (RK:859-901) from both fields at that level: the
            // reconstruction kept from the source half, the receiver field's previous slot
            if (neg && !first && I.cig && colint) {
                const float4 vxr = ld4((gcfl *)I.vx[po], o), vzr = ld4((gcfl *)I.vz[po], o);
                const float4 ps = cgk[0][r], vxs = cgk[1][r], vzs = cgk[2][r];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int zk = z + k;
                    if (!(FULL || (zk >= pmlc && zk < ntz - pmlc))) continue;
                    float F, a, b;
                    cig_products(elc(ps, k), elc(vxs, k), elc(vzs, k), elc(pold, k), elc(vxr, k), elc(vzr, k), F, a, b);
                    if (!cig_may_bin(F, a, b, I.cthr)) continue;
                    float *bins = I.cig + (static_cast<size_t>(x - pmlc) * g.nz + (zk - pmlc)) * I.thetan;
                    cig_bins(F, cig_norm(a, b), elc(ps, k) * elc(pold, k), I.thetan, I.cw1, I.cw2, I.cthr, bins);
                }
            }
        }
    }
}

// PAIR = false: one FWD wavefield per instance.  PAIR = true: instances
// 2k (SRC) and 2k+1 (RCV) of one shot share each workgroup; with BORN, 2k is
// the background and 2k+1 the scattered field of a Born forward (both FWD).
// WB > 0 holds the kernel to at least WB wavefronts per SIMD (a register
// bound the tile tuner tries).  Without it, the angle-gather variant at Lc 4
// and tiles up to 32 wide is held to four wavefronts per SIMD (128 VGPRs,
// 8-12 bytes of spill, faster than three wavefronts); elsewhere that bound
// would spill heavily.
template <int LC, int TX, int TQ, int WB, bool PAIR, bool CIG, bool BORN>
__global__ __launch_bounds__((fstep_threads<TX, TQ>()), (WB > 0 ? WB : (CIG && LC == 4 && TX <= 32) ? 4 : 1)) void k_fstep(
    KGrid g, const KStep *__restrict__ insts, int ninst, int nzb, int nxb, int it, int first)
{
    using G = FGeo<LC, TX, TQ>;
    constexpr int ROWS = 4 * TQ;
    __shared__ float4 sP[G::NPC * G::NPQ];
    __shared__ float4 sVX[G::NVX * TQ];
    __shared__ float4 sVZ[TX * G::VZQ];
    // the second set of a fast plain pair: the receiver tile's prefetched p, vx, vz (up to Lc 4, where
    // two sets leave the occupancy the registers allow)
    constexpr bool PFV = PAIR && !CIG && !BORN && LC <= 4;
    __shared__ float4 sP2[PFV ? G::NPC * G::NPQ : 1];
    __shared__ float4 sVXo[PFV ? G::NVX * TQ : 1];
    __shared__ float4 sVZo[PFV ? TX * G::VZQ : 1];

    const int L = xcd_block();
    if (L >= ninst * nzb * nxb) return;
    const int inst = L % ninst;
    const int zb = (L / ninst) % nzb, xb = L / (ninst * nzb);
    const int X0 = xb * TX, Z0 = zb * ROWS;
    const bool full = X0 - G::OWN >= g.fx0 && X0 + TX + 2 * LC - 2 <= g.fx1 && Z0 - 4 * G::RP >= g.fz0 &&
                      Z0 + ROWS + 4 * G::RP - 1 <= g.fz1;
    float4 dpa[G::NOWN], cgk[3][G::NOWN];
#pragma unroll
    for (int r = 0; r < G::NOWN; r++) dpa[r] = cgk[0][r] = cgk[1][r] = cgk[2][r] = zero4();
    if (!PAIR) {
        const KStep &I = insts[inst];
        if (full) fstep_tile<LC, TX, TQ, true, false>(g, I, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        else fstep_tile<LC, TX, TQ, false, false>(g, I, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        return;
    }
    const KStep &S = insts[2 * inst], &Rv = insts[2 * inst + 1];
    if constexpr (BORN) {
        // Born: background, then the scattered field with the background's sums in dpa
        if (full) fstep_tile<LC, TX, TQ, true, false, true>(g, S, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        else fstep_tile<LC, TX, TQ, false, false, true>(g, S, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        __syncthreads();   // LDS is staged again for the scattered field
        if (full) fstep_tile<LC, TX, TQ, true, false, true>(g, Rv, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        else fstep_tile<LC, TX, TQ, false, false, true>(g, Rv, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        return;
    }
    if constexpr (PFV) {
        if (full) {
            // a fast tile is interior: the reconstruction always runs, and prefetches the receiver tile
            FPre<LC, TX, TQ> pf;
            fstep_tile<LC, TX, TQ, true, false, false, 1>(g, S, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk, &pf, &Rv,
                                                          sP2, sVXo, sVZo);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this wavefront's direct-to-LDS loads
            __syncthreads();
            fstep_tile<LC, TX, TQ, true, false, false, 2>(g, Rv, X0, Z0, it, first, sP2, sVX, sVZ, dpa, cgk, &pf, nullptr,
                                                          nullptr, sVXo, sVZo);
            return;
        }
    }
    // the reconstruction writes only interior cells and the source cell
    const bool any = X0 + TX > g.pmlc && X0 < g.ntx - g.pmlc && Z0 + ROWS > g.pmlc && Z0 < g.ntz - g.pmlc;
    const bool src = S.s_ix >= X0 && S.s_ix < X0 + TX && S.s_iz >= Z0 && S.s_iz < Z0 + ROWS;
    if (any || src) {
        if (full) fstep_tile<LC, TX, TQ, true, CIG>(g, S, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        else fstep_tile<LC, TX, TQ, false, CIG>(g, S, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
        __syncthreads();   // LDS is staged again for the receiver wavefield
    }
    if (full) fstep_tile<LC, TX, TQ, true, CIG>(g, Rv, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
    else fstep_tile<LC, TX, TQ, false, CIG>(g, Rv, X0, Z0, it, first, sP, sVX, sVZ, dpa, cgk);
}

// Multi-grid launch: instances of different grid sizes (same Lc) in one
// launch, tiles 16 columns wide and 4 TQ rows high.  btab maps each logical
// workgroup to (instance, tile of that instance's grid, numbered z-fastest);
// each instance reads its own grid.
template <int LC, int TQ, int WB>
__global__ __launch_bounds__((fstep_threads<16, TQ>()), (WB > 0 ? WB : 1)) void k_fstep_multi(
    const KStep *__restrict__ insts, const int2 *__restrict__ btab, int nblk, int it)
{
    constexpr int TX = 16, ROWS = 4 * TQ;
    using G = FGeo<LC, TX, TQ>;
    __shared__ float4 sP[G::NPC * G::NPQ];
    __shared__ float4 sVX[G::NVX * TQ];
    __shared__ float4 sVZ[TX * G::VZQ];

    const int L = xcd_block();
    if (L >= nblk) return;
    const int2 bt = btab[L];
    const KStep &I = insts[bt.x];
    const KGrid &g = *I.grid;
    if (it >= g.nt) return;
    const int nzb = (g.ntz + ROWS - 1) / ROWS;
    const int zb = bt.y % nzb, xb = bt.y / nzb;
    const int X0 = xb * TX, Z0 = zb * ROWS;
    const bool full = X0 - G::OWN >= g.fx0 && X0 + TX + 2 * LC - 2 <= g.fx1 && Z0 - 4 * G::RP >= g.fz0 &&
                      Z0 + ROWS + 4 * G::RP - 1 <= g.fz1;
    float4 dpa[G::NOWN], cgk[3][G::NOWN];
#pragma unroll
    for (int r = 0; r < G::NOWN; r++) dpa[r] = zero4();
    if (full) fstep_tile<LC, TX, TQ, true, false>(g, I, X0, Z0, it, 0, sP, sVX, sVZ, dpa, cgk);
    else fstep_tile<LC, TX, TQ, false, false>(g, I, X0, Z0, it, 0, sP, sVX, sVZ, dpa, cgk);
}

// Launch of one k_fstep variant with tile configuration cfg (fstep_cfg) over
// n instances (pairs count once).  Every Lc has the 64-row tiles 16, 32 and 64
// wide; Lc 4 also 32- and 128-row tiles and the five-wavefront bound (for the
// forward and paired RTM variants).
template <bool PAIR, bool CIG, bool BORN>
hipError_t launch_fstep_variant(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s)
{
    const int tx = fstep_cfg_tx(cfg), tq = fstep_cfg_tq(cfg), wb = fstep_cfg_wb(cfg);
    const int nzb = (g.ntz + 4 * tq - 1) / (4 * tq);
    const int nxb = (g.ntx + tx - 1) / tx;
    const int nblocks = (n * nzb * nxb + 7) / 8 * 8;
#define LK(LV, TXV, TQV, WBV)                                                                                       \
    hipLaunchKernelGGL((k_fstep<LV, TXV, TQV, WBV, PAIR, CIG, BORN>), dim3(nblocks), dim3(FGeo<LV, TXV, TQV>::NT), 0, \
                       s, g, insts, n, nzb, nxb, it, first)
    if (tq == 16 && wb == 0) {
        switch (tx) {
        case 16: DISPATCH_LC(g.Lc, LK(L, 16, 16, 0)); break;
        case 32: DISPATCH_LC(g.Lc, LK(L, 32, 16, 0)); break;
        case 64: DISPATCH_LC(g.Lc, LK(L, 64, 16, 0)); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if constexpr (!CIG && !BORN) {
        if (g.Lc != 4 || (wb != 0 && wb != 5)) return hipErrorInvalidValue;
#define LW(TXV, TQV)                          \
    {                                         \
        if (wb == 5) LK(4, TXV, TQV, 5);      \
        else LK(4, TXV, TQV, 0);              \
    }
        switch (tq * 1000 + tx) {
        case 8016: LW(16, 8); break;
        case 8032: LW(32, 8); break;
        case 8064: LW(64, 8); break;
        case 16016: LK(4, 16, 16, 5); break;
        case 16032: LK(4, 32, 16, 5); break;
        case 16064: LK(4, 64, 16, 5); break;
        case 32016: LW(16, 32); break;
        case 32032: LW(32, 32); break;
        default: return hipErrorInvalidValue;
        }
#undef LW
        return hipGetLastError();
    }
#undef LK
    return hipErrorInvalidValue;
}

}  // namespace

// The variants, each compiled in its own translation unit (swd_fstep_*.hip).
hipError_t launch_fstep_fwd(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s);
hipError_t launch_fstep_pair(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s);
hipError_t launch_fstep_pair_cig(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s);
hipError_t launch_fstep_born(const KGrid &g, const KStep *insts, int n, int it, int first, int cfg, hipStream_t s);
