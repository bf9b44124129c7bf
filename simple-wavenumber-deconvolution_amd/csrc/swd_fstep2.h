// swd_fstep2.h -- fused leapfrog kernel that advances every wavefield of a
// batch by TWO time steps per launch (temporal blocking), Lc 4 only.
//
// Why: one step per launch reads and writes p, vx, vz, the CPML memories and
// (backward) the three image accumulators once per step, and the step is bound
// by those bytes.  Two steps per launch read and write them once per two
// steps; the price is recomputing a halo ring of the first step.
//
// A workgroup owns TX columns x 4 TQ rows (TQ float4 quads per column) of one
// wavefield.  With vx(x) reading p(x-3 .. x+4) and p(x) reading vx(x-4 .. x+3)
// (the same along z), level t+2 of the own tile needs level t+1 of p on a
// cross (own rows, columns -7 .. TX+6; own columns, quads -2 .. TQ+1), which
// needs velocities of level t+1/2 on a wider cross and p of level t on the
// widest one (columns -14 .. TX+13 on the own rows, quads -4 .. TQ+3 on the
// own columns).  The launch
//   1. stages p of level t on that cross in LDS and loads the old velocities,
//      the model and (frame tiles) the CPML memories of every cell it updates;
//   2. step 1: the velocities, then the pressure of the ring and the own
//      tile (the "core" items come first in each region's item order, so the
//      thread that updates a core cell in step 1 updates it again in step 2
//      and keeps its values in registers);
//   3. step 2: the velocities, then the pressure of the own tile, storing
//      them with the CPML memories, gathers, strips and image accumulators.
// Every cell update is the one-step kernel's (swd_fstep.h) operation for
// operation, so results are bit-identical to two one-step launches.
//
// Slots: a launch reads slot sa of p (sv of the velocities, sv = 1 - sa for
// the reconstruction, else sa) and writes the other slot, so no workgroup
// overwrites what a neighbour still reads; the pressure CPML memories are
// double-buffered too (pvx, pvz and pvxb, pvzb).  Launches come in pairs (ph =
// 0, 1): after a pair, level L is again in slot L & 1 and the pressure CPML
// memories in pvx, pvz, as the one-step kernel expects.
//
// Modes: FWD (forward with source, receivers, strips, AFor column capture),
// and the paired RTM backward: SRC (reconstruction with wavelet subtraction
// and frame restore; dp/dt of both levels stored for the up/down imaging if
// asked) then RCV (receiver back-propagation with hard replacement and the
// imaging condition; the first level's pressure stored if asked).  The MID
// variant of the paired kernel also stores both fields' first level (p and
// velocities) for the angle gathers, which k_cig bins after the launch.  Not
// here: Born, the first reconstruction step (which keeps the forward's
// velocities).
#pragma once
#include <algorithm>

#include "swd_device.h"
#include "swd_kernels.h"

namespace {

__device__ __forceinline__ float4 f2zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }

template <int TX, int TQ, int NW>
struct T2 {
    static constexpr int LC = 4;
    static constexpr int NT = 64 * NW;   // threads; thread t < TX TQ owns quad t
    static_assert(NT >= TX * TQ && NT <= 1024, "a thread for every own quad");
    // LDS boxes, column-major (a column's quads contiguous)
    static constexpr int PNQ = TQ + 8, PNC = TX + 28;   // sP: columns -14 .. TX+13, quads -4 .. TQ+3
    static constexpr int XNQ = TQ + 4, XNC = TX + 21;   // sVX: columns -11 .. TX+9, quads -2 .. TQ+1
    static constexpr int ZNQ = TQ + 6, ZNC = TX + 14;   // sVZ: columns -7 .. TX+6, quads -3 .. TQ+2
    // region item counts (core first)
    static constexpr int P0A = (TX + 28) * TQ, P0B = (TX + 14) * 4, P0C = TX * 4, NP0 = P0A + P0B + P0C;
    static constexpr int XCORE = (TX + 7) * TQ, XR1 = 14 * TQ, XR2 = (TX + 7) * 4, NX = XCORE + XR1 + XR2;
    static constexpr int ZOWN = TX * TQ, ZCORE = ZOWN + 2 * TX, ZR1A = 14 * TQ, ZR1 = 14 * (TQ + 2), ZR2 = TX * 4,
                         NZ = ZCORE + ZR1 + ZR2;
    static constexpr int PCORE = TX * TQ, PR1 = 14 * TQ, PR2 = TX * 4, NP = PCORE + PR1 + PR2;
    // rounds of NT items
    static constexpr int RP0 = (NP0 + NT - 1) / NT, RX = (NX + NT - 1) / NT, RZ = (NZ + NT - 1) / NT,
                         RP = (NP + NT - 1) / NT;
    static constexpr int RXC = (XCORE + NT - 1) / NT, RZC = (ZCORE + NT - 1) / NT;
    static constexpr int sp(int c, int q) { return (c + 14) * PNQ + q + 4; }
    static constexpr int sx(int c, int q) { return (c + 11) * XNQ + q + 2; }
    static constexpr int sz(int c, int q) { return (c + 7) * ZNQ + q + 3; }
};

// item -> (column, quad) relative to the tile origin, per region
template <int TX, int TQ>
__device__ __forceinline__ void f2_p0(int i, int &c, int &q)
{
    using G = T2<TX, TQ, 16>;   // region sizes do not depend on the thread count
    if (i < G::P0A) {
        c = i / TQ - 14;
        q = i % TQ;
    } else if (i < G::P0A + G::P0B) {
        const int j = i - G::P0A, h = j & 3;
        c = (j >> 2) - 7;
        q = h < 2 ? h - 2 : TQ + h - 2;
    } else {
        const int j = i - G::P0A - G::P0B, h = j & 3;
        c = j >> 2;
        q = h < 2 ? h - 4 : TQ + h;
    }
}
template <int TX, int TQ>
__device__ __forceinline__ void f2_vx(int i, int &c, int &q)
{
    using G = T2<TX, TQ, 16>;   // region sizes do not depend on the thread count
    if (i < G::XCORE) {
        c = i / TQ - 4;
        q = i % TQ;
    } else if (i < G::XCORE + G::XR1) {
        const int j = i - G::XCORE, cc = j / TQ;
        c = cc < 7 ? cc - 11 : TX + cc - 4;
        q = j % TQ;
    } else {
        const int j = i - G::XCORE - G::XR1, h = j & 3;
        c = (j >> 2) - 4;
        q = h < 2 ? h - 2 : TQ + h - 2;
    }
}
template <int TX, int TQ>
__device__ __forceinline__ void f2_vz(int i, int &c, int &q)
{
    using G = T2<TX, TQ, 16>;   // region sizes do not depend on the thread count
    if (i < G::ZOWN) {
        c = i / TQ;
        q = i % TQ;
    } else if (i < G::ZCORE) {
        const int j = i - G::ZOWN;
        c = j >> 1;
        q = (j & 1) ? TQ : -1;
    } else if (i < G::ZCORE + G::ZR1A) {
        const int j = i - G::ZCORE, cc = j / TQ;
        c = cc < 7 ? cc - 7 : TX + cc - 7;
        q = j % TQ;
    } else if (i < G::ZCORE + G::ZR1) {
        const int j = i - G::ZCORE - G::ZR1A, cc = j >> 1;
        c = cc < 7 ? cc - 7 : TX + cc - 7;
        q = (j & 1) ? TQ : -1;
    } else {
        const int j = i - G::ZCORE - G::ZR1, h = j & 3;
        c = j >> 2;
        q = h < 2 ? h - 3 : TQ + h - 1;
    }
}
template <int TX, int TQ>
__device__ __forceinline__ void f2_p(int i, int &c, int &q)
{
    using G = T2<TX, TQ, 16>;   // region sizes do not depend on the thread count
    if (i < G::PCORE) {
        c = i / TQ;
        q = i % TQ;
    } else if (i < G::PCORE + G::PR1) {
        const int j = i - G::PCORE, cc = j / TQ;
        c = cc < 7 ? cc - 7 : TX + cc - 7;
        q = j % TQ;
    } else {
        const int j = i - G::PCORE - G::PR1, h = j & 3;
        c = j >> 2;
        q = h < 2 ? h - 2 : TQ + h - 2;
    }
}

// Quads q-1, q, q+1 of an LDS column as 12 consecutive rows (whole 16-byte reads).
__device__ __forceinline__ void f2_window(const float4 *s, float (&a)[12])
{
    v4f t[3];
#pragma unroll
    for (int j = 0; j < 3; j++) t[j] = *reinterpret_cast<const v4f *>(s + j);
#pragma unroll
    for (int j = 0; j < 3; j++) asm volatile("" ::"v"(t[j]));
#pragma unroll
    for (int j = 0; j < 3; j++) {
        a[4 * j] = t[j].x;
        a[4 * j + 1] = t[j].y;
        a[4 * j + 2] = t[j].z;
        a[4 * j + 3] = t[j].w;
    }
}

__device__ __forceinline__ void f2_add_at(float4 &v, int k, float w)
{
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (j == k) el(v, j) += w;
}

// Velocity x of one quad from the staged pressure (pa: p at x+1 on the
// quad's rows, column stride PNQ; pb = pa - PNQ: p at x).  Returns the new
// value; psn is the new CPML memory (FWD/RCV frame tiles).
template <bool FULL, int MODE, int PNQ>
__device__ __forceinline__ float4 f2_vx_quad(const KGrid &g, int x, int z, const float4 *pa, float4 v, float4 B,
                                             float4 pso, float4 &psn, float &ah_out)
{
    constexpr int LC = 4;
    psn = pso;
    ah_out = 0.f;
    if (MODE == KSTEP_SRC) {
        const bool upd = FULL || (x >= g.pmlc && x < g.ntx - g.pmlc && z + 3 >= g.pmlc && z < g.ntz - g.pmlc);
        if (upd) {
            float4 d = f2zero();
#pragma unroll
            for (int ic = 0; ic < LC; ic++) {
                const float4 a = pa[ic * PNQ], b = pa[-(ic + 1) * PNQ];
#pragma unroll
                for (int k = 0; k < 4; k++) el(d, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), g.idx, el(d, k));
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float nv = __builtin_fmaf(-elc(B, k), elc(d, k), elc(v, k));
                el(v, k) = (FULL || (z + k >= g.pmlc && z + k < g.ntz - g.pmlc)) ? nv : elc(v, k);
            }
        }
        return v;
    }
    constexpr bool NEG = MODE == KSTEP_RCV;
    const float idx = NEG ? -g.idx : g.idx;
    if (FULL || (x >= LC - 1 && x <= g.ntx - LC - 1)) {
        float4 d = f2zero();
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float4 a = pa[ic * PNQ], b = pa[-(ic + 1) * PNQ];
#pragma unroll
            for (int k = 0; k < 4; k++) el(d, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), idx, el(d, k));
        }
        if (FULL) {
#pragma unroll
            for (int k = 0; k < 4; k++) el(v, k) = __builtin_fmaf(elc(B, k), elc(d, k), elc(v, k));
        } else {
            const float ah = ldg(g.axh, x), bh = ldg(g.bxh, x);
            ah_out = ah;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                const bool in = NEG ? zk < g.ntz : (zk >= LC && zk <= g.ntz - LC);
                const float ps = ah != 0.f ? __builtin_fmaf(bh, elc(pso, k), ah * elc(d, k)) : 0.f;
                el(psn, k) = in ? ps : elc(pso, k);
                const float nv = __builtin_fmaf(elc(B, k), elc(d, k) + ps, elc(v, k));
                el(v, k) = in ? nv : elc(v, k);
            }
        }
    }
    return v;
}

// Velocity z of one quad from the staged pressure window a (rows z-4 .. z+7).
template <bool FULL, int MODE>
__device__ __forceinline__ float4 f2_vz_quad(const KGrid &g, int x, int z, const float (&a)[12], float4 v, float4 B,
                                             float4 pso, float4 &psn, bool &anyps)
{
    constexpr int LC = 4;
    psn = pso;
    anyps = false;
    if (MODE == KSTEP_SRC) {
        const bool upd = FULL || (x >= g.pmlc && x < g.ntx - g.pmlc && z + 3 >= g.pmlc && z < g.ntz - g.pmlc);
        if (upd) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float d = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) d = __builtin_fmaf(g.rc[ic] * (a[4 + k + ic + 1] - a[4 + k - ic]), g.idz, d);
                const float nv = __builtin_fmaf(-elc(B, k), d, elc(v, k));
                el(v, k) = (FULL || (z + k >= g.pmlc && z + k < g.ntz - g.pmlc)) ? nv : elc(v, k);
            }
        }
        return v;
    }
    constexpr bool NEG = MODE == KSTEP_RCV;
    const float idz = NEG ? -g.idz : g.idz;
    if (FULL || NEG || (x >= LC && x <= g.ntx - LC)) {
        if (FULL) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float d = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) d = __builtin_fmaf(g.rc[ic] * (a[4 + k + ic + 1] - a[4 + k - ic]), idz, d);
                el(v, k) = __builtin_fmaf(elc(B, k), d, elc(v, k));
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                const int zc = zk < g.ntz ? zk : g.ntz - 1;
                const float ahz = ldg(g.azh, zc), bhz = ldg(g.bzh, zc);
                anyps |= ahz != 0.f;
                const bool in = zk >= LC - 1 && zk <= g.ntz - LC - 1;
                float d = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) d = __builtin_fmaf(g.rc[ic] * (a[4 + k + ic + 1] - a[4 + k - ic]), idz, d);
                const float ps = ahz != 0.f ? __builtin_fmaf(bhz, elc(pso, k), ahz * d) : 0.f;
                el(psn, k) = (in && ahz != 0.f) ? ps : elc(pso, k);
                const float nv = __builtin_fmaf(elc(B, k), d + ps, elc(v, k));
                el(v, k) = in ? nv : elc(v, k);
            }
        }
    }
    return v;
}

// Pressure of one quad at level lv from the velocities in LDS (vxa: vx at x
// on the quad's rows, column stride XNQ; vzw: rows z-4 .. z+7 of vz).  SRC:
// dp/dt into dp, the wavelet subtraction and the frame restore; FWD: the
// source; RCV: the hard replacement (rec: whether the receivers of this cell
// are replaced here).  qxo/qzo in, qxn/qzn out: the CPML memories (frame
// tiles).  Returns false where the reconstruction leaves the quad unchanged.
// srcnear: the source may lie in the quad (a workgroup-uniform test of the
// tile's region; false spares the per-quad comparison).
// BRN 2 (the Born pair's scattered field): sbv holds the background's sums S
// of the same quad and level; the Born source m S is added to the own sum on
// interior cells, in the one-step pair kernel's operation order.
template <bool FULL, int MODE, int XNQ, int BRN = 0>
__device__ __forceinline__ bool f2_p_quad(const KGrid &g, const KStep &I, int x, int z, int lv, const float4 *vxa,
                                          const float (&vzw)[12], float4 pold, float4 K, float4 K0, float4 qxo,
                                          float4 qzo, float4 &pv, float4 &dp, float4 &qxn, float4 &qzn, float &axo,
                                          bool &anyz, bool srcnear, float4 sbv = make_float4(0.f, 0.f, 0.f, 0.f))
{
    constexpr int LC = 4;
    pv = pold;
    qxn = qxo;
    qzn = qzo;
    axo = 0.f;
    anyz = false;
    float4 dvx = f2zero();
    if (MODE == KSTEP_SRC) {
        const bool colint = FULL || (x >= g.pmlc && x < g.ntx - g.pmlc);
        const bool rowint = FULL || (z + 3 >= g.pmlc && z < g.ntz - g.pmlc);
        const bool hassrc = srcnear && x == I.s_ix && I.s_iz >= z && I.s_iz < z + 4;
        if (!(colint && rowint) && !hassrc) return false;
        float rst[4];
        int fr[4];
        if (!FULL) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                fr[k] = frame_index(g, x, z + k);
                rst[k] = fr[k] >= 0 ? ((gcfl *)I.strips)[static_cast<size_t>(lv) * g.nframe + fr[k]] : 0.f;
            }
        }
        if (colint && rowint) {
#pragma unroll
            for (int ic = 0; ic < LC; ic++) {
                const float4 a = vxa[ic * XNQ], b = vxa[-(ic + 1) * XNQ];
#pragma unroll
                for (int k = 0; k < 4; k++) el(dvx, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), g.idx, el(dvx, k));
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                const bool in = FULL || (zk >= g.pmlc && zk < g.ntz - g.pmlc);
                float dvz = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) dvz = __builtin_fmaf(g.rc[ic] * (vzw[4 + k + ic] - vzw[4 + k - ic - 1]), g.idz, dvz);
                const float S = elc(dvx, k) + dvz;
                const float np = __builtin_fmaf(-elc(K, k), S, elc(pv, k));
                el(pv, k) = in ? np : elc(pv, k);
                el(dp, k) = in ? elc(K0, k) * S : 0.f;
            }
        }
        if (hassrc) f2_add_at(pv, I.s_iz - z, -ldg(I.wavelet, lv + 1));   // RK:478-481
        if (!FULL) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (fr[k] >= 0) el(pv, k) = rst[k];   // RK:575-620
        }
        return true;
    }
    constexpr bool NEG = MODE == KSTEP_RCV;
    const float idx = NEG ? -g.idx : g.idx, idz = NEG ? -g.idz : g.idz;
    if (FULL || (x >= LC && x <= g.ntx - LC)) {
#pragma unroll
        for (int ic = 0; ic < LC; ic++) {
            const float4 a = vxa[ic * XNQ], b = vxa[-(ic + 1) * XNQ];
#pragma unroll
            for (int k = 0; k < 4; k++) el(dvx, k) = __builtin_fmaf(g.rc[ic] * (elc(a, k) - elc(b, k)), idx, el(dvx, k));
        }
        if (FULL) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float dvz = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) dvz = __builtin_fmaf(g.rc[ic] * (vzw[4 + k + ic] - vzw[4 + k - ic - 1]), idz, dvz);
                float S = elc(dvx, k) + dvz;
                if (MODE == KSTEP_FWD) el(dp, k) = S;
                if constexpr (BRN == 2)   // the tile is interior
                    S = __builtin_fmaf(ldg(I.born_m, static_cast<unsigned>((x - g.pmlc) * g.nz + (z + k - g.pmlc))),
                                       elc(sbv, k), S);
                el(pv, k) = __builtin_fmaf(elc(K, k), S, elc(pold, k));
            }
        } else {
            const float ax = ldg(g.ax, x), bx = ldg(g.bx, x);
            axo = ax;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int zk = z + k;
                const int zc = zk < g.ntz ? zk : g.ntz - 1;
                const float az = ldg(g.az, zc), bz = ldg(g.bz, zc);
                anyz |= az != 0.f;
                const bool in = zk >= LC && zk <= g.ntz - LC;
                float dvz = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) dvz = __builtin_fmaf(g.rc[ic] * (vzw[4 + k + ic] - vzw[4 + k - ic - 1]), idz, dvz);
                const float qx = ax != 0.f ? __builtin_fmaf(bx, elc(qxo, k), ax * elc(dvx, k)) : 0.f;
                const float qz = az != 0.f ? __builtin_fmaf(bz, elc(qzo, k), az * dvz) : 0.f;
                el(qxn, k) = (in && ax != 0.f) ? qx : elc(qxo, k);
                el(qzn, k) = (in && az != 0.f) ? qz : elc(qzo, k);
                float S = ((elc(dvx, k) + qx) + dvz) + qz;
                if (MODE == KSTEP_FWD) el(dp, k) = S;
                if constexpr (BRN == 2)
                    if (x >= g.pmlc && x < g.ntx - g.pmlc && zk >= g.pmlc && zk < g.ntz - g.pmlc)
                        S = __builtin_fmaf(ldg(I.born_m, static_cast<unsigned>((x - g.pmlc) * g.nz + (zk - g.pmlc))),
                                           elc(sbv, k), S);
                const float np = __builtin_fmaf(elc(K, k), S, elc(pold, k));
                el(pv, k) = in ? np : elc(pold, k);
            }
        }
    }
    if (!NEG && srcnear && I.wavelet && x == I.s_ix && I.s_iz >= z && I.s_iz < z + 4)
        f2_add_at(pv, I.s_iz - z, ldg(I.wavelet, lv));   // RK:175-178
    return true;
}

// Receivers of one quad at level lv: RCV replaces (hard replacement,
// RK:765-780), FWD records (RK:182-197).
template <bool FULL, int MODE>
__device__ __forceinline__ void f2_receivers(const KGrid &g, const KStep &I, int cm, int z, int lv, float4 &pv)
{
    if (cm < 0) return;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!FULL && z + k >= g.ntz) break;
        const int tr = ldgi(I.rowmap, cm + z + k);
        if (tr >= 0) {
            const size_t go = static_cast<size_t>(tr) * g.nt + lv;
            if (MODE == KSTEP_RCV) el(pv, k) = ((gcfl *)I.adj)[go];
            else if (I.gather) ((gfl *)I.gather)[go] = elc(pv, k);
        }
    }
}

// The divergence sums S of one own quad's interior cells at level lv (FWD
// with s0_out: the Born adjoint's background sums, as the batched k_p stores them).
template <bool FULL>
__device__ __forceinline__ void f2_s0(const KGrid &g, const KStep &I, int x, int z, int lv, float4 S)
{
    if (!FULL && (x < g.pmlc || x >= g.ntx - g.pmlc)) return;
    gfl *dst = (gfl *)I.s0_out + static_cast<size_t>(lv - I.s0_t0) * g.nx * g.nz + static_cast<size_t>(x - g.pmlc) * g.nz;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int zk = z + k;
        if (FULL || (zk >= g.pmlc && zk < g.ntz - g.pmlc)) dst[zk - g.pmlc] = elc(S, k);
    }
}

// AFor column capture of one own quad at level lv (FK:324-327, 529-532): the pressure of a
// capture column into its kz-transform row, real or imaginary part by the instance.
template <bool FULL>
__device__ __forceinline__ void f2_capture(const KGrid &g, const KStep &I, int x, int z, int lv, float4 pv)
{
    const int cc = ldgi(I.colcap, x);
    if (cc < 0) return;
    gfl *dst = (gfl *)reinterpret_cast<float *>(I.capture + (static_cast<size_t>(lv) * I.ncap + cc) * I.nfftz + z);
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (FULL || z + k < g.ntz) dst[2 * k + I.capcomp] = elc(pv, k);
}

// Imaging condition of one own quad at one level (RK:785-817).
template <bool FULL>
__device__ __forceinline__ void f2_image(const KGrid &g, int z, float4 dp, float4 pv, float4 &im, float4 &is, float4 &ir)
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int zk = z + k;
        const bool in = FULL || (zk >= g.pmlc && zk < g.ntz - g.pmlc);
        const float a = __builtin_fmaf(elc(dp, k), elc(pv, k), elc(im, k));
        const float b = __builtin_fmaf(elc(dp, k), elc(dp, k), elc(is, k));
        const float c = __builtin_fmaf(elc(pv, k), elc(pv, k), elc(ir, k));
        el(im, k) = in ? a : elc(im, k);
        el(is, k) = in ? b : elc(is, k);
        el(ir, k) = in ? c : elc(ir, k);
    }
}

// Two steps of one wavefield on one tile.  dps: dp/dt of the own quad at the
// two levels (SRC writes them, RCV images with them).
// BRN: 0, or the Born pair's background (1: its sums of the first level's P1
// items and the second level's own quad kept in sb) and scattered field (2:
// adds m S from sb); the same thread handles an item in both fields.
template <int TX, int TQ, int NW, bool FULL, int MODE, bool MID = false, int BRN = 0>
__device__ __forceinline__ void f2_field(const KGrid &g, const KStep &I, int X0, int Z0, int it, int ph,
                                         float4 *__restrict__ sP, float4 *__restrict__ sVX, float4 *__restrict__ sVZ,
                                         float4 (&dps)[2], float4 *sb = nullptr)
{
    using G = T2<TX, TQ, NW>;
    constexpr int NT = G::NT, PNQ = G::PNQ, XNQ = G::XNQ, ZNQ = G::ZNQ;
    constexpr bool SRC = MODE == KSTEP_SRC, NEG = MODE == KSTEP_RCV, CPML = !FULL && !SRC;
    const int tid = threadIdx.x, lane0 = tid & 63;
    const int ntx = g.ntx, ntz = g.ntz;
    const unsigned ldz = static_cast<unsigned>(g.ldz);
    const int dir = MODE == KSTEP_FWD ? 1 : -1;
    const int lv1 = it, lv2 = it + dir;
    const int sa = (it + 1 + ph) & 1, sv = SRC ? 1 - sa : sa;
    gcfl *Pr = (gcfl *)I.p[sa];
    gfl *Pw = (gfl *)I.p[1 - sa];
    gcfl *VXr = (gcfl *)I.vx[sv];
    gfl *VXw = (gfl *)I.vx[1 - sv];
    gcfl *VZr = (gcfl *)I.vz[sv];
    gfl *VZw = (gfl *)I.vz[1 - sv];
    gcfl *PSXr = (gcfl *)I.psx[sv];
    gfl *PSXw = (gfl *)I.psx[1 - sv];
    gcfl *PSZr = (gcfl *)I.psz[sv];
    gfl *PSZw = (gfl *)I.psz[1 - sv];
    gcfl *PVXr = (gcfl *)(ph ? I.pvxb : I.pvx);
    gfl *PVXw = (gfl *)(ph ? I.pvx : I.pvxb);
    gcfl *PVZr = (gcfl *)(ph ? I.pvzb : I.pvz);
    gfl *PVZw = (gfl *)(ph ? I.pvz : I.pvzb);
    auto inb = [&](int x, int z) { return FULL || (x >= 0 && x < ntx && z >= 0 && z < ntz); };
    // element offset of a cell of the tile: the tile origin's (uniform) plus a
    // full-rate 24-bit product (|c| ldz stays far below 2^23)
    const unsigned obase = static_cast<unsigned>(X0) * ldz + static_cast<unsigned>(Z0);
    auto offc = [&](int c, int q) { return obase + static_cast<unsigned>(__mul24(c, static_cast<int>(ldz)) + 4 * q); };
    auto off = [&](int x, int z) { return offc(x - X0, (z - Z0) >> 2); };
    // dp/dt is stored where the reconstruction's pressure update runs (interior columns, quads reaching the interior rows)
    auto dpint = [&](int x, int z) {
        return FULL || (x >= g.pmlc && x < ntx - g.pmlc && z + 3 >= g.pmlc && z < ntz - g.pmlc);
    };

    // ---- loads: staged p (level t), old velocities and model of every updated cell
    float4 st[G::RP0];
#pragma unroll
    for (int r = 0; r < G::RP0; r++) {
        const int i = tid + NT * r;
        int c, q;
        f2_p0<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if constexpr (!FULL) st[r] = f2zero();
        if ((r + 1) * NT <= G::NP0 || i < G::NP0)
            if (inb(x, z)) st[r] = ld4(Pr, offc(c, q));
    }
    float4 vxv[G::RX], bxv[G::RX];
#pragma unroll
    for (int r = 0; r < G::RX; r++) {
        const int i = tid + NT * r;
        int c, q;
        f2_vx<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if constexpr (!FULL) vxv[r] = bxv[r] = f2zero();
        if ((r + 1) * NT <= G::NX || i < G::NX)
            if (inb(x, z)) {
                vxv[r] = ld4(VXr, offc(c, q));
                bxv[r] = ld4((gcfl *)g.Bx, offc(c, q));
            }
    }
    float4 vzv[G::RZ], bzv[G::RZ];
#pragma unroll
    for (int r = 0; r < G::RZ; r++) {
        const int i = tid + NT * r;
        int c, q;
        f2_vz<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if constexpr (!FULL) vzv[r] = bzv[r] = f2zero();
        if ((r + 1) * NT <= G::NZ || i < G::NZ)
            if (inb(x, z)) {
                vzv[r] = ld4(VZr, offc(c, q));
                bzv[r] = ld4((gcfl *)g.Bz, offc(c, q));
            }
    }
    float4 kv[G::RP];
#pragma unroll
    for (int r = 0; r < G::RP; r++) {
        const int i = tid + NT * r;
        int c, q;
        f2_p<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if constexpr (!FULL) kv[r] = f2zero();
        if ((r + 1) * NT <= G::NP || i < G::NP)
            if (inb(x, z)) kv[r] = ld4((gcfl *)g.K, offc(c, q));
    }
    // own quad (item tid of the core regions): receiver map, K0, image accumulators
    const int oc = tid / TQ, oq = tid % TQ, ox = X0 + oc, oz = Z0 + 4 * oq;
    const bool own = NT == TX * TQ || tid < TX * TQ;
    const bool oin = own && (FULL || (ox < ntx && oz < ntz));
    const bool ocolint = FULL || (ox >= g.pmlc && ox < ntx - g.pmlc);
    int ocm = -1;
    if (!SRC && I.colmap && oin) ocm = ldgi(I.colmap, ox);
    // whether the source or a receiver column lies in the pressure region (columns -7 .. TX+6, rows
    // -8 .. 4TQ+7): uniform tests that spare the others the per-quad checks
    const bool srcnear = I.s_ix >= X0 - 7 && I.s_ix < X0 + TX + 7 && I.s_iz >= Z0 - 8 && I.s_iz < Z0 + 4 * TQ + 8;
    bool rcvnear = false;
    if (NEG && I.colmap) {
        bool hit = false;
        for (int j = lane0; j < TX + 14; j += 64) {
            const int xc = X0 - 7 + j;
            if (xc >= 0 && xc < ntx) hit |= ldgi(I.colmap, xc) >= 0;
        }
        rcvnear = __ballot(hit) != 0;
    }
    float4 k0 = f2zero(), im = f2zero(), is = f2zero(), ir = f2zero();
    const bool doimg = NEG && I.image && oin && ocolint;
    auto own_loads = [&]() {
        if (SRC && oin) k0 = ld4((gcfl *)g.K0, off(ox, oz));
        if (doimg) {
            im = ld4((gcfl *)I.image, off(ox, oz));
            is = ld4((gcfl *)I.illum_s, off(ox, oz));
            ir = ld4((gcfl *)I.illum_r, off(ox, oz));
        }
    };

#pragma unroll
    for (int r = 0; r < G::RP0; r++) {
        const int i = tid + NT * r;
        int c, q;
        f2_p0<TX, TQ>(i, c, q);
        if ((r + 1) * NT <= G::NP0 || i < G::NP0) sP[G::sp(c, q)] = st[r];
    }
    __syncthreads();

    // ---- step 1, velocities (level t+1/2) on the V1 regions
    float4 psxv[G::RXC], pszv[G::RZC];
#pragma unroll
    for (int r = 0; r < G::RX; r++) {
        const int i = tid + NT * r;
        if (!((r + 1) * NT <= G::NX || i < G::NX)) break;
        int c, q;
        f2_vx<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        float4 v = vxv[r];
        if (inb(x, z)) {
            float4 pso = f2zero(), psn;
            float ah;
            if (CPML && (x < g.fx0 || x > g.fx1)) pso = ld4(PSXr, off(x, z));
            v = f2_vx_quad<FULL, MODE, PNQ>(g, x, z, &sP[G::sp(c + 1, q)], v, bxv[r], pso, psn, ah);
            if (r < G::RXC) psxv[r] = psn;
        }
        vxv[r] = v;
        sVX[G::sx(c, q)] = v;
    }
#pragma unroll
    for (int r = 0; r < G::RZ; r++) {
        const int i = tid + NT * r;
        if (!((r + 1) * NT <= G::NZ || i < G::NZ)) break;
        int c, q;
        f2_vz<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        float4 v = vzv[r];
        if (inb(x, z)) {
            float a[12];
            f2_window(&sP[G::sp(c, q - 1)], a);
            float4 pso = f2zero(), psn;
            bool anyps;
            if (CPML && (z < g.fz0 || z + 3 > g.fz1)) pso = ld4(PSZr, off(x, z));
            v = f2_vz_quad<FULL, MODE>(g, x, z, a, v, bzv[r], pso, psn, anyps);
            if (r < G::RZC) pszv[r] = psn;
        }
        vzv[r] = v;
        sVZ[G::sz(c, q)] = v;
    }
    __syncthreads();
    // the own quad's K0 and image accumulators, first needed by the pressure loop below (which
    // takes the own quad last), loaded here rather than with the others to keep registers free
    own_loads();

    // ---- step 1, pressure (level t+1) on the P1 region
    float4 p1 = f2zero(), qx1 = f2zero(), qz1 = f2zero();
#pragma unroll
    for (int rr = 0; rr < G::RP; rr++) {
        const int r = G::RP - 1 - rr;   // the own quad (round 0) last
        const int i = tid + NT * r;
        if (!((r + 1) * NT <= G::NP || i < G::NP)) continue;
        int c, q;
        f2_p<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if (!inb(x, z)) continue;
        const float4 pold = sP[G::sp(c, q)];
        float w[12];
        f2_window(&sVZ[G::sz(c, q - 1)], w);
        float4 qxo = f2zero(), qzo = f2zero(), pv, dp = f2zero(), qxn, qzn;
        float axo;
        bool anyz;
        if (CPML) {
            if (x < g.fx0 || x > g.fx1) qxo = ld4(PVXr, off(x, z));
            if (z < g.fz0 || z + 3 > g.fz1) qzo = ld4(PVZr, off(x, z));
        }
        const bool wr = f2_p_quad<FULL, MODE, XNQ, BRN>(g, I, x, z, lv1, &sVX[G::sx(c, q)], w, pold, kv[r], k0, qxo,
                                                        qzo, pv, dp, qxn, qzn, axo, anyz, srcnear,
                                                        BRN == 2 ? sb[r] : f2zero());
        if constexpr (BRN == 1) sb[r] = dp;
        if (i < G::PCORE) {   // the own quad (core = item tid)
            if (SRC) {
                dps[0] = dp;
                if (I.dpdt && wr && dpint(x, z)) st4((gfl *)I.dpdt, off(x, z), dp);   // for the up/down imaging
            }
            if (!SRC) f2_receivers<FULL, MODE>(g, I, ocm, z, lv1, pv);
            if (doimg) f2_image<FULL>(g, z, dps[0], pv, im, is, ir);
            if (MODE == KSTEP_FWD && !FULL && I.strips) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int f = frame_index(g, x, z + k);
                    if (f >= 0) ((gfl *)I.strips)[static_cast<size_t>(lv1) * g.nframe + f] = elc(pv, k);   // RK:200-219
                }
            }
            if (MODE == KSTEP_FWD && I.capture) f2_capture<FULL>(g, I, x, z, lv1, pv);
            if (MODE == KSTEP_FWD && I.s0_out) f2_s0<FULL>(g, I, x, z, lv1, dp);
            if (NEG && I.pmid) st4((gfl *)I.pmid, off(x, z), pv);
            p1 = pv;
            qx1 = qxn;
            qz1 = qzn;
        } else if (NEG && rcvnear) {
            f2_receivers<FULL, MODE>(g, I, ldgi(I.colmap, x), z, lv1, pv);
        }
        if (wr) sP[G::sp(c, q)] = pv;
    }
    __syncthreads();
    if constexpr (MID) {
        // the middle level for the angle gathers: the own quad's first-level velocities
        // (still in LDS) and, for the reconstruction, its first-level pressure
        if (oin) {
            if (SRC) st4((gfl *)I.pmid, off(ox, oz), sP[G::sp(oc, oq)]);
            st4((gfl *)I.vxmid, off(ox, oz), sVX[G::sx(oc, oq)]);
            st4((gfl *)I.vzmid, off(ox, oz), sVZ[G::sz(oc, oq)]);
        }
    }

    // ---- step 2, velocities (level t+3/2) on the core regions; own cells stored
#pragma unroll
    for (int r = 0; r < G::RXC; r++) {
        const int i = tid + NT * r;
        if (!((r + 1) * NT <= G::XCORE || i < G::XCORE)) break;
        int c, q;
        f2_vx<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if (!inb(x, z)) continue;
        float4 psn = f2zero();
        float ah = 0.f;
        const float4 bx2 = ld4((gcfl *)g.Bx, off(x, z));   // reloaded (L2): fewer registers held across the steps
        const float4 v = f2_vx_quad<FULL, MODE, PNQ>(g, x, z, &sP[G::sp(c + 1, q)], vxv[r], bx2,
                                                     CPML ? psxv[r] : f2zero(), psn, ah);
        sVX[G::sx(c, q)] = v;
        if (c >= 0 && c < TX) {
            if (SRC) {
                const bool upd = FULL || (x >= g.pmlc && x < ntx - g.pmlc && z + 3 >= g.pmlc && z < ntz - g.pmlc);
                if (upd) st4(VXw, off(x, z), v);
            } else {
                st4(VXw, off(x, z), v);
                if (CPML && ah != 0.f) st4(PSXw, off(x, z), psn);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < G::RZC; r++) {
        const int i = tid + NT * r;
        if (!((r + 1) * NT <= G::ZCORE || i < G::ZCORE)) break;
        int c, q;
        f2_vz<TX, TQ>(i, c, q);
        const int x = X0 + c, z = Z0 + 4 * q;
        if (!inb(x, z)) continue;
        float a[12];
        f2_window(&sP[G::sp(c, q - 1)], a);
        float4 psn = f2zero();
        bool anyps = false;
        const float4 bz2 = ld4((gcfl *)g.Bz, off(x, z));
        const float4 v = f2_vz_quad<FULL, MODE>(g, x, z, a, vzv[r], bz2, CPML ? pszv[r] : f2zero(), psn, anyps);
        sVZ[G::sz(c, q)] = v;
        if (q >= 0 && q < TQ) {
            if (SRC) {
                const bool upd = FULL || (x >= g.pmlc && x < ntx - g.pmlc && z + 3 >= g.pmlc && z < ntz - g.pmlc);
                if (upd) st4(VZw, off(x, z), v);
            } else {
                st4(VZw, off(x, z), v);
                if (CPML && anyps) st4(PSZw, off(x, z), psn);
            }
        }
    }
    __syncthreads();

    // ---- step 2, pressure (level t+2) of the own quad
    if (oin) {
        const int x = ox, z = oz;
        float w[12];
        f2_window(&sVZ[G::sz(oc, oq - 1)], w);
        float4 pv, dp = f2zero(), qxn, qzn;
        float axo;
        bool anyz;
        const bool wr = f2_p_quad<FULL, MODE, XNQ, BRN>(g, I, x, z, lv2, &sVX[G::sx(oc, oq)], w, p1, kv[0], k0, qx1,
                                                        qz1, pv, dp, qxn, qzn, axo, anyz, srcnear,
                                                        BRN == 2 ? sb[G::RP] : f2zero());
        if constexpr (BRN == 1) sb[G::RP] = dp;
        if (SRC) {
            dps[1] = dp;
            if (I.dpdt_b && wr && dpint(x, z)) st4((gfl *)I.dpdt_b, off(x, z), dp);
        }
        if (!SRC) f2_receivers<FULL, MODE>(g, I, ocm, z, lv2, pv);
        if (wr) st4(Pw, off(x, z), pv);
        if (CPML) {
            if (axo != 0.f) st4(PVXw, off(x, z), qxn);
            if (anyz) st4(PVZw, off(x, z), qzn);
        }
        if (MODE == KSTEP_FWD && !FULL && I.strips) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int f = frame_index(g, x, z + k);
                if (f >= 0) ((gfl *)I.strips)[static_cast<size_t>(lv2) * g.nframe + f] = elc(pv, k);
            }
        }
        if (MODE == KSTEP_FWD && I.capture) f2_capture<FULL>(g, I, x, z, lv2, pv);
        if (MODE == KSTEP_FWD && I.s0_out) f2_s0<FULL>(g, I, x, z, lv2, dp);
        if (doimg) {
            f2_image<FULL>(g, z, dps[1], pv, im, is, ir);
            st4((gfl *)I.image, off(x, z), im);
            st4((gfl *)I.illum_s, off(x, z), is);
            st4((gfl *)I.illum_r, off(x, z), ir);
        }
    }
}

// One launch: two steps of every instance (PAIR = false: FWD instances; PAIR
// = true: (SRC, RCV) pairs of one shot per workgroup, the reconstruction
// first).  it: the first step's time index; ph: the launch's place in its pair.
// The tiles whose stencils reach no damping or frame cell (the "fast" tiles)
// form a rectangle of tile columns fr.x .. fr.x+fr.y-1 and tile rows fr.z ..
// fr.z+fr.w-1; SEL = 1 launches cover those, SEL = 2 launches the others (the
// two kernels differ in registers: the fast one holds more wavefronts), SEL = 0
// launches every tile and picks the path per tile.
#ifndef SWD_F2_WAVES_FWD
#define SWD_F2_WAVES_FWD 4
#endif
#ifndef SWD_F2_WAVES_FAST
#define SWD_F2_WAVES_FAST 4
#endif
#ifndef SWD_F2_WAVES_FRAME
#define SWD_F2_WAVES_FRAME 4   // the paired kernel over every tile: spills a few dwords, 6% faster than three
#endif
template <int TX, int TQ, int NW, bool PAIR, int SEL, bool MID = false, bool BORN = false>
__global__ __launch_bounds__(64 * NW, SEL == 1 ? (PAIR ? SWD_F2_WAVES_FAST : SWD_F2_WAVES_FWD) : (PAIR ? SWD_F2_WAVES_FRAME : SWD_F2_WAVES_FWD)) void k_fstep2(
    KGrid g, const KStep *__restrict__ insts, int ninst, int nzb, int nxb, int4 fr, int it, int ph)
{
    using G = T2<TX, TQ, NW>;
    __shared__ float4 sP[G::PNC * G::PNQ];
    __shared__ float4 sVX[G::XNC * G::XNQ];
    __shared__ float4 sVZ[G::ZNC * G::ZNQ];
    // Workgroups go to the 8 XCDs in turn (block b to XCD b % 8); each XCD takes a contiguous
    // share of the tiles (neighbours share halos in its L2).  Forward, SEL 0: every XCD gets an
    // equal share of the frame tiles, then of the fast ones (the frame tiles, slower, first), so
    // that the XCDs holding the absorbing columns do not finish last (c2, 16 shots: forward
    // -14%).  The paired backward keeps whole tile columns per XCD, which measured 2% faster
    // there (its frame tiles outside the interior skip the reconstruction).
    const int nfast = fr.y * fr.w, nfr = nzb * nxb - nfast;
    int L;
    bool fastitem = SEL == 1;
    if (SEL == 0 && PAIR) {
        // contiguous tile columns per XCD
        L = xcd_block();
        if (L >= ninst * nzb * nxb) return;
        const int t0 = L / ninst, x0 = t0 / nzb, z0 = t0 % nzb;
        fastitem = x0 >= fr.x && x0 < fr.x + fr.y && z0 >= fr.z && z0 < fr.z + fr.w;
    } else if (SEL == 0) {
        const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
        const int F = ninst * nfr, S = ninst * nfast;
        const int f0 = xcd * F / 8, nf = (xcd + 1) * F / 8 - f0;
        const int s0 = xcd * S / 8, ns = (xcd + 1) * S / 8 - s0;
        if (j >= nf + ns) return;
        fastitem = j >= nf;
        L = fastitem ? s0 + j - nf : f0 + j;
    } else {
        L = xcd_block();
        if (L >= ninst * (SEL == 1 ? nfast : nfr)) return;
    }
    const int inst = L % ninst, t = L / ninst;
    int xb, zb;
    if (SEL == 0 && PAIR) {
        zb = t % nzb;
        xb = t / nzb;
    } else if (fastitem) {
        zb = fr.z + t % fr.w;
        xb = fr.x + t / fr.w;
    } else {
        // tile columns left of the rectangle, the rectangle's columns above and below it, the columns right of it
        const int a = fr.x * nzb, m = nzb - fr.w, b = a + fr.y * m;
        if (t < a) {
            xb = t / nzb;
            zb = t % nzb;
        } else if (t < b) {
            const int u = t - a, k = u % m;
            xb = fr.x + u / m;
            zb = k < fr.z ? k : k + fr.w;
        } else {
            const int u = t - b;
            xb = fr.x + fr.y + u / nzb;
            zb = u % nzb;
        }
    }
    const int X0 = xb * TX, Z0 = zb * 4 * TQ;
    const bool full = fastitem;
    float4 dps[2] = {f2zero(), f2zero()};
    if (!PAIR) {
        if (full) f2_field<TX, TQ, NW, SEL != 2, KSTEP_FWD>(g, insts[inst], X0, Z0, it, ph, sP, sVX, sVZ, dps);
        else f2_field<TX, TQ, NW, SEL == 1, KSTEP_FWD>(g, insts[inst], X0, Z0, it, ph, sP, sVX, sVZ, dps);
        return;
    }
    if constexpr (BORN) {
        // Born pair: the background (2k), then the scattered field (2k+1) with the background's
        // sums of both levels in registers
        const KStep &B = insts[2 * inst], &U = insts[2 * inst + 1];
        float4 sb[T2<TX, TQ, NW>::RP + 1];
        if (full) f2_field<TX, TQ, NW, SEL != 2, KSTEP_FWD, false, 1>(g, B, X0, Z0, it, ph, sP, sVX, sVZ, dps, sb);
        else f2_field<TX, TQ, NW, SEL == 1, KSTEP_FWD, false, 1>(g, B, X0, Z0, it, ph, sP, sVX, sVZ, dps, sb);
        __syncthreads();   // LDS is staged again for the scattered field
        if (full) f2_field<TX, TQ, NW, SEL != 2, KSTEP_FWD, false, 2>(g, U, X0, Z0, it, ph, sP, sVX, sVZ, dps, sb);
        else f2_field<TX, TQ, NW, SEL == 1, KSTEP_FWD, false, 2>(g, U, X0, Z0, it, ph, sP, sVX, sVZ, dps, sb);
        return;
    }
    const KStep &S = insts[2 * inst], &Rv = insts[2 * inst + 1];
    // the reconstruction writes only interior cells and the source cell of its own tile
    const int rows = 4 * TQ;
    const bool any = X0 + TX > g.pmlc && X0 < g.ntx - g.pmlc && Z0 + rows > g.pmlc && Z0 < g.ntz - g.pmlc;
    const bool src = S.s_ix >= X0 && S.s_ix < X0 + TX && S.s_iz >= Z0 && S.s_iz < Z0 + rows;
    if (any || src) {
        if (full) f2_field<TX, TQ, NW, SEL != 2, KSTEP_SRC, MID>(g, S, X0, Z0, it, ph, sP, sVX, sVZ, dps);
        else f2_field<TX, TQ, NW, SEL == 1, KSTEP_SRC, MID>(g, S, X0, Z0, it, ph, sP, sVX, sVZ, dps);
        __syncthreads();   // LDS is staged again for the receiver wavefield
    }
    if (full) f2_field<TX, TQ, NW, SEL != 2, KSTEP_RCV, MID>(g, Rv, X0, Z0, it, ph, sP, sVX, sVZ, dps);
    else f2_field<TX, TQ, NW, SEL == 1, KSTEP_RCV, MID>(g, Rv, X0, Z0, it, ph, sP, sVX, sVZ, dps);
}

// Multi-grid form of the forward (each instance on its own grid, KStep::grid):
// btab[b] = (instance, tile column, tile row, fast) for 32 x 32-row tiles, four
// wavefronts, the fast flag computed on the host from the instance's grid.
__global__ __launch_bounds__(256, SWD_F2_WAVES_FWD) void k_fstep2_multi(const KStep *__restrict__ insts,
                                                                         const int4 *__restrict__ btab, int nblk,
                                                                         int it, int ph)
{
    constexpr int TX = 32, TQ = 8, NW = 4;
    using G = T2<TX, TQ, NW>;
    __shared__ float4 sP[G::PNC * G::PNQ];
    __shared__ float4 sVX[G::XNC * G::XNQ];
    __shared__ float4 sVZ[G::ZNC * G::ZNQ];
    if (static_cast<int>(blockIdx.x) >= nblk) return;
    const int4 e = btab[blockIdx.x];
    const KStep &I = insts[e.x];
    const KGrid &g = *I.grid;
    float4 dps[2] = {f2zero(), f2zero()};
    if (e.w) f2_field<TX, TQ, NW, true, KSTEP_FWD>(g, I, e.y * TX, e.z * 4 * TQ, it, ph, sP, sVX, sVZ, dps);
    else f2_field<TX, TQ, NW, false, KSTEP_FWD>(g, I, e.y * TX, e.z * 4 * TQ, it, ph, sP, sVX, sVZ, dps);
}

// The fast rectangle of a grid for TX x 4TQ tiles (empty: y = w = 0): a tile
// is fast when its staged region (columns X0-14 .. X0+TX+13, rows Z0-16 ..
// Z0+4TQ+15) lies inside the damping-free box.
inline int4 f2_fast_rect(const KGrid &g, int tx, int tq, int nxb, int nzb)
{
    const int rows = 4 * tq;
    const int xa = std::max(0, (g.fx0 + 14 + tx - 1) / tx), xz = (g.fx1 - 13 - tx) >= 0 ? (g.fx1 - 13 - tx) / tx : -1;
    const int za = std::max(0, (g.fz0 + 16 + rows - 1) / rows), zz = (g.fz1 - 15 - rows) >= 0 ? (g.fz1 - 15 - rows) / rows : -1;
    const int nx = std::min(xz, nxb - 1) - xa + 1, nz = std::min(zz, nzb - 1) - za + 1;
    if (g.fx1 < g.fx0 || g.fz1 < g.fz0 || nx <= 0 || nz <= 0) return make_int4(0, 0, 0, 0);
    return make_int4(xa, nx, za, nz);
}

}  // namespace
