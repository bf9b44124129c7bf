// swd_fused.hip -- fused, batched time-step kernels for gfx950.
//
// One launch advances every wavefield of a batch by one full leapfrog step
// (velocity then pressure, RK:24-221) so each cell's state crosses HBM once
// per step instead of three times.  Because a tile needs its neighbours'
// velocities of the same step, the state is double-buffered: the launch reads
// time level n-1 from one buffer set and writes level n to the other, so no
// workgroup ever reads a value another workgroup writes in the same launch.
// Cells a sub-step does not update are copied, which keeps both buffers whole.
//
// Work decomposition.  A wavefront owns 64 consecutive z rows (lane = row,
// the contiguous axis) and sweeps TX columns along x:
//   * x-derivatives come from register windows of the last 2*Lc+1 columns;
//   * z-derivatives come from neighbouring lanes through ds_bpermute, so a
//     wave produces 64 - 2*(2*Lc-1) output rows (50 for Lc=4) and recomputes
//     the velocities of its 2*Lc-1 halo rows above and below;
//   * the next column's loads are issued before the current column is
//     computed (software prefetch, one column ahead).
// Tiles are ordered instance-fastest and assigned XCD-contiguously (hardware
// workgroup b runs on XCD b%8): all shots of one tile, which share the model
// and the tile's halo columns, run on the same XCD and hit its L2.
//
// Arithmetic is the canonical order of the plain kernels (swd_kernels.hip) and
// of the oracle, so results are identical to them bit for bit.
#include <algorithm>
#include <cstring>
#include <vector>

#include "swd_fused.h"
#include "swd_launch.h"
#include "swd.h"
#include "swd_internal.h"

#define HIP_TRY(expr)                                                                                            \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return swd_fail(SWD_E_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

namespace {

// One wavefield as seen by a fused step: time n-1 (read) and n (write).
struct FField {
    const float *p0, *vx0, *vz0, *psx0, *psz0;
    float *p1, *vx1, *vz1, *psx1, *psz1;
    float *pvx, *pvz;   // memory variables of the pressure update, own cells only: in place
    const float *wavelet;
};

struct FInst {
    FField w[2];
    int s_ix, s_iz;
    const int *colmap, *rowmap;
    float *gather;
    int rec_w;            // recorded wavefield
    float *strips;        // frame of w[0]
    float2 *capture;      // AFor: w[0] -> .x, w[1] -> .y
    const int *colcap;
    int ncap, nfftz;
    const float *born_m;  // Born: w[1] gets m * S(w[0]) on the interior
    float *s0_out;        // background sums of w[0] per step slot
    int s0_t0;
};

__device__ __forceinline__ float shfl(float v, int addr)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

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

// Maps the hardware workgroup to a logical one so that each XCD (b % 8)
// receives a contiguous range of logical workgroups.  gridDim.x is a multiple of 8.
__device__ __forceinline__ int logical_block()
{
    const int b = blockIdx.x, per = gridDim.x >> 3;
    return (b & 7) * per + (b >> 3);
}

// Global-memory inputs of one sweep iteration, loaded one iteration ahead.
template <int NW>
struct ColIn {
    float p[NW];                   // p(n-1) at column j
    float vx[NW], psx[NW], Bx;     // at column j-LC (own rows)
    float vz[NW], psz[NW], Bz;     // at column j-LC (own columns, rows in the grid)
    float pp[NW], pvx[NW], pvz[NW], K;   // at column j-2LC+1 (own cells)
};

template <int LC, int NW, int TX>
__global__ __launch_bounds__(256) void k_fwd(KGrid g, const FInst *__restrict__ tab, int ninst, int nzt, int nxt,
                                             int it)
{
    constexpr int H = 2 * LC - 1;
    constexpr int W = 64 - 2 * H;
    const int lane = threadIdx.x & 63;
    const int tile = logical_block() * 4 + (threadIdx.x >> 6);
    if (tile >= ninst * nzt * nxt) return;
    const int inst = tile % ninst;
    const int zt = (tile / ninst) % nzt;
    const int xt = tile / (ninst * nzt);
    const FInst &F = tab[inst];
    const int ntx = g.ntx, ntz = g.ntz;
    const size_t ldz = g.ldz;
    const int row = zt * W - H + lane;
    const bool rowok = row >= 0 && row < ntz;
    const bool own = rowok && lane >= H && lane < 64 - H;
    const int x0 = xt * TX, x1 = min(x0 + TX, ntx);
    const int rowc = rowok ? row : 0;
    // row-dependent CPML coefficients and domains (RK:51, 100, 157)
    const float azh = g.azh[rowc], bzh = g.bzh[rowc], az = g.az[rowc], bz = g.bz[rowc];
    const bool vxrow = row >= LC && row <= ntz - LC;
    const bool vzrow = row >= LC - 1 && row <= ntz - LC - 1;
    const bool prow = vxrow;
    const bool introw = row >= g.pmlc && row < ntz - g.pmlc;
    const float idx = g.idx, idz = g.idz;
    int up[LC + 1], dn[LC + 1];
#pragma unroll
    for (int k = 0; k <= LC; k++) {
        up[k] = ((lane + k) & 63) << 2;
        dn[k] = ((lane - k) & 63) << 2;
    }
    float pw[NW][2 * LC + 1];
    float vxw[NW][2 * LC];
    float vzq[NW][LC];
#pragma unroll
    for (int w = 0; w < NW; w++) {
#pragma unroll
        for (int k = 0; k < 2 * LC + 1; k++) pw[w][k] = 0.f;
#pragma unroll
        for (int k = 0; k < 2 * LC; k++) vxw[w][k] = 0.f;
#pragma unroll
        for (int k = 0; k < LC; k++) vzq[w][k] = 0.f;
    }
    const int js = x0 - 2 * LC + 1, je = x1 + 2 * LC - 2;

    // issues every global load iteration j will consume
    auto load = [&](int j, ColIn<NW> &d) {
        {
            const bool ok = rowok && j >= 0 && j < ntx;
            const size_t o = static_cast<size_t>(j) * ldz + row;
#pragma unroll
            for (int w = 0; w < NW; w++) d.p[w] = ok ? F.w[w].p0[o] : 0.f;
        }
        const int c = j - LC;
        if (c >= 0 && c < ntx && c >= x0 - LC && c <= x1 + LC - 2) {
            const size_t o = static_cast<size_t>(c) * ldz + rowc;
            const bool psx_on = g.axh[c] != 0.f;
            if (own) {
                d.Bx = g.Bx[o];
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    d.vx[w] = F.w[w].vx0[o];
                    d.psx[w] = psx_on ? F.w[w].psx0[o] : 0.f;
                }
            }
            if (c >= x0 && c < x1 && rowok) {
                d.Bz = g.Bz[o];
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    d.vz[w] = F.w[w].vz0[o];
                    d.psz[w] = azh != 0.f ? F.w[w].psz0[o] : 0.f;
                }
            }
        }
        const int cp = j - 2 * LC + 1;
        if (cp >= x0 && cp < x1 && own) {
            const size_t o = static_cast<size_t>(cp) * ldz + rowc;
            const bool pvx_on = g.ax[cp] != 0.f;
            d.K = g.K[o];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                d.pvx[w] = pvx_on ? F.w[w].pvx[o] : 0.f;
                d.pvz[w] = az != 0.f ? F.w[w].pvz[o] : 0.f;
            }
        }
    };

    ColIn<NW> nx_, cur;
    load(js, nx_);
    for (int j = js; j <= je; j++) {
        cur = nx_;
        if (j + 1 <= je) load(j + 1, nx_);
        // (a) p window: columns j-2LC .. j
#pragma unroll
        for (int w = 0; w < NW; w++) {
#pragma unroll
            for (int k = 0; k < 2 * LC; k++) pw[w][k] = pw[w][k + 1];
            pw[w][2 * LC] = cur.p[w];
        }
        // (b) velocities at column c = j - LC
        const int c = j - LC;
        float vxn[NW], vzn[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) vxn[w] = vzn[w] = 0.f;
        if (c >= 0 && c < ntx && c >= x0 - LC && c <= x1 + LC - 2) {
            const size_t o = static_cast<size_t>(c) * ldz + rowc;
            const bool ownc = c >= x0 && c < x1;
            const bool vxcol = c >= LC - 1 && c <= ntx - LC - 1;
            const float ah = g.axh[c], bh = g.bxh[c];
            if (own) {
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    float vn = cur.vx[w];
                    if (vxcol && vxrow) {
                        float d = 0.f;
#pragma unroll
                        for (int ic = 0; ic < LC; ic++)
                            d = __builtin_fmaf(g.rc[ic] * (pw[w][LC + ic + 1] - pw[w][LC - ic]), idx, d);
                        float ps = 0.f;
                        if (ah != 0.f) {
                            ps = __builtin_fmaf(bh, cur.psx[w], ah * d);
                            if (ownc) F.w[w].psx1[o] = ps;
                        }
                        vn = __builtin_fmaf(cur.Bx, d + ps, vn);
                    }
                    if (ownc) F.w[w].vx1[o] = vn;
                    vxn[w] = vn;
                }
            }
            if (ownc) {
                const bool vzcol = c >= LC && c <= ntx - LC;
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    const float pc = pw[w][LC];
                    float d = 0.f;
#pragma unroll
                    for (int ic = 0; ic < LC; ic++) {
                        const float pa = shfl(pc, up[ic + 1]);
                        const float pb = ic == 0 ? pc : shfl(pc, dn[ic]);
                        d = __builtin_fmaf(g.rc[ic] * (pa - pb), idz, d);
                    }
                    if (rowok) {
                        float vn = cur.vz[w];
                        if (vzcol && vzrow) {
                            float ps = 0.f;
                            if (azh != 0.f) {
                                ps = __builtin_fmaf(bzh, cur.psz[w], azh * d);
                                if (own) F.w[w].psz1[o] = ps;
                            }
                            vn = __builtin_fmaf(cur.Bz, d + ps, vn);
                        }
                        if (own) F.w[w].vz1[o] = vn;
                        vzn[w] = vn;
                    }
                }
            }
        }
#pragma unroll
        for (int w = 0; w < NW; w++) {
#pragma unroll
            for (int k = 0; k < 2 * LC - 1; k++) vxw[w][k] = vxw[w][k + 1];
            vxw[w][2 * LC - 1] = vxn[w];
#pragma unroll
            for (int k = 0; k < LC - 1; k++) vzq[w][k] = vzq[w][k + 1];
            vzq[w][LC - 1] = vzn[w];
        }
        // (c) pressure at column cp = j - 2LC + 1
        const int cp = j - 2 * LC + 1;
        if (cp >= x0 && cp < x1) {
            const size_t o = static_cast<size_t>(cp) * ldz + rowc;
            float dvz[NW];
#pragma unroll
            for (int w = 0; w < NW; w++) {
                const float vc = vzq[w][0];
                float d = 0.f;
#pragma unroll
                for (int ic = 0; ic < LC; ic++) {
                    const float va = ic == 0 ? vc : shfl(vc, up[ic]);
                    const float vb = shfl(vc, dn[ic + 1]);
                    d = __builtin_fmaf(g.rc[ic] * (va - vb), idz, d);
                }
                dvz[w] = d;
            }
            if (own) {
                const bool pcol = cp >= LC && cp <= ntx - LC;
                const float ax = g.ax[cp], bx = g.bx[cp];
                const bool inter = introw && cp >= g.pmlc && cp < ntx - g.pmlc;
                const int im = (cp - g.pmlc) * g.nz + (row - g.pmlc);
                float S0 = 0.f;
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    float pn = pw[w][1];
                    if (pcol && prow) {
                        float dvx = 0.f;
#pragma unroll
                        for (int ic = 0; ic < LC; ic++)
                            dvx = __builtin_fmaf(g.rc[ic] * (vxw[w][LC + ic] - vxw[w][LC - 1 - ic]), idx, dvx);
                        float qx = 0.f, qz = 0.f;
                        if (ax != 0.f) {
                            qx = __builtin_fmaf(bx, cur.pvx[w], ax * dvx);
                            F.w[w].pvx[o] = qx;
                        }
                        if (az != 0.f) {
                            qz = __builtin_fmaf(bz, cur.pvz[w], az * dvz[w]);
                            F.w[w].pvz[o] = qz;
                        }
                        float S = ((dvx + qx) + dvz[w]) + qz;
                        if (w == 0) {
                            S0 = S;
                            if (F.s0_out && inter) F.s0_out[static_cast<size_t>(it - F.s0_t0) * g.nx * g.nz + im] = S;
                        }
                        if (w == 1 && F.born_m && inter) S = __builtin_fmaf(F.born_m[im], S0, S);
                        pn = __builtin_fmaf(cur.K, S, pn);
                    }
                    if (F.w[w].wavelet && cp == F.s_ix && row == F.s_iz) pn += F.w[w].wavelet[it];
                    F.w[w].p1[o] = pn;
                    if (w == F.rec_w && F.colmap) {
                        const int cm = F.colmap[cp];
                        if (cm >= 0) {
                            const int tr = F.rowmap[cm + row];
                            if (tr >= 0) F.gather[static_cast<size_t>(tr) * g.nt + it] = pn;
                        }
                    }
                    if (w == 0 && F.strips) {
                        const int f = frame_index(g, cp, row);
                        if (f >= 0) F.strips[static_cast<size_t>(it) * g.nframe + f] = pn;
                    }
                    if (F.capture) {
                        const int cc = F.colcap[cp];
                        if (cc >= 0) {
                            float *dst = reinterpret_cast<float *>(F.capture + (static_cast<size_t>(it) * F.ncap + cc) *
                                                                                   F.nfftz + row);
                            dst[w] = pn;
                        }
                    }
                }
            }
        }
    }
}

// fused-step geometry for a grid
struct FGeom {
    int nzt, nxt, nblocks;
};

template <int LC, int TX>
FGeom geom(const KGrid &g, int ninst)
{
    constexpr int W = 64 - 2 * (2 * LC - 1);
    FGeom f;
    f.nzt = (g.ntz + W - 1) / W;
    f.nxt = (g.ntx + TX - 1) / TX;
    const int tiles = ninst * f.nzt * f.nxt;
    const int blocks = (tiles + 3) / 4;
    f.nblocks = (blocks + 7) / 8 * 8;
    return f;
}

constexpr int kTX = 32;

template <int LC>
hipError_t launch_fwd_lc(const KGrid &g, const FInst *tab, int ninst, int nw, int it, hipStream_t s)
{
    const FGeom f = geom<LC, kTX>(g, ninst);
    if (nw == 1)
        hipLaunchKernelGGL((k_fwd<LC, 1, kTX>), dim3(f.nblocks), dim3(256), 0, s, g, tab, ninst, f.nzt, f.nxt, it);
    else
        hipLaunchKernelGGL((k_fwd<LC, 2, kTX>), dim3(f.nblocks), dim3(256), 0, s, g, tab, ninst, f.nzt, f.nxt, it);
    return hipGetLastError();
}

hipError_t launch_fwd(const KGrid &g, const FInst *tab, int ninst, int nw, int it, hipStream_t s)
{
    switch (g.Lc) {
    case 1: return launch_fwd_lc<1>(g, tab, ninst, nw, it, s);
    case 2: return launch_fwd_lc<2>(g, tab, ninst, nw, it, s);
    case 3: return launch_fwd_lc<3>(g, tab, ninst, nw, it, s);
    case 4: return launch_fwd_lc<4>(g, tab, ninst, nw, it, s);
    case 5: return launch_fwd_lc<5>(g, tab, ninst, nw, it, s);
    case 6: return launch_fwd_lc<6>(g, tab, ninst, nw, it, s);
    case 7: return launch_fwd_lc<7>(g, tab, ninst, nw, it, s);
    case 8: return launch_fwd_lc<8>(g, tab, ninst, nw, it, s);
    default: return hipErrorInvalidValue;
    }
}

FField field_of(const KInst &I, int parity)
{
    FField f;
    const float *A[5] = {I.p, I.vx, I.vz, I.psx, I.psz};
    const float *B[5] = {I.p2, I.vx2, I.vz2, I.psx2, I.psz2};
    const float *const *rd = parity ? B : A;
    const float *const *wr = parity ? A : B;
    f.p0 = rd[0]; f.vx0 = rd[1]; f.vz0 = rd[2]; f.psx0 = rd[3]; f.psz0 = rd[4];
    f.p1 = const_cast<float *>(wr[0]); f.vx1 = const_cast<float *>(wr[1]); f.vz1 = const_cast<float *>(wr[2]);
    f.psx1 = const_cast<float *>(wr[3]); f.psz1 = const_cast<float *>(wr[4]);
    f.pvx = I.pvx; f.pvz = I.pvz;
    f.wavelet = I.wavelet;
    return f;
}

}  // namespace

size_t fused_table_bytes(int nshot) { return sizeof(FInst) * 2 * static_cast<size_t>(nshot); }

int fused_build_table(const KInst *h_inst, int n, int per_shot, void *d_table, hipStream_t s)
{
    if (per_shot < 1 || per_shot > 2 || n % per_shot) return swd_fail(SWD_E_ARG, "fused step: bad grouping");
    const int nshot = n / per_shot;
    std::vector<FInst> tab(2 * static_cast<size_t>(nshot));
    for (int par = 0; par < 2; par++)
        for (int k = 0; k < nshot; k++) {
            FInst &F = tab[static_cast<size_t>(par) * nshot + k];
            std::memset(&F, 0, sizeof F);
            int rec_w = -1;
            for (int w = 0; w < per_shot; w++) {
                const KInst &I = h_inst[k * per_shot + w];
                if (!I.p2 || !I.vx2 || !I.vz2 || !I.psx2 || !I.psz2)
                    return swd_fail(SWD_E_STATE, "fused step needs the second buffer set");
                F.w[w] = field_of(I, par);
                if (I.gather) { F.gather = I.gather; F.colmap = I.colmap; F.rowmap = I.rowmap; rec_w = w; }
                if (I.strips && w == 0) F.strips = I.strips;
                if (I.capture) { F.capture = I.capture; F.colcap = I.colcap; F.ncap = I.ncap; F.nfftz = I.nfftz; }
                if (I.born_m && w == 1) F.born_m = I.born_m;
                if (I.s0_out && w == 0) { F.s0_out = I.s0_out; F.s0_t0 = I.s0_t0; }
            }
            const KInst &I0 = h_inst[k * per_shot];
            F.s_ix = I0.s_ix;
            F.s_iz = I0.s_iz;
            F.rec_w = rec_w;
            if (rec_w < 0) F.colmap = nullptr;
        }
    HIP_TRY(hipMemcpyAsync(d_table, tab.data(), sizeof(FInst) * tab.size(), hipMemcpyHostToDevice, s));
    return SWD_OK;
}

int fused_forward_steps(const KGrid &g, const void *d_table, int nshot, int per_shot, int it0, int it1,
                        hipStream_t s, long long *launches)
{
    const FInst *dt = static_cast<const FInst *>(d_table);
    for (int it = it0; it < it1; it++) {
        HIP_TRY(launch_fwd(g, dt + static_cast<size_t>(it & 1) * nshot, nshot, per_shot, it, s));
        *launches += 1;
    }
    return SWD_OK;
}

int fused_backward_loop(const KGrid &g, const KInst *dS, const KInst *dR, const KInst *hS, const KInst *hR, int n,
                        hipStream_t s, long long *launches)
{
    (void)hS;
    (void)hR;
    for (int it = g.nt - 2; it >= 0; it--) {
        HIP_TRY(launch_src_p(g, dS, n, it, s));
        HIP_TRY(launch_src_v(g, dS, n, 3, s));
        HIP_TRY(launch_v(g, dR, n, true, 3, s));
        HIP_TRY(launch_p(g, dR, n, true, it, s));
        *launches += 4;
    }
    return SWD_OK;
}
