// swd_cig.hip -- angle-domain common-image gathers: the update that follows
// each backward step (the reference's sum_image_GPU_kernel_imageCIG,
// RK:819-905, whose launch at RK:1149-1157 is commented out).
//
// One thread per interior cell (cells in z-fastest order, spread over the
// XCDs in contiguous ranges).  The gather wants the reconstruction's
// velocities at the step's own level; the fused backward leaves them one step
// behind (its next launch updates them), so the thread forms them here with
// the same stencil and arithmetic as that update (RK:488-573):
//   vx(it) = fma(-Bx, D+x p(it), vx(it+1)),  vz(it) = fma(-Bz, D+z p(it), vz(it+1)).
// Then, per cell,
//   source Poynting vector   (fx, fz)   = (-p_s vx_s, -p_s vz_s)
//   receiver Poynting vector (fxr, fzr) = (p_r vx_r, p_r vz_r)
// (the receiver field is stepped with (-dx, -dz), so +p v is the Poynting
// vector of the time-reversed wave; the reference's -p v there would measure
// 90 degrees minus the reflection angle),
//   F = fma(fx, fxr, fz fzr),  G = |(fx, fz)| |(fxr, fzr)|,
// averaged over the last delt steps through rings, tht = F/G,
// ftheta = 0.5 acos(tht) 180 / PI in degrees (PI = 3.1415926), and the five
// bins ceil(ftheta + ii - 0.5), ii = -2..2, inside [2, thetan) take
// fma(p_s p_r, exp(-ii^2/18), bin).  sqrt and acos are evaluated in double and
// rounded to float; an undefined angle (G = 0) adds nothing.
#include "swd_device.h"
#include "swd_kernels.h"
#include "swd_launch.h"

namespace {

template <int LC>
__global__ __launch_bounds__(256) void k_cig(KGrid g, const KCig *__restrict__ shots, int thetan, int delt, float w1,
                                             float w2, float thr, int it)
{
    const int nint = g.nx * g.nz;
    // consecutive cells on one XCD: the x stencil's neighbouring columns then
    // come from that XCD's L2
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= nint) return;
    const KCig &K = shots[blockIdx.y];
    const int ix = g.pmlc + c / g.nz, iz = g.pmlc + c % g.nz;
    const size_t ld = static_cast<size_t>(g.ldz), o = static_cast<size_t>(ix) * ld + iz;
    const int s = it & 1, s1 = (it + 1) & 1;
    const float *P = K.ps[s];
    const float ps = P[o];
    const float vxs = __builtin_fmaf(-g.Bx[o], dplus<LC>(P, o, ld, g.rc, g.idx), K.vxs[s1][o]);
    const float vzs = __builtin_fmaf(-g.Bz[o], dplus<LC>(P, o, 1, g.rc, g.idz), K.vzs[s1][o]);
    const float pr = K.pr[s][o], vxr = K.vxr[s][o], vzr = K.vzr[s][o];
    float F, a, b;
    cig_products(ps, vxs, vzs, pr, vxr, vzr, F, a, b);
    if (delt == 1 && !cig_may_bin(F, a, b, thr)) return;
    const float G = cig_norm(a, b);
    float pre = F, pst = G;
    if (delt > 1) {
        float *rF = K.ring, *rG = K.ring + static_cast<size_t>(delt) * nint;
        rF[static_cast<size_t>(it % delt) * nint + c] = F;
        rG[static_cast<size_t>(it % delt) * nint + c] = G;
        if (g.nt - 2 - it + 1 < delt) return;
        pre = 0.f;
        pst = 0.f;
        for (int k = 0; k < delt; k++) {
            pre += rF[static_cast<size_t>(k) * nint + c];
            pst += rG[static_cast<size_t>(k) * nint + c];
        }
    }
    cig_bins(pre, pst, ps * pr, thetan, w1, w2, thr, K.cig + static_cast<size_t>(c) * thetan);
}

// Two levels of one cell in one launch (delt 1): step ita from the descriptors
// a, then step itb from b (the same gathers), the order of two k_cig
// launches; the model's Bx, Bz read once.
template <int LC>
__global__ __launch_bounds__(256) void k_cig2(KGrid g, const KCig *__restrict__ sa, const KCig *__restrict__ sb,
                                              int thetan, float w1, float w2, float thr, int ita, int itb)
{
    const int nint = g.nx * g.nz;
    const int c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= nint) return;
    const int ix = g.pmlc + c / g.nz, iz = g.pmlc + c % g.nz;
    const size_t ld = static_cast<size_t>(g.ldz), o = static_cast<size_t>(ix) * ld + iz;
    const float bx = g.Bx[o], bz = g.Bz[o];
#pragma unroll
    for (int l = 0; l < 2; l++) {
        const KCig &K = l ? sb[blockIdx.y] : sa[blockIdx.y];
        const int it = l ? itb : ita, s = it & 1, s1 = (it + 1) & 1;
        const float *P = K.ps[s];
        const float ps = P[o];
        const float vxs = __builtin_fmaf(-bx, dplus<LC>(P, o, ld, g.rc, g.idx), K.vxs[s1][o]);
        const float vzs = __builtin_fmaf(-bz, dplus<LC>(P, o, 1, g.rc, g.idz), K.vzs[s1][o]);
        const float pr = K.pr[s][o], vxr = K.vxr[s][o], vzr = K.vzr[s][o];
        float F, a, b;
        cig_products(ps, vxs, vzs, pr, vxr, vzr, F, a, b);
        if (!cig_may_bin(F, a, b, thr)) continue;
        cig_bins(F, cig_norm(a, b), ps * pr, thetan, w1, w2, thr, K.cig + static_cast<size_t>(c) * thetan);
    }
}

}  // namespace

hipError_t launch_cig2(const KGrid &g, const KCig *a, const KCig *b, int nshot, int thetan, float w1, float w2, float thr,
                       int ita, int itb, hipStream_t s)
{
    const dim3 grid(((g.nx * g.nz + 255) / 256 + 7) / 8 * 8, nshot);
    DISPATCH_LC(g.Lc, hipLaunchKernelGGL((k_cig2<L>), grid, dim3(256), 0, s, g, a, b, thetan, w1, w2, thr, ita, itb));
    return hipGetLastError();
}

hipError_t launch_cig(const KGrid &g, const KCig *shots, int nshot, int thetan, int delt, float w1, float w2, float thr,
                      int it, hipStream_t s)
{
    const dim3 grid(((g.nx * g.nz + 255) / 256 + 7) / 8 * 8, nshot);   // a multiple of 8 for xcd_block
    DISPATCH_LC(g.Lc,
                hipLaunchKernelGGL((k_cig<L>), grid, dim3(256), 0, s, g, shots, thetan, delt, w1, w2, thr, it));
    return hipGetLastError();
}
