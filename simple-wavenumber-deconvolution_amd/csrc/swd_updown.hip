// swd_updown.hip -- up/down imaging of one backward step: the kz split of
// the AFor separation (FK:52-138, there on receiver columns only) applied to
// whole columns of the backward's wavefields, and the cross-correlation of
// the separated fields, on the Stockham FFT of swd_fft.h.
//
// A workgroup takes one interior column of one shot.  Its first N/16 threads
// transform the analytic source column S = ds + i dh (the reconstructions'
// dp/dt, zero off the interior), the other N/16 the analytic receiver column
// R = pr + i ph (the receiver pressures of step it), over the whole padded
// column when it fits in N = nfftz, else over the nz interior rows; both are
// zero padded to N and split with the normalised inverse.  The images of the
// interior rows then take
//   du += Sd Ru,  ud += Su Rd,  uu += Su Ru,  dd += Sd Rd.
//
// (The AFor separation itself, k_kzsep, keeps a radix-2 LDS transform that
// follows the oracle's operation order and is bit-exact: its scattered-field
// gathers are differences of two runs, which magnify any rounding change.)
#include "swd_device.h"
#include "swd_fft.h"
#include "swd_kernels.h"
#include "swd_launch.h"

namespace {

// The twiddle half table e^{-2 pi i k/N}, k < N/2, staged into LDS once per
// workgroup: every pass then reads it from LDS instead of waiting on L2 (the
// kernel runs one workgroup's dependent chain of passes; the launches of the
// backward loop leave every workgroup resident at once).  Visible to the
// first twiddled pass through the first pass's barrier.
template <int LOGN, int V>
__device__ __forceinline__ void stage_twiddles(const float2 *__restrict__ tw, float2 *stw)
{
    constexpr int N = 1 << LOGN;
    for (int i = threadIdx.x; i < N / 2; i += 2 * N / V) stw[i] = tw[i];
}

// kz split of one column of one shot at step it: the analytic source column
// in L[0, NP), the receiver's in L[NP, 2 NP), both split into (Re up, Re
// down) in natural order of the rows z0 + i; tw: the twiddles in LDS.
template <int LOGN, int V>
__device__ __forceinline__ void ud_split(const KGrid &g, const KUpDown &U, int it, size_t c0, int z0, int nzw, float2 *lds,
                                         const float2 *tw)
{
    constexpr int N = 1 << LOGN, T = N / V, NP = N + N / 16;
    const int f = threadIdx.x / T, t = threadIdx.x - f * T;   // field 0: source, 1: receiver
    float2 *L = lds + f * NP;
    const int slot = it & 1;
    const float *re = (f ? U.pr[slot] : U.ds) + c0 + z0;
    const float *im = (f ? U.ph[slot] : U.dh) + c0 + z0;

    // transform of the column, zero padded to N, split, inverse transforms
    float2 v[V];
#pragma unroll
    for (int r = 0; r < V; r++) {
        const int i = t + r * T;
        v[r] = i < nzw ? make_float2(re[i], im[i]) : make_float2(0.f, 0.f);
    }
    first_pass<N, V>(v, t, L);
    passes<N, V, V, false>(v, t, L, tw);
    kz_pack<N, V>(v, t, L, U.coef, 0.5f / static_cast<float>(N));   // 1/(2N): a power of two, exact
    first_pass<N, V>(v, t, L);
    passes<N, V, V, true>(v, t, L, tw);
}

template <int LOGN, int V>
__global__ __launch_bounds__(2 * (1 << LOGN) / V) void k_updown(KGrid g, const KUpDown *__restrict__ shots, int it)
{
    constexpr int N = 1 << LOGN, T = N / V, NP = N + N / 16;
    __shared__ float2 lds[2 * NP];
    __shared__ float2 stw[N / 2];
    const int col = g.pmlc + blockIdx.x;
    const KUpDown &U = shots[blockIdx.y];
    const size_t c0 = static_cast<size_t>(col) * U.ld;
    // the whole padded column when it fits in N, else its interior rows
    const int z0 = N >= g.ntz ? 0 : g.pmlc, nzw = z0 ? g.nz : g.ntz;
    stage_twiddles<LOGN, V>(U.tw, stw);
    ud_split<LOGN, V>(g, U, it, c0, z0, nzw, lds, stw);

    // imaging of the interior rows
    const float2 *S = lds, *Rv = lds + NP;
    float *du = U.du + c0, *ud = U.ud + c0;
    float *uu = U.uu ? U.uu + c0 : nullptr, *dd = U.dd ? U.dd + c0 : nullptr;
    for (int iz = g.pmlc + static_cast<int>(threadIdx.x); iz < g.ntz - g.pmlc; iz += 2 * T) {
        const float2 s = S[pad(iz - z0)], r = Rv[pad(iz - z0)];
        du[iz] = __builtin_fmaf(s.y, r.x, du[iz]);
        ud[iz] = __builtin_fmaf(s.x, r.y, ud[iz]);
        if (uu) {
            uu[iz] = __builtin_fmaf(s.x, r.x, uu[iz]);
            dd[iz] = __builtin_fmaf(s.y, r.y, dd[iz]);
        }
    }
}

// Two steps of one shot's column in one workgroup: step ita (descriptors
// A), then step itb (descriptors B, the same images).  The images take
// level a's products and then level b's, the order of two k_updown launches,
// read and written once.  The interior rows are at most N (V/2 rows a thread).
template <int LOGN, int V>
__global__ __launch_bounds__(2 * (1 << LOGN) / V) void k_updown2(KGrid g, const KUpDown *__restrict__ sa,
                                                                 const KUpDown *__restrict__ sb, int ita, int itb)
{
    constexpr int N = 1 << LOGN, T = N / V, NP = N + N / 16, NR = V / 2;
    __shared__ float2 lds[2 * NP];
    __shared__ float2 stw[N / 2];
    const int col = g.pmlc + blockIdx.x;
    const KUpDown &A = sa[blockIdx.y], &B = sb[blockIdx.y];
    const size_t c0 = static_cast<size_t>(col) * A.ld;
    const int z0 = N >= g.ntz ? 0 : g.pmlc, nzw = z0 ? g.nz : g.ntz;
    stage_twiddles<LOGN, V>(A.tw, stw);
    ud_split<LOGN, V>(g, A, ita, c0, z0, nzw, lds, stw);

    const float2 *S = lds, *Rv = lds + NP;
    float *du = A.du + c0, *ud = A.ud + c0;
    float *uu = A.uu ? A.uu + c0 : nullptr, *dd = A.dd ? A.dd + c0 : nullptr;
    const int iz0 = g.pmlc + static_cast<int>(threadIdx.x), iz1 = g.ntz - g.pmlc;
    float a_du[NR], a_ud[NR], a_uu[NR], a_dd[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int iz = iz0 + r * 2 * T;
        if (iz < iz1) {
            const float2 s = S[pad(iz - z0)], q = Rv[pad(iz - z0)];
            a_du[r] = __builtin_fmaf(s.y, q.x, du[iz]);
            a_ud[r] = __builtin_fmaf(s.x, q.y, ud[iz]);
            if (uu) {
                a_uu[r] = __builtin_fmaf(s.x, q.x, uu[iz]);
                a_dd[r] = __builtin_fmaf(s.y, q.y, dd[iz]);
            }
        }
    }
    __syncthreads();   // level a's spectra read before level b's first pass overwrites them
    ud_split<LOGN, V>(g, B, itb, c0, z0, nzw, lds, stw);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int iz = iz0 + r * 2 * T;
        if (iz < iz1) {
            const float2 s = S[pad(iz - z0)], q = Rv[pad(iz - z0)];
            du[iz] = __builtin_fmaf(s.y, q.x, a_du[r]);
            ud[iz] = __builtin_fmaf(s.x, q.y, a_ud[r]);
            if (uu) {
                uu[iz] = __builtin_fmaf(s.x, q.x, a_uu[r]);
                dd[iz] = __builtin_fmaf(s.y, q.y, a_dd[r]);
            }
        }
    }
}

}  // namespace

hipError_t launch_updown(const KGrid &g, const KUpDown *shots, int nshot, int nfftz, int it, hipStream_t s)
{
    const dim3 grid(g.nx, nshot);
#define UD(LOGN, V)                                                                                   \
    case 1 << LOGN:                                                                                       \
        hipLaunchKernelGGL((k_updown<LOGN, V>), grid, dim3(2 * (1 << LOGN) / V), 0, s, g, shots, it); \
        break;
    switch (nfftz) {
        UD(7, 8) UD(8, 8) UD(9, 8) UD(10, 16) UD(11, 16) UD(12, 16)
    default:
        return hipErrorInvalidValue;
    }
#undef UD
    return hipGetLastError();
}

hipError_t launch_updown2(const KGrid &g, const KUpDown *a, const KUpDown *b, int nshot, int nfftz, int ita, int itb,
                          hipStream_t s)
{
    if (g.nz > nfftz) return hipErrorInvalidValue;
    // The two-level kernel takes 8 values a thread at every size: at 1024
    // points 16 values a thread held about 190 VGPRs (two wavefronts per SIMD;
    // c5, 54.2 us a launch against 2 x 23.4 us for two one-level launches),
    // 8 hold 104 (four wavefronts of 256-thread workgroups; c2 up/down RTM over
    // whole columns at nfftz 1024: backward 1198-1206 ms against 1280-1289 ms
    // with two one-level launches).  From 2048 points two one-level launches.
    if (updown2_launches(nfftz) == 2) {
        const hipError_t e = launch_updown(g, a, nshot, nfftz, ita, s);
        return e != hipSuccess ? e : launch_updown(g, b, nshot, nfftz, itb, s);
    }
    const dim3 grid(g.nx, nshot);
#define UD(LOGN, V)                                                                                                \
    case 1 << LOGN:                                                                                                    \
        hipLaunchKernelGGL((k_updown2<LOGN, V>), grid, dim3(2 * (1 << LOGN) / V), 0, s, g, a, b, ita, itb);            \
        break;
    switch (nfftz) {
        UD(7, 8) UD(8, 8) UD(9, 8) UD(10, 8)
    default:
        return hipErrorInvalidValue;
    }
#undef UD
    return hipGetLastError();
}
