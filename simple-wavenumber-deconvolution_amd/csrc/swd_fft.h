// swd_fft.h -- radix-16 Stockham FFT along a column in registers and LDS, and
// the kz up/down split built on it (device code; included by .hip files).
//
// A transform of N points (a power of two, at least V) runs on N/V threads
// that hold V = 8 or 16 values each.  Each thread keeps the values of its
// butterflies in registers; a pass reads them from LDS, waits, applies the
// twiddles and an in-register DFT (radix V, 16 = 4 x 4; a last pass of a
// smaller radix when log2 N is not a multiple of log2 V), and writes its
// outputs in place; the last pass writes natural order.  An N-point transform
// thus makes ceil(log2 N / log2 V) trips through LDS; V = 8 gives twice the
// threads for the same transform, for launches too small to fill the GPU.
// Twiddles come from the table e^{-2 pi i q/N} (q < N/2, evaluated in double
// and rounded): a radix-R butterfly loads w, w^2, w^4, w^8 and forms the other
// powers by at most three complex products.  The inverse is
// conj(FFT(conj(Y))).  LDS holds N + N/16 values: one pad element per 16.
#pragma once
#include <hip/hip_runtime.h>

namespace {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }   // a * (-i)
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// LDS index with one pad element per 16: the first pass writes 16 consecutive
// values per thread, which would otherwise put every other thread on the same banks.
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

// e^{-2 pi i q/N} for 0 <= q < N from the half table
template <int N>
__device__ __forceinline__ float2 twf(const float2 *__restrict__ tw, int q)
{
    const float2 w = tw[q & (N / 2 - 1)];
    return q < N / 2 ? w : make_float2(-w.x, -w.y);
}

// In-register DFTs (forward, e^{-2 pi i jk/R}), natural order in and out, on
// v[O .. O+R-1].
template <int V, int O>
__device__ __forceinline__ void dft2(float2 (&v)[V])
{
    const float2 a = v[O], b = v[O + 1];
    v[O] = cadd(a, b);
    v[O + 1] = csub(a, b);
}

__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3)
{
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_mi(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <int V, int O>
__device__ __forceinline__ void dft4(float2 (&v)[V])
{
    dft4(v[O], v[O + 1], v[O + 2], v[O + 3]);
}

constexpr float kC8 = 0.92387953251128674f;   // cos(pi/8)
constexpr float kS8 = 0.38268343236508977f;   // sin(pi/8)
constexpr float kR2 = 0.70710678118654752f;   // sqrt(1/2)

template <int V, int O>
__device__ __forceinline__ void dft8(float2 (&v)[V])
{
    // even and odd halves, then X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k]
    float2 e0 = v[O], e1 = v[O + 2], e2 = v[O + 4], e3 = v[O + 6];
    float2 o0 = v[O + 1], o1 = v[O + 3], o2 = v[O + 5], o3 = v[O + 7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = make_float2(kR2 * (o1.x + o1.y), kR2 * (o1.y - o1.x));    // * (1 - i)/sqrt2
    o2 = mul_mi(o2);                                                // * -i
    o3 = make_float2(kR2 * (o3.y - o3.x), -kR2 * (o3.x + o3.y));   // * (-1 - i)/sqrt2
    v[O] = cadd(e0, o0);
    v[O + 4] = csub(e0, o0);
    v[O + 1] = cadd(e1, o1);
    v[O + 5] = csub(e1, o1);
    v[O + 2] = cadd(e2, o2);
    v[O + 6] = csub(e2, o2);
    v[O + 3] = cadd(e3, o3);
    v[O + 7] = csub(e3, o3);
}

// 16 = 4 x 4: X[k1 + 4 k2] = sum_r2 W16^(r2 k1) W4^(r2 k2) sum_r1 x[4 r1 + r2] W4^(r1 k1)
__device__ __forceinline__ void dft16(float2 (&v)[16])
{
#pragma unroll
    for (int r2 = 0; r2 < 4; r2++) dft4(v[r2], v[4 + r2], v[8 + r2], v[12 + r2]);
    // slot 4 k1 + r2 holds the inner DFT; twiddle W16^(r2 k1)
    const float2 w1 = make_float2(kC8, -kS8), w2 = make_float2(kR2, -kR2), w3 = make_float2(kS8, -kC8);
    const float2 w6 = make_float2(-kR2, -kR2), w9 = make_float2(-kC8, kS8);
    v[5] = cmul(v[5], w1);
    v[6] = cmul(v[6], w2);
    v[7] = cmul(v[7], w3);
    v[9] = cmul(v[9], w2);
    v[10] = mul_mi(v[10]);
    v[11] = cmul(v[11], w6);
    v[13] = cmul(v[13], w3);
    v[14] = cmul(v[14], w6);
    v[15] = cmul(v[15], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    // slot 4 k1 + k2 holds X[k1 + 4 k2]: transpose to natural order
    float2 t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) t[i] = v[i];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = t[4 * k1 + k2];
}

template <int V, int R, int O>
__device__ __forceinline__ void dft_r(float2 (&v)[V])
{
    if constexpr (R == 2) dft2<V, O>(v);
    else if constexpr (R == 4) dft4<V, O>(v);
    else if constexpr (R == 8) dft8<V, O>(v);
    else dft16(v);
}

// Twiddles of butterfly j of a pass (span NS, radix R): x[r] *= W^(r k),
// W = e^{-2 pi i / (NS R)}, k = j mod NS.
template <int N, int V, int R, int NS, int O>
__device__ __forceinline__ void twiddle(float2 (&v)[V], int j, const float2 *__restrict__ tw)
{
    const int q1 = (j & (NS - 1)) * (N / (NS * R));
    const float2 w1 = twf<N>(tw, q1);
    v[O + 1] = cmul(v[O + 1], w1);
    if constexpr (R >= 4) {
        const float2 w2 = twf<N>(tw, 2 * q1), w3 = cmul(w1, w2);
        v[O + 2] = cmul(v[O + 2], w2);
        v[O + 3] = cmul(v[O + 3], w3);
        if constexpr (R >= 8) {
            const float2 w4 = twf<N>(tw, 4 * q1);
            v[O + 4] = cmul(v[O + 4], w4);
            v[O + 5] = cmul(v[O + 5], cmul(w4, w1));
            v[O + 6] = cmul(v[O + 6], cmul(w4, w2));
            v[O + 7] = cmul(v[O + 7], cmul(w4, w3));
            if constexpr (R >= 16) {
                const float2 w8 = twf<N>(tw, 8 * q1);
                v[8] = cmul(v[8], w8);
                v[9] = cmul(v[9], cmul(w8, w1));
                v[10] = cmul(v[10], cmul(w8, w2));
                v[11] = cmul(v[11], cmul(w8, w3));
                v[12] = cmul(v[12], cmul(w8, w4));
                v[13] = cmul(v[13], cmul(w8, cmul(w4, w1)));
                v[14] = cmul(v[14], cmul(w8, cmul(w4, w2)));
                v[15] = cmul(v[15], cmul(w8, cmul(w4, w3)));
            }
        }
    }
}

// One Stockham pass after the first (span NS > 1) of the thread's V/R
// butterflies j = t + (N/V) b: inputs x[j + r N/R], outputs at
// (j / NS) NS R + (j mod NS) + r NS.  CONJ conjugates what is stored.
template <int N, int V, int R, int NS, bool CONJ>
__device__ __forceinline__ void pass(float2 (&v)[V], int t, float2 *L, const float2 *__restrict__ tw)
{
    constexpr int T = N / V, NB = V / R;
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
        for (int r = 0; r < R; r++) v[b * R + r] = L[pad(t + T * b + r * (N / R))];
    __syncthreads();   // every input read before any output overwrites it
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = t + T * b;
        if constexpr (R == V) {
            twiddle<N, V, R, NS, 0>(v, j, tw);
            dft_r<V, R, 0>(v);
        } else {
            // b is a compile-time constant after unrolling; dispatch on it
            if (b == 0) { twiddle<N, V, R, NS, 0>(v, j, tw); dft_r<V, R, 0>(v); }
            if constexpr (NB > 1) if (b == 1) { twiddle<N, V, R, NS, R>(v, j, tw); dft_r<V, R, R>(v); }
            if constexpr (NB > 2) if (b == 2) { twiddle<N, V, R, NS, 2 * R>(v, j, tw); dft_r<V, R, 2 * R>(v); }
            if constexpr (NB > 3) if (b == 3) { twiddle<N, V, R, NS, 3 * R>(v, j, tw); dft_r<V, R, 3 * R>(v); }
            if constexpr (NB > 4) if (b == 4) { twiddle<N, V, R, NS, 4 * R>(v, j, tw); dft_r<V, R, 4 * R>(v); }
            if constexpr (NB > 5) if (b == 5) { twiddle<N, V, R, NS, 5 * R>(v, j, tw); dft_r<V, R, 5 * R>(v); }
            if constexpr (NB > 6) if (b == 6) { twiddle<N, V, R, NS, 6 * R>(v, j, tw); dft_r<V, R, 6 * R>(v); }
            if constexpr (NB > 7) if (b == 7) { twiddle<N, V, R, NS, 7 * R>(v, j, tw); dft_r<V, R, 7 * R>(v); }
        }
        const int base = (j / NS) * NS * R + (j & (NS - 1));
#pragma unroll
        for (int r = 0; r < R; r++) L[pad(base + r * NS)] = CONJ ? conjf2(v[b * R + r]) : v[b * R + r];
    }
    __syncthreads();
}

// Passes of span NS .. N after the first; the last one (span N/R) writes natural order.
template <int N, int V, int NS, bool CONJ>
__device__ __forceinline__ void passes(float2 (&v)[V], int t, float2 *L, const float2 *__restrict__ tw)
{
    if constexpr (NS < N) {
        constexpr int R = N / NS >= V ? V : N / NS;
        pass<N, V, R, NS, CONJ && NS * R == N>(v, t, L, tw);
        passes<N, V, NS * R, CONJ>(v, t, L, tw);
    }
}

// First pass (span 1, radix V) of the thread's butterfly j = t: inputs
// x[t + r N/V] already in v, outputs at V t + r.
template <int N, int V>
__device__ __forceinline__ void first_pass(float2 (&v)[V], int t, float2 *L)
{
    dft_r<V, V, 0>(v);
#pragma unroll
    for (int r = 0; r < V; r++) L[pad(V * t + r)] = v[r];
    __syncthreads();
}

// The kz up/down split of one column (the AFor separation, FK:52-138) is
//   first_pass, passes<N, V, V, false>   forward transform, X in L
//   kz_pack                              v = conj(Y) for the inverse
//   first_pass, passes<N, V, V, true>    L = (Re up, Re down), natural order
// The spectrum X is weighted by the taper c and split (bins 0..N/2 "up",
// the rest "down").  Only real parts are needed, so both halves share one
// inverse transform: with m = -k mod N,
//   Y[k] = (c X)_up[k] + conj((c X)_up[m])  +  i ((c X)_down[k] + conj((c X)_down[m])),
// scaled by 2 sc (sc = 1/2 gives the unnormalised inverse, 1/(2N) the
// normalised one), transforms back to Re(up) + i Re(down).  Every thread of
// the block takes part (barriers).  The kernels call the steps in sequence:
// one function around all of them made the compiler keep about 40 more
// registers live.
template <int N, int V>
__device__ __forceinline__ void kz_pack(float2 (&v)[V], int t, const float2 *L, const float *coef, float sc)
{
    constexpr int T = N / V;
#pragma unroll
    for (int r = 0; r < V; r++) {
        // k = t + r N/V lies below N/2 exactly when r < V/2
        const int k = t + r * T, m = (N - k) & (N - 1);
        const float2 xk = L[pad(k)], xm = L[pad(m)];
        const float ck = coef[k] * sc, cm = coef[m] * sc;
        float2 y;
        if (r < V / 2) {
            const float2 u = make_float2(ck * xk.x, ck * xk.y), d = make_float2(cm * xm.x, cm * xm.y);
            y = make_float2(u.x + d.y, u.y + d.x);   // c_k X_k + i conj(c_m X_m)
        } else {
            const float2 u = make_float2(cm * xm.x, cm * xm.y), d = make_float2(ck * xk.x, ck * xk.y);
            y = make_float2(u.x - d.y, d.x - u.y);   // conj(c_m X_m) + i c_k X_k
        }
        if ((r == 0 || r == V / 2) && t == 0) y = make_float2(2.f * ck * xk.x, 0.f);   // DC and Nyquist are "up"
        v[r] = conjf2(y);
    }
    __syncthreads();   // the spectrum is read before the inverse's first pass overwrites it
}

}  // namespace
