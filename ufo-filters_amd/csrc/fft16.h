// Register FFT building blocks for the ramp-filter kernels: radix-16 Stockham
// passes (with one radix-2/4/8 pass when log2 N is not a multiple of 4) where
// every thread keeps 16 (32 for N = 8192) complex values in registers.
//
//  * TF = min(256, N/16) threads work on one transform; a thread owns the
//    butterflies j = t + q*TF of every pass.
//  * The first forward pass reads its inputs straight from global memory and
//    the last inverse pass writes straight to global memory.
//  * The last forward pass and the first inverse pass use the same radix and
//    butterfly assignment, so the spectrum never leaves registers between
//    them: the filter multiply happens there.
//  * Passes exchange data through one LDS buffer of N complex values
//    (read all, barrier, write all, barrier).
// Twiddles exp(-+ 2 pi i m / M) come from the context's FP64-accurate table;
// a radix-16 pass loads w, w^4, w^8 and forms the other powers with at most
// three complex products.
#pragma once

#include "ufomi_internal.h"

namespace fft16 {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (DIR < 0) or +i (DIR > 0)
template <int DIR> __device__ __forceinline__ float2 rot(float2 a) { return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x); }

// N-point FFT in LDS (x[0..N)), exp(DIR * 2 pi i k n / N), cooperative over the
// TF threads with local index t.  Caller synchronises before (data ready); the
// function ends with a barrier (result ready).
template <int B, int DIR>
__device__ void lds_fft(float2 *x, int N, int log2N, int t, int TF, const float2 *__restrict__ tw)
{
    const int M = UFOMI_FFT_MAX;
    int Ns = 1;
    if (log2N & 1) {
        // radix-2 pass at Ns = 1: no twiddles; y[2j] = a + b, y[2j+1] = a - b.
        float2 a[2 * B], b[2 * B];
#pragma unroll
        for (int q = 0; q < 2 * B; q++) {
            const int j = t + q * TF;
            a[q] = x[j];
            b[q] = x[j + N / 2];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2 * B; q++) {
            const int j = t + q * TF;
            x[2 * j] = cadd(a[q], b[q]);
            x[2 * j + 1] = csub(a[q], b[q]);
        }
        __syncthreads();
        Ns = 2;
    }
    for (; Ns < N; Ns *= 4) {
        float2 v[B][4];
        const int quarter = N / 4;
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int j = t + q * TF;
#pragma unroll
            for (int r = 0; r < 4; r++)
                v[q][r] = x[j + r * quarter];
        }
        __syncthreads();
        const int tstep = M / (Ns * 4);  // table stride of one unit of angle 2 pi / (4 Ns)
#pragma unroll
        for (int q = 0; q < B; q++) {
            const int j = t + q * TF;
            const int k = j & (Ns - 1);
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < 4; r++) {
                    float2 w = tw[k * r * tstep];
                    if (DIR > 0)
                        w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            // 4-point DFT with W4 = exp(DIR * 2 pi i / 4)
            const float2 s02 = cadd(v[q][0], v[q][2]), d02 = csub(v[q][0], v[q][2]);
            const float2 s13 = cadd(v[q][1], v[q][3]), d13 = rot<DIR>(csub(v[q][1], v[q][3]));
            const int idx = (j - k) * 4 + k;
            x[idx] = cadd(s02, s13);
            x[idx + Ns] = cadd(d02, d13);
            x[idx + 2 * Ns] = csub(s02, s13);
            x[idx + 3 * Ns] = csub(d02, d13);
        }
        __syncthreads();
    }
}

constexpr float kS2 = 0.70710678118654752440f;   // cos(pi/4)
constexpr float kC8 = 0.92387953251128675613f;   // cos(pi/8)
constexpr float kS8 = 0.38268343236508977173f;   // sin(pi/8)

// exp(DIR * 2 pi i m / 16) for m = 0..9 (the products b*c of the 4x4 split)
template <int DIR> __device__ __forceinline__ float2 w16(int m)
{
    const float d = (float) DIR;
    switch (m) {
    case 0: return make_float2(1.0f, 0.0f);
    case 1: return make_float2(kC8, d * kS8);
    case 2: return make_float2(kS2, d * kS2);
    case 3: return make_float2(kS8, d * kC8);
    case 4: return make_float2(0.0f, d);
    case 6: return make_float2(-kS2, d * kS2);
    case 9: return make_float2(-kC8, -d * kS8);
    default: return make_float2(0.0f, 0.0f);  // m = 5, 7, 8 do not occur (b, c <= 3)
    }
}

template <int DIR> __device__ __forceinline__ void dft2(float2 *v)
{
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <int DIR> __device__ __forceinline__ void dft4(float2 &a, float2 &b, float2 &c, float2 &d)
{
    const float2 s02 = cadd(a, c), d02 = csub(a, c);
    const float2 s13 = cadd(b, d), d13 = rot<DIR>(csub(b, d));
    a = cadd(s02, s13);
    b = cadd(d02, d13);
    c = csub(s02, s13);
    d = csub(d02, d13);
}

template <int DIR> __device__ __forceinline__ void dft8(float2 *v)
{
    // split r = 2a + b: 4-point DFTs over a, twiddle W8^(b c), 2-point over b
    dft4<DIR>(v[0], v[2], v[4], v[6]);
    dft4<DIR>(v[1], v[3], v[5], v[7]);
    const float d = (float) DIR;
    const float2 w1 = make_float2(kS2, d * kS2), w3 = make_float2(-kS2, d * kS2);
    v[3] = cmul(v[3], w1);
    v[5] = rot<DIR>(v[5]);
    v[7] = cmul(v[7], w3);
    float2 o[8];
    for (int c = 0; c < 4; c++) {
        o[c] = cadd(v[2 * c], v[2 * c + 1]);
        o[c + 4] = csub(v[2 * c], v[2 * c + 1]);
    }
    for (int i = 0; i < 8; i++)
        v[i] = o[i];
}

template <int DIR> __device__ __forceinline__ void dft16(float2 *v)
{
    // r = 4a + b, q = c + 4d: y[b][c] = DFT4_a(v[4a+b]); y *= W16^(bc); V[c+4d] = DFT4_b(y[b][c])
#pragma unroll
    for (int b = 0; b < 4; b++)
        dft4<DIR>(v[b], v[4 + b], v[8 + b], v[12 + b]);
    // after dft4 over a, slot 4c + b holds y[b][c]
#pragma unroll
    for (int b = 1; b < 4; b++)
#pragma unroll
        for (int c = 1; c < 4; c++)
            v[4 * c + b] = cmul(v[4 * c + b], w16<DIR>(b * c));
#pragma unroll
    for (int c = 0; c < 4; c++)
        dft4<DIR>(v[4 * c + 0], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
    // slot 4c + d holds V[c + 4d]: transpose to natural order
    float2 o[16];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int d = 0; d < 4; d++)
            o[c + 4 * d] = v[4 * c + d];
#pragma unroll
    for (int i = 0; i < 16; i++)
        v[i] = o[i];
}

template <int R, int DIR> __device__ __forceinline__ void dft(float2 *v)
{
    if constexpr (R == 2) dft2<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<DIR>(v);
    else dft16<DIR>(v);
}

// Twiddle w^m for the butterfly's base w = exp(DIR 2 pi i k / (R Ns)); table
// entry t[i] = exp(-2 pi i i / M).
template <int DIR> __device__ __forceinline__ float2 tw_at(const float2 *__restrict__ t, int idx)
{
    // the table pointer is uniform: a 32-bit lane offset from it (no 64-bit address)
    const uint64_t raw = *(const __attribute__((address_space(1))) uint64_t *) (
        (const __attribute__((address_space(1))) char *) t + (unsigned) idx * 8u);
    float2 w = make_float2(__uint_as_float((unsigned) raw), __uint_as_float((unsigned) (raw >> 32)));
    if (DIR > 0) w.y = -w.y;
    return w;
}

template <int R, int DIR> __device__ __forceinline__ void twiddle(float2 *v, int k, int step, const float2 *__restrict__ t)
{
    if constexpr (R == 2) {
        v[1] = cmul(v[1], tw_at<DIR>(t, k * step));
    } else if constexpr (R == 4) {
        v[1] = cmul(v[1], tw_at<DIR>(t, k * step));
        v[2] = cmul(v[2], tw_at<DIR>(t, 2 * k * step));
        v[3] = cmul(v[3], tw_at<DIR>(t, 3 * k * step));
    } else if constexpr (R == 8) {
        const float2 w1 = tw_at<DIR>(t, k * step), w2 = tw_at<DIR>(t, 2 * k * step), w4 = tw_at<DIR>(t, 4 * k * step);
        const float2 w3 = cmul(w1, w2);
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        v[4] = cmul(v[4], w4);
        v[5] = cmul(v[5], cmul(w4, w1));
        v[6] = cmul(v[6], cmul(w4, w2));
        v[7] = cmul(v[7], cmul(w4, w3));
    } else {
        const float2 w1 = tw_at<DIR>(t, k * step), w4 = tw_at<DIR>(t, 4 * k * step), w8 = tw_at<DIR>(t, 8 * k * step);
        const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1), w12 = cmul(w8, w4);
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        v[4] = cmul(v[4], w4);
        v[5] = cmul(v[5], cmul(w4, w1));
        v[6] = cmul(v[6], cmul(w4, w2));
        v[7] = cmul(v[7], cmul(w4, w3));
        v[8] = cmul(v[8], w8);
        v[9] = cmul(v[9], cmul(w8, w1));
        v[10] = cmul(v[10], cmul(w8, w2));
        v[11] = cmul(v[11], cmul(w8, w3));
        v[12] = cmul(v[12], w12);
        v[13] = cmul(v[13], cmul(w12, w1));
        v[14] = cmul(v[14], cmul(w12, w2));
        v[15] = cmul(v[15], cmul(w12, w3));
    }
}

// Compile-time plan: log2 N = 4 P16 + rem bits; forward = [2^rem?, 16 x P16],
// inverse = reversed.  V = values per thread.
template <int N> struct Plan {
    static constexpr int L = N == 16 ? 4 : N == 32 ? 5 : N == 64 ? 6 : N == 128 ? 7 : N == 256 ? 8 : N == 512 ? 9
                           : N == 1024 ? 10 : N == 2048 ? 11 : N == 4096 ? 12 : 13;
    static constexpr int REM = L % 4 == 0 ? 0 : (1 << (L % 4));  // 0, 2, 4 or 8
    static constexpr int P16 = L / 4;
    static constexpr int NPASS = P16 + (REM ? 1 : 0);
    static constexpr int TF = N / 16 < 256 ? N / 16 : 256;  // threads per transform
    static constexpr int G = 256 / TF;                       // transforms per workgroup
    static constexpr int V = N / TF;                          // values per thread (16 or 32)
    // radix of forward pass p (inverse pass p uses forward pass NPASS-1-p)
    static constexpr int radix(int p) { return (REM && p == 0) ? REM : 16; }
};

// LDS slot of element i of a transform's exchange buffer.  A radix-16 pass
// stores at stride 16 (the first pass: 16 j + r), which puts the lanes of a
// wave 32 dwords apart, on the same few banks; one pad element after every 16
// spreads them.  N = 8192 stays unpadded (its buffer would pass 64 KiB).
template <int N> struct Lds {
    static constexpr bool PAD = N <= 4096;
    static constexpr int STRIDE = PAD ? N + N / 16 : N;  // elements per transform
    static __device__ __forceinline__ int at(int i) { return PAD ? i + (i >> 4) : i; }
};

// One Stockham pass in registers: v holds this thread's butterflies of radix R
// (values v[q*R + r] = x[j + r N/R], j = t + q TF); stride Ns of the pass.
template <int N, int R, int DIR>
__device__ __forceinline__ void pass_compute(float2 *v, int t, int Ns, const float2 *__restrict__ tw)
{
    using P = Plan<N>;
    constexpr int NB = P::V / R;
    const int step = UFOMI_FFT_MAX / (R * Ns);
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int j = t + q * P::TF;
        const int k = j & (Ns - 1);
        if (Ns > 1)
            twiddle<R, DIR>(v + q * R, k, step, tw);
        dft<R, DIR>(v + q * R);
    }
}

// Write the outputs of a pass at stride Ns to LDS (Stockham expand).
template <int N, int R>
__device__ __forceinline__ void pass_store(float2 *x, const float2 *v, int t, int Ns)
{
    using P = Plan<N>;
    constexpr int NB = P::V / R;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int j = t + q * P::TF;
        const int k = j & (Ns - 1);
        const int idx = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++)
            x[Lds<N>::at(idx + r * Ns)] = v[q * R + r];
    }
}

// Read the inputs of a radix-R pass from LDS.
template <int N, int R>
__device__ __forceinline__ void pass_load(const float2 *x, float2 *v, int t)
{
    using P = Plan<N>;
    constexpr int NB = P::V / R;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int j = t + q * P::TF;
#pragma unroll
        for (int r = 0; r < R; r++)
            v[q * R + r] = x[Lds<N>::at(j + r * (N / R))];
    }
}

// Element index held in register slot s before the first pass of radix R
// (and after the last pass of radix R): j + r N/R.
template <int N, int R> __device__ __forceinline__ int slot_index(int t, int s)
{
    using P = Plan<N>;
    const int q = s / R, r = s % R;
    return t + q * P::TF + r * (N / R);
}

// Forward transform: v holds slot_index<N, R0> inputs on entry and
// slot_index<N, RL> outputs on exit (R0 = first, RL = last forward radix).
template <int N, int DIR>
__device__ __forceinline__ void transform_forward_order(float2 *v, float2 *x, int t, const float2 *__restrict__ tw)
{
    using P = Plan<N>;
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < P::NPASS; p++) {
        const int R = P::radix(p);
        if (p > 0) {
            // exchange through LDS: previous pass outputs -> this pass inputs
            __syncthreads();
            if (R == 16) pass_load<N, 16>(x, v, t);
            else if (R == 8) pass_load<N, 8>(x, v, t);
            else if (R == 4) pass_load<N, 4>(x, v, t);
            else pass_load<N, 2>(x, v, t);
        }
        if (R == 16) pass_compute<N, 16, DIR>(v, t, Ns, tw);
        else if (R == 8) pass_compute<N, 8, DIR>(v, t, Ns, tw);
        else if (R == 4) pass_compute<N, 4, DIR>(v, t, Ns, tw);
        else pass_compute<N, 2, DIR>(v, t, Ns, tw);
        if (p + 1 < P::NPASS) {
            __syncthreads();  // everyone has read the previous contents
            if (R == 16) pass_store<N, 16>(x, v, t, Ns);
            else if (R == 8) pass_store<N, 8>(x, v, t, Ns);
            else if (R == 4) pass_store<N, 4>(x, v, t, Ns);
            else pass_store<N, 2>(x, v, t, Ns);
        }
        Ns *= R;
    }
    // the last pass (Ns_in = N / RL) leaves x[j + r N/RL] in slot q*RL + r: natural order by slot_index
}

// Same with the pass order reversed (used for the inverse so that its first
// radix equals the forward's last radix).
template <int N, int DIR>
__device__ __forceinline__ void transform_reverse_order(float2 *v, float2 *x, int t, const float2 *__restrict__ tw)
{
    using P = Plan<N>;
    int Ns = 1;
#pragma unroll
    for (int pp = 0; pp < P::NPASS; pp++) {
        const int R = P::radix(P::NPASS - 1 - pp);
        if (pp > 0) {
            __syncthreads();
            if (R == 16) pass_load<N, 16>(x, v, t);
            else if (R == 8) pass_load<N, 8>(x, v, t);
            else if (R == 4) pass_load<N, 4>(x, v, t);
            else pass_load<N, 2>(x, v, t);
        }
        if (R == 16) pass_compute<N, 16, DIR>(v, t, Ns, tw);
        else if (R == 8) pass_compute<N, 8, DIR>(v, t, Ns, tw);
        else if (R == 4) pass_compute<N, 4, DIR>(v, t, Ns, tw);
        else pass_compute<N, 2, DIR>(v, t, Ns, tw);
        if (pp + 1 < P::NPASS) {
            __syncthreads();
            if (R == 16) pass_store<N, 16>(x, v, t, Ns);
            else if (R == 8) pass_store<N, 8>(x, v, t, Ns);
            else if (R == 4) pass_store<N, 4>(x, v, t, Ns);
            else pass_store<N, 2>(x, v, t, Ns);
        }
        Ns *= R;
    }
}

}  // namespace fft16
