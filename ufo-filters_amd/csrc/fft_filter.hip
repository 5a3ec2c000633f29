// Ramp filtering: `fft ! filter ! ifft` as batched FFTs on 256-thread
// workgroups.
//
// 16 <= N <= 8192: the register radix-16 kernels of fft16.h (ramp16_kernel,
// fft16_rows_kernel, ifft16_rows_kernel): first pass straight from HBM, filter
// multiply in registers, last pass straight to HBM, passes exchanged through
// one LDS buffer.
// N = 4, 8: a radix-4 Stockham FFT in LDS (lds_fft): TF = N/4 threads per
// transform, each keeping its butterflies in registers between the read and
// write halves of a pass.
// Twiddles come from the context's FP64-accurate table (read through L2).
#include "ufomi_internal.h"
#include "fft16.h"

namespace {

constexpr int kThreads = 256;

using fft16::lds_fft;

struct Layout {
    int TF, G, f, t;  // threads per transform, transforms per WG, this thread's transform and lane
};

__device__ __forceinline__ Layout layout(int N)
{
    Layout L;
    L.TF = N / 4 < kThreads ? N / 4 : kThreads;
    L.G = kThreads / L.TF;
    L.f = threadIdx.x / L.TF;
    L.t = threadIdx.x % L.TF;
    return L;
}

// Fused `fft ! filter ! ifft`.  The coefficient class comes from the device
// flag written by classify_kernel in the same stream (no host round trip):
//  * real, even spectrum (c[2k] == c[2k+1], c_k == c_{N-k}; every filter but
//    faris-byer): rows 2m and 2m+1 ride in the real and imaginary parts of one
//    complex transform.  A real even spectrum maps real rows to real rows, so
//    the two results come back separated in Re and Im.  Blocks past the last
//    pair exit at once (the grid is sized for one row per transform).
//  * otherwise: one row per transform and Y_k = Re X_k c[2k] + i Im X_k c[2k+1],
//    element-wise exactly as filter.cl does.
template <int B, typename OutT>
__global__ __launch_bounds__(kThreads) void ramp_kernel(
    const float *__restrict__ in, int W, long rows, int N, int log2N, const float *__restrict__ coef,
    const int *__restrict__ pair_flag, int crop, OutT *__restrict__ out, float scale, const float2 *__restrict__ tw)
{
    extern __shared__ float2 lds[];
    const Layout L = layout(N);
    float2 *x = lds + (size_t) L.f * N;
    const bool pair = *pair_flag != 0;
    const long unit = (long) blockIdx.x * L.G + L.f;
    const long r0 = pair ? 2 * unit : unit, r1 = r0 + 1;
    if (pair && 2 * (long) blockIdx.x * L.G >= rows)
        return;  // whole block beyond the last pair
    const bool v0 = r0 < rows, v1 = pair && r1 < rows;
    const float *a = in + r0 * W, *b = in + r1 * W;

    for (int i = L.t; i < N; i += L.TF) {
        float re = 0.0f, im = 0.0f;
        if (i < W) {
            if (v0) re = a[i];
            if (v1) im = b[i];
        }
        x[i] = make_float2(re, im);
    }
    __syncthreads();
    lds_fft<B, -1>(x, N, log2N, L.t, L.TF, tw);
    for (int k = L.t; k < N; k += L.TF) {
        const float2 z = x[k];
        const float cr = coef[2 * k];
        const float ci = pair ? cr : coef[2 * k + 1];
        x[k] = make_float2(z.x * cr, z.y * ci);
    }
    __syncthreads();
    lds_fft<B, +1>(x, N, log2N, L.t, L.TF, tw);
    for (int i = L.t; i < crop; i += L.TF) {
        const float2 z = x[i];
        if (v0) ufomi_store(out, r0 * crop + i, z.x * scale);
        if (v1) ufomi_store(out, r1 * crop + i, z.y * scale);
    }
}

// flag = 1 when c[2k] == c[2k+1] and c[2k] == c[2(N-k)] for all k (one block).
__global__ __launch_bounds__(kThreads) void classify_kernel(const float *__restrict__ c, int N, int *__restrict__ flag)
{
    int ok = 1;
    for (int k = threadIdx.x; k < N; k += kThreads) {
        const float re = c[2 * k];
        if (re != c[2 * k + 1] || (k && re != c[2 * (N - k)]))
            ok = 0;
    }
    ok = __syncthreads_and(ok);
    if (threadIdx.x == 0)
        *flag = ok;
}

// `fft`: real rows -> zero-padded, forward-transformed interleaved complex rows.
template <int B>
__global__ __launch_bounds__(kThreads) void fft_rows_kernel(
    const float *__restrict__ in, int W, long rows, int N, int log2N, float2 *__restrict__ out,
    const float2 *__restrict__ tw)
{
    extern __shared__ float2 lds[];
    const Layout L = layout(N);
    float2 *x = lds + (size_t) L.f * N;
    const long r = (long) blockIdx.x * L.G + L.f;
    const bool valid = r < rows;
    const float *a = in + r * W;
    for (int i = L.t; i < N; i += L.TF)
        x[i] = make_float2((valid && i < W) ? a[i] : 0.0f, 0.0f);
    __syncthreads();
    lds_fft<B, -1>(x, N, log2N, L.t, L.TF, tw);
    if (valid)
        for (int i = L.t; i < N; i += L.TF)
            out[r * N + i] = x[i];
}

// `ifft`: interleaved complex rows -> Re(backward transform)[0:crop] * scale.
template <int B>
__global__ __launch_bounds__(kThreads) void ifft_rows_kernel(
    const float2 *__restrict__ in, long rows, int N, int log2N, int crop, float *__restrict__ out,
    float scale, const float2 *__restrict__ tw)
{
    extern __shared__ float2 lds[];
    const Layout L = layout(N);
    float2 *x = lds + (size_t) L.f * N;
    const long r = (long) blockIdx.x * L.G + L.f;
    const bool valid = r < rows;
    for (int i = L.t; i < N; i += L.TF)
        x[i] = valid ? in[r * N + i] : make_float2(0.0f, 0.0f);
    __syncthreads();
    lds_fft<B, +1>(x, N, log2N, L.t, L.TF, tw);
    if (valid)
        for (int i = L.t; i < crop; i += L.TF)
            out[r * crop + i] = x[i].x * scale;
}

// `filter`: element-wise product with the 2N coefficients, 4 floats per lane.
__global__ __launch_bounds__(kThreads) void filter_apply_kernel(
    const float *__restrict__ in, float *__restrict__ out, const float *__restrict__ coef, int width, long total)
{
    const long stride = (long) gridDim.x * kThreads;
    for (long i = (long) blockIdx.x * kThreads + threadIdx.x; i < total; i += stride) {
        const int c = (int) (i % width);
        out[i] = in[i] * coef[c];
    }
}

int fft_geometry(int N, int *B, int *G, size_t *lds_bytes)
{
    const int TF = N / 4 < kThreads ? N / 4 : kThreads;
    *G = kThreads / TF;
    *B = N / (4 * TF);
    *lds_bytes = (size_t) (*G) * N * sizeof(float2);
    return UFOMI_OK;
}

int check_n(int N)
{
    if (!ufomi_is_pow2(N) || N < 4 || N > UFOMI_FFT_MAX_N)
        return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "FFT size %d must be a power of two in [4, %d]", N, UFOMI_FFT_MAX_N);
    return UFOMI_OK;
}

template <typename K>
int launch_by_b(int B, K k)
{
    switch (B) {
    case 1: return k(std::integral_constant<int, 1>());
    case 2: return k(std::integral_constant<int, 2>());
    case 4: return k(std::integral_constant<int, 4>());
    case 8: return k(std::integral_constant<int, 8>());
    case 16: return k(std::integral_constant<int, 16>());
    default: return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "no FFT kernel for %d butterflies per thread", B);
    }
}

// ------------------------------------------------------------------------
// Register radix-16 kernels (fft16.h) for 16 <= N <= 8192.  The kernels above
// remain for N = 4 and 8.
// ------------------------------------------------------------------------

// Wave-uniform row pointers in SGPRs, so that a row's loads and stores take the
// scalar-base + 32-bit lane-offset form (no 64-bit address per element).
template <typename T> __device__ __forceinline__ T *uniform_row(T *p)
{
    const uint64_t v = (uint64_t) p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned) v), hi = __builtin_amdgcn_readfirstlane((unsigned) (v >> 32));
    return (T *) (((uint64_t) hi << 32) | lo);
}
using GlbChar = __attribute__((address_space(1))) char;
// Element i of a row (rows stay below 4 GiB): a 32-bit byte offset from its base.
template <typename R> __device__ __forceinline__ R glb_ld(const void *base, unsigned i)
{
    return *(const __attribute__((address_space(1))) R *) ((const GlbChar *) base + i * (unsigned) sizeof(R));
}
template <typename R> __device__ __forceinline__ void glb_st(void *base, unsigned i, R v)
{
    *(__attribute__((address_space(1))) R *) ((GlbChar *) base + i * (unsigned) sizeof(R)) = v;
}
__device__ __forceinline__ void glb_store(float *p, unsigned i, float v) { glb_st<float>(p, i, v); }
__device__ __forceinline__ void glb_store(__half *p, unsigned i, float v)
{
    glb_st<unsigned short>(p, i, __builtin_bit_cast(unsigned short, (_Float16) v));
}
__device__ __forceinline__ void glb_store(__hip_bfloat16 *p, unsigned i, float v)
{
    glb_st<unsigned short>(p, i, __builtin_bit_cast(unsigned short, __float2bfloat16(v)));
}

// Block `blk` of a ramp-filter launch over `rows` rows of width W: G transforms
// (2 G rows when `pair`, the real-even coefficient case) of the rows' block.
template <int N, typename OutT>
__device__ __forceinline__ void ramp16_block(
    const float *__restrict__ in, int W, long rows, const float *__restrict__ coef, bool pair,
    int crop, OutT *__restrict__ out, float scale, const float2 *__restrict__ tw, long blk)
{
    using P = fft16::Plan<N>;
    constexpr int R0 = P::radix(0), RL = P::radix(P::NPASS - 1);
    extern __shared__ float2 lds[];
    const int f = threadIdx.x / P::TF, t = threadIdx.x % P::TF;
    float2 *x = lds + (size_t) f * fft16::Lds<N>::STRIDE;
    if (pair && 2 * blk * P::G >= rows)
        return;  // whole block beyond the last pair
    const long unit = blk * P::G + f;
    const long r0 = pair ? 2 * unit : unit, r1 = r0 + 1;
    const bool v0 = r0 < rows, v1 = pair && r1 < rows;
    const float *a = in + r0 * W, *b = in + r1 * W;
    OutT *o0 = out + r0 * crop, *o1 = out + r1 * crop;

    // A transform of N >= 1024 points is whole waves (TF >= 64): its rows are
    // wave-uniform, and elements are addressed by 32-bit offsets from them.
    constexpr bool kUniform = P::TF >= 64;
    const float *ua = nullptr, *ub = nullptr;
    OutT *uo0 = nullptr, *uo1 = nullptr;
    if constexpr (kUniform) {
        ua = uniform_row(a);
        ub = uniform_row(b);
        uo0 = uniform_row(o0);
        uo1 = uniform_row(o1);
    }
    float2 v[P::V];
#pragma unroll
    for (int s = 0; s < P::V; s++) {
        const int e = fft16::slot_index<N, R0>(t, s);
        float re = 0.0f, im = 0.0f;
        if (e < W) {
            if constexpr (kUniform) {
                if (v0) re = glb_ld<float>(ua, (unsigned) e);
                if (v1) im = glb_ld<float>(ub, (unsigned) e);
            } else {
                if (v0) re = a[e];
                if (v1) im = b[e];
            }
        }
        v[s] = make_float2(re, im);
    }
    fft16::transform_forward_order<N, -1>(v, x, t, tw);
#pragma unroll
    for (int s = 0; s < P::V; s++) {
        const unsigned e = (unsigned) fft16::slot_index<N, RL>(t, s);
        // (c[2e], c[2e + 1]) in one 8-byte load, no branch on `pair`
        const uint64_t cc = glb_ld<uint64_t>(coef, e);
        const float cr = __uint_as_float((unsigned) cc);
        const float ci = pair ? cr : __uint_as_float((unsigned) (cc >> 32));
        v[s] = make_float2(v[s].x * cr, v[s].y * ci);
    }
    fft16::transform_reverse_order<N, +1>(v, x, t, tw);
#pragma unroll
    for (int s = 0; s < P::V; s++) {
        const int e = fft16::slot_index<N, R0>(t, s);
        if (e < crop) {
            if constexpr (kUniform) {
                if (v0) glb_store(uo0, (unsigned) e, v[s].x * scale);
                if (v1) glb_store(uo1, (unsigned) e, v[s].y * scale);
            } else {
                if (v0) ufomi_store(o0, e, v[s].x * scale);
                if (v1) ufomi_store(o1, e, v[s].y * scale);
            }
        }
    }
}

template <int N, typename OutT>
__global__ __launch_bounds__(kThreads) void ramp16_kernel(
    const float *__restrict__ in, int W, long rows, const float *__restrict__ coef, const int *__restrict__ pair_flag,
    int crop, OutT *__restrict__ out, float scale, const float2 *__restrict__ tw)
{
    ramp16_block<N, OutT>(in, W, rows, coef, *pair_flag != 0, crop, out, scale, tw, blockIdx.x);
}

// Ramp filter over segments of uneven width sharing N and the coefficients:
// block b belongs to the last segment whose block0 <= b; each segment's rows
// are cropped to its own width.
template <int N, typename OutT>
__global__ __launch_bounds__(kThreads) void ramp16_segments_kernel(
    const ufomi_ramp_seg *__restrict__ segs, int nseg, const float *__restrict__ coef, int pair, float scale,
    const float2 *__restrict__ tw)
{
    const long b = blockIdx.x;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].block0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    const ufomi_ramp_seg sg = segs[lo];
    ramp16_block<N, OutT>(sg.in, sg.W, sg.rows, coef, pair != 0, sg.W, (OutT *) sg.out, scale, tw, b - sg.block0);
}

template <int N>
__global__ __launch_bounds__(kThreads) void fft16_rows_kernel(const float *__restrict__ in, int W, long rows,
                                                              float2 *__restrict__ out, const float2 *__restrict__ tw)
{
    using P = fft16::Plan<N>;
    constexpr int R0 = P::radix(0), RL = P::radix(P::NPASS - 1);
    extern __shared__ float2 lds[];
    const int f = threadIdx.x / P::TF, t = threadIdx.x % P::TF;
    float2 *x = lds + (size_t) f * fft16::Lds<N>::STRIDE;
    const long r = (long) blockIdx.x * P::G + f;
    const bool valid = r < rows;
    const float *a = in + r * W;
    float2 v[P::V];
#pragma unroll
    for (int s = 0; s < P::V; s++) {
        const int e = fft16::slot_index<N, R0>(t, s);
        v[s] = make_float2((valid && e < W) ? a[e] : 0.0f, 0.0f);
    }
    fft16::transform_forward_order<N, -1>(v, x, t, tw);
    if (valid) {
#pragma unroll
        for (int s = 0; s < P::V; s++)
            out[r * N + fft16::slot_index<N, RL>(t, s)] = v[s];
    }
}

template <int N>
__global__ __launch_bounds__(kThreads) void ifft16_rows_kernel(const float2 *__restrict__ in, long rows, int crop,
                                                               float *__restrict__ out, float scale,
                                                               const float2 *__restrict__ tw)
{
    using P = fft16::Plan<N>;
    constexpr int R0 = P::radix(0), RL = P::radix(P::NPASS - 1);
    extern __shared__ float2 lds[];
    const int f = threadIdx.x / P::TF, t = threadIdx.x % P::TF;
    float2 *x = lds + (size_t) f * fft16::Lds<N>::STRIDE;
    const long r = (long) blockIdx.x * P::G + f;
    const bool valid = r < rows;
    float2 v[P::V];
#pragma unroll
    for (int s = 0; s < P::V; s++)
        v[s] = valid ? in[r * N + fft16::slot_index<N, RL>(t, s)] : make_float2(0.0f, 0.0f);
    fft16::transform_reverse_order<N, +1>(v, x, t, tw);
    if (valid) {
#pragma unroll
        for (int s = 0; s < P::V; s++) {
            const int e = fft16::slot_index<N, R0>(t, s);
            if (e < crop)
                out[r * crop + e] = v[s].x * scale;
        }
    }
}

template <typename K>
int launch_by_n(int N, K k)
{
    switch (N) {
    case 16: return k(std::integral_constant<int, 16>());
    case 32: return k(std::integral_constant<int, 32>());
    case 64: return k(std::integral_constant<int, 64>());
    case 128: return k(std::integral_constant<int, 128>());
    case 256: return k(std::integral_constant<int, 256>());
    case 512: return k(std::integral_constant<int, 512>());
    case 1024: return k(std::integral_constant<int, 1024>());
    case 2048: return k(std::integral_constant<int, 2048>());
    case 4096: return k(std::integral_constant<int, 4096>());
    case 8192: return k(std::integral_constant<int, 8192>());
    default: return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "no radix-16 FFT kernel for N = %d", N);
    }
}

template <int N> size_t lds16_bytes()
{
    using P = fft16::Plan<N>;
    return P::NPASS > 1 ? (size_t) P::G * fft16::Lds<N>::STRIDE * sizeof(float2) : 0;
}

}  // namespace

extern "C" int ufomi_fft_rows(ufomi_ctx *ctx, const float *d_in, int W, int P, int N, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if ((rc = check_n(N))) return rc;
    if (!d_in || !d_out || W <= 0 || P <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid fft arguments");
    if (N < W)
        return ufomi_fail(UFOMI_ERR_SIZE, "Specified size must be greater than or equal to input size");
    if (N >= 16)
        return launch_by_n(N, [&](auto n) {
            constexpr int NN = decltype(n)::value;
            using PL = fft16::Plan<NN>;
            const unsigned blocks = (unsigned) (((long) P + PL::G - 1) / PL::G);
            hipLaunchKernelGGL(fft16_rows_kernel<NN>, dim3(blocks), dim3(kThreads), lds16_bytes<NN>(), ctx->stream,
                               d_in, W, (long) P, (float2 *) d_out, ctx->twiddles);
            UFOMI_CHECK_LAUNCH();
            return (int) UFOMI_OK;
        });
    int B, G;
    size_t lds;
    fft_geometry(N, &B, &G, &lds);
    const unsigned blocks = (unsigned) ((P + G - 1) / G);
    return launch_by_b(B, [&](auto b) {
        constexpr int BB = decltype(b)::value;
        hipLaunchKernelGGL(fft_rows_kernel<BB>, dim3(blocks), dim3(kThreads), lds, ctx->stream,
                           d_in, W, (long) P, N, ufomi_log2(N), (float2 *) d_out, ctx->twiddles);
        UFOMI_CHECK_LAUNCH();
        return (int) UFOMI_OK;
    });
}

extern "C" int ufomi_ifft_rows(ufomi_ctx *ctx, const float *d_in, int N, int P, int crop, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if ((rc = check_n(N))) return rc;
    if (!d_in || !d_out || P <= 0 || crop < 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid ifft arguments");
    if (crop > N)
        return ufomi_fail(UFOMI_ERR_SIZE, "Cropped size must be less than or equal to input size");
    if (crop == 0) crop = N;
    const float scale = 1.0f / (float) N;  // fft_pack scale of ufo-ifft-task.c:296-303
    if (N >= 16)
        return launch_by_n(N, [&](auto n) {
            constexpr int NN = decltype(n)::value;
            using PL = fft16::Plan<NN>;
            const unsigned blocks = (unsigned) (((long) P + PL::G - 1) / PL::G);
            hipLaunchKernelGGL(ifft16_rows_kernel<NN>, dim3(blocks), dim3(kThreads), lds16_bytes<NN>(), ctx->stream,
                               (const float2 *) d_in, (long) P, crop, d_out, scale, ctx->twiddles);
            UFOMI_CHECK_LAUNCH();
            return (int) UFOMI_OK;
        });
    int B, G;
    size_t lds;
    fft_geometry(N, &B, &G, &lds);
    const unsigned blocks = (unsigned) ((P + G - 1) / G);
    return launch_by_b(B, [&](auto b) {
        constexpr int BB = decltype(b)::value;
        hipLaunchKernelGGL(ifft_rows_kernel<BB>, dim3(blocks), dim3(kThreads), lds, ctx->stream,
                           (const float2 *) d_in, (long) P, N, ufomi_log2(N), crop, d_out, scale, ctx->twiddles);
        UFOMI_CHECK_LAUNCH();
        return (int) UFOMI_OK;
    });
}

extern "C" int ufomi_filter_apply(ufomi_ctx *ctx, const float *d_in, float *d_out, const float *d_coef,
                                  int width, int P)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_in || !d_out || !d_coef || P <= 0 || width <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid filter arguments");
    if (!ufomi_is_pow2(width))
        return ufomi_fail(UFOMI_ERR_NOT_POW2, "Invalid filter input width %d, must be a power of two", width);
    const long total = (long) width * P;
    long blocks = (total + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(filter_apply_kernel, dim3((unsigned) blocks), dim3(kThreads), 0, ctx->stream,
                       d_in, d_out, d_coef, width, total);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

extern "C" int ufomi_ramp_filter(ufomi_ctx *ctx, const float *d_in, int W, int P, int N,
                                 const float *d_coef, int crop, void *d_out, int out_dtype)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if ((rc = check_n(N))) return rc;
    if (!d_in || !d_out || !d_coef || W <= 0 || P <= 0 || crop < 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid ramp filter arguments");
    if (N < W)
        return ufomi_fail(UFOMI_ERR_SIZE, "Specified size must be greater than or equal to input size");
    if (crop > N)
        return ufomi_fail(UFOMI_ERR_SIZE, "Cropped size must be less than or equal to input size");
    if (crop == 0) crop = N;
    if (out_dtype != UFOMI_F32 && out_dtype != UFOMI_F16 && out_dtype != UFOMI_BF16)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown output dtype %d", out_dtype);

    // Classify the coefficients on the device, in stream order.
    int *d_flag = ctx->flags;
    hipLaunchKernelGGL(classify_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, d_coef, N, d_flag);
    UFOMI_CHECK_LAUNCH();

    const float scale = 1.0f / (float) N;
    if (N >= 16) {
        auto go16 = [&](auto n, auto *out) {
            constexpr int NN = decltype(n)::value;
            using OutT = std::remove_pointer_t<decltype(out)>;
            using PL = fft16::Plan<NN>;
            const unsigned blocks = (unsigned) (((long) P + PL::G - 1) / PL::G);
            hipLaunchKernelGGL((ramp16_kernel<NN, OutT>), dim3(blocks), dim3(kThreads), lds16_bytes<NN>(),
                               ctx->stream, d_in, W, (long) P, d_coef, d_flag, crop, out, scale, ctx->twiddles);
            UFOMI_CHECK_LAUNCH();
            return (int) UFOMI_OK;
        };
        return launch_by_n(N, [&](auto n) {
            switch (out_dtype) {
            case UFOMI_F16: return go16(n, (__half *) d_out);
            case UFOMI_BF16: return go16(n, (__hip_bfloat16 *) d_out);
            default: return go16(n, (float *) d_out);
            }
        });
    }
    int B, G;
    size_t lds;
    fft_geometry(N, &B, &G, &lds);
    const unsigned blocks = (unsigned) (((long) P + G - 1) / G);
    const int log2N = ufomi_log2(N);
    auto go = [&](auto b, auto *out) {
        constexpr int BB = decltype(b)::value;
        using OutT = std::remove_pointer_t<decltype(out)>;
        hipLaunchKernelGGL((ramp_kernel<BB, OutT>), dim3(blocks), dim3(kThreads), lds, ctx->stream,
                           d_in, W, (long) P, N, log2N, d_coef, d_flag, crop, out, scale, ctx->twiddles);
        UFOMI_CHECK_LAUNCH();
        return (int) UFOMI_OK;
    };
    return launch_by_b(B, [&](auto b) {
        switch (out_dtype) {
        case UFOMI_F16: return go(b, (__half *) d_out);
        case UFOMI_BF16: return go(b, (__hip_bfloat16 *) d_out);
        default: return go(b, (float *) d_out);
        }
    });
}

long ufomi_ramp_rows_per_block(int N, int pair)
{
    const int g = launch_by_n(N, [](auto n) { return fft16::Plan<decltype(n)::value>::G; });
    return g < 0 ? g : (long) g * (pair ? 2 : 1);
}

int ufomi_ramp_segments_launch(hipStream_t stream, const float2 *tw, int N, int out_dtype,
                               const ufomi_ramp_seg *d_segs, int nseg, long nblocks, const float *d_coef, int pair)
{
    if (nblocks <= 0 || nblocks > 0x7fffffffL || nseg <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid segmented ramp launch (%ld blocks, %d segments)", nblocks, nseg);
    const float scale = 1.0f / (float) N;
    return launch_by_n(N, [&](auto n) {
        constexpr int NN = decltype(n)::value;
        auto go = [&](auto *tag) {
            using OutT = std::remove_pointer_t<decltype(tag)>;
            hipLaunchKernelGGL((ramp16_segments_kernel<NN, OutT>), dim3((unsigned) nblocks), dim3(kThreads),
                               lds16_bytes<NN>(), stream, d_segs, nseg, d_coef, pair, scale, tw);
            UFOMI_CHECK_LAUNCH();
            return (int) UFOMI_OK;
        };
        switch (out_dtype) {
        case UFOMI_F16: return go((__half *) nullptr);
        case UFOMI_BF16: return go((__hip_bfloat16 *) nullptr);
        default: return go((float *) nullptr);
        }
    });
}
