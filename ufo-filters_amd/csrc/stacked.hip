// `stacked-backproject`: parallel-beam backprojection of a stack of sinograms
// in three storage precisions (src/ufo-stacked-backproject-task.c:300-430,
// src/kernels/stacked-backproject.cl).  The reference samples
// h = bx cos - by sin + axis (the sign of the sin term is the opposite of
// `backproject`'s) with
//   single: float32 texels, linear filter, zero border;
//   half:   texels rounded to half, linear filter in float, zero border;
//   int8:   texels (v - min) * 255 / (max - min) truncated to 8 bits over the
//           whole stack's min and max, nearest texel clamped to the edge,
//           integer sums, then (sum * (max - min) / 255 + min) * pi / P.
// single and half run the tiled backprojector of bp_parallel.hip with a
// negated sin table (half through f16 storage with float interpolation);
// int8 has its own kernel: 4 slices per 32-bit texel, one uchar4 gather per
// (pixel, angle), integer sums exact in uint32.
#include "ufomi_internal.h"

#include <cmath>

// Every float operation rounds on its own, as in the reference kernels' source
// (hipcc would otherwise fuse a * b + c into one FMA).
#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr float kPiF = 3.14159274101257324219f;  // M_PI_F
constexpr int kMinMaxBlocks = 1024;

__global__ __launch_bounds__(kThreads) void minmax_partial_kernel(const float *__restrict__ x, long n,
                                                                   float2 *__restrict__ part)
{
    float lo = INFINITY, hi = -INFINITY;
    for (long i = (long) blockIdx.x * kThreads + threadIdx.x; i < n; i += (long) gridDim.x * kThreads) {
        const float v = x[i];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    __shared__ float slo[kThreads], shi[kThreads];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = fminf(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = fmaxf(shi[threadIdx.x], shi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = make_float2(slo[0], shi[0]);
}

__global__ __launch_bounds__(kThreads) void minmax_final_kernel(const float2 *__restrict__ part, int n,
                                                                 float2 *__restrict__ mm)
{
    float lo = INFINITY, hi = -INFINITY;
    for (int i = threadIdx.x; i < n; i += kThreads) {
        lo = fminf(lo, part[i].x);
        hi = fmaxf(hi, part[i].y);
    }
    __shared__ float slo[kThreads], shi[kThreads];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = fminf(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = fmaxf(shi[threadIdx.x], shi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        mm[0] = make_float2(slo[0], shi[0]);
}

// interleave_uint (stacked-backproject.cl): q = (v - min) * (255 / (max - min))
// converted to an unsigned integer (truncation) and saturated to 8 bits.
__device__ __forceinline__ unsigned quantize(float v, float lo, float scale)
{
    const float q = ((v - lo) * scale);
    if (!(q > 0.0f))
        return 0u;
    return q >= 255.0f ? 255u : (unsigned) q;
}

// Slices 4g .. 4g+3 of texel (p, x) in one uchar4 (missing slices of the last group are 0).
__global__ __launch_bounds__(kThreads) void quantize_kernel(const float *__restrict__ sino, long plane, int n_slices,
                                                            const float2 *__restrict__ mm, uchar4 *__restrict__ q,
                                                            long total)
{
    const long i = (long) blockIdx.x * kThreads + threadIdx.x;
    if (i >= total)
        return;
    const long g = i / plane, e = i - g * plane;
    const float lo = mm[0].x, hi = mm[0].y;
    const float scale = (255.0f / (hi - lo));
    unsigned c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long s = 4 * g + k;
        c[k] = s < n_slices ? quantize(sino[s * plane + e], lo, scale) : 0u;
    }
    q[i] = make_uchar4((unsigned char) c[0], (unsigned char) c[1], (unsigned char) c[2], (unsigned char) c[3]);
}

// texture_uint + uninterleave_uint: one thread per pixel of a 16 x 16 tile,
// four slices per gather.  The sample position keeps the reference's float
// operations and their order without contraction, so the nearest texel is the
// same: bx = ((x - axis) + roi_x) + 0.5, and a work-item of the reference
// computes rows y0 + 4q of its 16-row block as h(y0) - 4q sin.
__global__ __launch_bounds__(kThreads) void bp_u8_kernel(const uchar4 *__restrict__ q, int W, int P,
                                                         const float2 *__restrict__ trig, float axis, float roi_x,
                                                         float roi_y, int roi_w, int roi_h, int n_slices,
                                                         const float2 *__restrict__ mm, float *__restrict__ out)
{
    __shared__ float2 sc[kThreads];
    const int r = threadIdx.x >> 4;  // row in the 16-row block = base row + 4 qr
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + r;
    const int qr = r >> 2, y0 = y - 4 * qr;
    const int g = blockIdx.z;
    const float bx = (((float) x - axis) + roi_x) + 0.5f;
    const float by = (((float) y0 - axis) + roi_y) + 0.5f;
    const float shift = (float) (4 * qr);
    const uchar4 *plane = q + (long) g * P * W;
    unsigned s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int p0 = 0; p0 < P; p0 += kThreads) {
        __syncthreads();
        if (p0 + (int) threadIdx.x < P)
            sc[threadIdx.x] = trig[p0 + threadIdx.x];  // (-sin, cos)
        __syncthreads();
        const int n = min(kThreads, P - p0);
        for (int k = 0; k < n; k++) {
            const float2 t = sc[k];
            // h = (bx * cos - by * sin + axis) - 4 qr sin with t.x = -sin
            const float h = (((bx * t.y) + (by * t.x)) + axis) + (shift * t.x);
            int i = (int) floorf(h);
            i = i < 0 ? 0 : (i >= W ? W - 1 : i);
            const uchar4 v = plane[(long) (p0 + k) * W + i];
            s0 += v.x;
            s1 += v.y;
            s2 += v.z;
            s3 += v.w;
        }
    }
    if (x >= roi_w || y >= roi_h)
        return;
    const float lo = mm[0].x, hi = mm[0].y;
    const float scale = ((hi - lo) / 255.0f);
    const unsigned s[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sl = 4 * g + k;
        if (sl < n_slices) {
            const float r = (((float) s[k] * scale) + lo) * kPiF;
            out[((long) sl * roi_h + y) * roi_w + x] = (r / (float) P);
        }
    }
}

__global__ __launch_bounds__(kThreads) void to_half_kernel(const float *__restrict__ in, __half *__restrict__ out,
                                                           long n)
{
    const long i = (long) blockIdx.x * kThreads + threadIdx.x;
    if (i < n)
        out[i] = __float2half_rn(in[i]);
}

}  // namespace

extern "C" int ufomi_bp_stacked(ufomi_ctx *ctx, const float *d_sino, int W, int P_in, int n_slices,
                                const ufomi_bp_params *prm, int precision, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_sino || !d_out || !prm || W <= 0 || P_in <= 0 || n_slices <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid stacked backprojection arguments");
    if (precision != UFOMI_STACKED_INT8 && precision != UFOMI_STACKED_HALF && precision != UFOMI_STACKED_SINGLE)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown precision mode %d", precision);
    if (prm->mode != UFOMI_BP_TEXTURE)
        return ufomi_fail(UFOMI_ERR_ARG, "stacked backprojection samples like texture mode");
    const long plane = (long) P_in * W, total = plane * n_slices;
    if (precision == UFOMI_STACKED_SINGLE)
        return ufomi_bp_parallel_ex(ctx, d_sino, UFOMI_F32, W, P_in, n_slices, 0, prm, d_out, 0, -1, false);
    if (precision == UFOMI_STACKED_HALF) {
        __half *h = nullptr;
        if ((rc = ufomi_scratch(ctx, (size_t) total * sizeof(__half), (void **) &h))) return rc;
        hipLaunchKernelGGL(to_half_kernel, dim3((unsigned) ((total + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           ctx->stream, d_sino, h, total);
        UFOMI_CHECK_LAUNCH();
        // float interpolation of half texels, as the reference's half image
        return ufomi_bp_parallel_ex(ctx, h, UFOMI_F16, W, P_in, n_slices, 0, prm, d_out, 0, -1, false);
    }

    // int8: the same LUT, axis and ROI rules as ufomi_bp_parallel
    const unsigned n_proj = prm->num_projections ? prm->num_projections : (unsigned) P_in;
    if ((unsigned) P_in > n_proj)
        return ufomi_fail(UFOMI_ERR_PROJECTIONS,
                          "Total number of projections (%u) must be greater than or equal to sinogram height (%u)",
                          n_proj, (unsigned) P_in);
    if ((unsigned long) prm->offset + (unsigned) P_in > n_proj)
        return ufomi_fail(UFOMI_ERR_ARG, "offset %u + sinogram height %d exceeds the %u LUT entries", prm->offset,
                          P_in, n_proj);
    const int roi_w = prm->roi_width ? (int) prm->roi_width : W;
    const int roi_h = prm->roi_height ? (int) prm->roi_height : W;
    const double step = prm->angle_step <= 0.0 ? M_PI / (double) n_proj : prm->angle_step;
    const float axis = prm->axis_pos <= 0.0 ? (float) W / 2.0f : (float) prm->axis_pos;
    if (!std::isfinite(axis) || !std::isfinite(step) || !std::isfinite(prm->angle_offset))
        return ufomi_fail(UFOMI_ERR_ARG, "non-finite geometry");
    const int groups = (n_slices + 3) / 4;
    if ((long) groups > 65535 || (long) roi_w + 15 > (1L << 30) || (long) roi_h + 15 > (1L << 30))
        return ufomi_fail(UFOMI_ERR_ARG, "stack or region of interest too large for one launch");
    float2 *trig = nullptr;
    if ((rc = ufomi_bp_trig(ctx, prm->angle_offset, step, prm->offset, P_in, -1, &trig))) return rc;
    // scratch: min/max partials, the final pair, the quantized stack
    const size_t qbytes = (size_t) groups * plane * sizeof(uchar4);
    char *base = nullptr;
    if ((rc = ufomi_scratch(ctx, qbytes + (kMinMaxBlocks + 1) * sizeof(float2) + 512, (void **) &base))) return rc;
    uchar4 *q = (uchar4 *) base;
    float2 *part = (float2 *) (base + ((qbytes + 255) / 256) * 256);
    float2 *mm = part + kMinMaxBlocks;
    const int nb = (int) std::min<long>(kMinMaxBlocks, (total + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(minmax_partial_kernel, dim3(nb), dim3(kThreads), 0, ctx->stream, d_sino, total, part);
    UFOMI_CHECK_LAUNCH();
    hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(kThreads), 0, ctx->stream, (const float2 *) part, nb, mm);
    UFOMI_CHECK_LAUNCH();
    const long qn = (long) groups * plane;
    hipLaunchKernelGGL(quantize_kernel, dim3((unsigned) ((qn + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       ctx->stream, d_sino, plane, n_slices, (const float2 *) mm, q, qn);
    UFOMI_CHECK_LAUNCH();
    dim3 grid((unsigned) ((roi_w + 15) / 16), (unsigned) ((roi_h + 15) / 16), (unsigned) groups);
    hipLaunchKernelGGL(bp_u8_kernel, grid, dim3(kThreads), 0, ctx->stream, (const uchar4 *) q, W, P_in,
                       (const float2 *) trig, axis, (float) prm->roi_x, (float) prm->roi_y, roi_w, roi_h, n_slices,
                       (const float2 *) mm, d_out);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}
