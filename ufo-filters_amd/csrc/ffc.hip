// Adjacent steps before the reconstruction path:
//  * `flat-field-correct` (src/kernels/ffc.cl:20-48,
//    src/ufo-flat-field-correct-task.c:107-160): per pixel
//      d = dark * dark_scale, f = flat * flat_scale,
//      (data - d) / (f - d), or log((f - d) / (data - d)) with absorption,
//    NaN and Inf replaced by 0 on request; with sinogram input the dark and
//    flat are one row indexed by x.  One launch corrects a batch of items; an
//    item past the last dark (flat) item uses the last one, as a finished
//    input stream keeps feeding its last buffer.
//  * `average` (src/ufo-average-task.c:102-151): the float32 sum of the items
//    in arrival order divided by their count.
// Both are elementwise and bound by HBM.
#include "ufomi_internal.h"

// Every float operation rounds on its own, as in the reference kernels' source
// (hipcc would otherwise fuse a * b + c into one FMA).
#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float ffc_one(float v, float dk, float fl, bool absorb, bool fix)
{
    float r = absorb ? logf(((fl - dk) / (v - dk)))
                     : ((v - dk) / (fl - dk));
    if (fix && (isnan(r) || isinf(r)))
        r = 0.0f;
    return r;
}

__global__ __launch_bounds__(kThreads) void ffc_kernel(const float *__restrict__ data, const float *__restrict__ dark,
                                                       const float *__restrict__ flat, float *__restrict__ out,
                                                       long plane, int W, int n, int n_dark, int n_flat,
                                                       int sino_in, int absorb, int fix, float dark_scale,
                                                       float flat_scale)
{
    const long i = (long) blockIdx.x * kThreads + threadIdx.x;
    if (i >= plane * n)
        return;
    const long item = i / plane, e = i - item * plane;
    const long ce = sino_in ? e % W : e;
    const long cplane = sino_in ? W : plane;
    const float dk = dark[min(item, (long) n_dark - 1) * cplane + ce] * dark_scale;
    const float fl = flat[min(item, (long) n_flat - 1) * cplane + ce] * flat_scale;
    out[i] = ffc_one(data[i], dk, fl, absorb != 0, fix != 0);
}

__global__ __launch_bounds__(kThreads) void average_kernel(const float *__restrict__ in, long elems, int n,
                                                           int copies, float *__restrict__ out)
{
    const long e = (long) blockIdx.x * kThreads + threadIdx.x;
    if (e >= elems)
        return;
    float s = 0.0f;
    for (int k = 0; k < n; k++)
        s = (s + in[(long) k * elems + e]);
    const float m = (s / (float) n);
    for (int c = 0; c < copies; c++)
        out[(long) c * elems + e] = m;
}

}  // namespace

extern "C" int ufomi_flat_field_correct(ufomi_ctx *ctx, const float *d_data, int W, int H, int n,
                                        const float *d_dark, int n_dark, const float *d_flat, int n_flat,
                                        int sinogram_input, int absorption_correct, int fix_nan_and_inf,
                                        float dark_scale, float flat_scale, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_data || !d_dark || !d_flat || !d_out || W <= 0 || H <= 0 || n <= 0 || n_dark <= 0 || n_flat <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid flat-field arguments");
    const long plane = (long) W * H, total = plane * n;
    const long blocks = (total + kThreads - 1) / kThreads;
    if (blocks >= (1L << 31))
        return ufomi_fail(UFOMI_ERR_ARG, "flat-field batch too large for one launch");
    hipLaunchKernelGGL(ffc_kernel, dim3((unsigned) blocks), dim3(kThreads), 0, ctx->stream, d_data, d_dark, d_flat,
                       d_out, plane, W, n, n_dark, n_flat, sinogram_input, absorption_correct, fix_nan_and_inf,
                       dark_scale, flat_scale);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

extern "C" int ufomi_average(ufomi_ctx *ctx, const float *d_items, size_t elems, int n, int copies, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_items || !d_out || !elems || n <= 0 || copies <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid average arguments");
    const long blocks = ((long) elems + kThreads - 1) / kThreads;
    if (blocks >= (1L << 31))
        return ufomi_fail(UFOMI_ERR_ARG, "average item too large for one launch");
    hipLaunchKernelGGL(average_kernel, dim3((unsigned) blocks), dim3(kThreads), 0, ctx->stream, d_items,
                       (long) elems, n, copies, d_out);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}
