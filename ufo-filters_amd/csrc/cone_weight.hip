// FDK pre-weighting (`cone-beam-projection-weight` task).
//
// out = in * SOD * cos(axis_x) / sqrt(L^2 + x^2 + y^2) with the pixel centre
// (x, y) scaled to the isocentre (conebeam.cl:20-36) and the per-projection
// argument set-up of ufo-cone-beam-projection-weight-task.c:116-158.  One
// launch weights n_proj projections; the scalar arguments of every projection
// are precomputed on the host (FP64 inputs rounded to float like the task).
#include "ufomi_internal.h"

#include <cmath>
#include <vector>

namespace {

struct WeightArgs {
    float cx, cz, sd, overall, mag_recip, cos_angle;
};

__global__ __launch_bounds__(256) void cone_weight_kernel(const float *__restrict__ in, float *__restrict__ out,
                                                          int W, int H, const WeightArgs *__restrict__ args)
{
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int p = blockIdx.z;
    if (i >= W || j >= H)
        return;
    const WeightArgs a = args[p];
    const float x = (((float) i) - a.cx + 0.5f) * a.mag_recip;
    const float y = (((float) j) - a.cz + 0.5f) * a.mag_recip;
    const long o = ((long) p * H + j) * W + i;
    const float r = sqrtf(a.overall * a.overall + x * x + y * y);
    out[o] = in[o] * a.sd * a.cos_angle / r;
}

double pick(const double *a, int n, int i) { return n == 1 ? a[0] : a[i]; }

}  // namespace

extern "C" int ufomi_cone_weight(ufomi_ctx *ctx, const float *d_in, float *d_out, int W, int H, int n_proj,
                                 const double *center_x, const double *center_z, const double *source_distance,
                                 const double *detector_distance, const double *axis_angle_x, int n_param)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_in || !d_out || W <= 0 || H <= 0 || n_proj <= 0 || n_proj > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid cone weight arguments");
    if (!center_x || !center_z || !source_distance || !detector_distance || n_param <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "center-position-x/z, source-distance and detector-distance must be set");
    if (n_param != 1 && n_param < n_proj)
        return ufomi_fail(UFOMI_ERR_ARG, "per-projection arrays need %d values, got %d", n_proj, n_param);

    std::vector<WeightArgs> h(n_proj);
    for (int p = 0; p < n_proj; p++) {
        WeightArgs &a = h[p];
        a.cos_angle = (float) cos(axis_angle_x ? pick(axis_angle_x, n_param, p) : 0.0);
        a.cx = (float) pick(center_x, n_param, p);
        a.cz = (float) pick(center_z, n_param, p);
        a.sd = (float) pick(source_distance, n_param, p);
        const float dd = (float) pick(detector_distance, n_param, p);
        a.overall = a.sd + dd;
        a.mag_recip = a.sd / a.overall;
        if (a.cos_angle > 0.9999999f)
            a.overall = a.sd;
    }
    WeightArgs *d_args = nullptr;
    if ((rc = ufomi_scratch(ctx, n_proj * sizeof(WeightArgs), (void **) &d_args))) return rc;
    if ((rc = ufomi_upload(ctx, d_args, h.data(), n_proj * sizeof(WeightArgs)))) return rc;
    dim3 grid((W + 63) / 64, (H + 3) / 4, n_proj);
    hipLaunchKernelGGL(cone_weight_kernel, grid, dim3(256), 0, ctx->stream, d_in, d_out, W, H, d_args);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}
