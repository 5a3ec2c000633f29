// Roofline denominators measured on the running GPU: LDS read bandwidth
// (ds_read_b128, conflict-free, every CU busy), HBM copy bandwidth
// (streaming float4 copy) and dense BF16 MFMA throughput
// (v_mfma_f32_32x32x16_bf16 chains held in registers).  Used by bench.py to express kernel throughput as a
// fraction of what this device sustains.
#include "ufomi_internal.h"

namespace {

__global__ __launch_bounds__(256) void lds_read_kernel(int iters, float *__restrict__ sink)
{
    __shared__ float4 buf[1024];  // 16 KiB
    const int t = threadIdx.x;
    for (int i = t; i < 1024; i += 256)
        buf[i] = make_float4((float) i, 1.0f, 2.0f, 3.0f);
    __syncthreads();
    float4 a0 = make_float4(0, 0, 0, 0), a1 = a0, a2 = a0, a3 = a0;
    int idx = t;
    for (int it = 0; it < iters; it++) {
        // 4 independent 16-byte reads per lane; a wave reads 1 KiB contiguous each time.
        const float4 q0 = buf[idx & 1023];
        const float4 q1 = buf[(idx + 256) & 1023];
        const float4 q2 = buf[(idx + 512) & 1023];
        const float4 q3 = buf[(idx + 768) & 1023];
        a0.x += q0.x; a0.y += q0.y; a0.z += q0.z; a0.w += q0.w;
        a1.x += q1.x; a1.y += q1.y; a1.z += q1.z; a1.w += q1.w;
        a2.x += q2.x; a2.y += q2.y; a2.z += q2.z; a2.w += q2.w;
        a3.x += q3.x; a3.y += q3.y; a3.z += q3.z; a3.w += q3.w;
        idx += 64;
    }
    const float s = a0.x + a0.y + a0.z + a0.w + a1.x + a1.y + a1.z + a1.w + a2.x + a2.y + a2.z + a2.w +
                    a3.x + a3.y + a3.z + a3.w;
    if (s == -1.0f)  // never true; keeps the reads alive
        sink[blockIdx.x * 256 + t] = s;
}

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ in, float4 *__restrict__ out, long n4)
{
    const long stride = (long) gridDim.x * 256;
    for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < n4; i += stride)
        out[i] = in[i];
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// 8 independent v_mfma_f32_32x32x16_bf16 accumulator chains per wave, operands
// in registers.  (Fewer independent chains leave the pipe waiting on each
// chain's previous result: 8 chains of the 16x16x32 form reach about half the
// rate that 16 chains of it or the 32x32x16 form reach.)
__global__ __launch_bounds__(256) void mfma_kernel(int iters, float *__restrict__ sink)
{
    bf16x8_t a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        a[i] = (__bf16) (1.0f + 0.001f * (float) (threadIdx.x + i));
        b[i] = (__bf16) (0.5f - 0.001f * (float) i);
    }
    f32x16_t acc[8];
#pragma unroll
    for (int c = 0; c < 8; c++)
#pragma unroll
        for (int e = 0; e < 16; e++) acc[c][e] = 0.0f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) s += acc[c][0] + acc[c][15];
    if (s == -1.0f)  // never true; keeps the chains alive
        sink[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

extern "C" int ufomi_bench_mfma(ufomi_ctx *ctx, int blocks, int iters, float *d_sink, double *flops)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (blocks <= 0 || iters <= 0 || !d_sink)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid microbenchmark arguments");
    hipLaunchKernelGGL(mfma_kernel, dim3(blocks), dim3(256), 0, ctx->stream, iters, d_sink);
    UFOMI_CHECK_LAUNCH();
    if (flops)  // 4 waves x 8 chains x 32 x 32 x 16 x 2 per iteration
        *flops = (double) blocks * 4.0 * iters * 8.0 * 32768.0;
    return UFOMI_OK;
}

extern "C" int ufomi_bench_lds_read(ufomi_ctx *ctx, int blocks, int iters, float *d_sink, double *bytes)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (blocks <= 0 || iters <= 0 || !d_sink)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid microbenchmark arguments");
    hipLaunchKernelGGL(lds_read_kernel, dim3(blocks), dim3(256), 0, ctx->stream, iters, d_sink);
    UFOMI_CHECK_LAUNCH();
    if (bytes)
        *bytes = (double) blocks * 256.0 * iters * 64.0;
    return UFOMI_OK;
}

extern "C" int ufomi_bench_copy(ufomi_ctx *ctx, const float *d_in, float *d_out, size_t n)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_in || !d_out || n % 4)
        return ufomi_fail(UFOMI_ERR_ARG, "copy needs a multiple of 4 floats");
    const long n4 = (long) (n / 4);
    long blocks = (n4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned) blocks), dim3(256), 0, ctx->stream, (const float4 *) d_in,
                       (float4 *) d_out, n4);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}
