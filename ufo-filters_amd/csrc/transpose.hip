// `transpose-projections`: projections [n][H][W] -> sinograms [H][number][W].
//
// src/ufo-transpose-projections-task.c:64-99 copies row h of projection p to
// row p of sinogram h; rows of projections that never arrive stay zero (the
// store is g_malloc0'd at :143).  Rows stay contiguous under this
// permutation, so the kernel moves whole rows: a block owns a run of output
// rows (about 32 KiB), each thread keeps U independent 16-byte loads in
// flight before it stores them, and both sides are coalesced.  It is bound
// by HBM at 8 bytes moved per float (4 for a zero row).
#include "ufomi_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kU = 4;              // vectors in flight per thread
constexpr int kBlockVectors = 2048;  // vectors per block (32 KiB of float4)

// Output rows r in [0, n_rows) are (h, p) = (r / n, r % n); row p of
// sinogram h is row first + p of the [H][number] sinogram block.  `in` null
// writes zeros.
template <typename V>
__global__ __launch_bounds__(kThreads) void transpose_rows_kernel(const V *__restrict__ in, V *__restrict__ out,
                                                                   unsigned Wv, unsigned H, unsigned n,
                                                                   unsigned first, unsigned number, unsigned n_rows,
                                                                   unsigned rows_per_block)
{
    const unsigned r0 = blockIdx.x * rows_per_block;
    const unsigned rows = min(rows_per_block, n_rows - r0);
    const unsigned total = rows * Wv;
    for (unsigned j0 = threadIdx.x; j0 < total; j0 += kThreads * kU) {
        V v[kU];
        size_t dst[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const unsigned j = j0 + u * kThreads;
            dst[u] = ~(size_t) 0;
            if (j < total) {
                const unsigned k = j / Wv, x = j - k * Wv;
                const unsigned r = r0 + k;
                const unsigned h = r / n, p = r - h * n;
                v[u] = in ? in[((size_t) p * H + h) * Wv + x] : V{};
                dst[u] = ((size_t) h * number + first + p) * Wv + x;
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
            if (dst[u] != ~(size_t) 0)
                out[dst[u]] = v[u];
    }
}

template <typename V>
int launch(hipStream_t s, const float *in, float *out, int W, int H, int n, int first, int number)
{
    const unsigned Wv = (unsigned) (W * sizeof(float) / sizeof(V));
    const unsigned n_rows = (unsigned) H * (unsigned) n;
    const unsigned rpb = Wv >= (unsigned) kBlockVectors ? 1u : (unsigned) kBlockVectors / Wv;
    const unsigned blocks = (n_rows + rpb - 1) / rpb;
    hipLaunchKernelGGL(transpose_rows_kernel<V>, dim3(blocks), dim3(kThreads), 0, s, (const V *) in, (V *) out, Wv,
                       (unsigned) H, (unsigned) n, (unsigned) first, (unsigned) number, n_rows, rpb);
    return UFOMI_OK;
}

bool aligned(const void *p, size_t a) { return ((uintptr_t) p % a) == 0; }

}  // namespace

extern "C" int ufomi_transpose_projections(ufomi_ctx *ctx, const float *d_proj, int W, int H, int n, int first,
                                           int number, float *d_sino)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_sino || W <= 0 || H <= 0 || n < 0 || first < 0 || number <= 0 || (long) first + n > number)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid transpose arguments (width %d, height %d, %d projections at %d "
                                         "of %d)", W, H, n, first, number);
    if (!n)
        return UFOMI_OK;
    // 32-bit indices inside the kernel: rows and vectors per row
    if ((long) H * n >= (1L << 31) || (long) H * number >= (1L << 31) || (long) W * kBlockVectors >= (1L << 31))
        return ufomi_fail(UFOMI_ERR_ARG, "transpose of %d x %d x %d rows is too large for one launch", n, H, W);
    const bool a16 = W % 4 == 0 && aligned(d_sino, 16) && (!d_proj || aligned(d_proj, 16));
    const bool a8 = W % 2 == 0 && aligned(d_sino, 8) && (!d_proj || aligned(d_proj, 8));
    if (a16)
        rc = launch<float4>(ctx->stream, d_proj, d_sino, W, H, n, first, number);
    else if (a8)
        rc = launch<float2>(ctx->stream, d_proj, d_sino, W, H, n, first, number);
    else
        rc = launch<float>(ctx->stream, d_proj, d_sino, W, H, n, first, number);
    if (rc) return rc;
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}
