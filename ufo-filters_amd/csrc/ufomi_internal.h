// Internal declarations shared by the HIP translation units of libufomi.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "ufomi.h"

struct ufomi_ctx {
    int device;
    hipStream_t stream;
    float2 *twiddles;        // exp(-2 pi i k / UFOMI_FFT_MAX), k < UFOMI_FFT_MAX
    int *flags;              // small device words written by classification kernels
    void *scratch;           // grow-only device scratch
    size_t scratch_bytes;
    void *hscratch;          // grow-only pinned host staging
    size_t hscratch_bytes;
    // backprojection LUT cache: float2 (sin, cos) per projection
    float2 *lut;
    int lut_n;
    double lut_offset, lut_step;
    unsigned lut_first;
    int lut_sign;            // +1: (sin, cos); -1: (-sin, cos)
    // parallel backprojection tile configuration (0 = automatic), ufomi_bp_configure
    int bp_slices, bp_chunk, bp_half;
    // partial sums of angle-split backprojections of few tiles (grow-only)
    float *bp_part;
    size_t bp_part_bytes;
};

#define UFOMI_FFT_MAX 16384  // twiddle table length
#define UFOMI_FFT_MAX_N 8192 // largest transform (N complex values in <= 64 KiB of LDS)

// Thread-local error message; returns `code` for `return ufomi_fail(...)`.
int ufomi_fail(int code, const char *fmt, ...);

#define UFOMI_HIP(call)                                                                \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ufomi_fail(UFOMI_ERR_HIP, "%s failed: %s (%s:%d)", #call,           \
                              hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

#define UFOMI_CHECK_LAUNCH() UFOMI_HIP(hipGetLastError())

// Device scratch owned by the context (grow-only).
int ufomi_scratch(ufomi_ctx *ctx, size_t bytes, void **ptr);
int ufomi_host_scratch(ufomi_ctx *ctx, size_t bytes, void **ptr);

// Copy host -> device scratch in stream order: waits for earlier work on the
// context's stream (which may still read the scratch), then copies.
int ufomi_upload(ufomi_ctx *ctx, void *dst, const void *src, size_t bytes);

// Make the context's device current for the calling thread.
int ufomi_enter(ufomi_ctx *ctx);

static inline bool ufomi_is_pow2(long x) { return x > 0 && (x & (x - 1)) == 0; }

static inline int ufomi_log2(long x)
{
    int l = 0;
    while ((1L << l) < x)
        l++;
    return l;
}

// ---- parallel backprojection (bp_parallel.hip), shared with stacked.hip

// The (sin, cos) LUT of the context (sin_sign -1 stores -sin).
int ufomi_bp_trig(ufomi_ctx *ctx, double angle_offset, double step, unsigned offset, int P, int sin_sign,
                  float2 **out);
// ufomi_bp_parallel with h = axis + bx cos + sin_sign * by sin and half staging allowed or not.
int ufomi_bp_parallel_ex(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in, int n_slices,
                         size_t sino_pitch, const ufomi_bp_params *prm, float *d_out, size_t out_pitch, int sin_sign,
                         bool half_staging);

// Takes the calling thread's current device's address of the preallocated
// scratch of ufomi_bp_parallel_noalloc (once per device; ufomi_create).
int ufomi_bp_noalloc_setup(int device);

// ---- batched FBP (batch.hip) over the kernels of fft_filter.hip and bp_parallel.hip

// One segment of a segmented ramp-filter launch: `rows` rows of width W at `in`
// (float32), written cropped to W at `out` (the launch's storage type); the
// segment owns blocks [block0, next segment's block0).
struct ufomi_ramp_seg {
    const float *in;
    void *out;
    int W;
    int rows;
    long block0;
};

// Rows covered by one block of the radix-16 ramp kernel for transform size N
// (2 G when `pair`); < 0 (with the error set) when N has no such kernel.
long ufomi_ramp_rows_per_block(int N, int pair);
int ufomi_ramp_segments_launch(hipStream_t stream, const float2 *tw, int N, int out_dtype,
                               const ufomi_ramp_seg *d_segs, int nseg, long nblocks, const float *d_coef, int pair);

// One single-slice request of a persistent backprojection launch.
struct ufomi_bp_batch_req {
    const void *sino;   // filtered rows [P][W] in the batch's storage type
    float *out;         // [roi_h][roi_w]
    int W, P, roi_w, roi_h, tiles_x, trig0;
    float axis, roi_x, roi_y, pad;
};

// Tile / angle-chunk configurations of the persistent kernel.
bool ufomi_bp_batch_config_ok(int tile, int achunk);
int ufomi_bp_batch_blocks_per_cu(int tile, int achunk, int dtype, int *blocks);
int ufomi_bp_batch_launch(hipStream_t stream, int tile, int achunk, int dtype, int grid,
                          const ufomi_bp_batch_req *d_reqs, const int2 *d_items, int n_items, int *d_counter,
                          const float2 *d_trig);

// Storage conversions used by the producing / consuming kernels.
__device__ __forceinline__ float ufomi_load(const float *p, long i) { return p[i]; }
__device__ __forceinline__ float ufomi_load(const __half *p, long i) { return __half2float(p[i]); }
__device__ __forceinline__ float ufomi_load(const __hip_bfloat16 *p, long i) { return __bfloat162float(p[i]); }

__device__ __forceinline__ void ufomi_store(float *p, long i, float v) { p[i] = v; }
__device__ __forceinline__ void ufomi_store(__half *p, long i, float v) { p[i] = __float2half_rn(v); }
__device__ __forceinline__ void ufomi_store(__hip_bfloat16 *p, long i, float v) { p[i] = __float2bfloat16(v); }
