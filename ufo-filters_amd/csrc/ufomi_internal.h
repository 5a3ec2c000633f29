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

// Storage conversions used by the producing / consuming kernels.
__device__ __forceinline__ float ufomi_load(const float *p, long i) { return p[i]; }
__device__ __forceinline__ float ufomi_load(const __half *p, long i) { return __half2float(p[i]); }
__device__ __forceinline__ float ufomi_load(const __hip_bfloat16 *p, long i) { return __bfloat162float(p[i]); }

__device__ __forceinline__ void ufomi_store(float *p, long i, float v) { p[i] = v; }
__device__ __forceinline__ void ufomi_store(__half *p, long i, float v) { p[i] = __float2half_rn(v); }
__device__ __forceinline__ void ufomi_store(__hip_bfloat16 *p, long i, float v) { p[i] = __float2bfloat16(v); }
