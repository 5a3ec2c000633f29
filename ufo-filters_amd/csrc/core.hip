// Context, errors, scratch memory and the host-side filter coefficients.
#include "ufomi_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

static thread_local char g_error[512] = "";

int ufomi_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

// C entry for the host task layer (host/*.c): records a status and message.
extern "C" int ufomi_status_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *ufomi_last_error(void) { return g_error; }
extern "C" const char *ufomi_version(void) { return "ufomi 0.1 (gfx950)"; }

int ufomi_enter(ufomi_ctx *ctx)
{
    if (!ctx)
        return ufomi_fail(UFOMI_ERR_ARG, "null context");
    UFOMI_HIP(hipSetDevice(ctx->device));
    return UFOMI_OK;
}

int ufomi_scratch(ufomi_ctx *ctx, size_t bytes, void **ptr)
{
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) {
            UFOMI_HIP(hipStreamSynchronize(ctx->stream));
            UFOMI_HIP(hipFree(ctx->scratch));
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        UFOMI_HIP(hipMalloc(&ctx->scratch, want));
        ctx->scratch_bytes = want;
    }
    *ptr = ctx->scratch;
    return UFOMI_OK;
}

int ufomi_upload(ufomi_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    UFOMI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    return UFOMI_OK;
}

int ufomi_host_scratch(ufomi_ctx *ctx, size_t bytes, void **ptr)
{
    if (bytes > ctx->hscratch_bytes) {
        if (ctx->hscratch) {
            UFOMI_HIP(hipStreamSynchronize(ctx->stream));
            UFOMI_HIP(hipHostFree(ctx->hscratch));
            ctx->hscratch = nullptr;
            ctx->hscratch_bytes = 0;
        }
        size_t want = bytes < (1u << 16) ? (1u << 16) : bytes;
        UFOMI_HIP(hipHostMalloc(&ctx->hscratch, want, hipHostMallocDefault));
        ctx->hscratch_bytes = want;
    }
    *ptr = ctx->hscratch;
    return UFOMI_OK;
}

extern "C" int ufomi_create(int device, ufomi_ctx **out)
{
    if (!out)
        return ufomi_fail(UFOMI_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return ufomi_fail(UFOMI_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= count)
        return ufomi_fail(UFOMI_ERR_ARG, "device %d out of range (%d devices)", device, count);
    UFOMI_HIP(hipSetDevice(device));

    ufomi_ctx *ctx = (ufomi_ctx *) calloc(1, sizeof(ufomi_ctx));
    ctx->device = device;
    ctx->stream = nullptr;

    // Twiddles exp(-2 pi i k / M), evaluated in FP64 and rounded once to float.
    const int M = UFOMI_FFT_MAX;
    std::vector<float2> tw(M);
    for (int k = 0; k < M; k++) {
        double a = 2.0 * M_PI * (double) k / (double) M;
        tw[k] = make_float2((float) cos(a), (float) -sin(a));
    }
    if (hipMalloc(&ctx->flags, 256) != hipSuccess || hipMalloc(&ctx->twiddles, M * sizeof(float2)) != hipSuccess ||
        hipMemcpy(ctx->twiddles, tw.data(), M * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
        free(ctx);
        return ufomi_fail(UFOMI_ERR_HIP, "twiddle table upload failed");
    }
    if (int rc = ufomi_bp_noalloc_setup(device)) {
        (void) hipFree(ctx->flags);
        (void) hipFree(ctx->twiddles);
        free(ctx);
        return rc;
    }
    *out = ctx;
    return UFOMI_OK;
}

extern "C" int ufomi_destroy(ufomi_ctx *ctx)
{
    if (!ctx)
        return UFOMI_OK;
    // teardown: errors here have no one left to report to
    (void) hipSetDevice(ctx->device);
    if (ctx->stream)
        (void) hipStreamSynchronize(ctx->stream);
    else
        (void) hipDeviceSynchronize();
    (void) hipFree(ctx->twiddles);
    (void) hipFree(ctx->flags);
    if (ctx->lut)
        (void) hipFree(ctx->lut);
    if (ctx->bp_part)
        (void) hipFree(ctx->bp_part);
    if (ctx->scratch)
        (void) hipFree(ctx->scratch);
    if (ctx->hscratch)
        (void) hipHostFree(ctx->hscratch);
    free(ctx);
    return UFOMI_OK;
}

extern "C" int ufomi_set_stream(ufomi_ctx *ctx, void *stream)
{
    if (!ctx)
        return ufomi_fail(UFOMI_ERR_ARG, "null context");
    ctx->stream = (hipStream_t) stream;
    return UFOMI_OK;
}

extern "C" int ufomi_device_alloc(ufomi_ctx *ctx, size_t bytes, void **out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!out)
        return ufomi_fail(UFOMI_ERR_ARG, "null output pointer");
    *out = nullptr;
    const hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void) hipGetLastError();
        *out = nullptr;
        size_t free_b = 0, total_b = 0;
        (void) hipMemGetInfo(&free_b, &total_b);
        return ufomi_fail(UFOMI_ERR_MEMORY, "cannot allocate %zu bytes of device memory (%s; %zu of %zu bytes free)",
                          bytes, hipGetErrorString(e), free_b, total_b);
    }
    return UFOMI_OK;
}

extern "C" int ufomi_device_free(ufomi_ctx *ctx, void *ptr)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (ptr) {
        UFOMI_HIP(hipStreamSynchronize(ctx->stream));
        UFOMI_HIP(hipFree(ptr));
    }
    return UFOMI_OK;
}

extern "C" int ufomi_memcpy(ufomi_ctx *ctx, void *dst, const void *src, size_t bytes, int kind)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if ((!dst || !src) && bytes)
        return ufomi_fail(UFOMI_ERR_ARG, "null copy pointer");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                           : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown copy kind %d", kind);
    UFOMI_HIP(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
    if (kind != 2)
        UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    return UFOMI_OK;
}

extern "C" int ufomi_device_memory(ufomi_ctx *ctx, size_t *free_bytes, size_t *total_bytes)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    size_t f = 0, t = 0;
    UFOMI_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return UFOMI_OK;
}

extern "C" int ufomi_synchronize(ufomi_ctx *ctx)
{
    int rc = ufomi_enter(ctx);
    if (rc)
        return rc;
    UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    return UFOMI_OK;
}

// ----------------------------------------------------------------------------
// Filter coefficients (host, FP64): src/ufo-filter-task.c:162-336.
// ----------------------------------------------------------------------------

// In-place FP64 complex FFT, forward (exp(-2 pi i k n / N)), radix-2
// decimation-in-frequency Stockham autosort ping-ponging over an auxiliary
// buffer.  n is a power of two.
static void fft64_forward(std::vector<double> &re, std::vector<double> &im)
{
    const size_t n = re.size();
    std::vector<double> r2(n), i2(n);
    double *ar = re.data(), *ai = im.data(), *br = r2.data(), *bi = i2.data();
    for (size_t len = n, s = 1; len > 1; len >>= 1, s <<= 1) {
        const size_t m = len / 2;
        for (size_t p = 0; p < m; p++) {
            const double ang = -2.0 * M_PI * (double) p / (double) len;
            const double wr = cos(ang), wi = sin(ang);
            for (size_t q = 0; q < s; q++) {
                const size_t a = q + s * p, b = q + s * (p + m);
                const double dr = ar[a] - ar[b], di = ai[a] - ai[b];
                br[q + s * 2 * p] = ar[a] + ar[b];
                bi[q + s * 2 * p] = ai[a] + ai[b];
                br[q + s * (2 * p + 1)] = dr * wr - di * wi;
                bi[q + s * (2 * p + 1)] = dr * wi + di * wr;
            }
        }
        std::swap(ar, br);
        std::swap(ai, bi);
    }
    if (ar != re.data()) {
        memcpy(re.data(), ar, n * sizeof(double));
        memcpy(im.data(), ai, n * sizeof(double));
    }
}

static void mirror(float *f, unsigned width)
{
    for (unsigned k = width / 2 + 2; k < width; k += 2) {
        f[k] = f[width - k];
        f[k + 1] = f[width - k + 1];
    }
}

extern "C" int ufomi_filter_coefficients(int type, unsigned width, float scale, float cutoff,
                                         float order, float tau, float theta, float *c)
{
    if (!c)
        return ufomi_fail(UFOMI_ERR_ARG, "null coefficient buffer");
    if (width < 4 || !ufomi_is_pow2(width))
        return ufomi_fail(UFOMI_ERR_NOT_POW2, "Invalid filter input width %u, must be a power of two", width);
    if (width / 2 > (1u << 20))
        return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "filter width %u exceeds 2 * 2^20", width);

    memset(c, 0, width * sizeof(float));
    c[0] = c[1] = (float) (0.5 / width);
    const double step = 2.0 / width;
    const unsigned half = width / 4;  // N / 2
    bool fromreal = false;

    switch (type) {
    case UFOMI_FILTER_RAMP:
        for (unsigned k = 1; k <= half; k++)
            c[2 * k] = c[2 * k + 1] = (float) (k * step * scale);
        break;
    case UFOMI_FILTER_RAMP_FROMREAL:
    case UFOMI_FILTER_SHEPP_LOGAN:
        c[0] = c[1] = (float) (0.25 * scale);
        for (unsigned k = 1; k <= half; k++)
            c[2 * k] = c[2 * k + 1] = (k & 1) ? (float) (-scale / ((double) (k * k) * M_PI * M_PI)) : 0.0f;
        fromreal = true;
        break;
    case UFOMI_FILTER_BUTTERWORTH:
        for (unsigned k = 0; k <= half; k++) {
            const double f = k * step;
            c[2 * k] = c[2 * k + 1] = (float) (f / (1.0 + pow(f / cutoff, 2.0 * order)) * scale);
        }
        break;
    case UFOMI_FILTER_HAMMING:
        for (unsigned k = 0; k <= half; k++) {
            const double f = k * step;
            c[2 * k] = c[2 * k + 1] = f < cutoff ? (float) (f * (0.54 + 0.46 * cos(M_PI * f / cutoff)) * scale) : 0.0f;
        }
        break;
    case UFOMI_FILTER_BH3:
        for (unsigned k = 1; k <= half; k++) {
            const double f = k * step;
            c[2 * k] = c[2 * k + 1] = (float) (f * (0.42 + 0.5 * cos(f * M_PI) + 0.08 * cos(2.0 * f * M_PI)) * scale);
        }
        break;
    case UFOMI_FILTER_FARIS_BYER: {
        // Index arithmetic of ufo-filter-task.c:254-276 kept as written: the
        // second loop writes c[width - i - 1] (padding = 2 * width).
        const double pst = M_PI * M_PI * tau;
        const double st2 = -sin(theta) / 2;
        c[0] = 0;
        for (unsigned x = 1; x <= width / 2; x += 2)
            c[x] = (float) (1 / (pst * x));
        for (unsigned i = width / 2 + 1; i < width; i++) {
            const unsigned x = width + 1 - i;
            if (x & 1)
                c[width - i - 1] = (float) (st2 / ((double) (x * x) * pst));
        }
        break;
    }
    default:
        return ufomi_fail(UFOMI_ERR_ARG, "unknown filter type %d", type);
    }
    mirror(c, width);

    if (fromreal) {
        const unsigned n = width / 2;
        std::vector<double> re(n), im(n);
        for (unsigned k = 0; k < n; k++) {
            re[k] = c[2 * k];
            im[k] = c[2 * k + 1];
        }
        fft64_forward(re, im);
        for (unsigned k = 0; k < n; k++) {
            double wr = re[k], wi = im[k];
            if (type == UFOMI_FILTER_SHEPP_LOGAN && k) {
                // sinc window, f = k / N folded to [0, 1/2]
                const unsigned kk = k <= n / 2 ? k : n - k;
                const double x = M_PI * (double) kk / (double) n;
                const double w = sin(x) / x;
                wr *= w;
                wi *= w;
            }
            c[2 * k] = (float) wr;
            c[2 * k + 1] = (float) wi;
        }
    }
    return UFOMI_OK;
}
