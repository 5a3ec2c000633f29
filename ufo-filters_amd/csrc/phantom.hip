// Seeded ellipse / ellipsoid phantoms and their analytic parallel-beam
// sinograms, generated on the GPU in FP64 for benchmark-sized inputs.  Not
// part of the reference; it produces the synthetic data of the benchmark
// configurations.
//
// Geometry matches the backprojector: a detector pixel j samples the line
// x cos(theta) + y sin(theta) = t_j with t_j = j + 0.5 - axis, theta_p =
// p * angle_step, pixel units relative to the rotation axis.
#include "ufomi_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

// splitmix64 -> xoshiro256** (Blackman & Vigna), uniform doubles in [0, 1).
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        for (int i = 0; i < 4; i++)
            s[i] = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double) (next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// Line integral of one ellipse (x0, y0, a, b, cos phi, sin phi, density) along
// x cos(theta) + y sin(theta) = t: 2 rho a b sqrt(a'^2 - t'^2) / a'^2 with
// a'^2 = a^2 cos^2(theta - phi) + b^2 sin^2(theta - phi) and t' the offset of
// the line from the centre.
__device__ double ellipse_line_integral(const double *e, double c, double s, double t)
{
    const double tp = t - (e[0] * c + e[1] * s);
    const double ct = c * e[4] + s * e[5], st = s * e[4] - c * e[5];  // angle theta - phi
    const double a2 = e[2] * e[2] * ct * ct + e[3] * e[3] * st * st;
    const double d = a2 - tp * tp;
    return d > 0.0 ? 2.0 * e[6] * e[2] * e[3] * sqrt(d) / a2 : 0.0;
}

constexpr int kMaxEll = 64;

constexpr int kMaxWave = 64;

// Line integral of the texture wave (kx, ky, phase, amplitude) restricted to
// the disk |x| < R along x cos(theta) + y sin(theta) = t: the chord is
// t n + s d, |s| <= L = sqrt(R^2 - t^2), and
// int_{-L}^{L} a cos(k.n t + phi + (k.d) s) ds = 2 a cos(k.n t + phi) sin(b L) / b, b = k.d.
__device__ double wave_line_integral(const double *w, double c, double s, double t, double R)
{
    const double L2 = R * R - t * t;
    if (L2 <= 0.0) return 0.0;
    const double L = sqrt(L2);
    const double alpha = (w[0] * c + w[1] * s) * t + w[2];
    const double beta = -w[0] * s + w[1] * c;
    const double bl = beta * L;
    const double f = fabs(bl) < 1e-8 ? 2.0 * L : 2.0 * sin(bl) / beta;
    return w[3] * cos(alpha) * f;
}

__global__ __launch_bounds__(256) void sinogram_kernel(const double *__restrict__ ell, int n_ell,
                                                       const double *__restrict__ wav, int n_wav, double R, int W,
                                                       int P, double axis, double step, float *__restrict__ out)
{
    __shared__ double e[kMaxEll * 7];
    __shared__ double wv[kMaxWave * 4];
    const long sl = blockIdx.z;
    for (int k = threadIdx.x; k < n_ell; k += 256) {
        const double *src = ell + (sl * n_ell + k) * 6;
        e[7 * k + 0] = src[0];
        e[7 * k + 1] = src[1];
        e[7 * k + 2] = src[2];
        e[7 * k + 3] = src[3];
        e[7 * k + 4] = cos(src[4]);
        e[7 * k + 5] = sin(src[4]);
        e[7 * k + 6] = src[5];
    }
    for (int k = threadIdx.x; k < 4 * n_wav; k += 256) wv[k] = wav[sl * n_wav * 4 + k];
    __syncthreads();
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (j >= W)
        return;
    const double th = p * step;
    const double c = cos(th), s = sin(th);
    const double t = j + 0.5 - axis;
    double acc = 0.0;
    for (int k = 0; k < n_ell; k++)
        acc += ellipse_line_integral(e + 7 * k, c, s, t);
    for (int k = 0; k < n_wav; k++)
        acc += wave_line_integral(wv + 4 * k, c, s, t, R);
    out[(sl * P + p) * W + j] = (float) acc;
}

// The same phantom sampled at pixel centres (x, y) = (i + 0.5 - axis, j + 0.5 - axis).
__global__ __launch_bounds__(256) void phantom_image_kernel(const double *__restrict__ ell, int n_ell,
                                                            const double *__restrict__ wav, int n_wav, double R,
                                                            int N, double axis, float *__restrict__ out)
{
    const long sl = blockIdx.z;
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= N) return;
    const double x = i + 0.5 - axis, y = j + 0.5 - axis;
    double v = 0.0;
    for (int k = 0; k < n_ell; k++) {
        const double *q = ell + (sl * n_ell + k) * 6;
        const double dx = x - q[0], dy = y - q[1], cp = cos(q[4]), sp = sin(q[4]);
        const double u = (dx * cp + dy * sp) / q[2], w = (-dx * sp + dy * cp) / q[3];
        if (u * u + w * w <= 1.0) v += q[5];
    }
    if (x * x + y * y < R * R)
        for (int k = 0; k < n_wav; k++) {
            const double *q = wav + (sl * n_wav + k) * 4;
            v += q[3] * cos(q[0] * x + q[1] * y + q[2]);
        }
    out[(sl * N + j) * (long) N + i] = (float) v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based Gaussian noise: element i draws from hash(seed, i) (Box-Muller),
// independent of the launch shape.
__global__ __launch_bounds__(256) void noise_kernel(float *__restrict__ d, long n, float sigma, uint64_t seed)
{
    const long stride = (long) gridDim.x * 256;
    for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const uint64_t h1 = mix64(seed ^ mix64(2 * (uint64_t) i + 1));
        const uint64_t h2 = mix64(seed ^ mix64(2 * (uint64_t) i + 2));
        const double u1 = ((double) (h1 >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
        const double u2 = (double) (h2 >> 11) * 0x1.0p-53;
        const double g = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
        d[i] += sigma * (float) g;
    }
}

__global__ __launch_bounds__(256) void absmax_kernel(const float *__restrict__ d, long n, unsigned *__restrict__ out)
{
    __shared__ float red[256];
    float m = 0.0f;
    const long stride = (long) gridDim.x * 256;
    for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
        m = fmaxf(m, fabsf(d[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out, __float_as_uint(red[0]));  // non-negative floats order like their bits
}

// Projections of ellipsoids for a general geometry given by the same per-
// projection affine rows the general backprojector uses (u = U.X / D.X,
// v = V.X / D.X): the pixel (u, v) sees the line {X : (U - u D).X = 0,
// (V - v D).X = 0}.  The line integral of each ellipsoid (8 doubles: x0, y0,
// z0, a, b, c, phi, density) along it is density times the chord length, in
// voxel units.  FP64 throughout.
__global__ __launch_bounds__(256) void project_ellipsoids_kernel(const double *__restrict__ ell, int n_ell,
                                                                 const double *__restrict__ rows, int W, int H,
                                                                 float *__restrict__ out)
{
    __shared__ double e[kMaxEll * 10];
    const int p = blockIdx.z;
    for (int k = threadIdx.x; k < n_ell; k += 256) {
        const double *s = ell + 8 * k;
        double *o = e + 10 * k;
        o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
        o[3] = 1.0 / s[3]; o[4] = 1.0 / s[4]; o[5] = 1.0 / s[5];
        o[6] = cos(s[6]); o[7] = sin(s[6]); o[8] = s[7];
    }
    __syncthreads();
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= W || j >= H)
        return;
    const double *m = rows + 16 * (size_t) p;  // U, V, D rows of 4
    const double u = i + 0.5, v = j + 0.5;     // texel centre (read_imagef samples at u - 0.5)
    double n1[3], n2[3];
    for (int c = 0; c < 3; c++) {
        n1[c] = m[c] - u * m[8 + c];
        n2[c] = m[4 + c] - v * m[8 + c];
    }
    const double b1 = u * m[11] - m[3], b2 = v * m[11] - m[7];
    const double d[3] = {n1[1] * n2[2] - n1[2] * n2[1], n1[2] * n2[0] - n1[0] * n2[2], n1[0] * n2[1] - n1[1] * n2[0]};
    const double n11 = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    const double n22 = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
    const double n12 = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    const double det = n11 * n22 - n12 * n12;
    const double ca = (b1 * n22 - b2 * n12) / det, cb = (b2 * n11 - b1 * n12) / det;
    const double x0[3] = {ca * n1[0] + cb * n2[0], ca * n1[1] + cb * n2[1], ca * n1[2] + cb * n2[2]};
    const double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double acc = 0.0;
    for (int k = 0; k < n_ell; k++) {
        const double *q = e + 10 * k;
        // into the ellipsoid frame (rotation -phi about z), scaled to the unit sphere
        const double px = x0[0] - q[0], py = x0[1] - q[1], pz = x0[2] - q[2];
        const double ex = (q[6] * px + q[7] * py) * q[3], ey = (-q[7] * px + q[6] * py) * q[4], ez = pz * q[5];
        const double fx = (q[6] * d[0] + q[7] * d[1]) * q[3], fy = (-q[7] * d[0] + q[6] * d[1]) * q[4],
                     fz = d[2] * q[5];
        const double A = fx * fx + fy * fy + fz * fz, B = ex * fx + ey * fy + ez * fz;
        const double C = ex * ex + ey * ey + ez * ez - 1.0;
        const double disc = B * B - A * C;
        if (disc > 0.0)
            acc += q[8] * 2.0 * sqrt(disc) / A * dl;
    }
    out[((size_t) p * H + j) * W + i] = (float) acc;
}

}  // namespace

extern "C" int ufomi_project_ellipsoids(ufomi_ctx *ctx, const double *ellipsoids, int n_ellipsoids, int W, int H,
                                        int P, const double *rows, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!ellipsoids || !rows || !d_out || n_ellipsoids <= 0 || n_ellipsoids > kMaxEll || W <= 0 || H <= 0 ||
        P <= 0 || P > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid projection arguments");
    const size_t eb = (size_t) n_ellipsoids * 8 * sizeof(double), mb = (size_t) P * 16 * sizeof(double);
    char *d = nullptr;
    if ((rc = ufomi_scratch(ctx, eb + mb + 256, (void **) &d))) return rc;
    std::vector<char> h(eb + mb + 256);
    const size_t moff = (eb + 255) & ~(size_t) 255;
    memcpy(h.data(), ellipsoids, eb);
    memcpy(h.data() + moff, rows, mb);
    if ((rc = ufomi_upload(ctx, d, h.data(), moff + mb))) return rc;
    dim3 grid((W + 63) / 64, (H + 3) / 4, P);
    hipLaunchKernelGGL(project_ellipsoids_kernel, grid, dim3(256), 0, ctx->stream, (const double *) d, n_ellipsoids,
                       (const double *) (d + moff), W, H, d_out);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

extern "C" int ufomi_phantom_ellipses(uint64_t seed, int n, double radius_px, double *e)
{
    if (!e || n < 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid phantom arguments");
    Rng rng(seed);
    for (int k = 0; k < n; k++) {
        // centre uniform in the unit disk scaled by 0.7, semi-axes U[0.05, 0.4],
        // rotation U[0, pi), density U[-0.5, 1]
        const double r = sqrt(rng.uniform()) * 0.7, psi = 2.0 * M_PI * rng.uniform();
        e[6 * k + 0] = r * cos(psi) * radius_px;
        e[6 * k + 1] = r * sin(psi) * radius_px;
        e[6 * k + 2] = rng.uniform(0.05, 0.4) * radius_px;
        e[6 * k + 3] = rng.uniform(0.05, 0.4) * radius_px;
        e[6 * k + 4] = rng.uniform(0.0, M_PI);
        e[6 * k + 5] = rng.uniform(-0.5, 1.0);
    }
    return UFOMI_OK;
}

extern "C" int ufomi_phantom_ellipsoids(uint64_t seed, int n, double radius_px, double *e)
{
    if (!e || n < 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid phantom arguments");
    Rng rng(seed);
    for (int k = 0; k < n; k++) {
        // centre uniform in the unit ball scaled by 0.7, semi-axes U[0.05, 0.4],
        // rotation about z U[0, pi), density U[-0.5, 1]
        double x, y, z;
        do {
            x = rng.uniform(-1, 1);
            y = rng.uniform(-1, 1);
            z = rng.uniform(-1, 1);
        } while (x * x + y * y + z * z > 1.0);
        double *o = e + 8 * k;
        o[0] = 0.7 * x * radius_px;
        o[1] = 0.7 * y * radius_px;
        o[2] = 0.7 * z * radius_px;
        o[3] = rng.uniform(0.05, 0.4) * radius_px;
        o[4] = rng.uniform(0.05, 0.4) * radius_px;
        o[5] = rng.uniform(0.05, 0.4) * radius_px;
        o[6] = rng.uniform(0.0, M_PI);
        o[7] = rng.uniform(-0.5, 1.0);
    }
    return UFOMI_OK;
}

extern "C" int ufomi_slice_ellipsoids(const double *ell, int n, double z, double *out)
{
    // Plane section z of ellipsoids rotated about z: an ellipse with the same
    // centre and rotation, semi-axes scaled by sqrt(1 - ((z - z0) / c)^2);
    // ellipsoids the plane misses get density 0.
    if (!ell || !out || n < 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid slice arguments");
    for (int k = 0; k < n; k++) {
        const double *e = ell + 8 * k;
        double *o = out + 6 * k;
        const double q = 1.0 - ((z - e[2]) / e[5]) * ((z - e[2]) / e[5]);
        const double f = q > 0.0 ? sqrt(q) : 0.0;
        o[0] = e[0];
        o[1] = e[1];
        o[2] = q > 0.0 ? e[3] * f : 1.0;
        o[3] = q > 0.0 ? e[4] * f : 1.0;
        o[4] = e[6];
        o[5] = q > 0.0 ? e[7] : 0.0;
    }
    return UFOMI_OK;
}

extern "C" int ufomi_phantom_texture(uint64_t seed, int n, double period_min_px, double period_max_px, double rms,
                                     double *waves)
{
    if (!waves || n <= 0 || n > kMaxWave || !(period_min_px > 0.0) || !(period_max_px >= period_min_px))
        return ufomi_fail(UFOMI_ERR_ARG, "invalid texture arguments");
    Rng rng(seed);
    const double amp = rms * sqrt(2.0 / (double) n);
    for (int k = 0; k < n; k++) {
        // direction uniform, period log-uniform in [period_min, period_max], phase uniform
        const double psi = 2.0 * M_PI * rng.uniform();
        const double period = exp(rng.uniform(log(period_min_px), log(period_max_px)));
        const double kk = 2.0 * M_PI / period;
        waves[4 * k + 0] = kk * cos(psi);
        waves[4 * k + 1] = kk * sin(psi);
        waves[4 * k + 2] = 2.0 * M_PI * rng.uniform();
        waves[4 * k + 3] = amp;
    }
    return UFOMI_OK;
}

namespace {

int upload_phantom(ufomi_ctx *ctx, const double *ellipses, int n_ell, const double *waves, int n_wav, int n_slices,
                   double **d_e, double **d_w)
{
    const size_t eb = (size_t) n_slices * n_ell * 6 * sizeof(double);
    const size_t wb = (size_t) n_slices * n_wav * 4 * sizeof(double);
    const size_t woff = (eb + 255) & ~(size_t) 255;
    char *d = nullptr;
    int rc;
    if ((rc = ufomi_scratch(ctx, woff + wb + 8, (void **) &d))) return rc;
    std::vector<char> h(woff + wb + 8, 0);
    if (eb) memcpy(h.data(), ellipses, eb);
    if (wb) memcpy(h.data() + woff, waves, wb);
    if ((rc = ufomi_upload(ctx, d, h.data(), h.size()))) return rc;
    *d_e = (double *) d;
    *d_w = (double *) (d + woff);
    return UFOMI_OK;
}

int check_phantom(const double *ellipses, int n_ell, const double *waves, int n_wav, int n_slices)
{
    if (n_ell < 0 || n_ell > kMaxEll || (n_ell && !ellipses) || n_wav < 0 || n_wav > kMaxWave || (n_wav && !waves) ||
        n_slices <= 0 || n_slices > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid phantom arguments (%d ellipses, %d waves, %d slices)", n_ell, n_wav,
                          n_slices);
    return UFOMI_OK;
}

}  // namespace

extern "C" int ufomi_sinogram_phantom(ufomi_ctx *ctx, const double *ellipses, int n_ellipses, const double *waves,
                                      int n_waves, double radius_px, int n_slices, int W, int P, double axis_pos,
                                      double angle_step, float *d_sino)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if ((rc = check_phantom(ellipses, n_ellipses, waves, n_waves, n_slices))) return rc;
    if (!d_sino || W <= 0 || P <= 0 || P > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid sinogram arguments");
    double *d_e, *d_w;
    if ((rc = upload_phantom(ctx, ellipses, n_ellipses, waves, n_waves, n_slices, &d_e, &d_w))) return rc;
    dim3 grid((W + 255) / 256, P, n_slices);
    hipLaunchKernelGGL(sinogram_kernel, grid, dim3(256), 0, ctx->stream, d_e, n_ellipses, d_w, n_waves, radius_px, W,
                       P, axis_pos, angle_step, d_sino);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

extern "C" int ufomi_sinogram_ellipses(ufomi_ctx *ctx, const double *ellipses, int n_ellipses, int n_slices,
                                       int W, int P, double axis_pos, double angle_step, float *d_sino)
{
    if (!ellipses || n_ellipses <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid sinogram arguments");
    return ufomi_sinogram_phantom(ctx, ellipses, n_ellipses, nullptr, 0, 0.0, n_slices, W, P, axis_pos, angle_step,
                                  d_sino);
}

extern "C" int ufomi_phantom_image(ufomi_ctx *ctx, const double *ellipses, int n_ellipses, const double *waves,
                                   int n_waves, double radius_px, int n_slices, int N, double axis_pos, float *d_img)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if ((rc = check_phantom(ellipses, n_ellipses, waves, n_waves, n_slices))) return rc;
    if (!d_img || N <= 0 || N > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid image arguments");
    double *d_e, *d_w;
    if ((rc = upload_phantom(ctx, ellipses, n_ellipses, waves, n_waves, n_slices, &d_e, &d_w))) return rc;
    dim3 grid((N + 255) / 256, N, n_slices);
    hipLaunchKernelGGL(phantom_image_kernel, grid, dim3(256), 0, ctx->stream, d_e, n_ellipses, d_w, n_waves,
                       radius_px, N, axis_pos, d_img);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

extern "C" int ufomi_add_noise(ufomi_ctx *ctx, float *d, size_t n, double rel_sigma, uint64_t seed, float *max_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d || !n)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid noise arguments");
    unsigned *d_max = nullptr;
    if ((rc = ufomi_scratch(ctx, sizeof(unsigned), (void **) &d_max))) return rc;
    UFOMI_HIP(hipMemsetAsync(d_max, 0, sizeof(unsigned), ctx->stream));
    long blocks = ((long) n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(absmax_kernel, dim3((unsigned) blocks), dim3(256), 0, ctx->stream, d, (long) n, d_max);
    UFOMI_CHECK_LAUNCH();
    unsigned bits = 0;
    UFOMI_HIP(hipMemcpyAsync(&bits, d_max, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    float m;
    memcpy(&m, &bits, sizeof(m));
    if (max_out)
        *max_out = m;
    const float sigma = (float) (rel_sigma * m);
    if (sigma > 0.0f) {
        hipLaunchKernelGGL(noise_kernel, dim3((unsigned) blocks), dim3(256), 0, ctx->stream, d, (long) n, sigma, seed);
        UFOMI_CHECK_LAUNCH();
    }
    return UFOMI_OK;
}
