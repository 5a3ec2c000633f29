/*
 * ufo_oracle.c -- CPU restatement of ufo-filters' reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it, and only as the test oracle (or the CPU baseline),
 * never as the thing measured or shipped.
 *
 * Every function restates a piece of the reference (ufo-filters 0.17.0) and
 * cites the file:line it follows.  Two arithmetic modes exist where the
 * reference computes in float32 on the device:
 *   - "f64"  : every step in IEEE double (the FP64 oracle every result is
 *              reported against);
 *   - "ref32": float32 at the points where the reference rounds to float32
 *              (LUTs, kernel arithmetic, sequential float accumulation), which
 *              is the "reference semantics" the HIP path must match to 1e-5.
 *
 * Pinning (what each function is checked against; tests/test_oracle.py):
 *   orc_fft_c2c            numpy.fft vectors (tests/golden), naive O(N^2) DFT,
 *                          round trip, Parseval.
 *   orc_filter_coefficients closed forms of ufo-filter-task.c:171-276 evaluated
 *                          independently in numpy, ramp-fromreal against the
 *                          closed cosine sum H_k = 1/4 - 2/pi^2 sum_odd cos/n^2,
 *                          realness and evenness of H; shepp-logan (an
 *                          extension with no counterpart in the reference)
 *                          against sin(pi |f|)/pi and its 2/pi Nyquist ratio.
 *   orc_ramp_chain         numpy rfft/irfft filtering of the same rows,
 *                          fft->ifft identity (test_fft.py:114-129 semantics).
 *   orc_bp_texture_*       brute-force numpy on tiny inputs, constant sinogram
 *                          closed form, theta=0 reduces to a 1-D lerp,
 *                          linearity, FBP of an analytic disc.
 *   orc_bp_nearest_ref32   brute-force numpy restatement on tiny inputs.
 *   orc_cone_weight_ref32  closed-form per-pixel formula (numpy).
 *   orc_gbp_*              reduction to orc_bp_texture for untilted parallel
 *                          geometry, point-source closed form for cone,
 *                          lamino tilt 0 == tomography, brute force on tiny
 *                          inputs; every geometry branch (tilted detector,
 *                          shifted detector and source, volume and axis
 *                          rotation, cone and parallel) against an
 *                          independent ray-plane model with Rodrigues rotation
 *                          matrices and against hand-derived closed forms of
 *                          the projected coordinates
 *                          (tests/test_oracle_gbp_geometry.py), with planted
 *                          mistakes (orc_gbp_plant) shown to fail them.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define ORC_PI_F 3.14159274101257324219f /* M_PI_F of OpenCL, (float) pi */

/* ------------------------------------------------------------------------- */
/* FP64 FFT                                                                  */
/* ------------------------------------------------------------------------- */

/*
 * Unnormalised complex DFT, X_k = sum_n x_n exp(dir * 2 pi i k n / N), with
 * dir = -1 forward and +1 inverse: the convention of deps/oclfft/oclFFT.h:67-72
 * as called by src/common/ufo-fft.c:134-137.  Iterative radix-2, twiddles
 * from libm sin/cos in double.  n must be a power of two.
 */
void orc_fft_c2c(double *re, double *im, int n, int dir)
{
    int i, j, k, len;

    for (i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }

    for (len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (k = 0; k < half; k++) {
            double ang = dir * 2.0 * M_PI * (double) k / (double) len;
            double wr = cos (ang), wi = sin (ang);
            for (i = 0; i < n; i += len) {
                double ur = re[i + k], ui = im[i + k];
                double xr = re[i + k + half], xi = im[i + k + half];
                double vr = xr * wr - xi * wi;
                double vi = xr * wi + xi * wr;
                re[i + k] = ur + vr;
                im[i + k] = ui + vi;
                re[i + k + half] = ur - vr;
                im[i + k + half] = ui - vi;
            }
        }
    }
}

/* Naive O(N^2) DFT with exact integer phase reduction, used to pin orc_fft_c2c. */
void orc_dft_naive(const double *re, const double *im, double *out_re, double *out_im, int n, int dir)
{
    for (int k = 0; k < n; k++) {
        double sr = 0.0, si = 0.0;
        for (int m = 0; m < n; m++) {
            long ph = ((long) k * (long) m) % n;
            double ang = dir * 2.0 * M_PI * (double) ph / (double) n;
            double c = cos (ang), s = sin (ang);
            sr += re[m] * c - im[m] * s;
            si += re[m] * s + im[m] * c;
        }
        out_re[k] = sr;
        out_im[k] = si;
    }
}

/* ------------------------------------------------------------------------- */
/* Filter coefficients: src/ufo-filter-task.c:162-336                         */
/* ------------------------------------------------------------------------- */

enum { ORC_FILTER_RAMP = 0, ORC_FILTER_RAMP_FROMREAL, ORC_FILTER_BUTTERWORTH,
       ORC_FILTER_FARIS_BYER, ORC_FILTER_HAMMING, ORC_FILTER_BH3, ORC_FILTER_SHEPP_LOGAN };

/* ufo-filter-task.c:162-169 -- bin j in (N/2, N) copies bin N-j. */
static void orc_mirror(float *f, unsigned width)
{
    for (unsigned k = width / 2 + 2; k < width; k += 2) {
        f[k] = f[width - k];
        f[k + 1] = f[width - k + 1];
    }
}

/*
 * Fill out[width] (width = 2N floats, interleaved complex) exactly as
 * ufo_filter_task_get_requisition does (ufo-filter-task.c:290-335), including
 * the FFT of the real-space Ram-Lak kernel for ramp-fromreal (:317-334), which
 * the reference runs in float32 on the device and this oracle runs in FP64 and
 * rounds once.  Parameters are float like the task's private fields (:78-84).
 * Returns -1 when width is not a power of two (:295-300).
 */
int orc_filter_coefficients(int type, unsigned width, float scale, float cutoff,
                            float bw_order, float fb_tau, float fb_theta, float *out)
{
    double int_part;
    if (width < 4 || modf (log2 ((double) width), &int_part) != 0.0)
        return -1;

    memset (out, 0, width * sizeof (float));
    out[0] = (float) (0.5 / width);   /* :303-304 */
    out[1] = out[0];

    const double step = 2.0 / width;
    switch (type) {
        case ORC_FILTER_RAMP:               /* :171-182 */
            for (unsigned k = 1; k < width / 4 + 1; k++) {
                out[2 * k] = (float) (k * step * scale);
                out[2 * k + 1] = out[2 * k];
            }
            break;
        case ORC_FILTER_RAMP_FROMREAL:      /* :184-195 */
        case ORC_FILTER_SHEPP_LOGAN:        /* no counterpart: ramp-fromreal, windowed below */
            out[0] = out[1] = (float) (0.25 * scale);
            for (unsigned k = 1; k < width / 4 + 1; k++) {
                out[2 * k] = (k % 2) ? (float) (-scale / ((double) (k * k) * M_PI * M_PI)) : 0.0f;
                out[2 * k + 1] = out[2 * k];
            }
            break;
        case ORC_FILTER_BUTTERWORTH:        /* :197-209 */
            for (unsigned k = 0; k < width / 4 + 1; k++) {
                const double f = k * step;
                out[2 * k] = (float) (f / (1.0 + pow (f / cutoff, 2.0 * bw_order)) * scale);
                out[2 * k + 1] = out[2 * k];
            }
            break;
        case ORC_FILTER_HAMMING:            /* :211-224 */
            for (unsigned k = 0; k < width / 4 + 1; k++) {
                const double f = k * step;
                out[2 * k] = f < cutoff ? (float) (f * (0.54 + 0.46 * cos (M_PI * f / cutoff)) * scale) : 0.0f;
                out[2 * k + 1] = out[2 * k];
            }
            break;
        case ORC_FILTER_BH3:                /* :226-240 */
            for (unsigned k = 1; k < width / 4 + 1; k++) {
                const double f = k * step;
                out[2 * k] = (float) (f * (0.42 + 0.5 * cos (f * M_PI) + 0.08 * cos (2.0 * f * M_PI)) * scale);
                out[2 * k + 1] = out[2 * k];
            }
            break;
        case ORC_FILTER_FARIS_BYER: {       /* :242-276, indexing kept as written */
            const double pi_squared_tau = M_PI * M_PI * fb_tau;
            const double sin_theta_2 = -sin (fb_theta) / 2;
            unsigned padding = 1;
            while (padding < 2 * width)
                padding *= 2;
            out[0] = 0;
            for (unsigned x = 1; x <= width / 2; x++)
                if (x % 2 != 0)
                    out[x] = (float) (1 / (pi_squared_tau * x));
            for (unsigned i = width / 2 + 1; i < width; i++) {
                unsigned x = width + 1 - i;
                if (x % 2 != 0)
                    out[padding - width - i - 1] = (float) (sin_theta_2 / ((double) (x * x) * pi_squared_tau));
            }
            break;
        }
        default:
            return -2;
    }
    orc_mirror (out, width);

    if (type == ORC_FILTER_RAMP_FROMREAL || type == ORC_FILTER_SHEPP_LOGAN) {
        const int n = (int) (width / 2);
        double *re = malloc (n * sizeof (double)), *im = malloc (n * sizeof (double));
        for (int k = 0; k < n; k++) {
            re[k] = out[2 * k];
            im[k] = out[2 * k + 1];
        }
        orc_fft_c2c (re, im, n, -1);
        if (type == ORC_FILTER_SHEPP_LOGAN) {
            /* Shepp-Logan window sin(pi f)/(pi f) with f the bin frequency in cycles per
             * sample, |f| <= 1/2, so the window is 2/pi at Nyquist. */
            for (int k = 1; k < n; k++) {
                const double f = (double) (k <= n / 2 ? k : n - k) / n;
                const double w = sin (M_PI * f) / (M_PI * f);
                re[k] *= w;
                im[k] *= w;
            }
        }
        for (int k = 0; k < n; k++) {
            out[2 * k] = (float) re[k];
            out[2 * k + 1] = (float) im[k];
        }
        free (re);
        free (im);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* fft -> filter -> ifft chain                                               */
/* ------------------------------------------------------------------------- */

/*
 * fft task, 1-D, power-of-two N >= W (ufo-fft-task.c:122-196, 226-352):
 * fft_spread (fft.cl:20-48) pads each real row to N complex values with zero
 * imaginary part, then a forward unnormalised FFT.  out has P rows of 2N
 * doubles (interleaved re, im).
 */
void orc_fft_task(const float *in, int W, int P, int N, double *out)
{
    double *re = malloc (N * sizeof (double)), *im = malloc (N * sizeof (double));
    for (int r = 0; r < P; r++) {
        for (int i = 0; i < N; i++) {
            re[i] = i < W ? (double) in[(size_t) r * W + i] : 0.0;
            im[i] = 0.0;
        }
        orc_fft_c2c (re, im, N, -1);
        for (int i = 0; i < N; i++) {
            out[(size_t) r * 2 * N + 2 * i] = re[i];
            out[(size_t) r * 2 * N + 2 * i + 1] = im[i];
        }
    }
    free (re);
    free (im);
}

/* filter kernel (filter.cl:20-29): element-wise product over the 2N floats of every row. */
void orc_filter_task(double *data, int width, int P, const float *coef)
{
    for (int r = 0; r < P; r++)
        for (int i = 0; i < width; i++)
            data[(size_t) r * width + i] *= (double) coef[i];
}

/*
 * ifft task, 1-D, power-of-two N (ufo-ifft-task.c:111-177, 208-311): backward
 * unnormalised FFT in place on the input (:287-293), then fft_pack
 * (fft.cl:50-82) keeps the real part times 1/N of the first `crop` bins.
 */
void orc_ifft_task(double *data, int N, int P, int crop, double *out)
{
    double *re = malloc (N * sizeof (double)), *im = malloc (N * sizeof (double));
    for (int r = 0; r < P; r++) {
        for (int i = 0; i < N; i++) {
            re[i] = data[(size_t) r * 2 * N + 2 * i];
            im[i] = data[(size_t) r * 2 * N + 2 * i + 1];
        }
        orc_fft_c2c (re, im, N, +1);
        for (int i = 0; i < N; i++) {
            data[(size_t) r * 2 * N + 2 * i] = re[i];
            data[(size_t) r * 2 * N + 2 * i + 1] = im[i];
        }
        for (int i = 0; i < crop; i++)
            out[(size_t) r * crop + i] = re[i] * (1.0 / N);
    }
    free (re);
    free (im);
}

/* The whole chain `fft size-x=N ! filter ! ifft crop-width=crop` in FP64. */
void orc_ramp_chain(const float *in, int W, int P, int N, const float *coef, int crop, double *out)
{
    double *tmp = malloc ((size_t) P * 2 * N * sizeof (double));
    orc_fft_task (in, W, P, N, tmp);
    orc_filter_task (tmp, 2 * N, P, coef);
    orc_ifft_task (tmp, N, P, crop, out);
    free (tmp);
}

/*
 * The same chain in float32 throughout (the CPU baseline's restatement; the
 * reference's oclfft runs float32 transforms): per row, zero padding to N
 * complex values, an iterative radix-2 forward transform with float twiddles
 * (FP64 sin/cos rounded once), the coefficient product, the inverse transform,
 * Re / N cropped.  Rows run in parallel.  Checked against orc_ramp_chain.
 */
void orc_ramp_chain_f32(const float *in, int W, int P, int N, const float *coef, int crop, float *out)
{
    float *twr = malloc (sizeof (float) * (size_t) (N / 2)), *twi = malloc (sizeof (float) * (size_t) (N / 2));
    for (int k = 0; k < N / 2; k++) {
        twr[k] = (float) cos (2.0 * M_PI * k / N);
        twi[k] = (float) sin (2.0 * M_PI * k / N);
    }
    #pragma omp parallel
    {
        float *re = malloc (sizeof (float) * (size_t) N), *im = malloc (sizeof (float) * (size_t) N);
        #pragma omp for schedule(static)
        for (int r = 0; r < P; r++) {
            for (int pass = 0; pass < 2; pass++) {
                const float dir = pass == 0 ? -1.0f : 1.0f;
                if (pass == 0) {
                    for (int i = 0; i < N; i++) {
                        re[i] = i < W ? in[(size_t) r * W + i] : 0.0f;
                        im[i] = 0.0f;
                    }
                } else {
                    for (int i = 0; i < N; i++) {  /* filter: interleaved coefficients (re, im) of bin i */
                        re[i] *= coef[2 * i];
                        im[i] *= coef[2 * i + 1];
                    }
                }
                for (int i = 1, j = 0; i < N; i++) {
                    int bit = N >> 1;
                    for (; j & bit; bit >>= 1)
                        j ^= bit;
                    j ^= bit;
                    if (i < j) {
                        float t = re[i]; re[i] = re[j]; re[j] = t;
                        t = im[i]; im[i] = im[j]; im[j] = t;
                    }
                }
                for (int len = 2; len <= N; len <<= 1) {
                    const int half = len >> 1, step = N / len;
                    for (int i = 0; i < N; i += len) {
                        for (int k = 0; k < half; k++) {
                            const float wr = twr[k * step], wi = dir * twi[k * step];
                            const float xr = re[i + k + half], xi = im[i + k + half];
                            const float vr = xr * wr - xi * wi, vi = xr * wi + xi * wr;
                            re[i + k + half] = re[i + k] - vr;
                            im[i + k + half] = im[i + k] - vi;
                            re[i + k] += vr;
                            im[i + k] += vi;
                        }
                    }
                }
            }
            for (int i = 0; i < crop; i++)
                out[(size_t) r * crop + i] = re[i] / (float) N;
        }
        free (re);
        free (im);
    }
    free (twr);
    free (twi);
}

/* ------------------------------------------------------------------------- */
/* Parallel-beam backprojection: src/ufo-backproject-task.c, backproject.cl  */
/* ------------------------------------------------------------------------- */

/* LUT of ufo-backproject-task.c:174-196 (create_lut_buffer): FP64 sin/cos rounded to float. */
void orc_bp_lut(double angle_offset, double step, int n, float *sin_f, float *cos_f)
{
    for (int i = 0; i < n; i++) {
        sin_f[i] = (float) sin (angle_offset + i * step);
        cos_f[i] = (float) cos (angle_offset + i * step);
    }
}

/*
 * read_imagef with CLK_NORMALIZED_COORDS_FALSE | CLK_ADDRESS_CLAMP |
 * CLK_FILTER_LINEAR on a 1-pixel-high footprint (backproject.cl:20-22, 90):
 * the y coordinate proj + 0.5 lands exactly on row proj, so the OpenCL 1.2
 * linear filter reduces to (1-a) T[i0] + a T[i0+1] with u = x - 0.5,
 * i0 = floor(u), a = u - floor(u) and a zero border colour.
 */
static inline float orc_tex_row_f32(const float *row, int w, float x)
{
    float u = x - 0.5f;
    float fl = floorf (u);
    float a = u - fl;
    int i0 = (int) fl;
    float t0 = (i0 >= 0 && i0 < w) ? row[i0] : 0.0f;
    float t1 = (i0 + 1 >= 0 && i0 + 1 < w) ? row[i0 + 1] : 0.0f;
    return (1.0f - a) * t0 + a * t1;
}

static inline double orc_tex_row_f64(const double *row, int w, double x)
{
    double u = x - 0.5;
    double fl = floor (u);
    double a = u - fl;
    long i0 = (long) fl;
    double t0 = (i0 >= 0 && i0 < w) ? row[i0] : 0.0;
    double t1 = (i0 + 1 >= 0 && i0 + 1 < w) ? row[i0 + 1] : 0.0;
    return (1.0 - a) * t0 + a * t1;
}

/*
 * backproject_tex (backproject.cl:50-94) with the arguments that
 * ufo_backproject_task_process sets (ufo-backproject-task.c:96-150), float32
 * semantics: bx, by, h and the running sum in float, sequential over
 * projections, result sum * M_PI_F / n.  sino is [p_in][w], out [roi_h][roi_w].
 * Compile without FMA contraction (-ffp-contract=off) so every product rounds.
 */
void orc_bp_texture_ref32(const float *sino, int w, int p_in,
                          const float *sin_lut, const float *cos_lut, unsigned angle_offset,
                          float axis_pos, unsigned roi_x, unsigned roi_y, int roi_w, int roi_h,
                          float *out)
{
    #pragma omp parallel for schedule(static)
    for (int idy = 0; idy < roi_h; idy++) {
        for (int idx = 0; idx < roi_w; idx++) {
            const float bx = (float) idx - axis_pos + (float) roi_x + 0.5f;
            const float by = (float) idy - axis_pos + (float) roi_y + 0.5f;
            float sum = 0.0f;
            for (int proj = 0; proj < p_in; proj++) {
                float t1 = by * sin_lut[angle_offset + proj];
                float t2 = bx * cos_lut[angle_offset + proj];
                float h = t1 + t2;
                h = h + axis_pos;
                sum += orc_tex_row_f32 (sino + (size_t) proj * w, w, h);
            }
            float r = sum * ORC_PI_F;
            out[(size_t) idy * roi_w + idx] = r / (float) p_in;
        }
    }
}

/*
 * The same backprojection entirely in FP64: angles theta_i = angle_offset +
 * (offset + i) * step with exact double sin/cos, coordinates, interpolation
 * and sum in double, result * pi / n.  This is the FP64 oracle.
 */
void orc_bp_texture_f64(const double *sino, int w, int p_in,
                        double angle_offset, double angle_step, unsigned offset,
                        double axis_pos, unsigned roi_x, unsigned roi_y, int roi_w, int roi_h,
                        double *out)
{
    #pragma omp parallel for schedule(static)
    for (int idy = 0; idy < roi_h; idy++) {
        for (int idx = 0; idx < roi_w; idx++) {
            const double bx = idx - axis_pos + roi_x + 0.5;
            const double by = idy - axis_pos + roi_y + 0.5;
            double sum = 0.0;
            for (int proj = 0; proj < p_in; proj++) {
                double th = angle_offset + (offset + proj) * angle_step;
                double h = by * sin (th) + bx * cos (th) + axis_pos;
                sum += orc_tex_row_f64 (sino + (size_t) proj * w, w, h);
            }
            out[(size_t) idy * roi_w + idx] = sum * M_PI / p_in;
        }
    }
}

/*
 * backproject_nearest (backproject.cl:24-48), float32 semantics.  The reference
 * indexes the flat sinogram with the *output* width (global size 0) and does
 * no bounds check; reads outside [0, p_in * w_in) are undefined there and
 * return 0 here.
 */
void orc_bp_nearest_ref32(const float *sino, int w_in, int p_in,
                          const float *sin_lut, const float *cos_lut, unsigned angle_offset,
                          float axis_pos, unsigned roi_x, unsigned roi_y, int roi_w, int roi_h,
                          float *out)
{
    const long total = (long) p_in * w_in;
    #pragma omp parallel for schedule(static)
    for (int idy = 0; idy < roi_h; idy++) {
        for (int idx = 0; idx < roi_w; idx++) {
            const float bx = (float) idx - axis_pos + (float) roi_x + 0.5f;
            const float by = (float) idy - axis_pos + (float) roi_y + 0.5f;
            float sum = 0.0f;
            for (int proj = 0; proj < p_in; proj++) {
                float t1 = bx * cos_lut[angle_offset + proj];
                float t2 = by * sin_lut[angle_offset + proj];
                float h = axis_pos + t1;
                h = h + t2;
                float fidx = (float) (proj * roi_w) + h;
                long i = (long) (int) fidx;
                if (fidx >= -2147483648.0f && fidx < 2147483648.0f && i >= 0 && i < total)
                    sum += sino[i];
            }
            float r = sum * ORC_PI_F;
            out[(size_t) idy * roi_w + idx] = r / (float) p_in;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Cone-beam projection weight: conebeam.cl:20-36,                            */
/* ufo-cone-beam-projection-weight-task.c:116-158                             */
/* ------------------------------------------------------------------------- */

void orc_cone_weight_ref32(const float *in, float *out, int w, int h,
                           double center_x, double center_z, double source_distance,
                           double detector_distance, double axis_angle_x)
{
    const float cos_angle = (float) cos (axis_angle_x);
    const float cx = (float) center_x, cz = (float) center_z;
    const float sd = (float) source_distance, dd = (float) detector_distance;
    float overall = sd + dd;
    const float mag_recip = sd / overall;
    if (cos_angle > 0.9999999f)
        overall = sd;
    for (int j = 0; j < h; j++) {
        for (int i = 0; i < w; i++) {
            float x = (((float) i) - cx + 0.5f) * mag_recip;
            float y = (((float) j) - cz + 0.5f) * mag_recip;
            float r = sqrtf (overall * overall + x * x + y * y);
            float v = in[(size_t) j * w + i] * sd;
            v = v * cos_angle;
            out[(size_t) j * w + i] = v / r;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* General backprojection: ufo-general-backproject-task.c:807-1180 and        */
/* kernels/templates/general_bp_*.in                                          */
/* ------------------------------------------------------------------------- */

/*
 * Geometry per projection, already broadcast by the caller the way
 * ufo_scarray_get_double broadcasts one-element arrays (ufo-scarray.c:118-127).
 * All angles are in radians.  Units are detector pixels.
 */
typedef struct {
    double tomo;                 /* axis-angle-z, the tomographic angle */
    double axis_x, axis_y;       /* axis-angle-x (lamino tilt), axis-angle-y */
    double vol_x, vol_y, vol_z;  /* volume-angle-x/y/z */
    double det_ax, det_ay, det_az; /* detector-angle-x/y/z */
    double cx, cz;               /* center-position-x/z */
    double sx, sy, sz;           /* source-position-x/y/z (sy = -inf: parallel) */
    double dx, dy, dz;           /* detector-position-x/y/z */
} orc_gbp_geom;

/* Kernel-variant flags derived by node_setup (ufo-general-backproject-task.c:1305-1336). */
typedef struct {
    int with_axis, with_volume, perpendicular_detector, shifted_detector, shifted_source, parallel_beam;
} orc_gbp_flags;

typedef struct { double x, y, z; } v3d;
typedef struct { float x, y, z; } v3f;

/* general_bp_definitions.in:1-3, trig = (sin, cos) */
static inline v3d rx_d(double s, double c, v3d p) { v3d r = { p.x, c * p.y - s * p.z, s * p.y + c * p.z }; return r; }
static inline v3d ry_d(double s, double c, v3d p) { v3d r = { c * p.x + s * p.z, p.y, -s * p.x + c * p.z }; return r; }
static inline v3d rz_d(double s, double c, v3d p) { v3d r = { c * p.x - s * p.y, s * p.x + c * p.y, p.z }; return r; }
/* float versions with the mad() pairs of the template evaluated as fmaf */
static inline v3f rx_f(float s, float c, v3f p) { v3f r = { p.x, fmaf (c, p.y, -(s * p.z)), fmaf (s, p.y, c * p.z) }; return r; }
static inline v3f ry_f(float s, float c, v3f p) { v3f r = { fmaf (c, p.x, s * p.z), p.y, fmaf (-s, p.x, c * p.z) }; return r; }
static inline v3f rz_f(float s, float c, v3f p) { v3f r = { fmaf (c, p.x, -(s * p.y)), fmaf (s, p.x, c * p.y), p.z }; return r; }

/*
 * Round to the nearest IEEE binary16 value (ties to even), returned as a
 * double: 11 significant bits for normal halves, a quantum of 2^-24 below 2^-14,
 * infinity from 65520 on.  Pinned against numpy's float16 conversion.
 */
double orc_round_half(double x)
{
    if (!isfinite (x) || x == 0.0)
        return x;
    const double a = fabs (x);
    if (a >= 65520.0)
        return copysign (INFINITY, x);
    int e;
    frexp (a, &e);  /* a = m 2^e, m in [0.5, 1) */
    const double q = e - 1 >= -14 ? ldexp (1.0, e - 11) : ldexp (1.0, -24);
    return copysign (nearbyint (a / q) * q, x);
}

/* Bilinear read_imagef, unnormalised coordinates, addressing clamp (0 border),
 * clamp_to_edge, repeat or mirrored_repeat.  OpenCL defines the two repeat
 * modes for normalised coordinates only (OpenCL 1.2 spec, section 8.2); the
 * reference's sampler is unnormalised (ufo-general-backproject-task.c:1450),
 * so here they apply the spec's normalised formulas to s = x / w, t = y / h. */
enum { ORC_ADDR_CLAMP = 0, ORC_ADDR_CLAMP_TO_EDGE = 1, ORC_ADDR_REPEAT = 2, ORC_ADDR_MIRRORED_REPEAT = 3 };

/* texel indices and weight of one axis for the repeat modes, FP64 */
static inline void repeat_axis_d(double x, int w, int addr, long *i0, long *i1, double *a)
{
    double s = x / w, u;
    if (addr == ORC_ADDR_REPEAT) {
        u = (s - floor (s)) * w;
    } else {
        double sp = 2.0 * nearbyint (0.5 * s);
        u = fabs (s - sp) * w;
    }
    const double f = floor (u - 0.5);
    *i0 = (long) f;
    *i1 = *i0 + 1;
    *a = (u - 0.5) - f;
    if (addr == ORC_ADDR_REPEAT) {
        if (*i0 < 0) *i0 += w;
        if (*i1 > w - 1) *i1 -= w;
    } else {
        if (*i0 < 0) *i0 = 0;
        if (*i1 > w - 1) *i1 = w - 1;
    }
}

static inline void repeat_axis_f(float x, int w, int addr, int *i0, int *i1, float *a)
{
    float s = x / (float) w, u;
    if (addr == ORC_ADDR_REPEAT) {
        u = (s - floorf (s)) * (float) w;
    } else {
        float sp = 2.0f * rintf (0.5f * s);
        u = fabsf (s - sp) * (float) w;
    }
    const float f = floorf (u - 0.5f);
    *i0 = (int) f;
    *i1 = *i0 + 1;
    *a = (u - 0.5f) - f;
    if (addr == ORC_ADDR_REPEAT) {
        if (*i0 < 0) *i0 += w;
        if (*i1 > w - 1) *i1 -= w;
    } else {
        if (*i0 < 0) *i0 = 0;
        if (*i1 > w - 1) *i1 = w - 1;
    }
}

static inline double tex2_d(const float *img, int w, int h, double x, double y, int addr)
{
    if (!isfinite (x) || !isfinite (y))
        return 0.0;
    if (addr == ORC_ADDR_REPEAT || addr == ORC_ADDR_MIRRORED_REPEAT) {
        long i0, i1, j0, j1;
        double a, b;
        repeat_axis_d (x, w, addr, &i0, &i1, &a);
        repeat_axis_d (y, h, addr, &j0, &j1, &b);
        return (1 - a) * (1 - b) * img[j0 * w + i0] + a * (1 - b) * img[j0 * w + i1] +
               (1 - a) * b * img[j1 * w + i0] + a * b * img[j1 * w + i1];
    }
    double u = x - 0.5, v = y - 0.5;
    double fu = floor (u), fv = floor (v);
    double a = u - fu, b = v - fv;
    fu = fmax (fmin (fu, 1e9), -1e9);  /* far outside stays far outside, on its side */
    fv = fmax (fmin (fv, 1e9), -1e9);
    long i0 = (long) fu, j0 = (long) fv;
    double t[4];
    for (int q = 0; q < 4; q++) {
        long i = i0 + (q & 1), j = j0 + (q >> 1);
        if (addr == ORC_ADDR_CLAMP_TO_EDGE) {
            i = i < 0 ? 0 : (i >= w ? w - 1 : i);
            j = j < 0 ? 0 : (j >= h ? h - 1 : j);
            t[q] = img[(size_t) j * w + i];
        } else {
            t[q] = (i >= 0 && i < w && j >= 0 && j < h) ? img[(size_t) j * w + i] : 0.0;
        }
    }
    return (1 - a) * (1 - b) * t[0] + a * (1 - b) * t[1] + (1 - a) * b * t[2] + a * b * t[3];
}

static inline float tex2_f(const float *img, int w, int h, float x, float y, int addr)
{
    if (!isfinite (x) || !isfinite (y))
        return 0.0f;
    if (addr == ORC_ADDR_REPEAT || addr == ORC_ADDR_MIRRORED_REPEAT) {
        int i0, i1, j0, j1;
        float a, b;
        repeat_axis_f (x, w, addr, &i0, &i1, &a);
        repeat_axis_f (y, h, addr, &j0, &j1, &b);
        float r0 = (1 - a) * (1 - b) * img[(size_t) j0 * w + i0];
        float r1 = a * (1 - b) * img[(size_t) j0 * w + i1];
        float r2 = (1 - a) * b * img[(size_t) j1 * w + i0];
        float r3 = a * b * img[(size_t) j1 * w + i1];
        return ((r0 + r1) + r2) + r3;
    }
    float u = x - 0.5f, v = y - 0.5f;
    float fu = floorf (u), fv = floorf (v);
    float a = u - fu, b = v - fv;
    fu = fmaxf (fminf (fu, 1e9f), -1e9f);
    fv = fmaxf (fminf (fv, 1e9f), -1e9f);
    int i0 = (int) fu, j0 = (int) fv;
    float t[4];
    for (int q = 0; q < 4; q++) {
        int i = i0 + (q & 1), j = j0 + (q >> 1);
        if (addr == ORC_ADDR_CLAMP_TO_EDGE) {
            i = i < 0 ? 0 : (i >= w ? w - 1 : i);
            j = j < 0 ? 0 : (j >= h ? h - 1 : j);
            t[q] = img[(size_t) j * w + i];
        } else {
            t[q] = (i >= 0 && i < w && j >= 0 && j < h) ? img[(size_t) j * w + i] : 0.0f;
        }
    }
    float r0 = (1 - a) * (1 - b) * t[0];
    float r1 = a * (1 - b) * t[1];
    float r2 = (1 - a) * b * t[2];
    float r3 = a * b * t[3];
    return ((r0 + r1) + r2) + r3;
}

/*
 * Planted mistakes for the pinning tests (0 = the faithful restatement).  A test
 * sets a bit, checks that the independent geometry checks now FAIL, and clears it:
 *   bit 0: the inverse detector rotation about x keeps +sin (minus sign dropped);
 *   bit 1: the shifted-source term of the perpendicular cone branch has its sign flipped;
 *   bit 2: the volume rotation runs rotate_x, rotate_y, rotate_z (order reversed);
 *   bit 3: the detector normal is rotated x, y, z (order reversed);
 *   bit 4: the cone weight uses the source distance alone ((S_y - D_y) / S_y).
 * Only the FP64 mode reads it.
 */
int orc_gbp_plant = 0;

/*
 * One voxel of the generated `backproject` kernel for the z parameter
 * (general_bp_body.in:1-32 filled by make_kernel, :1069-1180), FP64.
 * (x, y, z) is voxel_0 before magnification: x_region.x + idx * x_region.y etc.
 * Returns the unnormalised sum over all projections.
 */
static double gbp_voxel_f64(const float *projs, int w, int h, int P, const orc_gbp_geom *g,
                            const orc_gbp_flags *fl, int addr, double x, double y, double z)
{
    double result = 0.0;
    for (int i = 0; i < P; i++) {
        const orc_gbp_geom *q = g + i;
        v3d v0 = { x, y, z }, v, n = { 0, -1, 0 };
        double det_off = 0, project_tmp = 0, tmp_tr = 0, coeff = 1;
        /* make_static_transformations (:807-849) */
        if (!fl->parallel_beam) {
            double m = -q->sy / (q->dy - q->sy);
            v0.x *= m; v0.y *= m; v0.z *= m;
        }
        if (!fl->perpendicular_detector) {
            if (orc_gbp_plant & 8) {
                n = rx_d (sin (q->det_ax), cos (q->det_ax), n);
                n = ry_d (sin (q->det_ay), cos (q->det_ay), n);
                n = rz_d (sin (q->det_az), cos (q->det_az), n);
            } else {
                n = rz_d (sin (q->det_az), cos (q->det_az), n);
                n = ry_d (sin (q->det_ay), cos (q->det_ay), n);
                n = rx_d (sin (q->det_ax), cos (q->det_ax), n);
            }
            det_off = -(q->dx * n.x + q->dy * n.y + q->dz * n.z);
        } else if (!fl->parallel_beam) {
            project_tmp = q->dy - q->sy;
        }
        if (fl->with_volume) {
            if (orc_gbp_plant & 4) {
                v0 = rx_d (sin (q->vol_x), cos (q->vol_x), v0);
                v0 = ry_d (sin (q->vol_y), cos (q->vol_y), v0);
                v0 = rz_d (sin (q->vol_z), cos (q->vol_z), v0);
            } else {
                v0 = rz_d (sin (q->vol_z), cos (q->vol_z), v0);
                v0 = ry_d (sin (q->vol_y), cos (q->vol_y), v0);
                v0 = rx_d (sin (q->vol_x), cos (q->vol_x), v0);
            }
        }
        if (!(fl->perpendicular_detector || fl->parallel_beam))
            tmp_tr = -(det_off + (q->sx * n.x + q->sy * n.y + q->sz * n.z));
        /* make_transformations (:897-1048) */
        v = rz_d (sin (q->tomo), cos (q->tomo), v0);
        if (fl->with_axis) {
            v = ry_d (sin (q->axis_y), cos (q->axis_y), v);
            v = rx_d (sin (q->axis_x), cos (q->axis_x), v);
        }
        if (!fl->parallel_beam)
            coeff = (orc_gbp_plant & 16) ? (q->sy - q->dy) / q->sy : (q->sy - q->dy) / (q->sy - v.y);
        if (fl->perpendicular_detector) {
            if (!fl->parallel_beam) {
                double t = project_tmp / (v.y - q->sy);
                if (fl->shifted_source && (orc_gbp_plant & 2)) {
                    v.x = t * (v.x + q->sx) + q->sx; v.y = t * (v.y - q->sy) + q->sy; v.z = t * (v.z + q->sz) + q->sz;
                } else if (fl->shifted_source) {
                    v.x = t * (v.x - q->sx) + q->sx; v.y = t * (v.y - q->sy) + q->sy; v.z = t * (v.z - q->sz) + q->sz;
                } else {
                    v.x = t * v.x + q->sx; v.y = t * v.y + q->sy; v.z = t * v.z + q->sz;
                }
            }
        } else {
            if (fl->parallel_beam) {
                v.y = -(v.z * n.z + (v.x * n.x + det_off)) / n.y;
            } else {
                v.x -= q->sx; v.y -= q->sy; v.z -= q->sz;
                double t = tmp_tr / (v.x * n.x + v.y * n.y + v.z * n.z);
                v.x = t * v.x + q->sx; v.y = t * v.y + q->sy; v.z = t * v.z + q->sz;
            }
        }
        if (!fl->perpendicular_detector || fl->shifted_detector) {
            v.x -= q->dx; v.y -= q->dy; v.z -= q->dz;
            if (!fl->perpendicular_detector) {
                v = rx_d ((orc_gbp_plant & 1) ? sin (q->det_ax) : -sin (q->det_ax), cos (q->det_ax), v);
                v = ry_d (-sin (q->det_ay), cos (q->det_ay), v);
                v = rz_d (-sin (q->det_az), cos (q->det_az), v);
            }
        }
        double val = tex2_d (projs + (size_t) i * w * h, w, h, v.x + q->cx, v.z + q->cz, addr);
        result += fl->parallel_beam ? val : val * coeff * coeff;
    }
    return result;
}

/* Same voxel in float32 with the template's operation order and the burst-wise
 * accumulation of ufo_general_backproject_task_process (:1610-1706):
 * volume = norm*sum(burst 0); volume += norm*sum(burst k).
 * result_type (general_bp_body.in:7, 21-30 with rtype): 0 half -- the running
 * sum `half result += float` is the float sum rounded to half after every
 * projection; 1 float; 2 double -- norm * result then in double, cast to the
 * float volume. */
static float gbp_voxel_ref32(const float *projs, int w, int h, int P, const orc_gbp_geom *g,
                             const orc_gbp_flags *fl, int addr, float x, float y, float z,
                             float norm, int burst, int result_type)
{
    float volume = 0.0f;
    double result = 0.0;
    for (int i = 0; i < P; i++) {
        const orc_gbp_geom *q = g + i;
        const float sx = (float) q->sx, sy = (float) q->sy, sz = (float) q->sz;
        const float dx = (float) q->dx, dy = (float) q->dy, dz = (float) q->dz;
        v3f v0 = { x, y, z }, v, n = { 0.0f, -1.0f, 0.0f };
        float det_off = 0, project_tmp = 0, tmp_tr = 0, coeff = 1;
        if (!fl->parallel_beam) {
            float m = -sy / (dy - sy);
            v0.x *= m; v0.y *= m; v0.z *= m;
        }
        if (!fl->perpendicular_detector) {
            n = rz_f ((float) sin (q->det_az), (float) cos (q->det_az), n);
            n = ry_f ((float) sin (q->det_ay), (float) cos (q->det_ay), n);
            n = rx_f ((float) sin (q->det_ax), (float) cos (q->det_ax), n);
            det_off = -(dx * n.x + dy * n.y + dz * n.z);
        } else if (!fl->parallel_beam) {
            project_tmp = dy - sy;
        }
        if (fl->with_volume) {
            v0 = rz_f ((float) sin (q->vol_z), (float) cos (q->vol_z), v0);
            v0 = ry_f ((float) sin (q->vol_y), (float) cos (q->vol_y), v0);
            v0 = rx_f ((float) sin (q->vol_x), (float) cos (q->vol_x), v0);
        }
        if (!(fl->perpendicular_detector || fl->parallel_beam))
            tmp_tr = -(det_off + (sx * n.x + sy * n.y + sz * n.z));
        v = rz_f ((float) sin (q->tomo), (float) cos (q->tomo), v0);
        if (fl->with_axis) {
            v = ry_f ((float) sin (q->axis_y), (float) cos (q->axis_y), v);
            v = rx_f ((float) sin (q->axis_x), (float) cos (q->axis_x), v);
        }
        if (!fl->parallel_beam)
            coeff = (sy - dy) / (sy - v.y);
        if (fl->perpendicular_detector) {
            if (!fl->parallel_beam) {
                float t = project_tmp / (v.y - sy);
                if (fl->shifted_source) {
                    v.x = fmaf (t, v.x - sx, sx); v.y = fmaf (t, v.y - sy, sy); v.z = fmaf (t, v.z - sz, sz);
                } else {
                    v.x = fmaf (t, v.x, sx); v.y = fmaf (t, v.y, sy); v.z = fmaf (t, v.z, sz);
                }
            }
        } else {
            if (fl->parallel_beam) {
                v.y = -(fmaf (v.z, n.z, fmaf (v.x, n.x, det_off)) / n.y);
            } else {
                v.x -= sx; v.y -= sy; v.z -= sz;
                float t = tmp_tr / (v.x * n.x + v.y * n.y + v.z * n.z);
                v.x = fmaf (t, v.x, sx); v.y = fmaf (t, v.y, sy); v.z = fmaf (t, v.z, sz);
            }
        }
        if (!fl->perpendicular_detector || fl->shifted_detector) {
            v.x -= dx; v.y -= dy; v.z -= dz;
            if (!fl->perpendicular_detector) {
                v = rx_f (-(float) sin (q->det_ax), (float) cos (q->det_ax), v);
                v = ry_f (-(float) sin (q->det_ay), (float) cos (q->det_ay), v);
                v = rz_f (-(float) sin (q->det_az), (float) cos (q->det_az), v);
            }
        }
        float val = tex2_f (projs + (size_t) i * w * h, w, h, v.x + (float) q->cx, v.z + (float) q->cz, addr);
        const float contrib = fl->parallel_beam ? val : val * coeff * coeff;
        if (result_type == 0)
            result = orc_round_half ((double) ((float) result + contrib));
        else if (result_type == 1)
            result = (double) ((float) result + contrib);
        else
            result += (double) contrib;
        if ((i + 1) % burst == 0 || i + 1 == P) {
            const float part = result_type == 2 ? (float) ((double) norm * result) : norm * (float) result;
            if (i + 1 <= burst)
                volume = part;
            else
                volume += part;
            result = 0.0;
        }
    }
    return volume;
}

/*
 * Volume [nz][ny][nx] for the z parameter: voxel_0 = (x0 + ix*dx, y0 + iy*dy,
 * z_k) with z_k = region_start + k * region_step (create_regions, :38-76) and
 * x0/dx, y0/dy from compute_slice_region (:246-259).  norm = |overall| / P
 * (set_static_args, :292).  mode 0: FP64 result * norm; mode 1: ref32 with the
 * running sum in result_type (0 half, 1 float, 2 double).
 */
void orc_gbp(const float *projs, int w, int h, int P, const orc_gbp_geom *g, const orc_gbp_flags *fl,
             int addr, int nx, int ny, int nz, double x0, double xs, double y0, double ys,
             double z0, double zs, double overall_angle, int mode, int burst, int result_type, double *out)
{
    const double norm = fabs (overall_angle) / P;
    #pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int iz = 0; iz < nz; iz++) {
        for (int iy = 0; iy < ny; iy++) {
            for (int ix = 0; ix < nx; ix++) {
                size_t o = ((size_t) iz * ny + iy) * nx + ix;
                if (mode == 0) {
                    double x = x0 + ix * xs, y = y0 + iy * ys, z = z0 + iz * zs;
                    out[o] = norm * gbp_voxel_f64 (projs, w, h, P, g, fl, addr, x, y, z);
                } else {
                    float x = fmaf ((float) ix, (float) xs, (float) x0);
                    float y = fmaf ((float) iy, (float) ys, (float) y0);
                    float z = (float) (z0 + iz * zs);
                    out[o] = gbp_voxel_ref32 (projs, w, h, P, g, fl, addr, x, y, z, (float) norm, burst,
                                              result_type);
                }
            }
        }
    }
}

/*
 * lamino-backproject (src/ufo-lamino-backproject-task.c:285-480 with the
 * kernel templates src/kernels/templates/common.in, z_template.in,
 * center_template.in, lamino_template.in, roll_template.in and the rotate()
 * macro of definitions.in:1-6), one voxel at a time.
 *   parameter 0 (z):            slice iz at z = region[0] + iz region[2]
 *   parameter 1 (x-center):     x centre region[0] + iz region[2] - offset_x,
 *                               rotate() pivots on the first value
 *   parameter 2 (lamino-angle): lamino angle region[0] + iz region[2]
 *   parameter 3 (roll-angle):   roll angle region[0] + iz region[2]
 * mode 0: FP64 throughout; mode 1: float32 with the kernel's mad()s as fmaf,
 * sums of `burst` projections added to the volume one burst at a time.
 */
void orc_lamino(const float *projs, int w, int h, int P, unsigned num_projections, float overall_angle,
                float tomo_angle, float lamino_angle, float roll_angle, float center_x, float center_y,
                int offset_x, int offset_y, const int *x_region, const int *y_region, const float *region,
                float z, int parameter, int addr, int nx, int ny, int nz, int mode, int burst, double *out)
{
    const float norm32 = fabsf (overall_angle) / num_projections;
    float *tomo = malloc (sizeof (float) * (size_t) P);
    for (int i = 0; i < P; i++)
        tomo[i] = tomo_angle > -FLT_MAX ? tomo_angle : overall_angle * (unsigned) i / num_projections;
    #pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int iz = 0; iz < nz; iz++) {
        for (int iy = 0; iy < ny; iy++) {
            for (int ix = 0; ix < nx; ix++) {
                const size_t o = ((size_t) iz * ny + iy) * nx + ix;
                if (mode == 0) {
                    const double swept = (double) region[0] + iz * (double) region[2];
                    const double x = x_region[0] + (double) ix * x_region[2];
                    const double y = y_region[0] + (double) iy * y_region[2];
                    const double vz = parameter == 0 ? swept : z;
                    const double xc0 = (parameter == 1 ? (double) region[0] : (double) center_x) - offset_x;
                    const double xc = parameter == 1 ? swept - offset_x : xc0;
                    const double yc = (double) center_y - offset_y;
                    const double l = parameter == 2 ? swept : lamino_angle;
                    const double r = parameter == 3 ? swept : roll_angle;
                    const double sl = sin (l), cl = cos (l), sr = sin (-r), cr = cos (-r);
                    double acc = 0.0;
                    for (int i = 0; i < P; i++) {
                        const double t = tomo_angle > -FLT_MAX ? tomo_angle
                                                               : (double) overall_angle * i / num_projections;
                        double px = x * cos (t) + y * sin (t) + xc;
                        double py = x * cl * sin (t) - y * cl * cos (t) + vz * sl + yc;
                        px -= xc0;
                        py -= yc;
                        px = px * cr + py * sr;
                        py = -px * sr + py * cr;
                        px += xc0;
                        py += yc;
                        acc += tex2_d (projs + (size_t) i * w * h, w, h, px, py, addr);
                    }
                    out[o] = acc * fabs ((double) overall_angle) / num_projections;
                } else {
                    const float vx = fmaf ((float) ix, (float) x_region[2], (float) x_region[0]);
                    const float vy = fmaf ((float) iy, (float) y_region[2], (float) y_region[0]);
                    const float xc0 = (parameter == 1 ? region[0] : center_x) - offset_x;
                    const float xc = parameter == 1 ? fmaf ((float) iz, region[2], xc0) : xc0;
                    const float yc = center_y - offset_y;
                    float sl = sinf (lamino_angle), cl = cosf (lamino_angle);
                    float sr = sinf (-roll_angle), cr = cosf (-roll_angle);
                    float vz = z;
                    if (parameter == 0) vz = fmaf ((float) iz, region[2], region[0]);
                    if (parameter == 2) {
                        const float a = fmaf ((float) iz, region[2], region[0]);
                        sl = sinf (a);
                        cl = cosf (a);
                    }
                    if (parameter == 3) {
                        const float a = fmaf ((float) iz, -region[2], -region[0]);
                        sr = sinf (a);
                        cr = cosf (a);
                    }
                    const float tmp = fmaf (vz, sl, yc), tx = vx * cl, ty = -vy * cl;
                    float vol = 0.0f, result = 0.0f;
                    const int full = (int) (num_projections / burst * burst);
                    for (int i = 0; i < P; i++) {
                        const float st = (float) sin ((double) tomo[i]), ct = (float) cos ((double) tomo[i]);
                        float px = fmaf (vx, ct, fmaf (vy, st, xc));
                        float py = fmaf (tx, st, fmaf (ty, ct, tmp));
                        px -= xc0;
                        py -= yc;
                        px = px * cr + py * sr;
                        py = -px * sr + py * cr;
                        px += xc0;
                        py += yc;
                        const float s = tex2_f (projs + (size_t) i * w * h, w, h, px, py, addr);
                        if (i >= full) {
                            vol = i == 0 ? s * norm32 : vol + s * norm32;
                        } else {
                            result = i % burst == 0 ? s : result + s;
                            if (i % burst == burst - 1)
                                vol = i == burst - 1 ? result * norm32 : vol + result * norm32;
                        }
                    }
                    out[o] = vol;
                }
            }
        }
    }
    free (tomo);
}
