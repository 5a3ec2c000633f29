/*
 * ufomi.h -- C ABI of the MI355X-native FBP hot path of ufo-filters.
 *
 * Plain C: no GLib, no torch types, no HIP types in the signatures.  Device
 * pointers are caller-owned (hipMalloc'd or any framework's device memory on
 * the context's device); streams are passed as opaque `void *` hipStream_t.
 * Every entry point returns UFOMI_OK (0) or a negative ufomi_status; the
 * message of the last failure on the calling thread is ufomi_last_error().
 *
 * Each entry point cites the reference interface (ufo-filters 0.17.0) whose
 * computation it replaces.  The GObject task shims (include/ufo_tasks.h)
 * translate task properties into these calls.
 */
#ifndef UFOMI_H
#define UFOMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    UFOMI_OK = 0,
    UFOMI_ERR_ARG = -1,          /* invalid argument (null pointer, size out of range) */
    UFOMI_ERR_HIP = -2,          /* HIP runtime error */
    UFOMI_ERR_NOT_POW2 = -3,     /* filter width not a power of two (ufo-filter-task.c:295-300) */
    UFOMI_ERR_UNSUPPORTED = -4,  /* option the reference accepts that is not implemented here */
    UFOMI_ERR_PROJECTIONS = -5,  /* sinogram taller than num-projections (ufo-backproject-task.c:231-237) */
    UFOMI_ERR_SIZE = -6,         /* fft size-x smaller than input (ufo-fft-task.c:139-143), crop larger than input (ufo-ifft-task.c:130-134) */
    UFOMI_ERR_NO_DEVICE = -7,    /* no HIP device / extension not usable */
    UFOMI_ERR_GRAY_MAP = -8,     /* gray-map-min >= gray-map-max for integer store (ufo-general-backproject-task.c:1448-1453) */
    UFOMI_ERR_MEMORY = -9,       /* volume does not fit (ufo-general-backproject-task.c:1542-1546) */
    UFOMI_ERR_NUM_PROJECTIONS = -10, /* general-backproject without num-projections (:1429-1433) */
    UFOMI_ERR_LAYOUT = -11,      /* ifft input is not complex (ufo-ifft-task.c:121-125) */
    UFOMI_ERR_INCOMPLETE = -12,  /* fewer projections than declared (ufo-general-backproject-task.c:1733-1738) */
    UFOMI_ERR_IO = -13           /* a file of the read / write tasks could not be opened, parsed, read or written */
} ufomi_status;

/* Storage type of a filtered sinogram / projection handed to a backprojector. */
typedef enum { UFOMI_F32 = 0, UFOMI_F16 = 1, UFOMI_BF16 = 2 } ufomi_dtype;

/* Filter enum of ufo-filter-task.c:51-68, same order and nicks. */
typedef enum {
    UFOMI_FILTER_RAMP = 0,
    UFOMI_FILTER_RAMP_FROMREAL = 1,
    UFOMI_FILTER_BUTTERWORTH = 2,
    UFOMI_FILTER_FARIS_BYER = 3,
    UFOMI_FILTER_HAMMING = 4,
    UFOMI_FILTER_BH3 = 5,
    /* Extension without a reference counterpart: Shepp-Logan window on the ramp,
       c_k = |f| sinc(f / 2), f = k/N.  No reference parity exists for it. */
    UFOMI_FILTER_SHEPP_LOGAN = 6
} ufomi_filter_type;

typedef struct ufomi_ctx ufomi_ctx;

/* ---------------------------------------------------------------- context */

/* One context per task thread and device; owns a twiddle table and scratch. */
int ufomi_create(int device, ufomi_ctx **out);
int ufomi_destroy(ufomi_ctx *ctx);
/* Launch subsequent work on `hip_stream` (NULL = the legacy default stream). */
int ufomi_set_stream(ufomi_ctx *ctx, void *hip_stream);
int ufomi_synchronize(ufomi_ctx *ctx);
const char *ufomi_last_error(void);
const char *ufomi_version(void);

/* Device memory of the context's device, for hosts that do not link HIP
 * (the task layer of include/ufo_tasks.h uses these).  ufomi_memcpy kinds:
 * 0 host->device, 1 device->host, 2 device->device, ordered on the context's
 * stream; copies touching host memory return after completion. */
int ufomi_device_alloc(ufomi_ctx *ctx, size_t bytes, void **out);
int ufomi_device_free(ufomi_ctx *ctx, void *ptr);
int ufomi_memcpy(ufomi_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);
int ufomi_device_memory(ufomi_ctx *ctx, size_t *free_bytes, size_t *total_bytes);

/* --------------------------------------------------------- ramp filtering */

/*
 * Host-side coefficients of the `filter` task, float32[width], width = 2N,
 * exactly the buffer ufo_filter_task_get_requisition uploads
 * (src/ufo-filter-task.c:290-335), including the FFT of the real-space Ram-Lak
 * kernel for ramp-fromreal (:317-334), done here in FP64 and rounded once.
 * Parameters are the task properties (:481-521): cutoff, order, tau, theta, scale.
 */
int ufomi_filter_coefficients(int filter_type, unsigned width, float scale, float cutoff,
                              float order, float tau, float theta, float *coefficients);

/*
 * `fft dimensions=1 size-x=N`: P real rows of W floats (row pitch W) ->
 * P rows of N interleaved complex (2N floats), zero padded, forward
 * unnormalised.  Replaces fft_spread (src/kernels/fft.cl:20-48) + oclfft
 * (src/ufo-fft-task.c:226-352 -> src/common/ufo-fft.c:123-139).  N: power of
 * two, W <= N <= 16384.
 */
int ufomi_fft_rows(ufomi_ctx *ctx, const float *d_in, int W, int P, int N, float *d_out);

/*
 * `filter`: out[r][i] = in[r][i] * coefficients[i] over the 2N floats of each
 * of P rows (src/kernels/filter.cl:20-29, launched by ufo-filter-task.c:117-144).
 * d_coefficients: device float[2N].  d_in may equal d_out.
 */
int ufomi_filter_apply(ufomi_ctx *ctx, const float *d_in, float *d_out, const float *d_coefficients,
                       int width, int P);

/*
 * `ifft dimensions=1 crop-width=crop`: P rows of N interleaved complex ->
 * P real rows of `crop` floats = Re(IFFT(row))[0:crop] / N
 * (src/ufo-ifft-task.c:208-311 + fft_pack, src/kernels/fft.cl:50-82).
 * crop = 0 means N.  The input is left unchanged (the reference transforms it
 * in place as a side effect).
 */
int ufomi_ifft_rows(ufomi_ctx *ctx, const float *d_in, int N, int P, int crop, float *d_out);

/*
 * The fft and ifft tasks in full (src/ufo-fft-task.c:122-346,
 * src/ufo-ifft-task.c:111-306, src/common/ufo-fft.c:161-354): 1-D, 2-D or 3-D
 * transforms over the first `dimensions` axes of an item, any size.
 * ufomi_fft_nd: the input [n2][n1][n0] (real, or interleaved complex when
 * in_complex) zero-padded to m[0..2] on every axis, forward DFT over the
 * transformed axes, interleaved complex [m2][m1][m0] out (m[k] >= n[k]).
 * ufomi_ifft_nd: interleaved complex [m2][m1][m0] in, inverse DFT over the
 * transformed axes normalised by their sizes, real part cropped to
 * crop_width x crop_height (0 = no crop) out, [m2][crop_height][crop_width].
 * Powers of two up to 8192 and other sizes up to 4096 (chirp-z) take one
 * LDS pass per transformed axis.  Longer axes run as two passes of a
 * four-step split (powers of two up to 2^20 points) or as Bluestein over that
 * split (other sizes up to 2^19): along the rows of 1-D transforms directly,
 * along any axis of 2-D and 3-D transforms after gathering the lines into
 * rows (and scattering them back); such transforms allocate two intermediates
 * of the output's size for the call.
 */
int ufomi_fft_nd(ufomi_ctx *ctx, const float *d_in, int in_complex, const int n[3], int dimensions, const int m[3],
                 float *d_out);
int ufomi_ifft_nd(ufomi_ctx *ctx, const float *d_in, const int m[3], int dimensions, int crop_width, int crop_height,
                  float *d_out);

/*
 * The fused chain `fft size-x=N ! filter ! ifft crop-width=crop` in one
 * launch: rows [P][W] (row pitch W) -> [P][crop] in out_dtype (f32, f16 or
 * bf16 storage).  d_coefficients: device float[2N] as produced by
 * ufomi_filter_coefficients.  When c[2k] == c[2k+1] and c is even in k (every
 * filter except faris-byer) two rows share one complex FFT; the test runs on
 * the device in stream order, so the call never blocks the host.
 */
int ufomi_ramp_filter(ufomi_ctx *ctx, const float *d_in, int W, int P, int N,
                      const float *d_coefficients, int crop, void *d_out, int out_dtype);

/* ------------------------------------------------ parallel backprojection */

typedef enum { UFOMI_BP_NEAREST = 0, UFOMI_BP_TEXTURE = 1 } ufomi_bp_mode;

/* Properties of the `backproject` task (src/ufo-backproject-task.c:442-510, 520-541). */
typedef struct {
    unsigned num_projections;   /* 0 -> sinogram height (:225-227) */
    unsigned offset;            /* first LUT entry used (kernel arg 6) */
    double axis_pos;            /* <= 0 -> W / 2 (:125-134) */
    double angle_step;          /* <= 0 -> pi / num_projections (:246-251) */
    double angle_offset;        /* radians */
    int mode;                   /* ufomi_bp_mode, default texture */
    unsigned roi_x, roi_y, roi_width, roi_height;  /* width/height 0 -> W */
} ufomi_bp_params;

void ufomi_bp_params_init(ufomi_bp_params *p);

/*
 * Tile configuration of ufomi_bp_parallel for this context: slices sharing
 * one voxel tile (0 = automatic: 8, 4, 2 or 1 as n_slices allows; or 1, 2, 4,
 * 8), angles per LDS chunk (0 = automatic, 8, 16), and, for f16 storage,
 * half-precision (v, d) staging interpolated with v_dot2_f32_f16 (half the
 * LDS bytes; adds one f16 rounding of the sample difference and of the
 * weight): 0 or positive = on (the default), negative = off.  Slices and
 * chunk do not change results; half staging does.  No counterpart in the
 * reference.
 */
int ufomi_bp_configure(ufomi_ctx *ctx, int slices_per_tile, int angle_chunk, int half_staging);

/*
 * `backproject`: n_slices sinograms [P_in][W] (slice pitch in elements, 0 =
 * P_in * W) -> n_slices slices [roi_h][roi_w] (out slice pitch 0 = roi_w *
 * roi_h).  Replaces backproject_tex / backproject_nearest
 * (src/kernels/backproject.cl:24-94) and the LUT set-up of
 * src/ufo-backproject-task.c:174-267.  Storage dtype: f32, f16 or bf16
 * (accumulation always f32; nearest mode requires f32).
 */
int ufomi_bp_parallel(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in,
                      int n_slices, size_t sino_slice_pitch, const ufomi_bp_params *params,
                      float *d_out, size_t out_slice_pitch);

/*
 * ufomi_bp_parallel without any device allocation: same arguments, kernels
 * and results, but the (sin, cos) table (computed on the device in FP64 and
 * rounded to float, as the LUT of ufomi_bp_parallel) and the partial volumes
 * of an angle split live in one preallocated array per GPU (64.25 MiB, held
 * until the process ends), and no host-to-device copy is made.  It has no
 * counterpart in the reference; it is for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Limit: P_in <= 32768 (UFOMI_ERR_SIZE beyond); every other
 * shape ufomi_bp_parallel takes fits.  All calls on one GPU share the array,
 * so calls that overlap on one GPU (other streams, other threads) race.
 */
int ufomi_bp_parallel_noalloc(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in,
                              int n_slices, size_t sino_slice_pitch, const ufomi_bp_params *params,
                              float *d_out, size_t out_slice_pitch);

/*
 * `stacked-backproject` (src/ufo-stacked-backproject-task.c:300-430,
 * src/kernels/stacked-backproject.cl): n_slices float32 sinograms [P_in][W]
 * (consecutive) -> slices [roi_h][roi_w] with the `backproject` LUT, axis and
 * ROI rules but h = axis + bx cos - by sin, in one of the task's precision
 * modes (enum order and nicks of precision_values, :37-42):
 *   UFOMI_STACKED_SINGLE  float32 texels, linear interpolation, zero border;
 *   UFOMI_STACKED_HALF    texels rounded to half, interpolated in float;
 *   UFOMI_STACKED_INT8    texels (v - min) * 255 / (max - min) truncated to
 *                         8 bits over the stack, nearest texel clamped to the
 *                         edge, integer sums, (sum (max - min) / 255 + min) pi / P.
 * params->mode must be texture.  Every slice is computed (the reference fills
 * slices only in groups of 2 or 4).
 */
enum { UFOMI_STACKED_INT8 = 0, UFOMI_STACKED_HALF = 1, UFOMI_STACKED_SINGLE = 2 };
int ufomi_bp_stacked(ufomi_ctx *ctx, const float *d_sino, int W, int P_in, int n_slices,
                     const ufomi_bp_params *params, int precision, float *d_out);

/* --------------------------------------------------- batched FBP requests */

/*
 * Uneven batches of single-slice FBP requests, run as one HIP graph: one
 * segmented ramp-filter launch per FFT size, then ONE persistent
 * backprojection launch whose workgroups take (request, voxel tile) items from
 * a device counter, largest angle counts first.  No counterpart in the
 * reference, which reconstructs one stream of equal-size sinograms per
 * pipeline: each request computes what `fft size-x=fft_size ! filter !
 * ifft crop-width=width ! backproject` (src/ufo-fft-task.c:226-352,
 * src/ufo-filter-task.c:117-144, src/ufo-ifft-task.c:208-311,
 * src/ufo-backproject-task.c:125-267) computes for its sinogram, with the
 * properties in `params` (texture mode).  For callers that queue many small
 * reconstructions (region-of-interest previews, parameter scans).
 */
typedef struct {
    const float *sinogram;      /* device [num_angles][width] float32 */
    float *slice;               /* device [roi_height][roi_width] float32 */
    int width;                  /* detector pixels W */
    int num_angles;             /* sinogram height P_in */
    int fft_size;               /* power of two >= width, 16..8192; 0 -> smallest >= 2 width */
    ufomi_bp_params params;     /* `backproject` properties; mode must be texture */
} ufomi_fbp_request;

typedef struct ufomi_fbp_batch ufomi_fbp_batch;

typedef struct {
    double voxel_updates;       /* sum over requests of roi_w * roi_h * P_in */
    int tile;                   /* voxel tile side of the persistent kernel */
    int angle_chunk;            /* angles staged per LDS pass */
    int workgroups;             /* persistent grid size */
    int items;                  /* (request, tile) work items */
    int filter_launches;        /* segmented ramp launches (one per FFT size) */
    size_t filtered_bytes;      /* device buffer of filtered rows owned by the batch */
} ufomi_fbp_batch_info;

/*
 * Validates every request (the errors of ufomi_bp_parallel / ufomi_ramp_filter,
 * naming the request index), computes the filter coefficients
 * (filter_type, scale as ufomi_filter_coefficients, other parameters at their
 * defaults) and the FP64 trig tables, allocates the filtered-row buffer in
 * `storage_dtype` (f32, f16, bf16) and builds the graph with tile 32, angle
 * chunk 16.  The batch keeps the request pointers: run it while they are valid.
 */
int ufomi_fbp_batch_create(ufomi_ctx *ctx, const ufomi_fbp_request *requests, int n, int filter_type,
                           float filter_scale, int storage_dtype, ufomi_fbp_batch **out);
/* Choose tile (32, 64) and angle chunk (8, 16, 32; 8, 16 for tile 64); rebuilds the graph. */
int ufomi_fbp_batch_configure(ufomi_ctx *ctx, ufomi_fbp_batch *batch, int tile, int angle_chunk);
/*
 * Autotuner: filters once, then times every tile / angle-chunk configuration
 * of the persistent backprojection `reps` times on this batch (HIP events on
 * the context's stream) and configures the fastest.  Results are cached per
 * process under the batch's shape signature; a cached choice is applied
 * without timing unless `force`.  Outputs (may be NULL): the choice and its
 * mean backprojection time in ms (0 when taken from the cache).
 */
int ufomi_fbp_batch_autotune(ufomi_ctx *ctx, ufomi_fbp_batch *batch, int reps, int force, int *tile,
                             int *angle_chunk, float *ms);
/* One replay of the graph on the context's stream (asynchronous). */
int ufomi_fbp_batch_run(ufomi_ctx *ctx, ufomi_fbp_batch *batch);
/* The same work launched directly, stage 1 = filters, 2 = backprojection, 3 = both. */
int ufomi_fbp_batch_launch(ufomi_ctx *ctx, ufomi_fbp_batch *batch, int stages);
int ufomi_fbp_batch_get_info(const ufomi_fbp_batch *batch, ufomi_fbp_batch_info *info);
int ufomi_fbp_batch_destroy(ufomi_fbp_batch *batch);

/* ------------------------------------- learned sparse-view stage (U-Net) */

/*
 * A residual U-Net that maps a sparse-view FBP slice to an estimate of the
 * full-view FBP slice: 4 levels of 32, 64, 128, 256 channels, two 3x3 convs +
 * ReLU per level, 2x2 max pooling, 2x2 stride-2 transposed convs, skip
 * concatenation, 1x1 output conv added to the input; about 1.93M parameters.
 * BF16 weights and activations, FP32 accumulation (MFMA implicit GEMM), FP32
 * master weights and Adam.  No counterpart in the reference: it follows
 * `backproject` for callers that reconstruct from few views.
 *
 * Images are float32 [batch][height][width] device arrays, height and width
 * powers of two (16..8192; width >= 64 with `training`, which allocates
 * gradient and optimiser state).  Parameters are a flat float32 array in the order of
 * ufomi_unet_tensor (index 0 .. ufomi_unet_num_tensors - 1), each tensor at
 * its offset (16-byte aligned gaps are zero).  Weights: 3x3 conv
 * [cout][3][3][cin], transposed conv [2x2][cout][1][cin], final [1][1][1][32].
 */
typedef struct ufomi_unet ufomi_unet;

int ufomi_unet_create(ufomi_ctx *ctx, int max_batch, int height, int width, int training, uint64_t seed,
                      ufomi_unet **out);
int ufomi_unet_destroy(ufomi_unet *net);
/* count: parameters; padded: length of the flat parameter array. */
int ufomi_unet_num_params(const ufomi_unet *net, size_t *count, size_t *padded);
int ufomi_unet_num_tensors(const ufomi_unet *net);
int ufomi_unet_tensor(const ufomi_unet *net, int index, const char **name, size_t *offset, int *dims4);
int ufomi_unet_get_params(ufomi_ctx *ctx, ufomi_unet *net, float *host_params);
/* Also resets the optimiser state. */
int ufomi_unet_set_params(ufomi_ctx *ctx, ufomi_unet *net, const float *host_params);
/* d_out = net(d_in), batch <= max_batch (asynchronous). */
int ufomi_unet_forward(ufomi_ctx *ctx, ufomi_unet *net, const float *d_in, int batch, float *d_out);
/* Mean squared error against d_target and its gradient for every parameter
 * (host_grads, the flat layout; may be NULL).  loss may be NULL (then no
 * synchronisation). */
int ufomi_unet_gradients(ufomi_ctx *ctx, ufomi_unet *net, const float *d_in, const float *d_target, int batch,
                         float *loss, float *host_grads);
/* One Adam step (beta 0.9, 0.999, eps 1e-8) on the MSE loss. */
int ufomi_unet_train_step(ufomi_ctx *ctx, ufomi_unet *net, const float *d_in, const float *d_target, int batch,
                          float lr, float *loss);
/* One 3x3 convolution of the network's GEMM core (test and benchmark hook):
 * NHWC bf16 [batch][height][width][cin] -> [..][cout], weights [cout][3][3][cin]
 * bf16, optional float bias, optional ReLU; cin a power of two >= 32. */
int ufomi_nn_conv3x3(ufomi_ctx *ctx, const void *d_in, int batch, int height, int width, int cin, const void *d_w,
                     const float *d_bias, int relu, int cout, void *d_out);

/* ---------------------------------------------------- cone-beam weighting */

/*
 * `cone-beam-projection-weight` for one projection [H][W]
 * (src/kernels/conebeam.cl:20-36, src/ufo-cone-beam-projection-weight-task.c:116-158).
 * n_proj consecutive projections may be weighted in one call; per-projection
 * values are given as arrays of length n_param (1 broadcasts, else n_proj).
 */
int ufomi_cone_weight(ufomi_ctx *ctx, const float *d_in, float *d_out, int W, int H, int n_proj,
                      const double *center_x, const double *center_z, const double *source_distance,
                      const double *detector_distance, const double *axis_angle_x, int n_param);

/*
 * `flat-field-correct` (src/kernels/ffc.cl:20-48,
 * src/ufo-flat-field-correct-task.c:107-160) of n items [H][W] at d_data:
 * with d = dark * dark_scale and f = flat * flat_scale per pixel,
 * (data - d) / (f - d), or log((f - d) / (data - d)) when absorption_correct;
 * NaN and Inf become 0 when fix_nan_and_inf.  With sinogram_input the dark and
 * flat items are one row of W.  Item i uses dark item min(i, n_dark - 1) and
 * flat item min(i, n_flat - 1): a finished input stream keeps feeding its last
 * buffer.  d_out may equal d_data.
 */
int ufomi_flat_field_correct(ufomi_ctx *ctx, const float *d_data, int W, int H, int n, const float *d_dark,
                             int n_dark, const float *d_flat, int n_flat, int sinogram_input, int absorption_correct,
                             int fix_nan_and_inf, float dark_scale, float flat_scale, float *d_out);

/*
 * `average` (src/ufo-average-task.c:102-151): the float32 sum of n items of
 * `elems` floats in order, divided by n, written `copies` times.
 */
int ufomi_average(ufomi_ctx *ctx, const float *d_items, size_t elems, int n, int copies, float *d_out);

/*
 * `transpose-projections` (src/ufo-transpose-projections-task.c:64-150): n projections [n][H][W] become rows first .. first + n - 1 of
 * the H sinograms [H][number][W] at d_sino (row h of projection p is row
 * first + p of sinogram h).  d_proj NULL writes zero rows there instead: the
 * reference's store is zero-initialised, so rows of projections that never
 * arrive read as zeros.  Requires first + n <= number.
 */
int ufomi_transpose_projections(ufomi_ctx *ctx, const float *d_proj, int W, int H, int n, int first, int number,
                                float *d_sino);

/* ------------------------------------------------- general backprojection */

/* Per-projection geometry, the arrays of the general-backproject task
 * (src/ufo-general-backproject-task.c:2087-2311); one value broadcasts. */
typedef struct {
    const double *axis_angle_x, *axis_angle_y, *axis_angle_z;
    const double *volume_angle_x, *volume_angle_y, *volume_angle_z;
    const double *detector_angle_x, *detector_angle_y, *detector_angle_z;
    const double *center_position_x, *center_position_z;
    const double *source_position_x, *source_position_y, *source_position_z;
    const double *detector_position_x, *detector_position_y, *detector_position_z;
    int n_axis_angle_x, n_axis_angle_y, n_axis_angle_z;
    int n_volume_angle_x, n_volume_angle_y, n_volume_angle_z;
    int n_detector_angle_x, n_detector_angle_y, n_detector_angle_z;
    int n_center_position_x, n_center_position_z;
    int n_source_position_x, n_source_position_y, n_source_position_z;
    int n_detector_position_x, n_detector_position_y, n_detector_position_z;
} ufomi_gbp_geometry;

typedef enum { UFOMI_ST_HALF = 0, UFOMI_ST_FLOAT, UFOMI_ST_DOUBLE, UFOMI_ST_UCHAR, UFOMI_ST_USHORT, UFOMI_ST_UINT } ufomi_store_type;
typedef enum { UFOMI_ADDR_NONE = 0, UFOMI_ADDR_CLAMP_TO_EDGE = 1, UFOMI_ADDR_CLAMP = 2,
               UFOMI_ADDR_REPEAT = 3, UFOMI_ADDR_MIRRORED_REPEAT = 4 } ufomi_addressing;

typedef struct {
    unsigned num_projections;
    double overall_angle;        /* default 2 pi */
    double z;                    /* slice z when region step is 0 */
    double region[3];            /* z region (from, to, step); step 0 -> one slice at 0 */
    double x_region[3], y_region[3];  /* step 0 -> (-W/2, 1) over W voxels */
    int store_type;              /* ufomi_store_type, default float */
    int addressing_mode;         /* default clamp */
    double gray_map_min, gray_map_max;
    int compute_type;            /* UFOMI_CT_FLOAT (default) or UFOMI_CT_DOUBLE */
    int result_type;             /* UFOMI_RT_HALF, UFOMI_RT_FLOAT (default) or UFOMI_RT_DOUBLE */
    unsigned burst;              /* projections per running sum (0 = 8); matters for result-type half */
} ufomi_gbp_params;

/* compute-type and result-type nicks of ufo-general-backproject-task.c:367-378 */
enum { UFOMI_CT_FLOAT = 0, UFOMI_CT_DOUBLE = 1 };
enum { UFOMI_RT_HALF = 0, UFOMI_RT_FLOAT = 1, UFOMI_RT_DOUBLE = 2 };

void ufomi_gbp_params_init(ufomi_gbp_params *p);
void ufomi_gbp_geometry_init(ufomi_gbp_geometry *g);

/* The per-projection maps the general backprojector applies, composed in FP64
 * from the transformation chain of make_static_transformations /
 * make_transformations (ufo-general-backproject-task.c:807-1048): 16 doubles
 * per projection, rows U, V, D, C over X = (x, y, z, 1) with u = U.X / D.X,
 * v = V.X / D.X (detector coordinates including the centre) and the FDK
 * factor coeff = 1 / C.X.  *parallel_beam = 1 when every source-position-y is
 * infinite (then D = C = (0, 0, 0, 1)). */
int ufomi_gbp_projection_rows(const ufomi_gbp_geometry *geometry, const ufomi_gbp_params *params, double *rows,
                              int *parallel_beam);

/* Volume dimensions the task's get_requisition would produce (:1464-1587). */
int ufomi_gbp_volume_shape(const ufomi_gbp_params *p, int W, int *nx, int *ny, int *nz);

/*
 * `general-backproject` (parameter z; other parameters: ufomi_gbp_sweep):
 * P projections [H][W] float32 (stacked,
 * projection pitch H*W) -> volume [nz][ny][nx] in store_type.  Each
 * projection's transformation chain (ufo-general-backproject-task.c:807-1048)
 * is composed in FP64 into a projective 4x4 map on the host; the kernel
 * accumulates all P projections per voxel in registers (no per-burst volume
 * read-modify-write).
 * compute-type double, result-type half or double (:367-378, 1075-1150) and the
 * addressing modes repeat / mirrored_repeat run a second kernel:
 *   - compute-type double evaluates the projective maps, the cone weight and
 *     the bilinear weights in FP64 (the reference's double kernel rounds the
 *     sample coordinate to float and interpolates in float);
 *   - the running sum of every burst of params->burst projections (0 = 8) is
 *     rounded to the result type after each projection, and norm * sum is
 *     added to the volume in the store type once per burst, as the reference's
 *     burst launches do (integer stores map the whole sum once);
 *   - repeat / mirrored_repeat apply OpenCL's normalised-coordinate rules
 *     (OpenCL 1.2, section 8.2) to u / W and v / H; the reference's sampler is
 *     unnormalised, where OpenCL leaves these modes undefined.
 */
int ufomi_gbp(ufomi_ctx *ctx, const float *d_projections, int W, int H, int P,
              const ufomi_gbp_geometry *geometry, const ufomi_gbp_params *params, void *d_volume);

/*
 * `general-backproject` with any `parameter` (the enum of
 * ufo-general-backproject-task.c:345-364, same order): the volume's third axis
 * sweeps that parameter over region (from, to, step) while voxels lie at
 * z = params->z; slice k replaces the parameter by region[0] + k region[2]
 * for every projection (make_parameter_assignment, :755-775).  parameter z
 * is ufomi_gbp.  axis-angle-z follows the reference's generated kernel,
 * which never reads the swept value: without per-projection geometry arrays
 * that kernel does not build (UFOMI_ERR_UNSUPPORTED here); with them every
 * slice is the reconstruction at z.
 */
enum { UFOMI_GBP_PARAMETER_Z = 17 };
int ufomi_gbp_sweep(ufomi_ctx *ctx, const float *d_projections, int W, int H, int P,
                    const ufomi_gbp_geometry *geometry, const ufomi_gbp_params *params, int parameter,
                    void *d_volume);

/* ------------------------------------------------------ lamino-backproject */

/* Parameter swept along the volume's third axis (parameter_values,
 * src/ufo-lamino-backproject-task.c:48-60). */
enum { UFOMI_LAMINO_Z = 0, UFOMI_LAMINO_X_CENTER = 1, UFOMI_LAMINO_LAMINO_ANGLE = 2, UFOMI_LAMINO_ROLL_ANGLE = 3 };

/* Properties of `lamino-backproject` (:688-783, defaults of init :796-831). */
typedef struct {
    unsigned num_projections;
    double overall_angle;        /* default pi; normalisation |overall| / num_projections */
    double tomo_angle;           /* -FLT_MAX (default): projection i at overall * i / num_projections */
    double lamino_angle, roll_angle;
    double center[2];            /* (x, y) rotation axes on the projection */
    int projection_offset[2];    /* (x, y) of a cropped projection */
    int x_region[3], y_region[3];/* (from, to, step), integers */
    double region[3];            /* the swept parameter (from, to, step) */
    double z;                    /* slice z when the parameter is not z */
    int parameter;               /* UFOMI_LAMINO_* */
    int addressing_mode;         /* ufomi_addressing, default clamp */
} ufomi_lamino_params;

void ufomi_lamino_params_init(ufomi_lamino_params *p);
/* Volume dimensions of get_requisition (:247-262). */
int ufomi_lamino_volume_shape(const ufomi_lamino_params *p, int *nx, int *ny, int *nz);
/*
 * `lamino-backproject` (src/ufo-lamino-backproject-task.c:285-480,
 * src/kernels/templates/{z,center,lamino,roll}_template.in, definitions.in):
 * P projections [H][W] (the first P of num_projections) -> volume
 * [nz][ny][nx] float32.  Each projection's parallel-beam map (tomographic
 * rotation, laminographic tilt, roll about the centre) is composed in FP64 and
 * run by the general backprojector's kernel.
 */
int ufomi_lamino_bp(ufomi_ctx *ctx, const float *d_projections, int W, int H, int P,
                    const ufomi_lamino_params *params, float *d_volume);

/* ------------------------------------------------- synthetic phantom data */

/*
 * Seeded ellipse phantoms and their analytic parallel-beam sinograms, for
 * benchmarks and tests (not part of the reference).  Ellipses: 6 doubles each
 * (x0, y0, a, b, phi, density); ellipsoids (rotated about z): 8 doubles
 * (x0, y0, z0, a, b, c, phi, density); pixel units relative to the rotation
 * axis.  Centres uniform in the unit disk (ball) x 0.7, semi-axes U[0.05, 0.4],
 * rotation U[0, pi), density U[-0.5, 1], all scaled by radius_px.
 */
int ufomi_phantom_ellipses(uint64_t seed, int n, double radius_px, double *ellipses);
int ufomi_phantom_ellipsoids(uint64_t seed, int n, double radius_px, double *ellipsoids);
/* Plane section z of n ellipsoids -> n ellipses (missed ones get density 0). */
int ufomi_slice_ellipsoids(const double *ellipsoids, int n, double z, double *ellipses);
/* n_slices sinograms [P][W] of n_ellipses ellipses each (host array
 * [n_slices][n_ellipses][6]); detector pixel j at t = j + 0.5 - axis_pos,
 * angle p * angle_step; FP64 line integrals rounded to float. */
int ufomi_sinogram_ellipses(ufomi_ctx *ctx, const double *ellipses, int n_ellipses, int n_slices,
                            int W, int P, double axis_pos, double angle_step, float *d_sino);
/* P projections [H][W] of n ellipsoids for the geometry given by
 * ufomi_gbp_projection_rows (texel (i, j) sees the ray through u = i + 0.5,
 * v = j + 0.5): density times chord length in voxel units, FP64. */
int ufomi_project_ellipsoids(ufomi_ctx *ctx, const double *ellipsoids, int n_ellipsoids, int W, int H, int P,
                             const double *rows, float *d_projections);
/*
 * Band-limited texture: n (<= 64) plane waves (kx, ky, phase, amplitude) with
 * uniform directions, periods log-uniform in [period_min_px, period_max_px],
 * uniform phases and amplitude rms * sqrt(2 / n) (RMS `rms`), restricted to
 * the disk of radius radius_px around the axis (a random-phase spectral
 * synthesis of the band, the Fourier counterpart of Perlin noise).
 */
int ufomi_phantom_texture(uint64_t seed, int n, double period_min_px, double period_max_px, double rms,
                          double *waves);
/* Sinograms of ellipses plus texture waves ([n_slices][n_waves][4], may be
 * NULL with n_waves 0): exact FP64 line integrals of both. */
int ufomi_sinogram_phantom(ufomi_ctx *ctx, const double *ellipses, int n_ellipses, const double *waves,
                           int n_waves, double radius_px, int n_slices, int W, int P, double axis_pos,
                           double angle_step, float *d_sino);
/* The same phantoms sampled at the backprojector's pixel centres
 * (i + 0.5 - axis_pos, j + 0.5 - axis_pos): n_slices images [N][N]. */
int ufomi_phantom_image(ufomi_ctx *ctx, const double *ellipses, int n_ellipses, const double *waves, int n_waves,
                        double radius_px, int n_slices, int N, double axis_pos, float *d_images);
/* Adds N(0, (rel_sigma * max|d|)^2) noise from a counter-based generator;
 * returns max|d| before the noise in *max_out (may be NULL). */
int ufomi_add_noise(ufomi_ctx *ctx, float *d, size_t n, double rel_sigma, uint64_t seed, float *max_out);

/* ------------------------------------------------ roofline microbenchmarks */

/* Measurement helpers (no reference counterpart): a conflict-free LDS read
 * loop (ds_read_b128) over `blocks` 256-thread workgroups, `iters` iterations
 * (*bytes = LDS bytes read), and a streaming float4 copy of n floats; time
 * them with events on the context's stream to get the device's sustained
 * LDS and HBM bandwidth. */
int ufomi_bench_lds_read(ufomi_ctx *ctx, int blocks, int iters, float *d_sink, double *bytes);
int ufomi_bench_copy(ufomi_ctx *ctx, const float *d_in, float *d_out, size_t n);
/* Dense BF16 MFMA: blocks x 256 threads, each wave `iters` x 8 independent
 * v_mfma_f32_32x32x16_bf16; *flops = the work issued (2 flops per MAC). */
int ufomi_bench_mfma(ufomi_ctx *ctx, int blocks, int iters, float *d_sink, double *flops);

#ifdef __cplusplus
}
#endif

#endif
