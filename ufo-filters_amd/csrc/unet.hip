// Learned sparse-view stage: a residual U-Net on FBP slices, BF16 weights and
// activations with FP32 accumulation, trained and run by the kernels below.
// No counterpart in the reference.
//
// Network (H, W powers of two, levels l = 0..3, C_l = 32 << l, NHWC bf16):
//   e1[0]  = relu(conv3x3(img, 1 -> 32))                      first_conv_kernel (VALU)
//   skip_l = relu(conv3x3(relu(conv3x3(in_l))))               Conv3 GEMMs; in_l = maxpool(skip_{l-1})
//   bottom = relu(conv3x3(relu(conv3x3(maxpool(skip_2)))))    C_3 = 256
//   dec_l  = relu(conv3x3(relu(conv3x3([up_l(src) | skip_l]))))   src = bottom or dec_{l+1}
//   up_l   = 2x2 stride-2 transposed conv C_{l+1} -> C_l (linear) UpFwd GEMM
//   out    = img + conv1x1(dec_0, 32 -> 1)                    fp32, in dec_0's conv epilogue (or final_kernel)
// The concatenation is never copied: up_l and skip_l are written into the two
// channel halves of one [pixels][2 C_l] buffer.
//
// Training: MSE loss, backward through the same GEMM core (data gradients as
// convolutions with flipped / transposed weight copies, weight gradients as
// split-K GEMMs over pixels reduced in a fixed order), Adam on FP32 master
// weights, BF16 copies refreshed after every step.
#include "ufomi_internal.h"
#include "nn_gemm.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using nn::bf16;
using nn::View;

namespace {

constexpr int kLevels = 4;
constexpr int kBase = 32;

// -------------------------------------------------------------- VALU layers

// e1 = relu(conv3x3(bf16(img)) + b), 1 -> 32 channels; also stores bf16(img).
// One pixel per thread.  The 9 taps load from clamped addresses without a
// branch (zero padding by select), so the loads issue together instead of each
// waiting under its own bounds test; the 64-byte outputs of a workgroup's 256
// pixels go out through LDS as contiguous 16-byte stores.
__global__ __launch_bounds__(256) void first_conv_kernel(const float *__restrict__ img, long npix, int lh, int lw,
                                                         const bf16 *__restrict__ w, const float *__restrict__ b,
                                                         bf16 *__restrict__ img16, bf16 *__restrict__ out)
{
    __shared__ float ws[kBase * 9 + kBase];
    __shared__ nn::bf16x8 tile[256 * kBase / 8];
    const long p0 = (long) blockIdx.x * 256, p = p0 + threadIdx.x;
    const long pc = p < npix ? p : npix - 1;  // a pixel past the end computes pixel npix - 1, unstored
    const nn::Pix px = nn::pix_of(pc, lh, lw);
    const float *row0 = img + (pc - px.x - ((long) px.y << lw));  // this image
    const int H = 1 << lh, W = 1 << lw;
    float v[9];
    bool in[9];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int yy = px.y + t / 3 - 1, xx = px.x + t % 3 - 1;
        in[t] = yy >= 0 && yy < H && xx >= 0 && xx < W;
        const int yc = min(max(yy, 0), H - 1), xc = min(max(xx, 0), W - 1);
        v[t] = row0[((long) yc << lw) + xc];
    }
    // the weights, while the taps are in flight
    for (int i = threadIdx.x; i < kBase * 9; i += 256) ws[i] = (float) w[i];
    for (int i = threadIdx.x; i < kBase; i += 256) ws[kBase * 9 + i] = b[i];
    float x[9];
#pragma unroll
    for (int t = 0; t < 9; t++) x[t] = in[t] ? (float) (bf16) v[t] : 0.0f;
    if (p < npix) img16[p] = (bf16) v[4];  // tap 4 is the pixel itself
    __syncthreads();  // ws filled
    nn::bf16x8 o[kBase / 8];
#pragma unroll
    for (int c = 0; c < kBase; c++) {
        float a = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) a = fmaf(ws[c * 9 + t], x[t], a);
        o[c / 8][c % 8] = (bf16) fmaxf(a + ws[kBase * 9 + c], 0.0f);
    }
#pragma unroll
    for (int i = 0; i < kBase / 8; i++) tile[threadIdx.x * (kBase / 8) + i] = o[i];
    __syncthreads();
    nn::bf16x8 *dst = reinterpret_cast<nn::bf16x8 *>(out + p0 * kBase);
    const long nvec = (npix - p0 < 256 ? npix - p0 : 256) * (kBase / 8);
#pragma unroll
    for (int k = 0; k < kBase / 8; k++) {
        const int vi = threadIdx.x + 256 * k;
        if (vi < nvec) dst[vi] = tile[vi];
    }
}

// out = img + sum_c w[c] d[p][c] + b (fp32).
__global__ __launch_bounds__(256) void final_kernel(const bf16 *__restrict__ d, long npix, const bf16 *__restrict__ w,
                                                    const float *__restrict__ b, const float *__restrict__ img,
                                                    float *__restrict__ out)
{
    const long p = (long) blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    const nn::bf16x8 *src = reinterpret_cast<const nn::bf16x8 *>(d + p * kBase);
    float a = 0.0f;
#pragma unroll
    for (int i = 0; i < kBase / 8; i++) {
        const nn::bf16x8 v = src[i];
#pragma unroll
        for (int j = 0; j < 8; j++) a = fmaf((float) w[8 * i + j], (float) v[j], a);
    }
    out[p] = img[p] + (a + b[0]);
}

// Backward of final_kernel: gd[p][c] = g[p] w[c] where d[p][c] > 0 (the ReLU of
// dec_0), and per-block partial sums of g[p] d[p][c] (dw) and g[p] (db).
__global__ __launch_bounds__(256) void final_bwd_kernel(const float *__restrict__ g, const bf16 *__restrict__ d,
                                                        long npix, const bf16 *__restrict__ w, bf16 *__restrict__ gd,
                                                        float *__restrict__ part)
{
    __shared__ float red[256][kBase + 1];
    float acc[kBase + 1];
#pragma unroll
    for (int c = 0; c <= kBase; c++) acc[c] = 0.0f;
    for (long p = (long) blockIdx.x * 256 + threadIdx.x; p < npix; p += (long) gridDim.x * 256) {
        const float gp = g[p];
        const nn::bf16x8 *src = reinterpret_cast<const nn::bf16x8 *>(d + p * kBase);
        nn::bf16x8 *dst = reinterpret_cast<nn::bf16x8 *>(gd + p * kBase);
#pragma unroll
        for (int i = 0; i < kBase / 8; i++) {
            const nn::bf16x8 v = src[i];
            nn::bf16x8 o;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float dv = (float) v[j];
                acc[8 * i + j] = fmaf(gp, dv, acc[8 * i + j]);
                o[j] = (bf16) (dv > 0.0f ? gp * (float) w[8 * i + j] : 0.0f);
            }
            dst[i] = o;
        }
        acc[kBase] += gp;
    }
#pragma unroll
    for (int c = 0; c <= kBase; c++) red[threadIdx.x][c] = acc[c];
    __syncthreads();
    if (threadIdx.x <= kBase) {
        float s = 0.0f;
        for (int t = 0; t < 256; t++) s += red[t][threadIdx.x];
        part[(long) blockIdx.x * (kBase + 1) + threadIdx.x] = s;
    }
}

// Per-block column sums of a [npix][C] bf16 view (bias gradients), C <= 256
// dividing 256: thread t sums channel t % C over pixels t / C, t / C + 256 / C, ...
__global__ __launch_bounds__(256) void colsum_kernel(View<const bf16> g, long npix, int C, float *__restrict__ part)
{
    __shared__ float red[256];
    const int c = threadIdx.x % C, lane = threadIdx.x / C, lanes = 256 / C;
    float s = 0.0f;
    for (long p = (long) blockIdx.x * lanes + lane; p < npix; p += (long) gridDim.x * lanes)
        s += (float) *g.at(p, c);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (int k = 0; k < lanes; k++) t += red[k * C + threadIdx.x];
        part[(long) blockIdx.x * C + threadIdx.x] = t;
    }
}

// dst[i] = sum_z src[z * len + i] (+ dst[i] when accumulate), fixed order.
__global__ __launch_bounds__(256) void reduce_kernel(const float *__restrict__ src, int nz, long len,
                                                     float *__restrict__ dst)
{
    const long i = (long) blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    float s = 0.0f;
    for (int z = 0; z < nz; z++) s += src[(long) z * len + i];
    dst[i] = s;
}

// dst[g * len + i] = sum over z in [32 g, 32 g + 32) of src[z * len + i].
__global__ __launch_bounds__(256) void reduce_groups_kernel(const float *__restrict__ src, int nz, long len,
                                                            float *__restrict__ dst)
{
    const long i = (long) blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    const int z0 = blockIdx.y * 32, z1 = z0 + 32 < nz ? z0 + 32 : nz;
    float s = 0.0f;
    for (int z = z0; z < z1; z++) s += src[(long) z * len + i];
    dst[(long) blockIdx.y * len + i] = s;
}

// 2x2 max pooling of C channels (C % 8 == 0); out grid (H/2, W/2).
__global__ __launch_bounds__(256) void maxpool_kernel(View<const bf16> in, View<bf16> out, long nout, int lh, int lw,
                                                      int C)
{
    const long t = (long) blockIdx.x * 256 + threadIdx.x;
    const int groups = C / 8;
    if (t >= nout * groups) return;
    const long po = t / groups;
    const int c = (int) (t % groups) * 8;
    const long img = po >> (lh + lw - 2);
    const nn::Pix px = nn::pix_of(po, lh - 1, lw - 1);
    const long pi = (((img << lh) + 2 * px.y) << lw) + 2 * px.x;
    nn::bf16x8 m = nn::ld8(in.at(pi, c));
    const long off[3] = {1, 1L << lw, (1L << lw) + 1};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const nn::bf16x8 v = nn::ld8(in.at(pi + off[k], c));
#pragma unroll
        for (int j = 0; j < 8; j++) m[j] = (float) v[j] > (float) m[j] ? v[j] : m[j];
    }
    *reinterpret_cast<nn::bf16x8 *>(out.at(po, c)) = m;
}

// Gradient of skip_l: the pooled gradient goes to the first maximum of each
// window (row-major order), the concatenation's gradient is added, and the
// ReLU of skip_l masks the sum.
__global__ __launch_bounds__(256) void maxpool_bwd_kernel(View<const bf16> skip, View<const bf16> gpool,
                                                          View<const bf16> gcat, View<bf16> gout, long nout, int lh,
                                                          int lw, int C)
{
    const long t = (long) blockIdx.x * 256 + threadIdx.x;
    const int groups = C / 8;
    if (t >= nout * groups) return;
    const long po = t / groups;
    const int c = (int) (t % groups) * 8;
    const long img = po >> (lh + lw - 2);
    const nn::Pix px = nn::pix_of(po, lh - 1, lw - 1);
    const long pi = (((img << lh) + 2 * px.y) << lw) + 2 * px.x;
    const long off[4] = {0, 1, 1L << lw, (1L << lw) + 1};
    nn::bf16x8 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = nn::ld8(skip.at(pi + off[k], c));
    const nn::bf16x8 gp = nn::ld8(gpool.at(po, c));
    int arg[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float m = (float) v[0][j];
        arg[j] = 0;
#pragma unroll
        for (int k = 1; k < 4; k++)
            if ((float) v[k][j] > m) {
                m = (float) v[k][j];
                arg[j] = k;
            }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const nn::bf16x8 gc = nn::ld8(gcat.at(pi + off[k], c));
        nn::bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float s = (float) gc[j] + (arg[j] == k ? (float) gp[j] : 0.0f);
            o[j] = (bf16) ((float) v[k][j] > 0.0f ? s : 0.0f);
        }
        *reinterpret_cast<nn::bf16x8 *>(gout.at(pi + off[k], c)) = o;
    }
}

// g = 2 (out - target) / n; per-block partial sums of (out - target)^2.
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ out, const float *__restrict__ tgt, long n,
                                                  float *__restrict__ g, float *__restrict__ part)
{
    __shared__ float red[256];
    float s = 0.0f;
    const float k = 2.0f / (float) n;
    for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < n; i += (long) gridDim.x * 256) {
        const float d = out[i] - tgt[i];
        g[i] = k * d;
        s = fmaf(d, d, s);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ w, bf16 *__restrict__ w16,
                                                   const float *__restrict__ g, float *__restrict__ m,
                                                   float *__restrict__ v, long n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2)
{
    const long i = (long) blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = (bf16) wi;
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float *__restrict__ a, bf16 *__restrict__ b, long n)
{
    const long i = (long) blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (bf16) a[i];
}

// wt[c][tap][co] = w[co][8 - tap][c]: the data gradient of a 3x3 conv is the
// 3x3 conv of the output gradient with these weights.
__global__ __launch_bounds__(256) void conv_flip_kernel(const bf16 *__restrict__ w, bf16 *__restrict__ wt, int cout,
                                                        int cin)
{
    const long i = (long) blockIdx.x * 256 + threadIdx.x;
    if (i >= 9L * cin * cout) return;
    const int co = (int) (i % cout), tap = (int) ((i / cout) % 9), c = (int) (i / (9L * cout));
    wt[i] = w[((long) co * 9 + (8 - tap)) * cin + c];
}

// wt[c][ab][co] = w[ab][co][c].
__global__ __launch_bounds__(256) void up_transpose_kernel(const bf16 *__restrict__ w, bf16 *__restrict__ wt, int cout,
                                                           int cin)
{
    const long i = (long) blockIdx.x * 256 + threadIdx.x;
    if (i >= 4L * cin * cout) return;
    const int co = (int) (i % cout), ab = (int) ((i / cout) % 4), c = (int) (i / (4L * cout));
    wt[i] = w[((long) ab * cout + co) * cin + c];
}

// Channel-major copies for the weight gradients: out[(s C + c) npix + p] =
// in[p + dx_s][c] with dx_s = s - 1 (nshift = 3) or 0 (nshift = 1), zero where
// x + dx_s leaves the row.  A block moves 64 consecutive pixels (plus one on
// each side) through LDS.  With C a multiple of 8 both sides move 16-byte
// vectors: 8 channels of a pixel in, 8 pixels of a channel out.
constexpr int kTpPix = 64;
__global__ __launch_bounds__(256) void transpose_shift_kernel(View<const bf16> in, int C, long npix, int lw,
                                                              int nshift, bf16 *__restrict__ out)
{
    extern __shared__ bf16 tile[];  // [kTpPix + 2][pitch]
    const int pitch = C % 8 == 0 ? C + 8 : C + 2;
    const long p0 = (long) blockIdx.x * kTpPix;
    const int W = 1 << lw;
    if (C % 8 == 0) {
        const int groups = C / 8;
        for (int e = threadIdx.x; e < (kTpPix + 2) * groups; e += 256) {
            const int px = e / groups, cg = e % groups;
            const long p = p0 + px - 1;
            nn::bf16x8 v = (p >= 0 && p < npix) ? nn::ld8(in.at(p, 8 * cg)) : nn::zero8();
            *reinterpret_cast<nn::bf16x8 *>(tile + px * pitch + 8 * cg) = v;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nshift * C * (kTpPix / 8); e += 256) {
            const int px8 = e % (kTpPix / 8), c = (e / (kTpPix / 8)) % C, sh = e / ((kTpPix / 8) * C);
            const int dx = nshift == 3 ? sh - 1 : 0;
            const long p = p0 + 8 * px8;
            if (p >= npix) continue;
            const int x0 = (int) (p & (W - 1)) + dx;
            nn::bf16x8 v;
#pragma unroll
            for (int j = 0; j < 8; j++)
                v[j] = (x0 + j >= 0 && x0 + j < W) ? tile[(8 * px8 + j + 1 + dx) * pitch + c] : (bf16) 0.0f;
            *reinterpret_cast<nn::bf16x8 *>(out + ((long) sh * C + c) * npix + p) = v;
        }
        return;
    }
    for (int e = threadIdx.x; e < (kTpPix + 2) * C; e += 256) {
        const int px = e / C, c = e % C;
        const long p = p0 + px - 1;
        tile[px * pitch + c] = (p >= 0 && p < npix) ? *in.at(p, c) : (bf16) 0.0f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nshift * C * kTpPix; e += 256) {
        const int px = e % kTpPix, c = (e / kTpPix) % C, sh = e / (kTpPix * C);
        const int dx = nshift == 3 ? sh - 1 : 0;
        const long p = p0 + px;
        if (p >= npix) continue;
        const int x = (int) (p & (W - 1)) + dx;
        out[((long) sh * C + c) * npix + p] = (x >= 0 && x < W) ? tile[(px + 1 + dx) * pitch + c] : (bf16) 0.0f;
    }
}

// ------------------------------------------------------------ GEMM launches

template <int MT, int NT, int WM, int WN, typename Op>
int gemm_go(hipStream_t s, const Op &op, long M, long N, int splits)
{
    constexpr int BM = WM * MT * 16, BN = WN * NT * 16;
    const long gx = (M + BM - 1) / BM, gy = (N + BN - 1) / BN;
    if (gx > 0x7fffffffL || gy > 65535 || splits > 65535)
        return ufomi_fail(UFOMI_ERR_SIZE, "GEMM grid too large (%ld x %ld x %d)", gx, gy, splits);
    hipLaunchKernelGGL((nn::gemm_kernel<MT, NT, WM, WN, Op>), dim3((unsigned) gx, (unsigned) gy, splits), dim3(256), 0,
                       s, op);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

// Tile shape by the width of N: 256 x 32, 256 x 64 or 128 x 128 per workgroup.
template <typename Op>
int gemm_fwd(hipStream_t s, const Op &op, long M, long N)
{
    if (N <= 32) return gemm_go<4, 2, 4, 1>(s, op, M, N, 1);
    if (N <= 64) return gemm_go<4, 4, 4, 1>(s, op, M, N, 1);
    return gemm_go<4, 4, 2, 2>(s, op, M, N, 1);
}

// Weight gradients, split-K over pixels: 32 x 256 tiles for 32 output channels,
// 64 x 256 otherwise (each wave 32 or 64 rows by 64 columns; measured against
// 64 x 128 tiles: training step 100.4 -> 89.8 ms).
template <typename Op>
int gemm_wgrad(hipStream_t s, const Op &op, long M, long N, int splits)
{
    if (M <= 32) return gemm_go<2, 4, 1, 4>(s, op, M, N, splits);
    return gemm_go<4, 4, 1, 4>(s, op, M, N, splits);
}

int ilog2(long x)
{
    int l = 0;
    while ((1L << l) < x) l++;
    return l;
}

unsigned blocks_for(long n) { return (unsigned) ((n + 255) / 256); }

// splitmix64 -> xoshiro256** + Box-Muller, for the weight initialisation.
struct HostRng {
    uint64_t s[4];
    explicit HostRng(uint64_t seed)
    {
        for (int i = 0; i < 4; i++) {
            uint64_t z = (seed += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
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
    double uniform() { return ((double) (next() >> 11) + 0.5) * 0x1.0p-53; }
    double normal() { return sqrt(-2.0 * log(uniform())) * cos(2.0 * M_PI * uniform()); }
};

struct ConvP {
    size_t w, b;  // offsets into the flat parameter arrays
    size_t wt;    // offset of the flipped copy in the transformed-weight array
    int cin, cout;
};

struct TensorInfo {
    std::string name;
    size_t off, count;
    int dims[4];
};

}  // namespace

struct ufomi_unet {
    ufomi_ctx *ctx = nullptr;
    int max_batch = 0, H = 0, W = 0, lh = 0, lw = 0;
    bool training = false;
    std::vector<TensorInfo> tensors;
    size_t nparams = 0, ntrans = 0;
    float *master = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16 *w16 = nullptr, *wtrans = nullptr;
    ConvP first, enc1[kLevels], enc2[kLevels], dec1[kLevels - 1], dec2[kLevels - 1], up[kLevels - 1];
    size_t fin_w = 0, fin_b = 0;
    // activations (NHWC bf16), sized for max_batch
    bf16 *img16 = nullptr, *e1[kLevels] = {}, *cat[kLevels - 1] = {}, *bottom = nullptr, *pool[kLevels - 1] = {},
         *d1[kLevels - 1] = {}, *d2[kLevels - 1] = {};
    // gradients (pre-activation where a ReLU follows)
    bf16 *g_e1[kLevels] = {}, *g_e2[kLevels] = {}, *g_cat[kLevels - 1] = {}, *g_pool[kLevels - 1] = {},
         *g_d1[kLevels - 1] = {}, *g_d2[kLevels - 1] = {};
    float *out = nullptr, *gout = nullptr, *ws = nullptr;
    size_t ws_floats = 0;
    bf16 *wgT = nullptr;  // channel-major copies of a weight gradient's operands (training)
    size_t wgT_elems = 0;
    std::vector<void *> allocs;
    long step = 0;
};

namespace {

int dalloc(ufomi_unet *n, size_t bytes, void **p)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess)
        return ufomi_fail(UFOMI_ERR_MEMORY, "cannot allocate %zu bytes for the U-Net: %s", bytes,
                          hipGetErrorString(e));
    n->allocs.push_back(*p);
    return UFOMI_OK;
}

template <typename T>
int dalloc_t(ufomi_unet *n, size_t count, T **p)
{
    return dalloc(n, count * sizeof(T), (void **) p);
}

void free_net(ufomi_unet *n)
{
    for (void *p : n->allocs) (void) hipFree(p);
    delete n;
}

long npix_at(const ufomi_unet *n, int B, int l) { return (long) B * (n->H >> l) * (n->W >> l); }

View<const bf16> cv(const bf16 *p, int cs, int c0 = 0) { return View<const bf16>{p, cs, c0}; }
View<bf16> mv(bf16 *p, int cs, int c0 = 0) { return View<bf16>{p, cs, c0}; }

int maxpool(hipStream_t s, View<const bf16> in, View<bf16> out, long nout, int lh, int lw, int C);

// The 1x1 output conv out[p] = img[p] + sum_c w[c] d[p][c] + b[0] of the last
// 32-channel activation d; keep_d: d itself is stored too.
struct FinalConv {
    const bf16 *w;
    const float *b, *img;
    float *out;
    bool keep_d;
};

// relu(conv3x3(in) + b) into out (forward), or the plain conv (data gradient);
// with pool.ptr, also the 2x2 max pool of out into pool (fused into the LDS kernel).
// With fin, *fused tells whether the output conv was fused into the epilogue (8-row
// kernel, cout == 32); if not, the caller runs it.
int conv3(ufomi_unet *n, hipStream_t s, View<const bf16> in, long npix, int l, int cin, int cout, const bf16 *w,
          const float *bias, bool relu, View<const bf16> mask, int mask_from, View<bf16> out,
          View<bf16> pool = View<bf16>{nullptr, 0, 0}, const FinalConv *fin = nullptr, bool *fused = nullptr)
{
    nn::Conv3 op;
    op.fin_w = nullptr;
    op.fin_b = op.fin_img = nullptr;
    op.fin_out = nullptr;
    if (fused) *fused = false;
    op.in = in;
    op.lh = n->lh - l;
    op.lw = n->lw - l;
    op.lcin = ilog2(cin);
    op.cout = cout;
    op.npix = npix;
    op.w = w;
    op.bias = bias;
    op.relu = relu;
    op.mask = mask;
    op.mask_from = mask_from;
    op.out = out;
    op.pool = pool;
    const int lh = n->lh - l, lw = n->lw - l;
    if (lw >= 6 && lh >= 2) {
        // LDS-staged input patches (8 or 4 rows x 64 pixels per workgroup)
        const long tiles = npix >> 8;
        const int NT = cout <= 32 ? 2 : 4;
        const long gy = (cout + 16 * NT - 1) / (16 * NT);
        if (tiles > 0x7fffffffL || gy > 65535)
            return ufomi_fail(UFOMI_ERR_SIZE, "conv grid too large");
        // 32 output channels and pooled outputs leave through LDS in 16-byte vectors; for 64
        // channels per workgroup the direct 8-byte stores measured faster.  Images of at least
        // 8 rows take the 8-row tiles, 32 pixels wide for 32 output channels (measured: forward
        // 34.2 ms against 34.7 ms with 64-pixel tiles and L2 weight fragments).
        auto launch = [&](auto kern, long nt) {
            hipLaunchKernelGGL(kern, dim3((unsigned) nt, (unsigned) gy), dim3(256), 0, s, op);
        };
        // the 8-row kernel takes one flat grid of pixel tiles x channel tiles (at most `tiles` of
        // them, the 4-row count)
        if (lh >= 3 && tiles * gy > 0x7fffffffL)
            return ufomi_fail(UFOMI_ERR_SIZE, "conv grid too large");
        auto launch8 = [&](auto kern, long nt) {
            hipLaunchKernelGGL(kern, dim3((unsigned) (nt * gy)), dim3(256), 0, s, op);
        };
        if (lh >= 3) {
            if (fin && cout == 32) {
                op.fin_w = fin->w;
                op.fin_b = fin->b;
                op.fin_img = fin->img;
                op.fin_out = fin->out;
                if (!fin->keep_d) op.out.ptr = nullptr;
                *fused = true;
            }
            if (NT == 2) launch8(nn::conv3_t8_kernel<2, true, 2, 32>, tiles);
            else if (pool.ptr) launch8(nn::conv3_t8_kernel<4, true, 2>, tiles / 2);
            else launch8(nn::conv3_t8_kernel<4, false, 2>, tiles / 2);
        } else {
            if (NT == 2) launch(nn::conv3_lds_kernel<2, true>, tiles);
            else if (pool.ptr) launch(nn::conv3_lds_kernel<4, true>, tiles);
            else launch(nn::conv3_lds_kernel<4, false>, tiles);
        }
        UFOMI_CHECK_LAUNCH();
        return UFOMI_OK;
    }
    op.pool = View<bf16>{nullptr, 0, 0};
    int rc = gemm_fwd(s, op, npix, cout);
    if (rc || !pool.ptr) return rc;
    return maxpool(s, View<const bf16>{out.ptr, out.cs, out.c0}, pool, npix / 4, lh, lw, cout);
}

// Split count for a K of `pixels` over an M x N weight gradient: enough
// workgroups to fill the chip, bounded by the workspace.
int wgrad_splits(const ufomi_unet *n, long M, long N, long pixels, long *kchunk)
{
    const long tiles = ((M + 63) / 64) * ((N + 127) / 128);
    long splits = 2048 / tiles;
    if (splits < 1) splits = 1;
    const long max_ws = (long) (n->ws_floats / (size_t) (M * N)) * 32 / 34;  // room for the group sums
    if (splits > max_ws) splits = max_ws;
    const long groups = (pixels + 255) / 256;  // at least 256 pixels per split
    if (splits > groups) splits = groups;
    if (splits < 1) splits = 1;
    long chunk = (pixels + splits - 1) / splits;
    chunk = (chunk + 31) / 32 * 32;
    *kchunk = chunk;
    return (int) ((pixels + chunk - 1) / chunk);
}

// dst[i] = sum_z ws[z * len + i].  Many partials are first summed in groups of
// 32 (one thread per (i, group), in parallel) into the workspace after them, so
// no thread walks a long dependent chain; the order is fixed either way.
int reduce_into(hipStream_t s, const float *ws, int nz, long len, float *dst)
{
    if (nz > 32) {
        const int ng = (nz + 31) / 32;
        float *tmp = const_cast<float *>(ws) + (long) nz * len;
        hipLaunchKernelGGL(reduce_groups_kernel, dim3(blocks_for(len), ng), dim3(256), 0, s, ws, nz, len, tmp);
        UFOMI_CHECK_LAUNCH();
        ws = tmp;
        nz = ng;
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(blocks_for(len)), dim3(256), 0, s, ws, nz, len, dst);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

int conv3_wgrad(ufomi_unet *n, hipStream_t s, View<const bf16> g, View<const bf16> in, long npix, int l, int cin,
                int cout, float *dw)
{
    const long M = cout, N = 9L * cin;
    long chunk;
    const int splits = wgrad_splits(n, M, N, npix, &chunk);
    const int lh = n->lh - l, lw = n->lw - l;
    const size_t need = (size_t) (cout + 3 * cin) * (size_t) npix;
    if (lw >= 3 && n->wgT && need <= n->wgT_elems && npix % kTpPix == 0) {
        // channel-major copies (g, and the input at the three row shifts), then 16-byte fragments
        bf16 *gT = n->wgT, *inT = n->wgT + (size_t) cout * npix;
        const unsigned blocks = (unsigned) (npix / kTpPix);
        auto lds = [](int C) { return (size_t) (kTpPix + 2) * (C % 8 == 0 ? C + 8 : C + 2) * sizeof(bf16); };
        hipLaunchKernelGGL(transpose_shift_kernel, dim3(blocks), dim3(256), lds(cout), s, g, cout, npix, lw, 1, gT);
        hipLaunchKernelGGL(transpose_shift_kernel, dim3(blocks), dim3(256), lds(cin), s, in, cin, npix, lw, 3, inT);
        UFOMI_CHECK_LAUNCH();
        nn::Conv3WgradT op;
        op.gT = gT;
        op.inT = inT;
        op.lh = lh;
        op.lw = lw;
        op.cin = cin;
        op.cout = cout;
        op.npix = npix;
        op.part = nn::Partial{n->ws, M, N, chunk, npix};
        int rc = gemm_wgrad(s, op, M, N, splits);
        if (rc) return rc;
        return reduce_into(s, n->ws, splits, M * N, dw);
    }
    nn::Conv3Wgrad op;
    op.g = g;
    op.in = in;
    op.lh = lh;
    op.lw = lw;
    op.cin = cin;
    op.cout = cout;
    op.npix = npix;
    op.part = nn::Partial{n->ws, M, N, chunk, npix};
    int rc = gemm_wgrad(s, op, M, N, splits);
    if (rc) return rc;
    return reduce_into(s, n->ws, splits, M * N, dw);
}

int bias_grad(ufomi_unet *n, hipStream_t s, View<const bf16> g, long npix, int C, float *db)
{
    long blocks = (npix * C + 256L * 64 - 1) / (256L * 64);
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    if ((size_t) (blocks * C) > n->ws_floats) blocks = (long) (n->ws_floats / C);
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned) blocks), dim3(256), 0, s, g, npix, C, n->ws);
    UFOMI_CHECK_LAUNCH();
    return reduce_into(s, n->ws, (int) blocks, C, db);
}

int maxpool(hipStream_t s, View<const bf16> in, View<bf16> out, long nout, int lh, int lw, int C)
{
    hipLaunchKernelGGL(maxpool_kernel, dim3(blocks_for(nout * (C / 8))), dim3(256), 0, s, in, out, nout, lh, lw, C);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

int refresh_transforms(ufomi_unet *n, hipStream_t s)
{
    auto flip = [&](const ConvP &c) {
        hipLaunchKernelGGL(conv_flip_kernel, dim3(blocks_for(9L * c.cin * c.cout)), dim3(256), 0, s, n->w16 + c.w,
                           n->wtrans + c.wt, c.cout, c.cin);
    };
    for (int l = 1; l < kLevels; l++) flip(n->enc1[l]);
    for (int l = 0; l < kLevels; l++) flip(n->enc2[l]);
    for (int l = 0; l < kLevels - 1; l++) {
        flip(n->dec1[l]);
        flip(n->dec2[l]);
        const ConvP &u = n->up[l];
        hipLaunchKernelGGL(up_transpose_kernel, dim3(blocks_for(4L * u.cin * u.cout)), dim3(256), 0, s, n->w16 + u.w,
                           n->wtrans + u.wt, u.cout, u.cin);
    }
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

int forward(ufomi_unet *n, hipStream_t s, const float *img, int B, float *out)
{
    int rc;
    bool fused = false;
    const int C0 = kBase;
    {
        const long np = npix_at(n, B, 0);
        hipLaunchKernelGGL(first_conv_kernel, dim3(blocks_for(np)), dim3(256), 0, s, img, np, n->lh, n->lw,
                           n->w16 + n->first.w, n->master + n->first.b, n->img16, n->e1[0]);
        UFOMI_CHECK_LAUNCH();
    }
    for (int l = 0; l < kLevels; l++) {
        const int C = C0 << l;
        const long np = npix_at(n, B, l);
        if (l > 0) {
            const ConvP &c1 = n->enc1[l];
            if ((rc = conv3(n, s, cv(n->pool[l - 1], C / 2), np, l, C / 2, C, n->w16 + c1.w, n->master + c1.b, true,
                            cv(nullptr, 0), 0, mv(n->e1[l], C))))
                return rc;
        }
        const ConvP &c2 = n->enc2[l];
        View<bf16> dst = l < kLevels - 1 ? mv(n->cat[l], 2 * C, C) : mv(n->bottom, C);
        View<bf16> pooled = l < kLevels - 1 ? mv(n->pool[l], C) : View<bf16>{nullptr, 0, 0};
        if ((rc = conv3(n, s, cv(n->e1[l], C), np, l, C, C, n->w16 + c2.w, n->master + c2.b, true, cv(nullptr, 0), 0,
                        dst, pooled)))
            return rc;
    }
    for (int l = kLevels - 2; l >= 0; l--) {
        const int C = C0 << l;
        const long np = npix_at(n, B, l);
        const bf16 *src = l == kLevels - 2 ? n->bottom : n->d2[l + 1];
        nn::UpFwd up;
        up.in = cv(src, 2 * C);
        up.lh = n->lh - l - 1;
        up.lw = n->lw - l - 1;
        up.lcin = ilog2(2 * C);
        up.lcout = ilog2(C);
        up.npix = npix_at(n, B, l + 1);
        up.w = n->w16 + n->up[l].w;
        up.bias = n->master + n->up[l].b;
        up.out = mv(n->cat[l], 2 * C, 0);
        if ((8L << (up.lh + up.lw)) * up.out.cs >= (1L << 32))
            return ufomi_fail(UFOMI_ERR_SIZE, "transposed conv output of %ld pixels x %d channels exceeds 32-bit offsets",
                              4L << (up.lh + up.lw), up.out.cs);
        if ((rc = gemm_fwd(s, up, up.npix, 4L * C))) return rc;
        const ConvP &a = n->dec1[l], &b = n->dec2[l];
        if ((rc = conv3(n, s, cv(n->cat[l], 2 * C), np, l, 2 * C, C, n->w16 + a.w, n->master + a.b, true,
                        cv(nullptr, 0), 0, mv(n->d1[l], C))))
            return rc;
        // level 0: the output conv rides in this conv's epilogue where the 8-row kernel runs
        // (training keeps dec_0 for the backward pass)
        const FinalConv fc{n->w16 + n->fin_w, n->master + n->fin_b, img, out, n->training};
        if ((rc = conv3(n, s, cv(n->d1[l], C), np, l, C, C, n->w16 + b.w, n->master + b.b, true, cv(nullptr, 0), 0,
                        mv(n->d2[l], C), View<bf16>{nullptr, 0, 0}, l == 0 ? &fc : nullptr, &fused)))
            return rc;
    }
    if (!fused) {
        const long np = npix_at(n, B, 0);
        hipLaunchKernelGGL(final_kernel, dim3(blocks_for(np)), dim3(256), 0, s, n->d2[0], np, n->w16 + n->fin_w,
                           n->master + n->fin_b, img, out);
        UFOMI_CHECK_LAUNCH();
    }
    return UFOMI_OK;
}

// Loss and all parameter gradients into n->grad; loss (mean squared error) to
// *loss when not NULL (synchronises).
int backward(ufomi_unet *n, hipStream_t s, const float *img, const float *tgt, int B, float *loss)
{
    int rc;
    const int C0 = kBase;
    const long np0 = npix_at(n, B, 0);
    float *g = n->grad;
    {
        long blocks = (np0 + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(mse_kernel, dim3((unsigned) blocks), dim3(256), 0, s, n->out, tgt, np0, n->gout, n->ws);
        UFOMI_CHECK_LAUNCH();
        if (loss) {
            std::vector<float> part(blocks);
            UFOMI_HIP(hipMemcpyAsync(part.data(), n->ws, blocks * sizeof(float), hipMemcpyDeviceToHost, s));
            UFOMI_HIP(hipStreamSynchronize(s));
            double sum = 0.0;
            for (float v : part) sum += v;
            *loss = (float) (sum / (double) np0);
        }
    }
    {
        // final 1x1 layer: gradient of dec_0 (pre-activation) and dw, db
        // (stored contiguously: fin_b = fin_w + 32)
        long blocks = (np0 + 255) / 256;
        if (blocks > 1024) blocks = 1024;
        hipLaunchKernelGGL(final_bwd_kernel, dim3((unsigned) blocks), dim3(256), 0, s, n->gout, n->d2[0], np0,
                           n->w16 + n->fin_w, n->g_d2[0], n->ws);
        UFOMI_CHECK_LAUNCH();
        if ((rc = reduce_into(s, n->ws, (int) blocks, kBase + 1, g + n->fin_w))) return rc;
    }
    // decoder, level 0 upwards
    for (int l = 0; l < kLevels - 1; l++) {
        const int C = C0 << l;
        const long np = npix_at(n, B, l);
        const ConvP &a = n->dec1[l], &b = n->dec2[l], &u = n->up[l];
        // d2 = relu(conv(d1))
        if ((rc = conv3_wgrad(n, s, cv(n->g_d2[l], C), cv(n->d1[l], C), np, l, C, C, g + b.w)) ||
            (rc = bias_grad(n, s, cv(n->g_d2[l], C), np, C, g + b.b)) ||
            (rc = conv3(n, s, cv(n->g_d2[l], C), np, l, C, C, n->wtrans + b.wt, nullptr, false, cv(n->d1[l], C), 0,
                        mv(n->g_d1[l], C))))
            return rc;
        // d1 = relu(conv(cat)), cat = [up | skip]: plain data gradient of the concatenation
        if ((rc = conv3_wgrad(n, s, cv(n->g_d1[l], C), cv(n->cat[l], 2 * C), np, l, 2 * C, C, g + a.w)) ||
            (rc = bias_grad(n, s, cv(n->g_d1[l], C), np, C, g + a.b)) ||
            (rc = conv3(n, s, cv(n->g_d1[l], C), np, l, C, 2 * C, n->wtrans + a.wt, nullptr, false, cv(nullptr, 0),
                        0, mv(n->g_cat[l], 2 * C))))
            return rc;
        // up = upconv(src): src = bottom (l = 2) or d2[l + 1]
        const bf16 *src = l == kLevels - 2 ? n->bottom : n->d2[l + 1];
        bf16 *gsrc = l == kLevels - 2 ? n->g_e2[kLevels - 1] : n->g_d2[l + 1];
        const long nps = npix_at(n, B, l + 1);
        nn::UpWgrad uw;
        uw.g = cv(n->g_cat[l], 2 * C, 0);
        uw.in = cv(src, 2 * C);
        uw.lh = n->lh - l - 1;
        uw.lw = n->lw - l - 1;
        uw.lcout = ilog2(C);
        uw.cin = 2 * C;
        uw.npix = nps;
        {
            const long M = 4L * C, N = 2L * C;
            long chunk;
            const int splits = wgrad_splits(n, M, N, nps, &chunk);
            uw.part = nn::Partial{n->ws, M, N, chunk, nps};
            if ((rc = gemm_wgrad(s, uw, M, N, splits)) || (rc = reduce_into(s, n->ws, splits, M * N, g + u.w)))
                return rc;
        }
        if ((rc = bias_grad(n, s, cv(n->g_cat[l], 2 * C, 0), np, C, g + u.b))) return rc;
        nn::UpDgrad ud;
        ud.g = cv(n->g_cat[l], 2 * C, 0);
        ud.lh = n->lh - l - 1;
        ud.lw = n->lw - l - 1;
        ud.lcout = ilog2(C);
        ud.cin = 2 * C;
        ud.npix = nps;
        ud.wt = n->wtrans + u.wt;
        ud.mask = cv(src, 2 * C);
        ud.out = mv(gsrc, 2 * C);
        if ((rc = gemm_fwd(s, ud, nps, 2L * C))) return rc;
    }
    // encoder, bottom level down to 0
    for (int l = kLevels - 1; l >= 0; l--) {
        const int C = C0 << l;
        const long np = npix_at(n, B, l);
        if (l < kLevels - 1) {
            // gradient of skip_l: pooled path + concatenation, masked by its ReLU
            hipLaunchKernelGGL(maxpool_bwd_kernel, dim3(blocks_for(npix_at(n, B, l + 1) * (C / 8))), dim3(256), 0,
                               s, cv(n->cat[l], 2 * C, C), cv(n->g_pool[l], C), cv(n->g_cat[l], 2 * C, C),
                               mv(n->g_e2[l], C), npix_at(n, B, l + 1), n->lh - l, n->lw - l, C);
            UFOMI_CHECK_LAUNCH();
        }
        const ConvP &c2 = n->enc2[l], &c1 = n->enc1[l];
        // skip_l (or bottom) = relu(conv(e1))
        if ((rc = conv3_wgrad(n, s, cv(n->g_e2[l], C), cv(n->e1[l], C), np, l, C, C, g + c2.w)) ||
            (rc = bias_grad(n, s, cv(n->g_e2[l], C), np, C, g + c2.b)) ||
            (rc = conv3(n, s, cv(n->g_e2[l], C), np, l, C, C, n->wtrans + c2.wt, nullptr, false, cv(n->e1[l], C), 0,
                        mv(n->g_e1[l], C))))
            return rc;
        if (l > 0) {
            // e1 = relu(conv(pool_{l-1})); pool values are maxima of ReLU outputs, the
            // gradient goes on to the pooling (no mask here)
            if ((rc = conv3_wgrad(n, s, cv(n->g_e1[l], C), cv(n->pool[l - 1], C / 2), np, l, C / 2, C, g + c1.w)) ||
                (rc = bias_grad(n, s, cv(n->g_e1[l], C), np, C, g + c1.b)) ||
                (rc = conv3(n, s, cv(n->g_e1[l], C), np, l, C, C / 2, n->wtrans + c1.wt, nullptr, false,
                            cv(nullptr, 0), 0, mv(n->g_pool[l - 1], C / 2))))
                return rc;
        } else {
            const ConvP &f = n->first;
            if ((rc = conv3_wgrad(n, s, cv(n->g_e1[0], C), cv(n->img16, 1), np, 0, 1, C, g + f.w)) ||
                (rc = bias_grad(n, s, cv(n->g_e1[0], C), np, C, g + f.b)))
                return rc;
        }
    }
    return UFOMI_OK;
}

int check_net(ufomi_ctx *ctx, ufomi_unet *n, int B)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!n) return ufomi_fail(UFOMI_ERR_ARG, "U-Net is NULL");
    if (n->ctx != ctx) return ufomi_fail(UFOMI_ERR_ARG, "U-Net belongs to another context");
    if (B < 1 || B > n->max_batch)
        return ufomi_fail(UFOMI_ERR_SIZE, "batch %d outside 1..%d", B, n->max_batch);
    return UFOMI_OK;
}

}  // namespace

extern "C" int ufomi_unet_create(ufomi_ctx *ctx, int max_batch, int height, int width, int training, uint64_t seed,
                                 ufomi_unet **out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!out) return ufomi_fail(UFOMI_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (max_batch < 1 || !ufomi_is_pow2(height) || !ufomi_is_pow2(width) || height < 16 || width < 16 ||
        height > 8192 || width > 8192)
        return ufomi_fail(UFOMI_ERR_SIZE, "U-Net images must be powers of two in 16..8192 (got %d x %d), batch >= 1",
                          height, width);
    if (training && width < 64)
        return ufomi_fail(UFOMI_ERR_SIZE, "training needs width >= 64 (rows of 8 pixels at the bottom level), got %d",
                          width);
    ufomi_unet *n = new ufomi_unet;
    n->ctx = ctx;
    n->max_batch = max_batch;
    n->H = height;
    n->W = width;
    n->lh = ilog2(height);
    n->lw = ilog2(width);
    n->training = training != 0;

    // parameter layout, in forward order
    auto add = [&](const char *name, int d0, int d1, int d2, int d3) {
        TensorInfo t;
        t.name = name;
        t.off = n->nparams;
        t.dims[0] = d0;
        t.dims[1] = d1;
        t.dims[2] = d2;
        t.dims[3] = d3;
        t.count = (size_t) d0 * d1 * d2 * d3;
        n->nparams += (t.count + 7) / 8 * 8;  // keep every tensor 16-byte aligned in bf16
        n->tensors.push_back(t);
        return t.off;
    };
    auto conv = [&](ConvP &c, const std::string &name, int cin, int cout) {
        c.cin = cin;
        c.cout = cout;
        c.w = add((name + ".w").c_str(), cout, 3, 3, cin);
        c.b = add((name + ".b").c_str(), cout, 1, 1, 1);
        c.wt = n->ntrans;
        n->ntrans += 9L * cin * cout;
    };
    conv(n->first, "enc0.conv1", 1, kBase);
    n->enc1[0] = n->first;
    for (int l = 0; l < kLevels; l++) {
        const int C = kBase << l;
        if (l > 0) conv(n->enc1[l], "enc" + std::to_string(l) + ".conv1", C / 2, C);
        conv(n->enc2[l], "enc" + std::to_string(l) + ".conv2", C, C);
    }
    for (int l = kLevels - 2; l >= 0; l--) {
        const int C = kBase << l;
        ConvP &u = n->up[l];
        u.cin = 2 * C;
        u.cout = C;
        u.w = add(("up" + std::to_string(l) + ".w").c_str(), 4, C, 1, 2 * C);
        u.b = add(("up" + std::to_string(l) + ".b").c_str(), C, 1, 1, 1);
        u.wt = n->ntrans;
        n->ntrans += 8L * C * C;
        conv(n->dec1[l], "dec" + std::to_string(l) + ".conv1", 2 * C, C);
        conv(n->dec2[l], "dec" + std::to_string(l) + ".conv2", C, C);
    }
    {
        // final 1x1: w and b contiguous (the backward reduces both in one pass)
        TensorInfo t;
        t.name = "final.w";
        t.off = n->nparams;
        t.count = kBase;
        t.dims[0] = 1; t.dims[1] = 1; t.dims[2] = 1; t.dims[3] = kBase;
        n->tensors.push_back(t);
        t.name = "final.b";
        t.off = n->nparams + kBase;
        t.count = 1;
        t.dims[3] = 1;
        n->tensors.push_back(t);
        n->fin_w = n->nparams;
        n->fin_b = n->nparams + kBase;
        n->nparams += kBase + 8;
    }

    // He-normal weights, zero biases, zero final layer (the untrained net is the identity)
    std::vector<float> h(n->nparams, 0.0f);
    HostRng rng(seed);
    for (const TensorInfo &t : n->tensors) {
        if (t.name.size() < 2 || t.name.compare(t.name.size() - 2, 2, ".w") != 0 || t.name == "final.w") continue;
        const bool upw = t.name.compare(0, 2, "up") == 0;
        const double fan_in = upw ? (double) t.dims[3] : 9.0 * t.dims[3];
        const double sd = sqrt(2.0 / fan_in);
        for (size_t i = 0; i < t.count; i++) h[t.off + i] = (float) (sd * rng.normal());
    }

    const long B = max_batch;
    auto act = [&](bf16 **p, int l, int C) { return dalloc_t(n, (size_t) npix_at(n, (int) B, l) * C, p); };
    if ((rc = dalloc_t(n, n->nparams, &n->master)) || (rc = dalloc_t(n, n->nparams, &n->w16)) ||
        (rc = dalloc_t(n, n->ntrans, &n->wtrans)) || (rc = act(&n->img16, 0, 1)) ||
        (rc = dalloc_t(n, (size_t) npix_at(n, (int) B, 0), &n->out))) {
        free_net(n);
        return rc;
    }
    for (int l = 0; l < kLevels && !rc; l++) {
        const int C = kBase << l;
        rc = act(&n->e1[l], l, C);
        if (!rc && l < kLevels - 1) rc = act(&n->cat[l], l, 2 * C);
        if (!rc && l < kLevels - 1) rc = act(&n->pool[l], l + 1, C);
        if (!rc && l < kLevels - 1) rc = act(&n->d1[l], l, C);
        if (!rc && l < kLevels - 1) rc = act(&n->d2[l], l, C);
    }
    if (!rc) rc = act(&n->bottom, kLevels - 1, kBase << (kLevels - 1));
    if (!rc && n->training) {
        rc = dalloc_t(n, n->nparams, &n->grad);
        if (!rc) rc = dalloc_t(n, n->nparams, &n->adam_m);
        if (!rc) rc = dalloc_t(n, n->nparams, &n->adam_v);
        if (!rc) rc = dalloc_t(n, (size_t) npix_at(n, (int) B, 0), &n->gout);
        for (int l = 0; l < kLevels && !rc; l++) {
            const int C = kBase << l;
            rc = act(&n->g_e1[l], l, C);
            if (!rc) rc = act(&n->g_e2[l], l, C);
            if (!rc && l < kLevels - 1) rc = act(&n->g_cat[l], l, 2 * C);
            if (!rc && l < kLevels - 1) rc = act(&n->g_pool[l], l + 1, C);
            if (!rc && l < kLevels - 1) rc = act(&n->g_d1[l], l, C);
            if (!rc && l < kLevels - 1) rc = act(&n->g_d2[l], l, C);
        }
        n->ws_floats = 32u << 20;  // 128 MiB of split-K partial sums
        if (!rc) rc = dalloc_t(n, n->ws_floats, &n->ws);
        // channel-major operand copies of the largest weight gradient: (cout + 3 cin) x pixels
        size_t wg = 0;
        auto need = [&](const ConvP &c, int l) {
            wg = std::max(wg, (size_t) (c.cout + 3 * c.cin) * (size_t) npix_at(n, max_batch, l));
        };
        need(n->first, 0);
        for (int l = 0; l < kLevels; l++) {
            if (l > 0) need(n->enc1[l], l);
            need(n->enc2[l], l);
        }
        for (int l = 0; l < kLevels - 1; l++) {
            need(n->dec1[l], l);
            need(n->dec2[l], l);
        }
        n->wgT_elems = wg;
        if (!rc) rc = dalloc_t(n, wg, &n->wgT);
    }
    if (rc) {
        free_net(n);
        return rc;
    }
    hipError_t e = hipMemcpy(n->master, h.data(), n->nparams * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && n->training) {
        e = hipMemset(n->adam_m, 0, n->nparams * sizeof(float));
        if (e == hipSuccess) e = hipMemset(n->adam_v, 0, n->nparams * sizeof(float));
        if (e == hipSuccess) e = hipMemset(n->grad, 0, n->nparams * sizeof(float));
    }
    if (e != hipSuccess) {
        free_net(n);
        return ufomi_fail(UFOMI_ERR_HIP, "U-Net initialisation failed: %s", hipGetErrorString(e));
    }
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for((long) n->nparams)), dim3(256), 0, ctx->stream, n->master,
                       n->w16, (long) n->nparams);
    if ((rc = refresh_transforms(n, ctx->stream)) || (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
        free_net(n);
        return rc ? rc : ufomi_fail(UFOMI_ERR_HIP, "U-Net initialisation failed: %s", hipGetErrorString(e));
    }
    *out = n;
    return UFOMI_OK;
}

extern "C" int ufomi_unet_destroy(ufomi_unet *n)
{
    if (!n) return UFOMI_OK;
    (void) hipSetDevice(n->ctx->device);
    (void) hipDeviceSynchronize();
    free_net(n);
    return UFOMI_OK;
}

extern "C" int ufomi_unet_num_params(const ufomi_unet *n, size_t *count, size_t *padded)
{
    if (!n) return ufomi_fail(UFOMI_ERR_ARG, "U-Net is NULL");
    size_t c = 0;
    for (const TensorInfo &t : n->tensors) c += t.count;
    if (count) *count = c;
    if (padded) *padded = n->nparams;
    return UFOMI_OK;
}

extern "C" int ufomi_unet_num_tensors(const ufomi_unet *n) { return n ? (int) n->tensors.size() : 0; }

extern "C" int ufomi_unet_tensor(const ufomi_unet *n, int i, const char **name, size_t *offset, int *dims)
{
    if (!n || i < 0 || i >= (int) n->tensors.size()) return ufomi_fail(UFOMI_ERR_ARG, "no tensor %d", i);
    const TensorInfo &t = n->tensors[i];
    if (name) *name = t.name.c_str();
    if (offset) *offset = t.off;
    if (dims) memcpy(dims, t.dims, sizeof(t.dims));
    return UFOMI_OK;
}

extern "C" int ufomi_unet_get_params(ufomi_ctx *ctx, ufomi_unet *n, float *host)
{
    int rc = check_net(ctx, n, 1);
    if (rc) return rc;
    UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    UFOMI_HIP(hipMemcpy(host, n->master, n->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return UFOMI_OK;
}

extern "C" int ufomi_unet_set_params(ufomi_ctx *ctx, ufomi_unet *n, const float *host)
{
    int rc = check_net(ctx, n, 1);
    if (rc) return rc;
    UFOMI_HIP(hipStreamSynchronize(ctx->stream));
    UFOMI_HIP(hipMemcpy(n->master, host, n->nparams * sizeof(float), hipMemcpyHostToDevice));
    if (n->training) {
        UFOMI_HIP(hipMemset(n->adam_m, 0, n->nparams * sizeof(float)));
        UFOMI_HIP(hipMemset(n->adam_v, 0, n->nparams * sizeof(float)));
        n->step = 0;
    }
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for((long) n->nparams)), dim3(256), 0, ctx->stream, n->master,
                       n->w16, (long) n->nparams);
    UFOMI_CHECK_LAUNCH();
    return refresh_transforms(n, ctx->stream);
}

extern "C" int ufomi_unet_forward(ufomi_ctx *ctx, ufomi_unet *n, const float *d_in, int batch, float *d_out)
{
    int rc = check_net(ctx, n, batch);
    if (rc) return rc;
    if (!d_in || !d_out) return ufomi_fail(UFOMI_ERR_ARG, "input or output is NULL");
    return forward(n, ctx->stream, d_in, batch, d_out);
}

extern "C" int ufomi_unet_gradients(ufomi_ctx *ctx, ufomi_unet *n, const float *d_in, const float *d_target,
                                    int batch, float *loss, float *host_grads)
{
    int rc = check_net(ctx, n, batch);
    if (rc) return rc;
    if (!n->training) return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "U-Net was created without training state");
    if (!d_in || !d_target) return ufomi_fail(UFOMI_ERR_ARG, "input or target is NULL");
    if ((rc = forward(n, ctx->stream, d_in, batch, n->out)) ||
        (rc = backward(n, ctx->stream, d_in, d_target, batch, loss)))
        return rc;
    if (host_grads) {
        UFOMI_HIP(hipStreamSynchronize(ctx->stream));
        UFOMI_HIP(hipMemcpy(host_grads, n->grad, n->nparams * sizeof(float), hipMemcpyDeviceToHost));
    }
    return UFOMI_OK;
}

extern "C" int ufomi_unet_train_step(ufomi_ctx *ctx, ufomi_unet *n, const float *d_in, const float *d_target,
                                     int batch, float lr, float *loss)
{
    int rc = ufomi_unet_gradients(ctx, n, d_in, d_target, batch, loss, nullptr);
    if (rc) return rc;
    n->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = (float) (1.0 - pow(0.9, (double) n->step)), c2 = (float) (1.0 - pow(0.999, (double) n->step));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for((long) n->nparams)), dim3(256), 0, ctx->stream, n->master, n->w16,
                       n->grad, n->adam_m, n->adam_v, (long) n->nparams, lr, b1, b2, 1e-8f, c1, c2);
    UFOMI_CHECK_LAUNCH();
    return refresh_transforms(n, ctx->stream);
}

extern "C" int ufomi_nn_conv3x3(ufomi_ctx *ctx, const void *d_in, int batch, int height, int width, int cin,
                                const void *d_w, const float *d_bias, int relu, int cout, void *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_in || !d_w || !d_out || batch < 1 || cout < 1)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid conv3x3 arguments");
    if (!ufomi_is_pow2(height) || !ufomi_is_pow2(width) || !ufomi_is_pow2(cin) || cin < 32 || cout % 8)
        return ufomi_fail(UFOMI_ERR_SIZE, "conv3x3 needs power-of-two height, width and cin >= 32, cout % 8 == 0");
    ufomi_unet shape;  // only lh / lw are read by conv3
    shape.lh = ilog2(height);
    shape.lw = ilog2(width);
    const long np = (long) batch * height * width;
    return conv3(&shape, ctx->stream, cv((const bf16 *) d_in, cin), np, 0, cin, cout, (const bf16 *) d_w, d_bias,
                 relu != 0, cv(nullptr, 0), 0, mv((bf16 *) d_out, cout));
}
