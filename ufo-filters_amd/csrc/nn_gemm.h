// BF16 MFMA implicit-GEMM core of the learned sparse-view stage (unet.hip).
//
// One kernel template computes C[M][N] = sum_k A[m][k] B[k][n] with
// v_mfma_f32_16x16x32_bf16 (fp32 accumulation).  The operands are never
// materialised: an operation type `Op` supplies
//   Op::RowA row_a(m)            per-row context (pixel coordinates, ...)
//   Op::ColB col_b(n)            per-column context
//   bf16x8 load_a(RowA, k)       A[m][k .. k+7]
//   bf16x8 load_b(ColB, k)       B[k .. k+7][n]
//   void k_range(z, &k0, &k1)    the K slice of split index z (grid z)
//   void store(m, n, float, z)   epilogue for C[m][n]
// Convolutions (NHWC activations), their data gradients, the 2x2 transposed
// convolution and the weight gradients are all such operations.
//
// Lane layout of v_mfma_f32_16x16x32_bf16 (wave of 64, r = lane % 16,
// q = lane / 16): lane supplies A[r][8q .. 8q+7] and B[8q .. 8q+7][r] and
// receives D[4q + e][r], e < 4.  The order of the 8 k values inside a lane
// only has to agree between A and B.
#pragma once

#include <hip/hip_runtime.h>

namespace nn {

using bf16 = __bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; i++) z[i] = (bf16) 0.0f;
    return z;
}

__device__ __forceinline__ bf16x8 ld8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// Workgroup of 256 threads = WM x WN waves; each wave owns MT x NT tiles of
// 16 x 16.  Fragments of step k+32 are loaded before the MFMAs of step k.
template <int MT, int NT, int WM, int WN, typename Op>
__global__ __launch_bounds__(256) void gemm_kernel(Op op)
{
    static_assert(WM * WN == 4, "four waves per workgroup");
    constexpr int BM = WM * MT * 16, BN = WN * NT * 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const long m0 = (long) blockIdx.x * BM + (wave % WM) * MT * 16;
    const long n0 = (long) blockIdx.y * BN + (wave / WM) * NT * 16;
    long kb, ke;
    op.k_range(blockIdx.z, kb, ke);

    typename Op::RowA ra[MT];
    typename Op::ColB cb[NT];
#pragma unroll
    for (int i = 0; i < MT; i++) ra[i] = op.row_a(m0 + i * 16 + r);
#pragma unroll
    for (int j = 0; j < NT; j++) cb[j] = op.col_b(n0 + j * 16 + r);

    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    if (kb < ke) {
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int i = 0; i < MT; i++) a[i] = op.load_a(ra[i], kb + 8 * q);
#pragma unroll
        for (int j = 0; j < NT; j++) b[j] = op.load_b(cb[j], kb + 8 * q);
        for (long k = kb; k < ke; k += 32) {
            const bool more = k + 32 < ke;
            bf16x8 an[MT], bn[NT];
            if (more) {
#pragma unroll
                for (int i = 0; i < MT; i++) an[i] = op.load_a(ra[i], k + 32 + 8 * q);
#pragma unroll
                for (int j = 0; j < NT; j++) bn[j] = op.load_b(cb[j], k + 32 + 8 * q);
            }
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++) {
                    // kVecStore: B as the row operand, so a lane receives 4 consecutive columns of one row
                    if constexpr (Op::kVecStore)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
                }
            if (more) {
#pragma unroll
                for (int i = 0; i < MT; i++) a[i] = an[i];
#pragma unroll
                for (int j = 0; j < NT; j++) b[j] = bn[j];
            }
        }
    }
    if constexpr (Op::kVecStore) {
        // finished bf16 values through LDS (D[4q + e][r] of tile (i, j) is column n0 + 16 j + 4 q + e
        // of row m0 + 16 i + r: one 8-byte write per tile), then 16-byte stores of 8 consecutive
        // columns; a tile wholly inside M x N skips the per-element bounds tests, and the stores
        // address from a per-tile base with 32-bit offsets
        __shared__ bf16x8 tile[BM * BN / 8];
        bf16 *t = reinterpret_cast<bf16 *>(tile);
        const int mw = (wave % WM) * MT * 16, nw = (wave / WM) * NT * 16;
        const long mb = (long) blockIdx.x * BM, nb = (long) blockIdx.y * BN;
        const bool full = op.inside(mb + BM - 1, nb + BN - 1);
#pragma unroll
        for (int j = 0; j < NT; j++) {
            // the 4 columns' bias, loaded once and unconditionally (loads under the bounds
            // test would each wait on their own)
            const long n = n0 + j * 16 + 4 * q;
            const f32x4 bj = op.bias4(n);
#pragma unroll
            for (int i = 0; i < MT; i++) {
                const long m = m0 + i * 16 + r;
                bf16x4 o;
#pragma unroll
                for (int e = 0; e < 4; e++)
                    o[e] = (bf16) (full || op.valid(m, n + e) ? acc[i][j][e] + bj[e] : 0.0f);
                *reinterpret_cast<bf16x4 *>(t + (mw + i * 16 + r) * BN + nw + j * 16 + 4 * q) = o;
            }
        }
        __syncthreads();
        const typename Op::TileBase tb = op.tile_base(mb);
        for (int v = threadIdx.x; v < BM * BN / 8; v += 256) {
            const int lm = v / (BN / 8);
            const long n = nb + (v % (BN / 8)) * 8;
            if (full || op.valid(mb + lm, n)) op.store8(tb, lm, (int) n, tile[v]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int j = 0; j < NT; j++)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    op.store(m0 + i * 16 + 4 * q + e, n0 + j * 16 + r, acc[i][j][e], blockIdx.z);
    }
}

// A tensor of pixels in NHWC order: element (pixel p, channel c) at
// ptr[p * cs + c0 + c].
template <typename T>
struct View {
    T *ptr;
    int cs, c0;
    __device__ __forceinline__ T *at(long p, int c) const { return ptr + p * cs + c0 + c; }
};

// Pixel p of a [n][H][W] grid with power-of-two H and W.
struct Pix {
    int y, x;
};
__device__ __forceinline__ Pix pix_of(long p, int lh, int lw)
{
    Pix r;
    r.x = (int) (p & ((1 << lw) - 1));
    r.y = (int) ((p >> lw) & ((1 << lh) - 1));
    return r;
}

// ---------------------------------------------------------------- 3x3 conv
// out[p][n] = act(sum_{tap, c} in[p + d(tap)][c] w[n][tap][c] + bias[n]),
// zero padding, K ordered (tap, c); cin a multiple of 32.  With mask, channels
// n >= mask_from are zeroed where mask[p][n] <= 0 (a ReLU derivative).
struct Conv3 {
    static constexpr bool kVecStore = false;
    View<const bf16> in;
    int lh, lw, lcin, cout;
    long npix;
    const bf16 *w;        // [cout][9][cin]
    const float *bias;    // nullptr: none
    int relu;
    View<const bf16> mask;  // mask.ptr nullptr: none
    int mask_from;
    View<bf16> out;
    View<bf16> pool;        // pool.ptr not nullptr: also the 2x2 max pool of out (LDS kernel only)
    // fin_out not nullptr (8-row kernel with LDS_OUT, cout == 32): also the 1x1 output conv
    // fin_out[p] = fin_img[p] + (sum_c fin_w[c] out[p][c] + fin_b[0]) from the finished tile;
    // out.ptr may then be nullptr (the tile itself not stored)
    const bf16 *fin_w;
    const float *fin_b, *fin_img;
    float *fin_out;

    struct RowA {
        long p;
        int y, x;
        bool ok;
    };
    struct ColB {
        const bf16 *w;
        bool ok;
    };
    __device__ __forceinline__ RowA row_a(long m) const
    {
        RowA r;
        r.p = m;
        r.ok = m < npix;
        const Pix px = pix_of(m, lh, lw);
        r.y = px.y;
        r.x = px.x;
        return r;
    }
    __device__ __forceinline__ ColB col_b(long n) const
    {
        ColB c;
        c.ok = n < cout;
        c.w = w + (c.ok ? n : 0) * (9L << lcin);
        return c;
    }
    __device__ __forceinline__ void k_range(int, long &k0, long &k1) const
    {
        k0 = 0;
        k1 = 9L << lcin;
    }
    __device__ __forceinline__ bf16x8 load_a(const RowA &r, long k) const
    {
        const int tap = (int) (k >> lcin), c = (int) (k & ((1 << lcin) - 1));
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const int yy = r.y + dy, xx = r.x + dx;
        if (!r.ok || yy < 0 || yy >= (1 << lh) || xx < 0 || xx >= (1 << lw)) return zero8();
        return ld8(in.at(r.p + (long) dy * (1L << lw) + dx, c));
    }
    __device__ __forceinline__ bf16x8 load_b(const ColB &c, long k) const
    {
        return c.ok ? ld8(c.w + k) : zero8();
    }
    // bias, ReLU and mask of output (m, n) (m < npix, n < cout)
    __device__ __forceinline__ float finish(long m, long n, float v) const
    {
        if (bias) v += bias[n];
        if (relu) v = fmaxf(v, 0.0f);
        if (mask.ptr && n >= mask_from && !((float) *mask.at(m, (int) n) > 0.0f)) v = 0.0f;
        return v;
    }
    __device__ __forceinline__ void store(long m, long n, float v, int) const
    {
        if (m >= npix || n >= cout) return;
        *out.at(m, (int) n) = (bf16) finish(m, n, v);
    }
    // mask values at (m, n .. n + 3) (m < npix, n a multiple of 4, n + 3 < cout), read
    // unconditionally so that a tile's loads issue together
    __device__ __forceinline__ bf16x4 mask4(long m, long n) const
    {
        return *reinterpret_cast<const bf16x4 *>(mask.at(m, (int) n));
    }
    // bias of output channels n .. n + 3 (n a multiple of 4, n + 3 < cout; zeros without bias)
    __device__ __forceinline__ f32x4 bias4(long n) const
    {
        if (!bias) return f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        return f32x4{bias[n], bias[n + 1], bias[n + 2], bias[n + 3]};
    }
    // finish() of the four output channels n .. n + 3 of pixel m (n a multiple of 4,
    // n + 3 < cout) with their bias b (bias4(n)), rounded to bf16
    // and, with a mask, its values mk at (m, n .. n + 3) (mask4)
    __device__ __forceinline__ bf16x4 finish4(long n, f32x4 v, f32x4 b, bf16x4 mk) const
    {
#pragma unroll
        for (int e = 0; e < 4; e++) v[e] += b[e];
        if (relu) {
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = fmaxf(v[e], 0.0f);
        }
        if (mask.ptr) {
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (n + e >= mask_from && !((float) mk[e] > 0.0f)) v[e] = 0.0f;
        }
        bf16x4 o;
#pragma unroll
        for (int e = 0; e < 4; e++) o[e] = (bf16) v[e];
        return o;
    }
};

// ------------------------------------------------ 2x2 stride-2 transposed conv
// out[child(p, ab)][co] = sum_c in[p][c] w[ab][co][c] + bias[co]; M = input
// pixels, N = 4 cout ((ab, co)), K = cin (multiple of 32).  Output grid is
// (2h, 2w).
struct UpFwd {
    // outputs leave as 8 consecutive channels (cout a multiple of 8) of one child pixel
    static constexpr bool kVecStore = true;
    View<const bf16> in;
    int lh, lw, lcin, lcout;
    long npix;
    const bf16 *w;      // [4][cout][cin]
    const float *bias;  // [cout]
    View<bf16> out;

    struct RowA {
        long p;
        bool ok;
    };
    struct ColB {
        const bf16 *w;
        bool ok;
    };
    __device__ __forceinline__ RowA row_a(long m) const { return RowA{m, m < npix}; }
    __device__ __forceinline__ ColB col_b(long n) const
    {
        const bool ok = n < (4L << lcout);
        return ColB{w + ((ok ? n : 0) << lcin), ok};
    }
    __device__ __forceinline__ void k_range(int, long &k0, long &k1) const
    {
        k0 = 0;
        k1 = 1L << lcin;
    }
    __device__ __forceinline__ bf16x8 load_a(const RowA &r, long k) const
    {
        return r.ok ? ld8(in.at(r.p, (int) k)) : zero8();
    }
    __device__ __forceinline__ bf16x8 load_b(const ColB &c, long k) const { return c.ok ? ld8(c.w + k) : zero8(); }
    __device__ __forceinline__ bool valid(long m, long n) const { return m < npix && n < (4L << lcout); }
    __device__ __forceinline__ bool inside(long m, long n) const { return valid(m, n); }
    __device__ __forceinline__ float finish(long, long n, float v) const
    {
        return v + bias[n & ((1 << lcout) - 1)];
    }
    // bias of columns n .. n + 3 (n a multiple of 4; any n, as finish reads it)
    __device__ __forceinline__ f32x4 bias4(long n) const
    {
        const float *b = bias + (n & ((1 << lcout) - 1));
        return f32x4{b[0], b[1], b[2], b[3]};
    }
    // Output base of the image holding input pixel mb, and mb's (y, x): the children of input
    // pixel mb + lm are then 32-bit offsets from it (a tile's rows run on into the next image
    // contiguously; 8 h w cs < 2^32, checked on the host).
    struct TileBase {
        bf16 *img0;
        int y0, x0;
    };
    __device__ __forceinline__ TileBase tile_base(long mb) const
    {
        const long img = mb >> (lh + lw);
        const Pix px = pix_of(mb, lh, lw);
        return TileBase{out.ptr + (img << (lh + lw + 2)) * out.cs + out.c0, px.y, px.x};
    }
    __device__ __forceinline__ long child(long m, int ab) const
    {
        const long img = m >> (lh + lw);
        const Pix px = pix_of(m, lh, lw);
        return (((img << (lh + 1)) + 2 * px.y + (ab >> 1)) << (lw + 1)) + 2 * px.x + (ab & 1);
    }
    // columns n .. n + 7 of input pixel tb + lm (one child pixel)
    __device__ __forceinline__ void store8(const TileBase &tb, int lm, int n, bf16x8 v) const
    {
        const int ab = n >> lcout, co = n & ((1 << lcout) - 1);
        const int xx = tb.x0 + lm, x = xx & ((1 << lw) - 1), y = tb.y0 + (xx >> lw);
        const unsigned c = (unsigned) ((((2 * y + (ab >> 1)) << (lw + 1)) + 2 * x + (ab & 1)));
        *reinterpret_cast<bf16x8 *>(tb.img0 + c * (unsigned) out.cs + (unsigned) co) = v;
    }
    __device__ __forceinline__ void store(long m, long n, float v, int) const
    {
        if (!valid(m, n)) return;
        const int ab = (int) (n >> lcout), co = (int) (n & ((1 << lcout) - 1));
        *out.at(child(m, ab), co) = (bf16) finish(m, n, v);
    }
};

// Data gradient of UpFwd: gin[p][c] = sum_{ab, co} g[child(p, ab)][co] wT[c][ab][co],
// masked by the ReLU of the forward input.  M = input pixels, N = cin,
// K = 4 cout ((ab, co), cout a multiple of 32).
struct UpDgrad {
    static constexpr bool kVecStore = false;
    View<const bf16> g;
    int lh, lw, lcout, cin;
    long npix;
    const bf16 *wt;  // [cin][4][cout]
    View<const bf16> mask;
    View<bf16> out;

    struct RowA {
        long child0;
        bool ok;
    };
    struct ColB {
        const bf16 *w;
        bool ok;
    };
    __device__ __forceinline__ RowA row_a(long m) const
    {
        RowA r;
        r.ok = m < npix;
        const long img = m >> (lh + lw);
        const Pix px = pix_of(m, lh, lw);
        r.child0 = (((img << (lh + 1)) + 2 * px.y) << (lw + 1)) + 2 * px.x;
        return r;
    }
    __device__ __forceinline__ ColB col_b(long n) const
    {
        ColB c;
        c.ok = n < cin;
        c.w = wt + (c.ok ? n : 0) * (4L << lcout);
        return c;
    }
    __device__ __forceinline__ void k_range(int, long &k0, long &k1) const
    {
        k0 = 0;
        k1 = 4L << lcout;
    }
    __device__ __forceinline__ bf16x8 load_a(const RowA &r, long k) const
    {
        if (!r.ok) return zero8();
        const int ab = (int) (k >> lcout), co = (int) (k & ((1 << lcout) - 1));
        return ld8(g.at(r.child0 + ((long) (ab >> 1) << (lw + 1)) + (ab & 1), co));
    }
    __device__ __forceinline__ bf16x8 load_b(const ColB &c, long k) const
    {
        return c.ok ? ld8(c.w + k) : zero8();
    }
    __device__ __forceinline__ void store(long m, long n, float v, int) const
    {
        if (m >= npix || n >= cin) return;
        if (!((float) *mask.at(m, (int) n) > 0.0f)) v = 0.0f;
        *out.at(m, (int) n) = (bf16) v;
    }
};

// Split-K partial sums of a weight gradient: ws[z][m][n] (fp32).
struct Partial {
    float *ws;
    long M, N;
    long kchunk, K;
    __device__ __forceinline__ void range(int z, long &k0, long &k1) const
    {
        k0 = z * kchunk;
        k1 = k0 + kchunk < K ? k0 + kchunk : K;
    }
    __device__ __forceinline__ void put(long m, long n, float v, int z) const
    {
        if (m < M && n < N) ws[(z * M + m) * N + n] = v;
    }
};

// Weight gradient of Conv3: dW[co][tap][c] = sum_p g[p][co] in[p + d(tap)][c].
// M = cout, N = 9 cin ((tap, c)), K = pixels in groups of 8 along a row
// (W a multiple of 8).  Pixel-strided operands: 8 scalar loads per fragment.
struct Conv3Wgrad {
    static constexpr bool kVecStore = false;
    View<const bf16> g;
    View<const bf16> in;
    int lh, lw, cin, cout;
    long npix;
    Partial part;

    struct RowA {
        int co;
        bool ok;
    };
    struct ColB {
        int c, dy, dx;
        bool ok;
    };
    __device__ __forceinline__ RowA row_a(long m) const { return RowA{(int) m, m < cout}; }
    __device__ __forceinline__ ColB col_b(long n) const
    {
        ColB c;
        c.ok = n < 9L * cin;
        const int tap = c.ok ? (int) (n / cin) : 0;
        c.c = c.ok ? (int) (n % cin) : 0;
        c.dy = tap / 3 - 1;
        c.dx = tap % 3 - 1;
        return c;
    }
    __device__ __forceinline__ void k_range(int z, long &k0, long &k1) const { part.range(z, k0, k1); }
    __device__ __forceinline__ bf16x8 load_a(const RowA &r, long k) const
    {
        bf16x8 v = zero8();
        if (!r.ok || k >= npix) return v;
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = *g.at(k + i, r.co);
        return v;
    }
    __device__ __forceinline__ bf16x8 load_b(const ColB &c, long k) const
    {
        bf16x8 v = zero8();
        if (!c.ok || k >= npix) return v;
        const Pix px = pix_of(k, lh, lw);
        const int yy = px.y + c.dy;
        if (yy < 0 || yy >= (1 << lh)) return v;
        const long base = k + (long) c.dy * (1L << lw) + c.dx;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int xx = px.x + i + c.dx;
            if (xx >= 0 && xx < (1 << lw)) v[i] = *in.at(base + i, c.c);
        }
        return v;
    }
    __device__ __forceinline__ void store(long m, long n, float v, int z) const { part.put(m, n, v, z); }
};

// Weight gradient of Conv3 on channel-major copies: the same GEMM as Conv3Wgrad,
// but A and B rows are contiguous along pixels, so a fragment (8 pixels of one
// channel, in one image row: W >= 8) is one 16-byte load.
//   gT[co][p]                   the output gradient, transposed;
//   inT[(dx + 1) cin + c][p]    the input transposed and shifted by dx in {-1, 0, 1}
//                               along the row, zero where x + dx leaves the row.
// The tap's row shift dy is an offset of dy W pixels, zero where y + dy leaves
// the image.
struct Conv3WgradT {
    static constexpr bool kVecStore = false;
    const bf16 *gT;
    const bf16 *inT;
    int lh, lw, cin, cout;
    long npix;
    Partial part;

    struct RowA {
        const bf16 *p;
        bool ok;
    };
    struct ColB {
        const bf16 *p;
        int dy;
        bool ok;
    };
    __device__ __forceinline__ RowA row_a(long m) const
    {
        const bool ok = m < cout;
        return RowA{gT + (ok ? m : 0) * npix, ok};
    }
    __device__ __forceinline__ ColB col_b(long n) const
    {
        ColB c;
        c.ok = n < 9L * cin;
        const int tap = c.ok ? (int) (n / cin) : 4, ch = c.ok ? (int) (n % cin) : 0;
        c.dy = tap / 3 - 1;
        const int dx = tap % 3 - 1;
        c.p = inT + ((long) (dx + 1) * cin + ch) * npix + ((long) c.dy << lw);
        return c;
    }
    __device__ __forceinline__ void k_range(int z, long &k0, long &k1) const { part.range(z, k0, k1); }
    __device__ __forceinline__ bf16x8 load_a(const RowA &r, long k) const
    {
        return (r.ok && k < npix) ? ld8(r.p + k) : zero8();
    }
    __device__ __forceinline__ bf16x8 load_b(const ColB &c, long k) const
    {
        if (!c.ok || k >= npix) return zero8();
        const int yy = (int) ((k >> lw) & ((1 << lh) - 1)) + c.dy;
        if (yy < 0 || yy >= (1 << lh)) return zero8();
        return ld8(c.p + k);
    }
    __device__ __forceinline__ void store(long m, long n, float v, int z) const { part.put(m, n, v, z); }
};

// Weight gradient of UpFwd: dW[ab][co][c] = sum_p g[child(p, ab)][co] in[p][c].
// M = 4 cout ((ab, co)), N = cin, K = input pixels (groups of 8 along a row).
struct UpWgrad {
    static constexpr bool kVecStore = false;
    View<const bf16> g;
    View<const bf16> in;
    int lh, lw, lcout, cin;
    long npix;
    Partial part;

    struct RowA {
        int ab, co;
        bool ok;
    };
    struct ColB {
        int c;
        bool ok;
    };
    __device__ __forceinline__ RowA row_a(long m) const
    {
        RowA r;
        r.ok = m < (4L << lcout);
        r.ab = (int) (m >> lcout);
        r.co = (int) (m & ((1 << lcout) - 1));
        return r;
    }
    __device__ __forceinline__ ColB col_b(long n) const { return ColB{(int) n, n < cin}; }
    __device__ __forceinline__ void k_range(int z, long &k0, long &k1) const { part.range(z, k0, k1); }
    __device__ __forceinline__ bf16x8 load_a(const RowA &r, long k) const
    {
        bf16x8 v = zero8();
        if (!r.ok || k >= npix) return v;
        const long img = k >> (lh + lw);
        const Pix px = pix_of(k, lh, lw);
        const long child0 = (((img << (lh + 1)) + 2 * px.y + (r.ab >> 1)) << (lw + 1)) + 2 * px.x + (r.ab & 1);
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = *g.at(child0 + 2 * i, r.co);
        return v;
    }
    __device__ __forceinline__ bf16x8 load_b(const ColB &c, long k) const
    {
        bf16x8 v = zero8();
        if (!c.ok || k >= npix) return v;
#pragma unroll
        for (int i = 0; i < 8; i++) v[i] = *in.at(k + i, c.c);
        return v;
    }
    __device__ __forceinline__ void store(long m, long n, float v, int z) const { part.put(m, n, v, z); }
};

}  // namespace nn

namespace nn {

// 3x3 convolution (the Conv3 operation) with the input staged in LDS: a
// workgroup owns TY = 4 rows x 64 columns of one image and BN output channels;
// per 32-channel chunk the (TY + 2) x 66 pixel patch is staged once and all 9
// taps read their A fragments from it (9x reuse), with the next chunk's patch
// prefetched into registers.  Wave w computes row w of the tile (4 x NT MFMA
// tiles).  Needs W >= 64 (a multiple of 64), H a multiple of 4, cin a
// multiple of 32.
template <int NT, bool LDS_OUT>
__global__ __launch_bounds__(256) void conv3_lds_kernel(Conv3 op)
{
    constexpr int TX = 64, TY = 4, PX = TX + 2, PY = TY + 2, MT = 4, BN = NT * 16;
    constexpr int NVEC = PY * PX * 4;                   // 16-byte vectors per patch (32 channels)
    constexpr int PER = (NVEC + 255) / 256;
    // the patch, then (after the last chunk, LDS_OUT) the finished output tile [TY x TX pixels][BN channels]
    constexpr int SH = (!LDS_OUT || NVEC > TY * TX * BN / 8) ? NVEC : TY * TX * BN / 8;
    __shared__ bf16x8 sh[SH];
    bf16x8 *patch = sh;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int W = 1 << op.lw, H = 1 << op.lh;
    const int tiles_x = W / TX, tiles_y = H / TY;
    int b = blockIdx.x;
    const int tx = b % tiles_x;
    b /= tiles_x;
    const int ty = b % tiles_y;
    const long img = b / tiles_y;
    const int x0 = tx * TX, y0 = ty * TY;
    const long n0 = (long) blockIdx.y * BN;
    const long pix0 = (img << (op.lh + op.lw)) + ((long) y0 << op.lw) + x0;  // pixel (y0, x0)
    const int cin = 1 << op.lcin;

    // staging: vector v of the patch = (row, col, channel group)
    bf16x8 pre[PER];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int v = threadIdx.x + 256 * i;
            bf16x8 val = zero8();
            if (v < NVEC) {
                const int g = v & 3, pc = (v >> 2) % PX, pr = (v >> 2) / PX;
                const int yy = y0 + pr - 1, xx = x0 + pc - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    val = ld8(op.in.at(pix0 + (long) (pr - 1) * W + (pc - 1), c0 + 8 * g));
            }
            pre[i] = val;
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int v = threadIdx.x + 256 * i;
            if (v < NVEC) patch[v] = pre[i];
        }
    };

    const bf16 *wrow[NT];
    bool wok[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const long co = n0 + j * 16 + r;
        wok[j] = co < op.cout;
        wrow[j] = op.w + (wok[j] ? co : 0) * (9L << op.lcin);
    }
    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    const int nchunks = cin / 32;
    fetch(0);
    commit();
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
        const int c0 = ch * 32;
        if (ch + 1 < nchunks) fetch(c0 + 32);
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;  // patch offsets (0..2)
            bf16x8 bfr[NT];
#pragma unroll
            for (int j = 0; j < NT; j++)
                bfr[j] = wok[j] ? ld8(wrow[j] + ((long) tap << op.lcin) + c0 + 8 * q) : zero8();
            bf16x8 afr[MT];
#pragma unroll
            for (int i = 0; i < MT; i++)
                afr[i] = patch[((wave + dy) * PX + (i * 16 + r + dx)) * 4 + q];
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[i], bfr[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (ch + 1 < nchunks) {
            commit();
            __syncthreads();
        }
    }
    // D[4q + e][r] of tile (i, j): pixel (y0 + wave, x0 + 16 i + 4 q + e), channel n0 + 16 j + r.
    if constexpr (!LDS_OUT) {
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int j = 0; j < NT; j++)
#pragma unroll
                for (int e = 0; e < 4; e++)
                    op.store(pix0 + ((long) wave << op.lw) + i * 16 + 4 * q + e, n0 + j * 16 + r, acc[i][j][e], 0);
        return;
    }
    // LDS_OUT: finished values go through LDS (the loop ended with a barrier: the patch is
    // free), so the output leaves in 16-byte vectors of 8 channels, and the optional 2 x 2 max
    // pool of the tile (TY = 4 rows: 2 pooled rows) is formed from it without another pass over
    // the output.
    bf16 *tile = reinterpret_cast<bf16 *>(sh);
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int px = wave * TX + i * 16 + 4 * q + e, ch = j * 16 + r;
                const long n = n0 + ch;
                const float v = n < op.cout ? op.finish(pix0 + ((long) wave << op.lw) + i * 16 + 4 * q + e, n,
                                                        acc[i][j][e]) : 0.0f;
                tile[px * BN + ch] = (bf16) v;
            }
    __syncthreads();
    constexpr int VPP = BN / 8;  // 16-byte vectors per pixel
    for (int v = threadIdx.x; v < TY * TX * VPP; v += 256) {
        const int px = v / VPP, cg = v % VPP;
        const long n = n0 + cg * 8;
        if (n < op.cout)
            *reinterpret_cast<bf16x8 *>(op.out.at(pix0 + ((long) (px / TX) << op.lw) + px % TX, (int) n)) = sh[v];
    }
    if (op.pool.ptr) {
        // pooled pixel (y0 / 2 + pr, x0 / 2 + pc); window order (0,0), (0,1), (1,0), (1,1), first
        // maximum kept (a NaN never replaces the running value)
        const long pbase = (img << (op.lh + op.lw - 2)) + ((long) (y0 / 2) << (op.lw - 1)) + x0 / 2;
        for (int v = threadIdx.x; v < (TY / 2) * (TX / 2) * VPP; v += 256) {
            const int cg = v % VPP, pc = (v / VPP) % (TX / 2), pr = v / (VPP * (TX / 2));
            const long n = n0 + cg * 8;
            if (n >= op.cout) continue;
            const int p00 = (2 * pr) * TX + 2 * pc;
            bf16x8 m = sh[p00 * VPP + cg];
            const int offs[3] = {1, TX, TX + 1};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const bf16x8 w = sh[(p00 + offs[k]) * VPP + cg];
#pragma unroll
                for (int c = 0; c < 8; c++) m[c] = (float) w[c] > (float) m[c] ? w[c] : m[c];
            }
            *reinterpret_cast<bf16x8 *>(op.pool.at(pbase + ((long) pr << (op.lw - 1)) + pc, (int) n)) = m;
        }
    }
}


// 16 zero bytes in device memory: the source of patch lanes outside the image.
static __device__ __align__(16) unsigned char g_conv_zero16[16];

// 3x3 convolution (Conv3) on 8-row tiles: a workgroup owns 8 rows x TX pixels
// (TX = 64 or 32) of one image and BN = 16 NT output channels; wave w of the
// 8 / RPW waves computes rows RPW w .. RPW w + RPW - 1 (RPW TX / 16 x NT MFMA
// tiles), and the patch halo is 10/8 instead of the 4-row kernel's 6/4 of the
// rows.  Per 32-channel chunk the 10 x (TX + 2) pixel patch goes straight into
// LDS (global_load_lds_dwordx4: no staging registers; lanes outside the image or
// past cout load g_conv_zero16), and so do the chunk's BN x 9 x 32 weights, so
// the tap loop issues no global loads: 78 KB of LDS for 64 channels x 64 pixels
// (two workgroups per CU, as the registers allow), 40 KB for 32 channels x 32
// pixels (four).  Other shapes take the weight fragments from L2 a tap ahead.
// With LDS_OUT the finished tile leaves through LDS in 16-byte vectors, with the
// optional 2 x 2 max pool (4 pooled rows).  Needs W a multiple of TX, H a
// multiple of 8, cin a multiple of 32.  The accumulators stay in VGPRs (two
// waves per SIMD with RPW = 2, four with RPW = 1).
template <int NT, bool LDS_OUT, int RPW, int TXP = 64>
__global__ __launch_bounds__(512 / RPW) __attribute__((amdgpu_waves_per_eu(4 / RPW))) void conv3_t8_kernel(Conv3 op)
{
    constexpr int TX = TXP, CG = TX / 16, TY = 8, PX = TX + 2, PY = TY + 2, MT = CG * RPW, BN = NT * 16;
    constexpr int WAVES = TY / RPW, NTH = 64 * WAVES;
    constexpr int NVEC = PY * PX * 4;        // 16-byte vectors of a patch (32 channels)
    constexpr int PINS = (NVEC + 63) / 64;   // wave loads per patch
    constexpr bool WL = NT >= 4 || TX <= 32;  // weights staged in LDS
    constexpr int WV = WL ? BN * 36 : 0;    // 16-byte vectors of a chunk's weights: [BN][9 taps][4]
    constexpr int WINS = (WV + 63) / 64;     // wave loads per chunk of weights
    constexpr int STAGED = (PINS + WINS) * 64;
    constexpr int OUTV = TY * TX * BN / 8;   // 16-byte vectors of the finished tile
    constexpr int SH = (!LDS_OUT || STAGED > OUTV) ? STAGED : OUTV;
    __shared__ bf16x8 sh[SH];
    bf16x8 *patch = sh;
    [[maybe_unused]] bf16x8 *wts = sh + PINS * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, q = lane >> 4;
    const int W = 1 << op.lw, H = 1 << op.lh;
    const int tiles_x = W / TX, tiles_y = H / TY;
    // One flat grid of (pixel tile, channel tile) workgroups.  Workgroup w runs on XCD w % 8:
    // each XCD takes a contiguous eighth of the pixel tiles (a tile's upper neighbour, whose
    // patch shares two rows with it, ran on the same XCD shortly before), and the gy channel
    // tiles of a pixel tile take consecutive slots of one XCD, so all but the first find the
    // patch in that XCD's L2 (when the pixel tiles are a multiple of 8).
    const int gy = (op.cout + BN - 1) / BN, nw = (int) gridDim.x;
    int b, ny;
    if (((nw / gy) & 7) == 0) {
        const int xcd = (int) blockIdx.x & 7, slot = (int) blockIdx.x >> 3;
        ny = slot % gy;
        b = xcd * (nw / gy / 8) + slot / gy;
    } else {
        ny = (int) blockIdx.x % gy;
        b = (int) blockIdx.x / gy;
    }
    const int tx = b % tiles_x;
    b /= tiles_x;
    const int ty = b % tiles_y;
    const long img = b / tiles_y;
    const int x0 = tx * TX, y0 = ty * TY;
    const long n0 = (long) ny * BN;
    const long pix0 = (img << (op.lh + op.lw)) + ((long) y0 << op.lw) + x0;  // pixel (y0, x0)
    const int cin = 1 << op.lcin;
    typedef __attribute__((address_space(3))) void lds_void;
    const void *zero16 = g_conv_zero16;

    // The staged vectors' sources differ between chunks only by c0: their element offsets (or
    // kNone for a zero vector) are formed once.  Patch vector v = (pixel p, channel group g);
    // weight vector u = (output channel co, tap, channel group g) of w[n0 + co][tap][c0 + 8 g ..].
    constexpr int PJ = (PINS + WAVES - 1) / WAVES, WJ = (WINS + WAVES - 1) / WAVES, kNone = -1;
    const bf16 *ibase = op.in.ptr + pix0 * op.in.cs + op.in.c0;
    const bf16 *wbase = op.w + n0 * (9L << op.lcin);
    int poff[PJ];
    [[maybe_unused]] int woff[WJ > 0 ? WJ : 1];
#pragma unroll
    for (int k = 0; k < PJ; k++) {
        const int v = (wave + WAVES * k) * 64 + lane;
        poff[k] = kNone;
        if (v < NVEC) {
            const int p = v >> 2, g = v & 3, pc = p % PX, pr = p / PX;
            const int yy = y0 + pr - 1, xx = x0 + pc - 1;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) poff[k] = ((pr - 1) * W + (pc - 1)) * op.in.cs + 8 * g;
        }
    }
#pragma unroll
    for (int k = 0; k < WJ; k++) {
        const int u = (wave + WAVES * k) * 64 + lane;
        woff[k] = kNone;
        if (u < WV) {
            const int co = u / 36, rem = u - 36 * co, tap = rem >> 2, g = rem & 3;
            if (n0 + co < op.cout) woff[k] = co * (9 << op.lcin) + (tap << op.lcin) + 8 * g;
        }
    }
    auto stage = [&](int c0) {
#pragma unroll
        for (int k = 0; k < PJ; k++) {
            const int j = wave + WAVES * k;
            if (j < PINS) {
                const void *src = poff[k] != kNone ? (const void *) (ibase + poff[k] + c0) : zero16;
                __builtin_amdgcn_global_load_lds(src, (lds_void *) (patch + j * 64), 16, 0, 0);
            }
        }
#pragma unroll
        for (int k = 0; k < WJ; k++) {
            const int j = wave + WAVES * k;
            if (WL && j < WINS) {
                const void *src = woff[k] != kNone ? (const void *) (wbase + woff[k] + c0) : zero16;
                __builtin_amdgcn_global_load_lds(src, (lds_void *) (wts + j * 64), 16, 0, 0);
            }
        }
    };

    // weight fragments from L2 when they are not staged (32 channels on 64-pixel tiles): row
    // n0 + 16 j + r, loaded a tap ahead
    const bf16 *wrow[NT];
    bool wok[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const long co = n0 + j * 16 + r;
        wok[j] = co < op.cout;
        wrow[j] = op.w + (wok[j] ? co : 0) * (9L << op.lcin);
    }
    auto load_b = [&](bf16x8 *bfr, int tap, int c0) {
#pragma unroll
        for (int j = 0; j < NT; j++)
            bfr[j] = wok[j] ? ld8(wrow[j] + ((long) tap << op.lcin) + c0 + 8 * q) : zero8();
    };
    f32x4 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; i++)
#pragma unroll
        for (int j = 0; j < NT; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    const int nchunks = cin / 32;
    for (int ch = 0; ch < nchunks; ch++) {
        const int c0 = ch * 32;
        stage(c0);
        bf16x8 bcur[NT], bnext[NT];
        if constexpr (!WL) load_b(bcur, 0, c0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // patch (and weights) of chunk ch in LDS
        // rows of taps: dy in a loop, dx unrolled (the patch and weight offsets of a row are
        // immediates from one base)
#pragma unroll 1
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
            const int tap = 3 * dy + dx;  // patch offsets (dy, dx), 0..2
            bf16x8 bfr[NT];
            if constexpr (WL) {
#pragma unroll
                for (int j = 0; j < NT; j++) bfr[j] = wts[(16 * j + r) * 36 + tap * 4 + q];
            } else {
                if (tap < 8) load_b(bnext, tap + 1, c0);
#pragma unroll
                for (int j = 0; j < NT; j++) bfr[j] = bcur[j];
            }
            bf16x8 afr[MT];
#pragma unroll
            for (int i = 0; i < MT; i++)
                afr[i] = patch[((RPW * wave + i / CG + dy) * PX + (i % CG) * 16 + r + dx) * 4 + q];
            // weights as the row operand: lane (r, q) receives channels 4q .. 4q + 3 of pixel r
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j], afr[i], acc[i][j], 0, 0, 0);
            if constexpr (!WL) {
#pragma unroll
                for (int j = 0; j < NT; j++) bcur[j] = bnext[j];
            }
        }
        __syncthreads();  // every wave done with this chunk before the next is staged
    }
    // D[4q + e][r] of tile (i, j): channel n0 + 16 j + 4 q + e of pixel (y0 + RPW wave + i / 4, x0 + 16 (i % 4) + r)
    // the bias of this lane's channels, loaded once and unconditionally (clamped into cout)
    f32x4 bq[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const long n = n0 + j * 16 + 4 * q;
        bq[j] = op.bias4(n < op.cout ? n : 0);
    }
    // a pixel row's masks (data gradients), all NT loads issued before their use
    auto masks = [&](long m, bf16x4 *mk) {
        if (op.mask.ptr) {
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const long n = n0 + j * 16 + 4 * q;
                mk[j] = op.mask4(m, n < op.cout ? n : 0);
            }
        }
    };
    if constexpr (!LDS_OUT) {
#pragma unroll
        for (int i = 0; i < MT; i++) {
            const long m = pix0 + ((long) (RPW * wave + i / CG) << op.lw) + (i % CG) * 16 + r;
            bf16x4 mk[NT];
            masks(m, mk);
#pragma unroll
            for (int j = 0; j < NT; j++) {
                const long n = n0 + j * 16 + 4 * q;
                if (n < op.cout)
                    *reinterpret_cast<bf16x4 *>(op.out.at(m, (int) n)) = op.finish4(n, acc[i][j], bq[j], mk[j]);
            }
        }
        return;
    }
    bf16 *tile = reinterpret_cast<bf16 *>(sh);
#pragma unroll
    for (int i = 0; i < MT; i++) {
        const int row = RPW * wave + i / CG, col = (i % CG) * 16 + r;
        bf16x4 mk[NT];
        masks(pix0 + ((long) row << op.lw) + col, mk);
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const int ch = j * 16 + 4 * q;
            const long n = n0 + ch;
            bf16x4 v;
            if (n < op.cout) {
                v = op.finish4(n, acc[i][j], bq[j], mk[j]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] = (bf16) 0.0f;
            }
            *reinterpret_cast<bf16x4 *>(tile + (row * TX + col) * BN + ch) = v;
        }
    }
    __syncthreads();
    constexpr int VPP = BN / 8;  // 16-byte vectors per pixel
    if (op.out.ptr) {
        for (int v = threadIdx.x; v < OUTV; v += NTH) {
            const int px = v / VPP, cg = v % VPP;
            const long n = n0 + cg * 8;
            if (n < op.cout)
                *reinterpret_cast<bf16x8 *>(op.out.at(pix0 + ((long) (px / TX) << op.lw) + px % TX, (int) n)) =
                    sh[v];
        }
    }
    if (op.pool.ptr) {
        // pooled pixel (y0 / 2 + pr, x0 / 2 + pc); window order (0,0), (0,1), (1,0), (1,1), first
        // maximum kept (a NaN never replaces the running value)
        const long pbase = (img << (op.lh + op.lw - 2)) + ((long) (y0 / 2) << (op.lw - 1)) + x0 / 2;
        for (int v = threadIdx.x; v < (TY / 2) * (TX / 2) * VPP; v += NTH) {
            const int cg = v % VPP, pc = (v / VPP) % (TX / 2), pr = v / (VPP * (TX / 2));
            const long n = n0 + cg * 8;
            if (n >= op.cout) continue;
            const int p00 = (2 * pr) * TX + 2 * pc;
            bf16x8 m = sh[p00 * VPP + cg];
            const int offs[3] = {1, TX, TX + 1};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const bf16x8 w = sh[(p00 + offs[k]) * VPP + cg];
#pragma unroll
                for (int c = 0; c < 8; c++) m[c] = (float) w[c] > (float) m[c] ? w[c] : m[c];
            }
            *reinterpret_cast<bf16x8 *>(op.pool.at(pbase + ((long) pr << (op.lw - 1)) + pc, (int) n)) = m;
        }
    }
    if constexpr (BN == 32) {
        if (op.fin_out) {
            // the 1x1 output conv over the tile's 32 channels, in channel order as a separate pass
            // over the stored tile would sum them
            for (int px = threadIdx.x; px < TY * TX; px += NTH) {
                const long p = pix0 + ((long) (px / TX) << op.lw) + px % TX;
                float a = 0.0f;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const bf16x8 v = sh[px * 4 + i];
#pragma unroll
                    for (int j = 0; j < 8; j++) a = fmaf((float) op.fin_w[8 * i + j], (float) v[j], a);
                }
                op.fin_out[p] = op.fin_img[p] + (a + op.fin_b[0]);
            }
        }
    }
}

}  // namespace nn
