// `fft` / `ifft` with the reference's full semantics: 1-D, 2-D and 3-D
// transforms over the leading axes of an item, any transform size, zero
// padding to a user size and cropping (src/ufo-fft-task.c:122-346,
// src/ufo-ifft-task.c:111-306, src/common/ufo-fft.c:161-354).
//
// A transform is a sequence of axis passes.  One pass is one kernel launch
// that moves every line of the array along one axis through LDS:
//   load the line (zero beyond the input extent) -> FFT in LDS -> store
// the first n_out outputs.  Sizes that are not a power of two (and powers of
// two below 16) use Bluestein's algorithm (L. Bluestein, "A linear filtering
// approach to the computation of the discrete Fourier transform", 1970)
// fused into the same pass:
//   X_k = c_k * sum_n (x_n c_n) conj(c_{k-n}),   c_n = exp(dir i pi n^2 / M),
// a circular convolution of length L >= 2M - 1 done as FFT_L, a product with
// the precomputed spectrum of conj(c) (1/L folded in), and an inverse FFT_L,
// all in LDS.  The reference runs the same mathematics as a chain of
// separate kernels over global memory (spread, modulate, FFT, multiply,
// IFFT, multiply, pack); here each axis costs one HBM round trip.
//
// Lines along axis 0 are contiguous: a block takes G whole lines.  Lines
// along axes 1 and 2 are strided: a block takes G neighbouring lines, so
// one row of the block's tile is G consecutive complex values in memory
// (coalesced), and LDS lines are padded by one element against bank
// conflicts on the transposed write.
#include "ufomi_internal.h"
#include "fft16.h"

#include <cmath>
#include <complex>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kMaxL = 8192;  // longest line transform (64 KiB of LDS)

using fft16::cmul;
using fft16::lds_fft;

struct Pass {
    long in_a, in_b, in_e;     // input offsets of line (a, b) and element stride
    long out_a, out_b, out_e;  // output offsets and element stride
    int A, Bn;                 // lines: a in [0, A) (neighbours in memory on strided axes), b in [0, Bn)
    int n_in, n_out;           // input elements per line (zero beyond), outputs per line
    int L, log2L;              // transform length in LDS
    int G, Ls;                 // lines per block, LDS line stride
    float scale;
    long flat_in;              // > 0: element i of line a exists iff a in_a + i in_e < flat_in (not i < n_in)
    long flat_out;             // > 0: element i of line a is stored iff a out_a + i out_e < flat_out
    int tw_n;                  // > 0: output i of line a times exp(DIR 2 pi i (a i mod tw_n) / tw_n)
};

__device__ __forceinline__ float2 load_as_complex(const float *p, long o) { return make_float2(p[o], 0.0f); }
__device__ __forceinline__ float2 load_as_complex(const float2 *p, long o) { return p[o]; }

template <int BB, int DIR, bool CHIRP, typename InT, bool OUT_REAL>
__global__ __launch_bounds__(kThreads) void axis_dft_kernel(const InT *__restrict__ in, void *__restrict__ out_, Pass g,
                                                            const float2 *__restrict__ chirp,
                                                            const float2 *__restrict__ kern,
                                                            const float2 *__restrict__ tw)
{
    extern __shared__ float2 lds[];
    const int TF = g.L / (4 * BB);
    const int f = threadIdx.x / TF, t = threadIdx.x - f * TF;
    const int ablocks = (g.A + g.G - 1) / g.G;
    const int b = blockIdx.x / ablocks;
    const int a0 = (blockIdx.x - b * ablocks) * g.G;
    const bool strided_in = g.in_e != 1, strided_out = g.out_e != 1;

    const int total = g.G * g.L;
    for (int idx = threadIdx.x; idx < total; idx += kThreads) {
        int gl, i;
        if (strided_in) {
            gl = idx % g.G;
            i = idx / g.G;
        } else {
            gl = idx / g.L;
            i = idx - gl * g.L;
        }
        const int a = a0 + gl;
        float2 v = make_float2(0.0f, 0.0f);
        const bool inside = g.flat_in ? (long) a * g.in_a + (long) i * g.in_e < g.flat_in : i < g.n_in;
        if (a < g.A && inside) {
            v = load_as_complex(in, (long) a * g.in_a + (long) b * g.in_b + (long) i * g.in_e);
            if (CHIRP)
                v = cmul(v, chirp[i]);
        }
        lds[gl * g.Ls + i] = v;
    }
    __syncthreads();
    float2 *x = lds + f * g.Ls;
    if (CHIRP) {
        lds_fft<BB, -1>(x, g.L, g.log2L, t, TF, tw);
        for (int i = t; i < g.L; i += TF)
            x[i] = cmul(x[i], kern[i]);
        __syncthreads();
        lds_fft<BB, +1>(x, g.L, g.log2L, t, TF, tw);
    } else {
        lds_fft<BB, DIR>(x, g.L, g.log2L, t, TF, tw);
    }

    const int total_out = g.G * g.n_out;
    for (int idx = threadIdx.x; idx < total_out; idx += kThreads) {
        int gl, i;
        if (strided_out) {
            gl = idx % g.G;
            i = idx / g.G;
        } else {
            gl = idx / g.n_out;
            i = idx - gl * g.n_out;
        }
        const int a = a0 + gl;
        if (a >= g.A || (g.flat_out && (long) a * g.out_a + (long) i * g.out_e >= g.flat_out))
            continue;
        float2 y = lds[gl * g.Ls + i];
        if (CHIRP)
            y = cmul(y, chirp[i]);
        if (g.tw_n) {
            // (a i mod tw_n) / tw_n is exact in float for a power-of-two tw_n
            const long q = ((long) a * i) & (g.tw_n - 1);
            float sn, cs;
            sincospif(2.0f * (float) q / (float) g.tw_n, &sn, &cs);
            y = cmul(y, make_float2(cs, DIR * sn));
        }
        const long o = (long) a * g.out_a + (long) b * g.out_b + (long) i * g.out_e;
        if (OUT_REAL)
            ((float *) out_)[o] = y.x * g.scale;
        else
            ((float2 *) out_)[o] = make_float2(y.x * g.scale, y.y * g.scale);
    }
}

bool pow2(long x) { return x > 0 && (x & (x - 1)) == 0; }

long next_pow2(long x)
{
    long p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

// Transform length for a DFT of size M: M itself for powers of two >= 16,
// else the Bluestein length.
int line_length(int M, bool *chirp)
{
    *chirp = !(pow2(M) && M >= 16);
    return *chirp ? (int) std::max(16L, next_pow2(2L * M - 1)) : M;
}

// Radix-2 FFT in double on the host (for the Bluestein kernel spectrum).
void host_fft(std::vector<std::complex<double>> &a, int dir)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1)
            j ^= bit;
        j ^= bit;
        if (i < j)
            std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                const std::complex<double> w = std::polar(1.0, dir * 2.0 * M_PI * (double) k / (double) len);
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// chirp[n] = exp(dir i pi n^2 / M) (n^2 reduced mod 2M before the angle),
// kern = FFT_L(conj(chirp) placed at n and L - n) / L.
void bluestein_tables(int M, int L, int dir, float2 *chirp, float2 *kern)
{
    std::vector<std::complex<double>> c(M), b(L, 0.0);
    for (int n = 0; n < M; n++) {
        const long long q = ((long long) n * n) % (2LL * M);
        c[n] = std::polar(1.0, dir * M_PI * (double) q / (double) M);
        chirp[n] = make_float2((float) c[n].real(), (float) c[n].imag());
    }
    b[0] = std::conj(c[0]);
    for (int n = 1; n < M; n++)
        b[n] = b[L - n] = std::conj(c[n]);
    host_fft(b, -1);
    for (int k = 0; k < L; k++)
        kern[k] = make_float2((float) (b[k].real() / L), (float) (b[k].imag() / L));
}

struct PassSpec {
    Pass g;
    int BB;
    bool chirp;
    int M;
    size_t table_off;  // float2 offset of this pass's chirp table in the table block (kern follows)
};

// Lines along `axis` of a source array with extents S (x fastest) into a
// destination with extents D; E[] gives how many lines exist along each other
// axis (coordinates beyond E are not computed).
int make_pass(int axis, const int S[3], const int D[3], const int E[3], int n_in, int n_out, int M, PassSpec *ps)
{
    Pass &g = ps->g;
    bool chirp;
    const int L = line_length(M, &chirp);
    if (L > kMaxL)
        return ufomi_fail(UFOMI_ERR_UNSUPPORTED,
                          "transform size %d is not built (powers of two up to %d, other sizes up to %d)", M, kMaxL,
                          kMaxL / 2);
    ps->chirp = chirp;
    ps->M = M;
    g.L = L;
    g.log2L = ufomi_log2(L);
    ps->BB = L >= 8192 ? 8 : std::min(4, L / 16);
    g.G = 1024 * ps->BB / L;
    g.Ls = g.G > 1 ? L + 1 : L;
    g.n_in = n_in;
    g.n_out = n_out;
    g.scale = 1.0f;
    g.flat_in = g.flat_out = 0;
    g.tw_n = 0;
    const long s1 = S[0], s2 = (long) S[0] * S[1], d1 = D[0], d2 = (long) D[0] * D[1];
    if (axis == 0) {
        g.A = E[1];
        g.Bn = E[2];
        g.in_e = 1, g.in_a = s1, g.in_b = s2;
        g.out_e = 1, g.out_a = d1, g.out_b = d2;
    } else if (axis == 1) {
        g.A = E[0];
        g.Bn = E[2];
        g.in_e = s1, g.in_a = 1, g.in_b = s2;
        g.out_e = d1, g.out_a = 1, g.out_b = d2;
    } else {
        if (S[0] != D[0] || S[1] != D[1] || E[0] != S[0] || E[1] != S[1])
            return ufomi_fail(UFOMI_ERR_ARG, "internal: axis-2 pass needs whole planes");
        g.A = E[0] * E[1];
        g.Bn = 1;
        g.in_e = s2, g.in_a = 1, g.in_b = 0;
        g.out_e = d2, g.out_a = 1, g.out_b = 0;
    }
    const long blocks = (long) g.Bn * ((g.A + g.G - 1) / g.G);
    if (blocks <= 0 || blocks >= (1L << 31))
        return ufomi_fail(UFOMI_ERR_ARG, "transform of %d x %d lines does not fit one launch", g.A, g.Bn);
    return UFOMI_OK;
}

template <int DIR, bool CHIRP, typename InT, bool OUT_REAL>
int launch_bb(hipStream_t s, const PassSpec &ps, const void *in, void *out, const float2 *chirp, const float2 *kern,
              const float2 *tw)
{
    const Pass &g = ps.g;
    const unsigned blocks = (unsigned) ((long) g.Bn * ((g.A + g.G - 1) / g.G));
    const size_t lds = (size_t) g.G * g.Ls * sizeof(float2);
    switch (ps.BB) {
#define UFOMI_AXIS_CASE(BBV)                                                                                       \
    case BBV:                                                                                                      \
        hipLaunchKernelGGL((axis_dft_kernel<BBV, DIR, CHIRP, InT, OUT_REAL>), dim3(blocks), dim3(kThreads), lds, s, \
                           (const InT *) in, out, g, chirp, kern, tw);                                             \
        break;
        UFOMI_AXIS_CASE(1)
        UFOMI_AXIS_CASE(2)
        UFOMI_AXIS_CASE(4)
        UFOMI_AXIS_CASE(8)
#undef UFOMI_AXIS_CASE
    default: return ufomi_fail(UFOMI_ERR_ARG, "internal: no axis kernel for %d butterflies", ps.BB);
    }
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

template <int DIR, typename InT, bool OUT_REAL>
int launch_pass(hipStream_t s, const PassSpec &ps, const void *in, void *out, const float2 *tables, const float2 *tw)
{
    const float2 *chirp = tables + ps.table_off, *kern = chirp + ps.M;
    return ps.chirp ? launch_bb<DIR, true, InT, OUT_REAL>(s, ps, in, out, chirp, kern, tw)
                    : launch_bb<DIR, false, InT, OUT_REAL>(s, ps, in, out, nullptr, nullptr, tw);
}

// Scratch: two complex intermediates of `tmp_elems` and the Bluestein tables of every pass.
int setup_scratch(ufomi_ctx *ctx, PassSpec *ps, int np, int dir, size_t tmp_elems, float2 **tmp0, float2 **tmp1,
                  float2 **tables)
{
    size_t n_tab = 0;
    for (int p = 0; p < np; p++) {
        ps[p].table_off = n_tab;
        if (ps[p].chirp)
            n_tab += (size_t) ps[p].M + ps[p].g.L;
    }
    const size_t tmp_bytes = ((tmp_elems * sizeof(float2) + 255) / 256) * 256;
    void *base = nullptr;
    int rc = ufomi_scratch(ctx, 2 * tmp_bytes + n_tab * sizeof(float2) + 256, &base);
    if (rc) return rc;
    *tmp0 = (float2 *) base;
    *tmp1 = (float2 *) ((char *) base + tmp_bytes);
    *tables = (float2 *) ((char *) base + 2 * tmp_bytes);
    if (n_tab) {
        std::vector<float2> h(n_tab);
        for (int p = 0; p < np; p++)
            if (ps[p].chirp)
                bluestein_tables(ps[p].M, ps[p].g.L, dir, h.data() + ps[p].table_off,
                                 h.data() + ps[p].table_off + ps[p].M);
        if ((rc = ufomi_upload(ctx, *tables, h.data(), n_tab * sizeof(float2)))) return rc;
    }
    return UFOMI_OK;
}

// Power-of-two line lengths above kMaxL along axis 0 (1-D transforms):
// N = N1 N2 with n = N2 n1 + n2 and k = k1 + N1 k2 (the "four-step" split),
//   pass 1: for every n2, an N1-point transform over n1 (stride N2), times
//           exp(dir 2 pi i n2 k1 / N), into [k1][n2];
//   pass 2: for every k1, an N2-point transform over n2, written to k1 + N1 k2.
// Row r of the input holds `valid` elements at pitch `in_pitch`; row r of the
// output holds the first `keep` outputs at pitch `out_pitch`.
constexpr long kMaxSplit = 1L << 20;

void split_pass(Pass &g, int L)
{
    g.L = L;
    g.log2L = ufomi_log2(L);
    g.scale = 1.0f;
    g.flat_in = g.flat_out = 0;
    g.tw_n = 0;
}

int bb_of(int L) { return L >= 8192 ? 8 : std::min(4, L / 16); }

template <int DIR, typename InT, bool OUT_REAL>
int split_passes(ufomi_ctx *ctx, const void *in, long in_pitch, long valid, int rows, int N, long keep,
                 long out_pitch, float scale, float2 *tmp, void *out)
{
    const int lg = ufomi_log2(N);
    const int N1 = 1 << (lg / 2), N2 = N / N1;
    if (rows <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid transform rows");
    PassSpec p1{}, p2{};
    Pass &a = p1.g, &b = p2.g;
    split_pass(a, N1);
    p1.BB = bb_of(N1);
    a.G = 1024 * p1.BB / N1;
    a.Ls = a.G > 1 ? N1 + 1 : N1;
    a.A = N2, a.Bn = rows;
    a.in_a = 1, a.in_e = N2, a.in_b = in_pitch;
    a.n_in = N1, a.flat_in = valid;
    a.out_a = 1, a.out_e = N2, a.out_b = N;
    a.n_out = N1;
    a.tw_n = N;
    split_pass(b, N2);
    p2.BB = bb_of(N2);
    b.G = 1024 * p2.BB / N2;
    b.Ls = b.G > 1 ? N2 + 1 : N2;
    b.A = N1, b.Bn = rows;
    b.in_a = N2, b.in_e = 1, b.in_b = N;
    b.n_in = N2;
    b.out_a = 1, b.out_e = N1, b.out_b = out_pitch;
    b.n_out = N2, b.flat_out = keep;
    b.scale = scale;
    for (const Pass *g : {&a, &b}) {
        const long blocks = (long) g->Bn * ((g->A + g->G - 1) / g->G);
        if (blocks <= 0 || blocks >= (1L << 31))
            return ufomi_fail(UFOMI_ERR_ARG, "transform of %d rows does not fit one launch", rows);
    }
    int rc = launch_bb<DIR, false, InT, false>(ctx->stream, p1, in, tmp, nullptr, nullptr, ctx->twiddles);
    if (rc) return rc;
    return launch_bb<DIR, false, float2, OUT_REAL>(ctx->stream, p2, tmp, out, nullptr, nullptr, ctx->twiddles);
}

template <int DIR, typename InT, bool OUT_REAL>
int split_transform(ufomi_ctx *ctx, const void *in, long in_pitch, long valid, int rows, int N, long keep,
                    long out_pitch, float scale, void *out)
{
    void *tmp = nullptr;
    int rc = ufomi_scratch(ctx, (size_t) rows * N * sizeof(float2), &tmp);
    if (rc) return rc;
    return split_passes<DIR, InT, OUT_REAL>(ctx, in, in_pitch, valid, rows, N, keep, out_pitch, scale, (float2 *) tmp,
                                            out);
}

// Bluestein for rows longer than the LDS passes hold (the same mathematics as
// the fused pass): premultiply by the chirp into L = 2^k >= 2M - 1 points,
// forward split transform, product with the kernel spectrum, inverse split
// transform, postmultiply by the chirp.
constexpr int kEw = 256;

template <typename InT>
__global__ __launch_bounds__(kEw) void chirp_pre_kernel(const InT *__restrict__ in, long in_pitch, long valid, int M,
                                                        int L, const float2 *__restrict__ c, float2 *__restrict__ out,
                                                        long total)
{
    for (long idx = (long) blockIdx.x * kEw + threadIdx.x; idx < total; idx += (long) gridDim.x * kEw) {
        const long r = idx / L, n = idx - r * L;
        float2 v = make_float2(0.0f, 0.0f);
        if (n < M && n < valid)
            v = cmul(load_as_complex(in, r * in_pitch + n), c[n]);
        out[idx] = v;
    }
}

__global__ __launch_bounds__(kEw) void cmul_rows_kernel(float2 *__restrict__ x, const float2 *__restrict__ k, int L,
                                                        long total)
{
    for (long idx = (long) blockIdx.x * kEw + threadIdx.x; idx < total; idx += (long) gridDim.x * kEw)
        x[idx] = cmul(x[idx], k[idx % L]);
}

template <bool OUT_REAL>
__global__ __launch_bounds__(kEw) void chirp_post_kernel(const float2 *__restrict__ x, int L,
                                                         const float2 *__restrict__ c, long keep, long out_pitch,
                                                         float scale, void *__restrict__ out, long total)
{
    for (long idx = (long) blockIdx.x * kEw + threadIdx.x; idx < total; idx += (long) gridDim.x * kEw) {
        const long r = idx / keep, k = idx - r * keep;
        const float2 y = cmul(x[r * L + k], c[k]);
        if (OUT_REAL)
            ((float *) out)[r * out_pitch + k] = y.x * scale;
        else
            ((float2 *) out)[r * out_pitch + k] = make_float2(y.x * scale, y.y * scale);
    }
}

unsigned ew_blocks(long total) { return (unsigned) std::min<long>((total + kEw - 1) / kEw, 1L << 16); }

// Bytes of work space split_chirp_on needs for `rows` rows of an M-point transform.
size_t split_chirp_bytes(int rows, int M)
{
    const long L = next_pow2(2L * M - 1);
    return (((size_t) M + L) * sizeof(float2) + 255) / 256 * 256 + 3 * (size_t) rows * L * sizeof(float2);
}

template <int DIR, typename InT, bool OUT_REAL>
int split_chirp_on(ufomi_ctx *ctx, const void *in, long in_pitch, long valid, int rows, int M, long keep,
                   long out_pitch, float scale, void *out, void *base)
{
    const int L = (int) next_pow2(2L * M - 1);
    if (rows <= 0 || keep > M)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid transform rows");
    const size_t row_bytes = (size_t) rows * L * sizeof(float2);
    const size_t tab = (((size_t) M + L) * sizeof(float2) + 255) / 256 * 256;
    int rc;
    float2 *chirp = (float2 *) base, *kern = chirp + M;
    float2 *A = (float2 *) ((char *) base + tab), *B = (float2 *) ((char *) A + row_bytes),
           *T = (float2 *) ((char *) B + row_bytes);
    {
        std::vector<float2> h((size_t) M + L);
        bluestein_tables(M, L, DIR, h.data(), h.data() + M);
        if ((rc = ufomi_upload(ctx, chirp, h.data(), h.size() * sizeof(float2)))) return rc;
    }
    const long total = (long) rows * L;
    hipLaunchKernelGGL(chirp_pre_kernel<InT>, dim3(ew_blocks(total)), dim3(kEw), 0, ctx->stream, (const InT *) in,
                       in_pitch, valid, M, L, chirp, A, total);
    UFOMI_CHECK_LAUNCH();
    if ((rc = split_passes<-1, float2, false>(ctx, A, L, L, rows, L, L, L, 1.0f, T, B))) return rc;
    hipLaunchKernelGGL(cmul_rows_kernel, dim3(ew_blocks(total)), dim3(kEw), 0, ctx->stream, B, kern, L, total);
    UFOMI_CHECK_LAUNCH();
    if ((rc = split_passes<+1, float2, false>(ctx, B, L, L, rows, L, L, L, 1.0f, T, A))) return rc;
    const long tot_out = (long) rows * keep;
    hipLaunchKernelGGL(chirp_post_kernel<OUT_REAL>, dim3(ew_blocks(tot_out)), dim3(kEw), 0, ctx->stream, A, L, chirp,
                       keep, out_pitch, scale, out, tot_out);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

template <int DIR, typename InT, bool OUT_REAL>
int split_chirp(ufomi_ctx *ctx, const void *in, long in_pitch, long valid, int rows, int M, long keep, long out_pitch,
                float scale, void *out)
{
    void *base = nullptr;
    int rc = ufomi_scratch(ctx, split_chirp_bytes(rows, M), &base);
    if (rc) return rc;
    return split_chirp_on<DIR, InT, OUT_REAL>(ctx, in, in_pitch, valid, rows, M, keep, out_pitch, scale, out, base);
}

// ---- axes longer than one LDS pass holds, in 2-D and 3-D transforms
//
// The lines along the axis are gathered into contiguous rows, transformed by the
// four-step split (powers of two) or Bluestein over it (other sizes), and
// scattered back along the axis of the destination.  Line (b, c) runs over the
// two other axes; element i of it sits at b * sb + c * sc + i * sa.
struct LineMap {
    long sa, sb, sc;  // element, line-b and line-c strides (elements)
};

template <typename InT>
__global__ __launch_bounds__(kEw) void line_gather_kernel(const InT *__restrict__ in, LineMap m, int nb, int n,
                                                          float2 *__restrict__ rows, long total)
{
    for (long idx = (long) blockIdx.x * kEw + threadIdx.x; idx < total; idx += (long) gridDim.x * kEw) {
        const long l = idx / n, i = idx - l * n;
        const long b = l % nb, c = l / nb;
        rows[idx] = load_as_complex(in, b * m.sb + c * m.sc + i * m.sa);
    }
}

template <bool OUT_REAL>
__global__ __launch_bounds__(kEw) void line_scatter_kernel(const float2 *__restrict__ rows, LineMap m, int nb, int n,
                                                           void *__restrict__ out, long total)
{
    for (long idx = (long) blockIdx.x * kEw + threadIdx.x; idx < total; idx += (long) gridDim.x * kEw) {
        const long l = idx / n, k = idx - l * n;
        const long b = l % nb, c = l / nb;
        const long o = b * m.sb + c * m.sc + k * m.sa;
        if (OUT_REAL)
            ((float *) out)[o] = rows[idx].x;
        else
            ((float2 *) out)[o] = rows[idx];
    }
}

LineMap line_map(int axis, const int X[3])
{
    const long s[3] = {1, X[0], (long) X[0] * X[1]};
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    return LineMap{s[axis], s[o1], s[o2]};
}

// One axis of `M` points (n_in inputs per line, n_out outputs kept) from src
// (extents S) into dst (extents D); lines cover the other axes up to E.
template <int DIR, typename InT, bool OUT_REAL>
int long_axis(ufomi_ctx *ctx, int axis, const void *src, const int S[3], const int D[3], const int E[3], int n_in,
              int M, int n_out, float scale, void *dst)
{
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    const int nb = E[o1], nc = E[o2];
    const long lines = (long) nb * nc;
    if (lines <= 0 || lines >= (1L << 31))
        return ufomi_fail(UFOMI_ERR_ARG, "transform of %ld lines does not fit one launch", lines);
    const bool p2 = pow2(M);
    if (p2 ? (M > kMaxSplit) : (M > kMaxSplit / 2))
        return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "transform size %d is not built (powers of two up to %ld, "
                          "other sizes up to %ld)", M, kMaxSplit, kMaxSplit / 2);
    const size_t in_bytes = ((size_t) lines * n_in * sizeof(float2) + 255) / 256 * 256;
    const size_t out_bytes = ((size_t) lines * n_out * sizeof(float2) + 255) / 256 * 256;
    const size_t work = p2 ? (size_t) lines * M * sizeof(float2) : split_chirp_bytes((int) lines, M);
    void *base = nullptr;
    int rc = ufomi_scratch(ctx, in_bytes + out_bytes + work, &base);
    if (rc) return rc;
    float2 *rows_in = (float2 *) base, *rows_out = (float2 *) ((char *) base + in_bytes);
    void *wk = (char *) base + in_bytes + out_bytes;
    const long t_in = lines * n_in, t_out = lines * n_out;
    hipLaunchKernelGGL(line_gather_kernel<InT>, dim3(ew_blocks(t_in)), dim3(kEw), 0, ctx->stream, (const InT *) src,
                       line_map(axis, S), nb, n_in, rows_in, t_in);
    UFOMI_CHECK_LAUNCH();
    if (p2)
        rc = split_passes<DIR, float2, false>(ctx, rows_in, n_in, n_in, (int) lines, M, n_out, n_out, scale,
                                              (float2 *) wk, rows_out);
    else
        rc = split_chirp_on<DIR, float2, false>(ctx, rows_in, n_in, n_in, (int) lines, M, n_out, n_out, scale,
                                                rows_out, wk);
    if (rc) return rc;
    hipLaunchKernelGGL(line_scatter_kernel<OUT_REAL>, dim3(ew_blocks(t_out)), dim3(kEw), 0, ctx->stream, rows_out,
                       line_map(axis, D), nb, n_out, dst, t_out);
    UFOMI_CHECK_LAUNCH();
    return UFOMI_OK;
}

bool fits_lds(int M)
{
    bool chirp;
    return line_length(M, &chirp) <= kMaxL;
}

// A transform with an axis longer than one LDS pass: axis by axis through two
// complex intermediates of the library's own, each axis either one LDS pass
// or a long_axis gather / split transform / scatter.
//   forward: axes 0 .. dims-1, the last writing every output extent m;
//   inverse: axes dims-1 .. 0, the last (axis 0) writing real, cropped rows.
template <int DIR>
int axis_by_axis(ufomi_ctx *ctx, const float *d_in, bool in_complex, const int n[3], int dimensions, const int m[3],
                 int cw, int ch, float *d_out)
{
    int order[3], np = 0;
    for (int k = 0; k < dimensions; k++)
        order[np++] = DIR < 0 ? k : dimensions - 1 - k;
    // extents before / after each step and the largest intermediate
    int S[3] = {n[0], n[1], n[2]};
    size_t inter = 1;
    int Ss[3][3], Ds[3][3], Es[3][3], nin[3], nout[3];
    for (int p = 0; p < np; p++) {
        const int a = order[p];
        int D[3] = {S[0], S[1], S[2]}, E[3] = {S[0], S[1], S[2]};
        const bool last = p == np - 1;
        nin[p] = S[a];
        if (DIR < 0) {
            nout[p] = m[a];
            D[a] = m[a];
            if (last)
                for (int k = 0; k < 3; k++) D[k] = m[k];
        } else {
            nout[p] = a == 0 ? cw : (a == 1 ? ch : m[2]);
            D[a] = nout[p];
            if (a == 0) E[1] = D[1] = ch;
        }
        for (int k = 0; k < 3; k++) Ss[p][k] = S[k], Ds[p][k] = D[k], Es[p][k] = E[k];
        if (!last)
            inter = std::max(inter, (size_t) D[0] * D[1] * D[2]);
        for (int k = 0; k < 3; k++) S[k] = D[k];
    }
    float2 *buf[2] = {nullptr, nullptr};
    int rc = UFOMI_OK;
    if (np > 1) {
        if ((rc = ufomi_device_alloc(ctx, inter * sizeof(float2), (void **) &buf[0]))) return rc;
        if (np > 2 && (rc = ufomi_device_alloc(ctx, inter * sizeof(float2), (void **) &buf[1]))) {
            ufomi_device_free(ctx, buf[0]);
            return rc;
        }
    }
    if (DIR < 0) {
        bool padded_batch = false;
        for (int k = dimensions; k < 3; k++) padded_batch |= m[k] > n[k];
        if (padded_batch)
            rc = hipMemsetAsync(d_out, 0, (size_t) m[0] * m[1] * m[2] * sizeof(float2), ctx->stream) == hipSuccess
                     ? UFOMI_OK : ufomi_fail(UFOMI_ERR_HIP, "hipMemsetAsync failed");
    }
    double norm = 1.0;
    for (int k = 0; k < dimensions; k++) norm *= m[k];
    const void *src = d_in;
    for (int p = 0; p < np && !rc; p++) {
        const int a = order[p];
        const bool last = p == np - 1, real_in = p == 0 && DIR < 0 && !in_complex, real_out = last && DIR > 0;
        void *dst = last ? (void *) d_out : (void *) buf[p & 1];
        const float scale = real_out ? (float) (1.0 / norm) : 1.0f;
        if (fits_lds(m[a])) {
            PassSpec ps;
            if ((rc = make_pass(a, Ss[p], Ds[p], Es[p], nin[p], nout[p], m[a], &ps))) break;
            ps.g.scale = scale;
            float2 *t0, *t1, *tables;
            if ((rc = setup_scratch(ctx, &ps, 1, DIR, 1, &t0, &t1, &tables))) break;
            if (real_in)
                rc = launch_pass<DIR, float, false>(ctx->stream, ps, src, dst, tables, ctx->twiddles);
            else if (real_out)
                rc = launch_pass<DIR, float2, true>(ctx->stream, ps, src, dst, tables, ctx->twiddles);
            else
                rc = launch_pass<DIR, float2, false>(ctx->stream, ps, src, dst, tables, ctx->twiddles);
        } else if (real_in) {
            rc = long_axis<DIR, float, false>(ctx, a, src, Ss[p], Ds[p], Es[p], nin[p], m[a], nout[p], scale, dst);
        } else if (real_out) {
            rc = long_axis<DIR, float2, true>(ctx, a, src, Ss[p], Ds[p], Es[p], nin[p], m[a], nout[p], scale, dst);
        } else {
            rc = long_axis<DIR, float2, false>(ctx, a, src, Ss[p], Ds[p], Es[p], nin[p], m[a], nout[p], scale, dst);
        }
        src = dst;
    }
    if (buf[0] || buf[1]) {
        // the intermediates are the library's own: free them once the stream is done with them
        (void) hipStreamSynchronize(ctx->stream);
        if (buf[0]) ufomi_device_free(ctx, buf[0]);
        if (buf[1]) ufomi_device_free(ctx, buf[1]);
    }
    return rc;
}

}  // namespace

extern "C" int ufomi_fft_nd(ufomi_ctx *ctx, const float *d_in, int in_complex, const int *n, int dimensions,
                            const int *m, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_in || !d_out || !n || !m || dimensions < 1 || dimensions > 3)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid fft arguments");
    for (int a = 0; a < 3; a++) {
        if (n[a] < 1 || m[a] < n[a])
            return ufomi_fail(n[a] < 1 ? UFOMI_ERR_ARG : UFOMI_ERR_SIZE,
                              "Specified size must be greater than or equal to input size");
    }
    // the radix-16 row kernels for plain 1-D power-of-two transforms of real rows
    if (dimensions == 1 && !in_complex && pow2(m[0]) && m[0] >= 16 && m[0] <= kMaxL && m[1] == n[1] && m[2] == n[2])
        return ufomi_fft_rows(ctx, d_in, n[0], n[1] * n[2], m[0], d_out);
    // long non-power-of-two rows: Bluestein over the four-step split
    if (dimensions == 1 && !pow2(m[0]) && m[0] > kMaxL / 2 && m[0] <= kMaxSplit / 2 && m[1] == n[1] && m[2] == n[2]) {
        const int rows = n[1] * n[2];
        return in_complex ? split_chirp<-1, float2, false>(ctx, d_in, n[0], n[0], rows, m[0], m[0], m[0], 1.0f, d_out)
                          : split_chirp<-1, float, false>(ctx, d_in, n[0], n[0], rows, m[0], m[0], m[0], 1.0f, d_out);
    }
    // long power-of-two rows: two passes of the four-step split
    if (dimensions == 1 && pow2(m[0]) && m[0] > kMaxL && m[0] <= kMaxSplit && m[1] == n[1] && m[2] == n[2]) {
        const int rows = n[1] * n[2];
        return in_complex ? split_transform<-1, float2, false>(ctx, d_in, n[0], n[0], rows, m[0], m[0], m[0], 1.0f,
                                                               d_out)
                          : split_transform<-1, float, false>(ctx, d_in, n[0], n[0], rows, m[0], m[0], m[0], 1.0f,
                                                              d_out);
    }

    // an axis longer than one LDS pass: axis by axis through the long-axis path
    for (int a = 0; a < dimensions; a++)
        if (!fits_lds(m[a]))
            return axis_by_axis<-1>(ctx, d_in, in_complex != 0, n, dimensions, m, m[0], m[1], d_out);

    // passes from axis 0 up; the last one writes the output extents on every axis
    PassSpec ps[3];
    int S[3] = {n[0], n[1], n[2]};
    size_t tmp_elems = 1;
    for (int a = 0; a < dimensions; a++) {
        int D[3] = {S[0], S[1], S[2]};
        D[a] = m[a];
        if (a == dimensions - 1)
            for (int k = 0; k < 3; k++)
                D[k] = m[k];  // batch axes padded with zeros in the output
        if ((rc = make_pass(a, S, D, S, n[a], m[a], m[a], &ps[a]))) return rc;
        if (a < dimensions - 1)
            tmp_elems = std::max(tmp_elems, (size_t) D[0] * D[1] * D[2]);
        S[a] = m[a];
    }
    float2 *tmp[2], *tables;
    if ((rc = setup_scratch(ctx, ps, dimensions, -1, tmp_elems, &tmp[0], &tmp[1], &tables))) return rc;
    bool padded_batch = false;
    for (int k = dimensions; k < 3; k++)
        padded_batch |= m[k] > n[k];
    if (padded_batch)
        UFOMI_HIP(hipMemsetAsync(d_out, 0, (size_t) m[0] * m[1] * m[2] * sizeof(float2), ctx->stream));
    const void *src = d_in;
    for (int a = 0; a < dimensions; a++) {
        void *dst = a == dimensions - 1 ? (void *) d_out : (void *) tmp[a & 1];
        if (a == 0 && !in_complex)
            rc = launch_pass<-1, float, false>(ctx->stream, ps[a], src, dst, tables, ctx->twiddles);
        else
            rc = launch_pass<-1, float2, false>(ctx->stream, ps[a], src, dst, tables, ctx->twiddles);
        if (rc) return rc;
        src = dst;
    }
    return UFOMI_OK;
}

extern "C" int ufomi_ifft_nd(ufomi_ctx *ctx, const float *d_in, const int *m, int dimensions, int crop_width,
                             int crop_height, float *d_out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_in || !d_out || !m || dimensions < 1 || dimensions > 3 || m[0] < 1 || m[1] < 1 || m[2] < 1 ||
        crop_width < 0 || crop_height < 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid ifft arguments");
    if (crop_width > m[0] || crop_height > m[1])
        return ufomi_fail(UFOMI_ERR_SIZE, "Cropped size must be less than or equal to input size");
    const int cw = crop_width ? crop_width : m[0], ch = crop_height ? crop_height : m[1];
    if (dimensions == 1 && pow2(m[0]) && m[0] >= 16 && m[0] <= kMaxL && (ch == m[1] || m[2] == 1))
        return ufomi_ifft_rows(ctx, d_in, m[0], ch * m[2], cw, d_out);
    if (dimensions == 1 && !pow2(m[0]) && m[0] > kMaxL / 2 && m[0] <= kMaxSplit / 2 && (ch == m[1] || m[2] == 1))
        return split_chirp<+1, float2, true>(ctx, d_in, m[0], m[0], ch * m[2], m[0], cw, cw, (float) (1.0 / m[0]),
                                             d_out);
    if (dimensions == 1 && pow2(m[0]) && m[0] > kMaxL && m[0] <= kMaxSplit && (ch == m[1] || m[2] == 1))
        return split_transform<+1, float2, true>(ctx, d_in, m[0], m[0], ch * m[2], m[0], cw, cw,
                                                 (float) (1.0 / m[0]), d_out);

    for (int a = 0; a < dimensions; a++)
        if (!fits_lds(m[a]))
            return axis_by_axis<+1>(ctx, d_in, true, m, dimensions, m, cw, ch, d_out);

    // passes from the highest transformed axis down to axis 0, which writes the real, cropped rows
    PassSpec ps[3];
    int S[3] = {m[0], m[1], m[2]};
    size_t tmp_elems = 1;
    double norm = 1.0;
    int np = 0;
    for (int a = dimensions - 1; a >= 0; a--, np++) {
        const int n_out = a == 0 ? cw : (a == 1 ? ch : m[2]);
        int D[3] = {S[0], S[1], S[2]};
        D[a] = n_out;
        int E[3] = {S[0], S[1], S[2]};
        if (a == 0)
            E[1] = ch;  // rows cropped away are never transformed
        if (a == 0)
            D[1] = ch;
        if ((rc = make_pass(a, S, D, E, m[a], n_out, m[a], &ps[np]))) return rc;
        norm *= m[a];
        if (a > 0)
            tmp_elems = std::max(tmp_elems, (size_t) D[0] * D[1] * D[2]);
        for (int k = 0; k < 3; k++)
            S[k] = D[k];
    }
    ps[np - 1].g.scale = (float) (1.0 / norm);
    float2 *tmp[2], *tables;
    if ((rc = setup_scratch(ctx, ps, np, +1, tmp_elems, &tmp[0], &tmp[1], &tables))) return rc;
    const void *src = d_in;
    for (int p = 0; p < np; p++) {
        if (p == np - 1) {
            rc = launch_pass<+1, float2, true>(ctx->stream, ps[p], src, d_out, tables, ctx->twiddles);
        } else {
            rc = launch_pass<+1, float2, false>(ctx->stream, ps[p], src, tmp[p & 1], tables, ctx->twiddles);
            src = tmp[p & 1];
        }
        if (rc) return rc;
    }
    return UFOMI_OK;
}
