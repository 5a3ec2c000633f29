// Parallel-beam backprojection (`backproject` task) for gfx950.
//
// Texture mode (backproject.cl:50-94) is a gather: voxel (x, y) at angle p
// reads the filtered sinogram row p at h = by sin_p + bx cos_p + axis with a
// linear interpolation and a zero border.  Hardware texture filtering keeps
// 8-bit weights, too coarse for 1e-5 parity, so interpolation is done in
// float32 in registers from LDS:
//
//  * A 256-thread workgroup owns a 32x32 voxel tile of S slices that share the
//    geometry (S = 1, 2 or 4).  Thread (lx, ly) owns voxels
//    (lx + 16 kx, ly + 16 ky), kx, ky in {0, 1}; a wave covers 16x4 adjacent
//    voxels, so its 64 LDS reads fall inside ~18 consecutive entries.
//  * Angles are processed in chunks of A = 16.  For every angle of a chunk the
//    detector span the tile can touch (<= 31 sqrt 2 + 3 < 64 entries) is
//    staged into LDS as pairs (v_j, v_{j+1} - v_j) with the zero border applied,
//    so one interpolation costs one ds_read_b64 (b128 for two slices) and a
//    single fma: v + f * d.
//  * The next chunk is prefetched into registers while the current one is
//    consumed, so one LDS buffer suffices (2 barriers per chunk).
//  * Per voxel and angle: u = bx cos + (by sin + a'), i = (int) u,
//    f = fract(u): 4 VALU shared by the S slices, then 2 VALU per slice.
//  * Workgroup ids are remapped so each XCD walks a contiguous range of
//    (slice group, tile) work and keeps its sinograms in its own L2.
#include "ufomi_internal.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 32;       // voxels per tile side
constexpr float kPiF = 3.14159274101257324219f;  // M_PI_F

// Layout of one staged detector entry for S slices in LDS.
//  f32 (v, d) pairs: S = 1 one float2; S >= 2 S/2 float4 planes, each the
//  (v, d) of two slices.  H16 (f16 storage, S >= 4): S/4 uint4 planes, each
//  the half2 (v, d) of four slices, consumed by v_dot2 with (1, f).
template <int S, bool H16> struct LdsT {
    using E = float4;
    static constexpr int NP = S / 2;
};
template <> struct LdsT<1, false> {
    using E = float2;
    static constexpr int NP = 1;
};
template <int S> struct LdsT<S, true> {
    using E = uint4;
    static constexpr int NP = S / 4;
};

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// The value of lane + 1 of the wave (DPP wave_shl:1, a VALU move; lane 63 gets 0).
__device__ __forceinline__ float next_lane(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ unsigned pack_vd(float v, float d)
{
    const _Float16 hv = (_Float16) v, hd = (_Float16) d;
    return (unsigned) __builtin_bit_cast(unsigned short, hv) | ((unsigned) __builtin_bit_cast(unsigned short, hd) << 16);
}

__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int q = nb >> 3;
    if (b < (q << 3))
        return (b & 7) * q + (b >> 3);
    return b;
}

// Global-memory view of a type (loads with a uniform base pointer then take
// the SGPR-base + 32-bit VGPR-offset form: no 64-bit address per load).
template <typename T> using Glb = __attribute__((address_space(1))) T;

// Raw stored value at byte offset b of a slice's filtered sinogram.
template <typename R, typename Ptr> __device__ __forceinline__ R load_at(Ptr p, unsigned b)
{
    return *(const Glb<R> *) ((const Glb<char> *) p + b);
}

// Raw stored values of a filtered sinogram and their conversion to float.
template <typename T> struct Stage;
// Loads go through the caller's pointer type, which may carry the global
// address space.
template <> struct Stage<float> {
    using R = float;
    template <typename Ptr> static __device__ __forceinline__ R load(Ptr p, long i) { return p[i]; }
    static __device__ __forceinline__ float cvt(R v) { return v; }
};
template <> struct Stage<__half> {
    using R = unsigned short;
    template <typename Ptr> static __device__ __forceinline__ R load(Ptr p, long i)
    {
        return __builtin_bit_cast(R, p[i]);
    }
    static __device__ __forceinline__ float cvt(R v) { return (float) __builtin_bit_cast(_Float16, v); }
};
template <> struct Stage<__hip_bfloat16> {
    using R = unsigned short;
    template <typename Ptr> static __device__ __forceinline__ R load(Ptr p, long i)
    {
        return __builtin_bit_cast(R, p[i]);
    }
    static __device__ __forceinline__ float cvt(R v) { return __uint_as_float((unsigned) v << 16); }
};

// Per-angle staging values for this tile: (sin, cos, a' = axis - 0.5 - lo, lo).
struct AngleSpan {
    float s, c, a;
    int lo;
};

__device__ __forceinline__ AngleSpan angle_span(float2 sc, bool valid, float bx0, float bx1, float by0, float by1,
                                                float am)
{
    AngleSpan r;
    if (valid) {
        r.s = sc.x;
        r.c = sc.y;
        const float umin = am + fminf(by0 * r.s, by1 * r.s) + fminf(bx0 * r.c, bx1 * r.c);
        r.lo = (int) floorf(umin) - 1;
        r.a = am - (float) r.lo;
    } else {
        r.s = r.c = r.a = 0.0f;
        r.lo = 0;
    }
    return r;
}

// One tile of TILE x TILE voxels of S slices, angles in chunks of A.  Thread
// (lx, ly) owns voxels (lx + 16 kx, ly + 16 ky), kx, ky < K = TILE / 16.
// Every thread of the workgroup calls it (it synchronises); it leaves the LDS
// buffers free for the next call.  The detector span of a tile at one angle is
// at most (TILE - 1) sqrt 2 + 3 entries, inside STRIDE = 2 TILE.
// H16 (f16 storage only) stages (v, d) as half2 and interpolates with
// v_dot2_f32_f16: half the LDS bytes and one VALU op per update.
template <int S, int TILE, int A, typename InT, bool H16 = false, typename InPtr = const InT *,
          typename OutPtr = float *>
__device__ __forceinline__ void bp_tile(
    InPtr __restrict__ sino, long sino_pitch, int W, int P, const float2 *__restrict__ trig,
    float axis, float roi_x, float roi_y, int roi_w, int roi_h, int n_slices, int slice0,
    OutPtr __restrict__ out, long out_pitch, int x0, int y0, int p_norm = 0)
{
    static_assert(!H16 || (S % 4 == 0 && std::is_same<InT, __half>::value), "half staging: f16, 4 | S");
    using L = LdsT<S, H16>;
    using E = typename L::E;
    constexpr int K = TILE / 16;
    constexpr int STRIDE = 2 * TILE;
    constexpr int ROWS = kThreads / STRIDE;           // angles staged per pass
    constexpr int NPLANE = L::NP;
    constexpr int PASSES = A * STRIDE / kThreads;     // staging passes per chunk
    static_assert(A % ROWS == 0, "angle chunk must be a multiple of the staged rows per pass");
    __shared__ E lds[NPLANE][A][STRIDE];
    __shared__ float4 par[A];

    const int tid = threadIdx.x;
    const int lx = tid & 15, ly = tid >> 4;

    // Voxel-centre coordinates exactly as backproject.cl:63-64 writes them.
    float bx[K], by[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        bx[k] = (float) (x0 + lx + 16 * k) - axis + roi_x + 0.5f;
        by[k] = (float) (y0 + ly + 16 * k) - axis + roi_y + 0.5f;
    }
    // Tile corners for the span computation.
    const float cbx0 = (float) x0 - axis + roi_x + 0.5f, cbx1 = (float) (x0 + TILE - 1) - axis + roi_x + 0.5f;
    const float cby0 = (float) y0 - axis + roi_y + 0.5f, cby1 = (float) (y0 + TILE - 1) - axis + roi_y + 0.5f;
    const float am = axis - 0.5f;

    InPtr src[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int sl = slice0 + s < n_slices ? slice0 + s : n_slices - 1;
        src[s] = sino + (long) sl * sino_pitch;
    }

    // Staging registers: this thread fills entry e of angles arow + ROWS m.
    const int e = tid & (STRIDE - 1);
    const int arow = tid / STRIDE;
    float sv[S], sd[S];
    AngleSpan sp[PASSES];

    // fetch() issues the loads of the next chunk (clamped addresses, no
    // branches) and keeps the raw stored values; commit() converts them after
    // the current chunk's interpolation, so the loads stay in flight meanwhile.
    using R = typename Stage<InT>::R;
    // Each lane loads sample j = lo + e; sample j + 1 comes from the next lane
    // of the wave (a wave covers 64 consecutive entries of one angle).  The
    // interpolation reads entries below (TILE - 1) sqrt 2 + 2 < STRIDE - 1, so
    // the last entry of an angle is never read; with STRIDE = 128 the entry at
    // the wave boundary (e = 63) is, and that lane loads j + 1 itself.
    // Float staging of 8 slices keeps both loads: its register budget (3 waves
    // per SIMD) is what the extra lane moves would cost (measured: 1345 vs 1488 ms
    // for 2048 x 1800 x 2048).
    static_assert(STRIDE == 64 || STRIDE == 128, "a staged angle spans one or two waves");
    constexpr bool kDpp = H16 || S < 8;
    constexpr bool kSeam = kDpp && STRIDE == 128;
    R r0[PASSES][S], r1[(kSeam || !kDpp) ? PASSES : 1][S];
    bool in0[PASSES], in1[PASSES];
    // the staged span of an angle lies inside the detector (no border masking);
    // one angle row is one or two whole waves, so this is uniform per wave
    bool inner[PASSES];
    const bool wave_last = kSeam && e == 63;
    // (sin, cos) of the angles the next fetch() stages, loaded one chunk ahead
    // so that no fetch waits for the trig table before issuing its loads.
    float2 tq[PASSES];
    auto load_trig = [&](int p0) {
#pragma unroll
        for (int m = 0; m < PASSES; m++)
            tq[m] = trig[min(p0 + arow + ROWS * m, P - 1)];
    };
    auto fetch = [&](int p0) {
#pragma unroll
        for (int m = 0; m < PASSES; m++) {
            const int pa = arow + ROWS * m;
            const int p = p0 + pa;
            sp[m] = angle_span(tq[m], p < P, cbx0, cbx1, cby0, cby1, am);
            const int j = sp[m].lo + e;
            in0[m] = p < P && j >= 0 && j < W;
            in1[m] = p < P && j + 1 >= 0 && j + 1 < W;
            inner[m] = p < P && sp[m].lo >= 0 && sp[m].lo + STRIDE + 1 <= W;
            const int j0 = min(max(j, 0), W - 1), j1 = min(max(j + 1, 0), W - 1);
            // byte offsets in a slice's sinogram (< 4 GiB, checked on the host),
            // shared by the S slices, whose base pointers are uniform
            const unsigned rb = (unsigned) (p < P ? p : 0) * (unsigned) W;
            const unsigned b0 = (rb + (unsigned) j0) * (unsigned) sizeof(R);
            const unsigned b1 = (rb + (unsigned) j1) * (unsigned) sizeof(R);
#pragma unroll
            for (int s = 0; s < S; s++) r0[m][s] = load_at<R>(src[s], b0);
            if constexpr (!kDpp) {
#pragma unroll
                for (int s = 0; s < S; s++) r1[m][s] = load_at<R>(src[s], b1);
            } else if constexpr (kSeam) {
                if (wave_last) {
#pragma unroll
                    for (int s = 0; s < S; s++) r1[m][s] = load_at<R>(src[s], b1);
                }
            }
        }
        load_trig(p0 + A);
    };
    auto commit = [&]() {
#pragma unroll
        for (int m = 0; m < PASSES; m++) {
            const int pa = arow + ROWS * m;
            if (inner[m]) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const float v0 = Stage<InT>::cvt(r0[m][s]);
                    float v1;
                    if constexpr (!kDpp) {
                        v1 = Stage<InT>::cvt(r1[m][s]);
                    } else {
                        v1 = next_lane(v0);
                        if constexpr (kSeam) {
                            if (wave_last) v1 = Stage<InT>::cvt(r1[m][s]);
                        }
                    }
                    sv[s] = v0;
                    sd[s] = v1 - v0;
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const float v0 = in0[m] ? Stage<InT>::cvt(r0[m][s]) : 0.0f;
                    // entry e + 1 of the same angle sits in the next lane of this wave
                    float v1;
                    if constexpr (!kDpp) {
                        v1 = in1[m] ? Stage<InT>::cvt(r1[m][s]) : 0.0f;
                    } else {
                        v1 = next_lane(v0);
                        if constexpr (kSeam) {
                            if (wave_last) v1 = in1[m] ? Stage<InT>::cvt(r1[m][s]) : 0.0f;
                        }
                    }
                    sv[s] = v0;
                    sd[s] = v1 - v0;
                }
            }
            if constexpr (H16) {
#pragma unroll
                for (int pl = 0; pl < NPLANE; pl++)
                    lds[pl][pa][e] = make_uint4(pack_vd(sv[4 * pl], sd[4 * pl]), pack_vd(sv[4 * pl + 1], sd[4 * pl + 1]),
                                                pack_vd(sv[4 * pl + 2], sd[4 * pl + 2]),
                                                pack_vd(sv[4 * pl + 3], sd[4 * pl + 3]));
            } else if constexpr (S == 1) {
                lds[0][pa][e] = make_float2(sv[0], sd[0]);
            } else {
#pragma unroll
                for (int pl = 0; pl < NPLANE; pl++)
                    lds[pl][pa][e] = make_float4(sv[2 * pl], sd[2 * pl], sv[2 * pl + 1], sd[2 * pl + 1]);
            }
            if (e == 0)
                par[pa] = make_float4(sp[m].s, sp[m].c, sp[m].a, 0.0f);
        }
    };

    float acc[K][K][S];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int s = 0; s < S; s++)
                acc[i][k][s] = 0.0f;

    const int nchunks = (P + A - 1) / A;
    load_trig(0);
    fetch(0);
    commit();
    __syncthreads();
    for (int ch = 0; ch < nchunks; ch++) {
        if (ch + 1 < nchunks)
            fetch((ch + 1) * A);
#pragma unroll 2
        for (int pa = 0; pa < A; pa++) {
            const float4 pr = par[pa];
            // Addresses and weights of all K*K voxels first, then all LDS reads,
            // then the accumulation: keeps K*K (2 K*K) reads in flight per wave.
            constexpr int KK = K * K;
            int iv[KK];
            float fv[KK];
#pragma unroll
            for (int ky = 0; ky < K; ky++) {
                const float t = fmaf(by[ky], pr.x, pr.z);
#pragma unroll
                for (int kx = 0; kx < K; kx++) {
                    const float u = fmaf(bx[kx], pr.y, t);
                    iv[K * ky + kx] = (int) u;
                    fv[K * ky + kx] = __builtin_amdgcn_fractf(u);
                }
            }
            E q[NPLANE][KK];
#pragma unroll
            for (int pl = 0; pl < NPLANE; pl++)
#pragma unroll
                for (int v = 0; v < KK; v++)
                    q[pl][v] = lds[pl][pa][iv[v]];
            if constexpr (H16) {
#pragma unroll
                for (int v = 0; v < KK; v++) {
                    const half2_t w = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(1.0f, fv[v]));
#pragma unroll
                    for (int pl = 0; pl < NPLANE; pl++) {
                        const uint4 qq = q[pl][v];
                        float *a = &acc[v / K][v % K][4 * pl];
                        a[0] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, qq.x), w, a[0], false);
                        a[1] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, qq.y), w, a[1], false);
                        a[2] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, qq.z), w, a[2], false);
                        a[3] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, qq.w), w, a[3], false);
                    }
                }
            } else if constexpr (S == 1) {
#pragma unroll
                for (int v = 0; v < KK; v++)
                    acc[v / K][v % K][0] += fmaf(fv[v], q[0][v].y, q[0][v].x);
            } else {
#pragma unroll
                for (int v = 0; v < KK; v++)
#pragma unroll
                    for (int pl = 0; pl < NPLANE; pl++) {
                        acc[v / K][v % K][2 * pl] += fmaf(fv[v], q[pl][v].y, q[pl][v].x);
                        acc[v / K][v % K][2 * pl + 1] += fmaf(fv[v], q[pl][v].w, q[pl][v].z);
                    }
            }
        }
        __syncthreads();
        if (ch + 1 < nchunks) {
            commit();
            __syncthreads();
        }
    }

    // pi / (sinogram height); an angle-split part divides by the whole height
    const float fp = (float) (p_norm ? p_norm : P);
#pragma unroll
    for (int ky = 0; ky < K; ky++) {
        const int y = y0 + ly + 16 * ky;
#pragma unroll
        for (int kx = 0; kx < K; kx++) {
            const int x = x0 + lx + 16 * kx;
            if (x < roi_w && y < roi_h) {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    if (slice0 + s < n_slices) {
                        const float r = acc[ky][kx][s] * kPiF;
                        out[(long) (slice0 + s) * out_pitch + (long) y * roi_w + x] = r / fp;
                    }
                }
            }
        }
    }
}

// S = 8 (float staging) holds about twice the registers of S = 4: two waves per SIMD.
template <int S, int A, typename InT, bool H16>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(S == 8 && !H16 ? 2 : 4))) void
bp_texture_kernel(const InT *__restrict__ sino, long sino_pitch, int W, int P, const float2 *__restrict__ trig,
                  float axis, float roi_x, float roi_y, int roi_w, int roi_h, int n_slices,
                  float *__restrict__ out, long out_pitch, int tiles_x, int tiles_y, int p_chunk)
{
    const int nb = tiles_x * tiles_y * ((n_slices + S - 1) / S);
    const int b = xcd_remap(blockIdx.x, nb);
    const int tiles = tiles_x * tiles_y;
    const int g = b / tiles;
    const int ty = (b % tiles) / tiles_x, tx = b % tiles_x;
    // angle split j = blockIdx.y (p_chunk > 0): angles [j p_chunk, ...) into partial volume j
    const int p0 = p_chunk > 0 ? (int) blockIdx.y * p_chunk : 0;
    const int np = p_chunk > 0 ? min(p_chunk, P - p0) : P;
    float *o = p_chunk > 0 ? out + (long) blockIdx.y * n_slices * out_pitch : out;
    bp_tile<S, kTile, A, InT, H16>(sino + (long) p0 * W, sino_pitch, W, np, trig + p0, axis, roi_x, roi_y, roi_w,
                                   roi_h, n_slices, g * S, o, out_pitch, tx * kTile, ty * kTile, P);
}

// Sum of the partial volumes of an angle-split launch, in split order.
__global__ __launch_bounds__(kThreads) void bp_sum_parts_kernel(const float *__restrict__ part, int n_split,
                                                                long part_stride, long plane, int n_slices,
                                                                float *__restrict__ out, long out_pitch)
{
    const long total = plane * n_slices;
    for (long i = (long) blockIdx.x * kThreads + threadIdx.x; i < total; i += (long) gridDim.x * kThreads) {
        float s = part[i];
        for (int j = 1; j < n_split; j++)
            s += part[j * part_stride + i];
        const long sl = i / plane;
        out[sl * out_pitch + (i - sl * plane)] = s;
    }
}

// Persistent backprojection of a batch of single-slice requests of uneven
// size: each workgroup takes (request, tile) items from a global counter until
// the list is exhausted.  The list is sorted by decreasing cost, so the tail
// of the launch is made of the cheapest tiles.
template <int TILE, int A, typename InT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void bp_batch_kernel(
    const ufomi_bp_batch_req *__restrict__ reqs, const int2 *__restrict__ items, int n_items,
    int *__restrict__ counter, const float2 *__restrict__ trig)
{
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0)
            next = atomicAdd(counter, 1);
        __syncthreads();
        // uniform across the workgroup: the request's fields load into SGPRs
        const int it = __builtin_amdgcn_readfirstlane(next);
        if (it >= n_items)
            break;
        const int2 w = items[it];
        const ufomi_bp_batch_req &r = reqs[w.x];
        const int ty = w.y / r.tiles_x, tx = w.y % r.tiles_x;
        // The request's buffers are global memory: say so, so that the loads and
        // stores are global (not flat) instructions and do not hold up LDS waits.
        using GIn = __attribute__((address_space(1))) const InT;
        using GOut = __attribute__((address_space(1))) float;
        bp_tile<1, TILE, A, InT, false, GIn *, GOut *>((GIn *) r.sino, 0, r.W, r.P, trig + r.trig0, r.axis, r.roi_x,
                                                r.roi_y, r.roi_w, r.roi_h, 1, 0, (GOut *) r.out, 0, tx * TILE,
                                                ty * TILE);
    }
}

__global__ void reset_counter_kernel(int *counter)
{
    if (threadIdx.x == 0)
        *counter = 0;
}

// backproject_nearest (backproject.cl:24-48): flat index (int)(p * roi_w + h)
// into the sinogram, with the reference's operation order; indices outside the
// sinogram (undefined in the reference) read 0.
__global__ __launch_bounds__(kThreads) void bp_nearest_kernel(
    const float *__restrict__ sino, long sino_pitch, int W, int P, const float2 *__restrict__ trig,
    float axis, float roi_x, float roi_y, int roi_w, int roi_h, int n_slices,
    float *__restrict__ out, long out_pitch)
{
    // Every product rounds on its own, as written in the reference kernel: the
    // truncated index is sensitive to the last bit of h.
#pragma clang fp contract(off)
    const int idx = blockIdx.x * 16 + (threadIdx.x & 15);
    const int idy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int sl = blockIdx.z;
    if (idx >= roi_w || idy >= roi_h)
        return;
    const float *s = sino + (long) sl * sino_pitch;
    const long total = (long) P * W;
    const float bx = (float) idx - axis + roi_x + 0.5f;
    const float by = (float) idy - axis + roi_y + 0.5f;
    float sum = 0.0f;
    for (int p = 0; p < P; p++) {
        const float2 sc = trig[p];
        float h = axis + bx * sc.y;
        h = h + by * sc.x;
        const float fi = (float) (p * roi_w) + h;
        if (fi >= -2147483648.0f && fi < 2147483648.0f) {
            const long i = (long) (int) fi;
            if (i >= 0 && i < total)
                sum += s[i];
        }
    }
    const float r = sum * kPiF;
    out[(long) sl * out_pitch + (long) idy * roi_w + idx] = r / (float) P;
}

// ---- preallocated device memory of ufomi_bp_parallel_noalloc
//
// One file-scope array per GPU, held until the process ends: the (sin, cos)
// table of up to 32768 angles (the range of the num-projections property),
// then the partial volumes of an angle-split launch.  A split happens only
// below 1024 workgroups (nb tiles x slice groups), with n_split <= 1024 / nb + 1
// parts of at most S * 32 * 32 <= 8192 voxels per workgroup, so the partial
// volumes never exceed 2047 * 8192 floats < 64 MiB.
constexpr size_t kNoallocTrigMax = 32768;
constexpr size_t kNoallocTrigBytes = kNoallocTrigMax * sizeof(float2);
constexpr size_t kNoallocPartBytes = (size_t) 64 << 20;
constexpr int kNoallocDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocTrigBytes + kNoallocPartBytes];

// Host-side table of each device's address of `scratch` (hipGetSymbolAddress).
unsigned char *g_noalloc_ptr[kNoallocDevices];
std::mutex g_noalloc_mutex;

// The LUT of ufo-backproject-task.c:174-196 computed on the device: FP64
// sin/cos of angle_offset + (offset + i) step rounded to float (sin_sign -1
// stores -sin).  The device's FP64 sin/cos are within an ulp of the host's,
// so the rounded floats agree except when the FP64 value lies within an ulp
// of a float rounding boundary.
__global__ __launch_bounds__(kThreads) void trig_table_kernel(float2 *__restrict__ trig, int P, double angle_offset,
                                                              double step, unsigned offset, int sin_sign)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= P)
        return;
    const double a = angle_offset + (double) (offset + (unsigned) i) * step;
    const float sv = (float) sin(a);
    trig[i] = make_float2(sin_sign < 0 ? -sv : sv, (float) cos(a));
}

int bp_parallel_impl(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in, int n_slices,
                     size_t sino_pitch, const ufomi_bp_params *prm, float *d_out, size_t out_pitch, int sin_sign,
                     bool half_staging, unsigned char *fixed);

}  // namespace

// LUT of ufo-backproject-task.c:174-196: FP64 sin/cos of angle_offset + i*step,
// rounded to float; entries offset..offset+P-1 are uploaded, as (sin, cos).
// sin_sign -1 stores -sin (exact in float): the stacked kernels' h = bx cos - by sin.
int ufomi_bp_trig(ufomi_ctx *ctx, double angle_offset, double step, unsigned offset, int P, int sin_sign,
                  float2 **out)
{
    const bool cached = ctx->lut && ctx->lut_n == P && ctx->lut_offset == angle_offset &&
                        ctx->lut_step == step && ctx->lut_first == offset && ctx->lut_sign == sin_sign;
    if (!cached) {
        std::vector<float2> h(P);
        for (int i = 0; i < P; i++) {
            const double a = angle_offset + (double) (offset + i) * step;
            const float sv = (float) sin(a);
            h[i] = make_float2(sin_sign < 0 ? -sv : sv, (float) cos(a));
        }
        if (ctx->lut) {
            UFOMI_HIP(hipStreamSynchronize(ctx->stream));
            UFOMI_HIP(hipFree(ctx->lut));
            ctx->lut = nullptr;
        }
        UFOMI_HIP(hipMalloc(&ctx->lut, (size_t) P * sizeof(float2)));
        UFOMI_HIP(hipMemcpy(ctx->lut, h.data(), (size_t) P * sizeof(float2), hipMemcpyHostToDevice));
        ctx->lut_n = P;
        ctx->lut_offset = angle_offset;
        ctx->lut_step = step;
        ctx->lut_first = offset;
        ctx->lut_sign = sin_sign;
    }
    *out = ctx->lut;
    return UFOMI_OK;
}

extern "C" int ufomi_bp_configure(ufomi_ctx *ctx, int slices_per_tile, int angle_chunk, int half_staging)
{
    if (!ctx) return ufomi_fail(UFOMI_ERR_ARG, "context is NULL");
    if (slices_per_tile != 0 && slices_per_tile != 1 && slices_per_tile != 2 && slices_per_tile != 4 &&
        slices_per_tile != 8)
        return ufomi_fail(UFOMI_ERR_ARG, "slices per tile must be 0 (automatic), 1, 2, 4 or 8");
    if (angle_chunk != 0 && angle_chunk != 8 && angle_chunk != 16)
        return ufomi_fail(UFOMI_ERR_ARG, "angle chunk must be 0 (automatic), 8 or 16");
    ctx->bp_slices = slices_per_tile;
    ctx->bp_chunk = angle_chunk;
    ctx->bp_half = half_staging;
    return UFOMI_OK;
}

extern "C" void ufomi_bp_params_init(ufomi_bp_params *p)
{
    memset(p, 0, sizeof(*p));
    p->axis_pos = -1.0;
    p->angle_step = -1.0;
    p->angle_offset = 0.0;
    p->mode = UFOMI_BP_TEXTURE;
}

extern "C" int ufomi_bp_parallel(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in,
                                 int n_slices, size_t sino_pitch, const ufomi_bp_params *prm,
                                 float *d_out, size_t out_pitch)
{
    return ufomi_bp_parallel_ex(ctx, d_sino, dtype, W, P_in, n_slices, sino_pitch, prm, d_out, out_pitch, +1, true);
}

int ufomi_bp_parallel_ex(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in, int n_slices,
                         size_t sino_pitch, const ufomi_bp_params *prm, float *d_out, size_t out_pitch, int sin_sign,
                         bool half_staging)
{
    return bp_parallel_impl(ctx, d_sino, dtype, W, P_in, n_slices, sino_pitch, prm, d_out, out_pitch, sin_sign,
                            half_staging, nullptr);
}

namespace {

// One body for both entry points.  `fixed` = nullptr: the (sin, cos) table is
// the context's cached LUT and angle-split partial volumes live in the
// context's grow-only buffer.  Otherwise `fixed` is this device's address of
// the file-scope `scratch` array: the table is computed into its first
// kNoallocTrigBytes by a kernel and the partial volumes follow it, so nothing
// is allocated or copied from the host.
int bp_parallel_impl(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in, int n_slices,
                     size_t sino_pitch, const ufomi_bp_params *prm, float *d_out, size_t out_pitch, int sin_sign,
                     bool half_staging, unsigned char *fixed)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_sino || !d_out || !prm || W <= 0 || P_in <= 0 || n_slices <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid backprojection arguments");
    if (dtype != UFOMI_F32 && dtype != UFOMI_F16 && dtype != UFOMI_BF16)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown sinogram dtype %d", dtype);

    // ufo_backproject_task_get_requisition (ufo-backproject-task.c:212-267)
    const unsigned n_proj = prm->num_projections ? prm->num_projections : (unsigned) P_in;
    if ((unsigned) P_in > n_proj)
        return ufomi_fail(UFOMI_ERR_PROJECTIONS,
                          "Total number of projections (%u) must be greater than or equal to sinogram height (%u)",
                          n_proj, (unsigned) P_in);
    if ((unsigned long) prm->offset + (unsigned) P_in > n_proj)
        return ufomi_fail(UFOMI_ERR_ARG, "offset %u + sinogram height %d exceeds the %u LUT entries",
                          prm->offset, P_in, n_proj);
    const int roi_w = prm->roi_width ? (int) prm->roi_width : W;
    const int roi_h = prm->roi_height ? (int) prm->roi_height : W;
    const double step = prm->angle_step <= 0.0 ? M_PI / (double) n_proj : prm->angle_step;
    // ufo_backproject_task_process (:125-134): axis <= 0 means W / 2.
    const float axis = prm->axis_pos <= 0.0 ? (float) W / 2.0f : (float) prm->axis_pos;
    if (!std::isfinite(axis) || !std::isfinite(step) || !std::isfinite(prm->angle_offset))
        return ufomi_fail(UFOMI_ERR_ARG, "non-finite geometry");
    // Coordinates stay below 2^20 so float32 rounding of u stays far inside the
    // one-entry margin of the staged detector span.
    const double lim = (double) (1 << 20);
    if ((double) roi_w + prm->roi_x > lim || (double) roi_h + prm->roi_y > lim || W > (1 << 20) ||
        fabs((double) axis) > lim)
        return ufomi_fail(UFOMI_ERR_ARG, "region of interest or axis beyond 2^20 pixels");
    if (!sino_pitch) sino_pitch = (size_t) P_in * W;
    if (!out_pitch) out_pitch = (size_t) roi_w * roi_h;
    // the staging loads address one slice's sinogram with 32-bit byte offsets
    if ((double) P_in * W * 4.0 >= 4294967296.0)
        return ufomi_fail(UFOMI_ERR_ARG, "a sinogram of %d x %d exceeds 4 GiB", P_in, W);

    float2 *trig = nullptr;
    if (fixed) {
        if ((size_t) P_in > kNoallocTrigMax)
            return ufomi_fail(UFOMI_ERR_SIZE,
                              "sinogram height %d exceeds the %zu angles of the preallocated backprojection scratch",
                              P_in, kNoallocTrigMax);
        trig = (float2 *) fixed;
        hipLaunchKernelGGL(trig_table_kernel, dim3((P_in + kThreads - 1) / kThreads), dim3(kThreads), 0,
                           ctx->stream, trig, P_in, prm->angle_offset, step, prm->offset, sin_sign);
        UFOMI_CHECK_LAUNCH();
    } else if ((rc = ufomi_bp_trig(ctx, prm->angle_offset, step, prm->offset, P_in, sin_sign, &trig))) {
        return rc;
    }

    if (prm->mode == UFOMI_BP_NEAREST) {
        if (dtype != UFOMI_F32)
            return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "nearest mode needs a float32 sinogram");
        dim3 grid((roi_w + 15) / 16, (roi_h + 15) / 16, n_slices);
        hipLaunchKernelGGL(bp_nearest_kernel, grid, dim3(kThreads), 0, ctx->stream,
                           (const float *) d_sino, (long) sino_pitch, W, P_in, trig, axis,
                           (float) prm->roi_x, (float) prm->roi_y, roi_w, roi_h, n_slices, d_out, (long) out_pitch);
        UFOMI_CHECK_LAUNCH();
        return UFOMI_OK;
    }
    if (prm->mode != UFOMI_BP_TEXTURE)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown backprojection mode %d", prm->mode);

    // Tile configuration: slices sharing a tile (S), angles per LDS chunk (A),
    // half-precision staging for f16 storage; ufomi_bp_configure overrides.
    // Measured on MI355X (2048 slices of 1800 x 2048): S = 8, A = 8 is the
    // fastest f32 configuration; half staging speeds f16 storage up further.
    int S = n_slices >= 8 ? 8 : (n_slices >= 4 ? 4 : (n_slices >= 2 ? 2 : 1));
    if (ctx->bp_slices && n_slices >= ctx->bp_slices) S = ctx->bp_slices;
    const bool h16 = dtype == UFOMI_F16 && S >= 4 && ctx->bp_half >= 0 && half_staging;
    // S = 8 with half staging keeps its registers only with 8-angle chunks
    const int A = (S == 8 && h16) ? 8 : (ctx->bp_chunk ? ctx->bp_chunk : (S >= 4 ? 8 : 16));
    const int tiles_x = (roi_w + kTile - 1) / kTile, tiles_y = (roi_h + kTile - 1) / kTile;
    const long nb = (long) tiles_x * tiles_y * ((n_slices + S - 1) / S);
    if (nb > 0x7fffffffL)
        return ufomi_fail(UFOMI_ERR_ARG, "too many tiles");
    // Few tiles (one 512^2 slice is 256 workgroups for 256 CUs): split the
    // angles over up to 8 launches' worth of workgroups, at least 64 angles
    // each, into partial volumes summed afterwards in a fixed order.
    int n_split = 1, p_chunk = 0;
    if (nb < 1024 && prm->mode == UFOMI_BP_TEXTURE) {
        n_split = (int) std::max<long>(1, std::min<long>((1024 + nb - 1) / nb, std::min(8, P_in / 64)));
        if (n_split > 1) {
            p_chunk = (P_in + n_split - 1) / n_split;
            n_split = (P_in + p_chunk - 1) / p_chunk;
        }
    }
    const long plane = (long) roi_w * roi_h;
    float *part = nullptr;
    if (n_split > 1 && fixed) {
        const size_t bytes = (size_t) n_split * n_slices * plane * sizeof(float);
        if (bytes > kNoallocPartBytes)
            return ufomi_fail(UFOMI_ERR_SIZE,
                              "angle-split partial volumes of %zu bytes exceed the %zu bytes of the preallocated "
                              "backprojection scratch", bytes, kNoallocPartBytes);
        part = (float *) (fixed + kNoallocTrigBytes);
    } else if (n_split > 1) {
        const size_t bytes = (size_t) n_split * n_slices * plane * sizeof(float);
        if (bytes > ctx->bp_part_bytes) {
            if (ctx->bp_part) {
                UFOMI_HIP(hipStreamSynchronize(ctx->stream));
                UFOMI_HIP(hipFree(ctx->bp_part));
                ctx->bp_part = nullptr;
                ctx->bp_part_bytes = 0;
            }
            UFOMI_HIP(hipMalloc(&ctx->bp_part, bytes));
            ctx->bp_part_bytes = bytes;
        }
        part = ctx->bp_part;
    }
    auto go = [&](auto s_const, auto a_const, auto h_const, auto *in) {
        constexpr int SS = decltype(s_const)::value, AA = decltype(a_const)::value;
        constexpr bool HH = decltype(h_const)::value;
        using InT = std::remove_const_t<std::remove_pointer_t<decltype(in)>>;
        if constexpr (HH && !(std::is_same<InT, __half>::value && SS % 4 == 0)) {
            return ufomi_fail(UFOMI_ERR_UNSUPPORTED, "half staging needs f16 storage and 4 | S");
        } else {
            hipLaunchKernelGGL((bp_texture_kernel<SS, AA, InT, HH>), dim3((unsigned) nb, (unsigned) n_split),
                               dim3(kThreads), 0, ctx->stream, in, (long) sino_pitch, W, P_in, trig, axis,
                               (float) prm->roi_x, (float) prm->roi_y, roi_w, roi_h, n_slices,
                               part ? part : d_out, part ? plane : (long) out_pitch, tiles_x, tiles_y, p_chunk);
            UFOMI_CHECK_LAUNCH();
            if (part) {
                const long total = plane * n_slices;
                const unsigned blocks = (unsigned) std::min<long>((total + kThreads - 1) / kThreads, 4096);
                hipLaunchKernelGGL(bp_sum_parts_kernel, dim3(blocks), dim3(kThreads), 0, ctx->stream, part, n_split,
                                   (long) n_slices * plane, plane, n_slices, d_out, (long) out_pitch);
                UFOMI_CHECK_LAUNCH();
            }
            return (int) UFOMI_OK;
        }
    };
    using I8 = std::integral_constant<int, 8>;
    using I16 = std::integral_constant<int, 16>;
    auto by_a = [&](auto s_const, auto h_const, auto *in) {
        return A == 8 ? go(s_const, I8(), h_const, in) : go(s_const, I16(), h_const, in);
    };
    auto by_s = [&](auto *in) {
        using F = std::false_type;
        using T = std::true_type;
        switch (S) {
        case 8: return h16 ? by_a(std::integral_constant<int, 8>(), T(), in) : by_a(std::integral_constant<int, 8>(), F(), in);
        case 4: return h16 ? by_a(std::integral_constant<int, 4>(), T(), in) : by_a(std::integral_constant<int, 4>(), F(), in);
        case 2: return by_a(std::integral_constant<int, 2>(), F(), in);
        default: return by_a(std::integral_constant<int, 1>(), F(), in);
        }
    };
    switch (dtype) {
    case UFOMI_F16: return by_s((const __half *) d_sino);
    case UFOMI_BF16: return by_s((const __hip_bfloat16 *) d_sino);
    default: return by_s((const float *) d_sino);
    }
}

}  // namespace

// Set-up of ufomi_bp_parallel_noalloc, called by ufomi_create with `device`
// current: this device's address of `scratch`, taken once per device.
int ufomi_bp_noalloc_setup(int device)
{
    if (device < 0 || device >= kNoallocDevices)
        return ufomi_fail(UFOMI_ERR_ARG, "device %d beyond the %d devices of the preallocated scratch table", device,
                          kNoallocDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc_ptr[device])
        return UFOMI_OK;
    void *ptr = nullptr;
    UFOMI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr)
        return ufomi_fail(UFOMI_ERR_HIP, "hipGetSymbolAddress returned no address for the backprojection scratch");
    g_noalloc_ptr[device] = (unsigned char *) ptr;
    return UFOMI_OK;
}

extern "C" int ufomi_bp_parallel_noalloc(ufomi_ctx *ctx, const void *d_sino, int dtype, int W, int P_in,
                                         int n_slices, size_t sino_pitch, const ufomi_bp_params *prm, float *d_out,
                                         size_t out_pitch)
{
    if (!ctx) return ufomi_fail(UFOMI_ERR_ARG, "context is NULL");
    if (ctx->device < 0 || ctx->device >= kNoallocDevices)
        return ufomi_fail(UFOMI_ERR_ARG, "device %d beyond the preallocated scratch table", ctx->device);
    unsigned char *fixed;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        fixed = g_noalloc_ptr[ctx->device];
    }
    if (!fixed)
        return ufomi_fail(UFOMI_ERR_ARG, "the preallocated scratch of device %d was not set up (ufomi_create)",
                          ctx->device);
    return bp_parallel_impl(ctx, d_sino, dtype, W, P_in, n_slices, sino_pitch, prm, d_out, out_pitch, +1, true, fixed);
}

// ---- persistent batch kernel launchers

namespace {

template <typename F>
int by_batch_config(int tile, int achunk, int dtype, F f)
{
    auto with_t = [&](auto *tag) {
        using InT = std::remove_pointer_t<decltype(tag)>;
        if (tile == 32 && achunk == 8) return f(bp_batch_kernel<32, 8, InT>);
        if (tile == 32 && achunk == 16) return f(bp_batch_kernel<32, 16, InT>);
        if (tile == 32 && achunk == 32) return f(bp_batch_kernel<32, 32, InT>);
        if (tile == 64 && achunk == 8) return f(bp_batch_kernel<64, 8, InT>);
        if (tile == 64 && achunk == 16) return f(bp_batch_kernel<64, 16, InT>);
        return ufomi_fail(UFOMI_ERR_ARG, "no persistent backprojection kernel for tile %d, angle chunk %d",
                          tile, achunk);
    };
    switch (dtype) {
    case UFOMI_F16: return with_t((__half *) nullptr);
    case UFOMI_BF16: return with_t((__hip_bfloat16 *) nullptr);
    case UFOMI_F32: return with_t((float *) nullptr);
    default: return ufomi_fail(UFOMI_ERR_ARG, "unknown sinogram dtype %d", dtype);
    }
}

}  // namespace

bool ufomi_bp_batch_config_ok(int tile, int achunk)
{
    return (tile == 32 && (achunk == 8 || achunk == 16 || achunk == 32)) ||
           (tile == 64 && (achunk == 8 || achunk == 16));
}

int ufomi_bp_batch_blocks_per_cu(int tile, int achunk, int dtype, int *blocks)
{
    return by_batch_config(tile, achunk, dtype, [&](auto kern) {
        int n = 0;
        UFOMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kThreads, 0));
        *blocks = n > 0 ? n : 1;
        return (int) UFOMI_OK;
    });
}

int ufomi_bp_batch_launch(hipStream_t stream, int tile, int achunk, int dtype, int grid,
                          const ufomi_bp_batch_req *d_reqs, const int2 *d_items, int n_items, int *d_counter,
                          const float2 *d_trig)
{
    if (grid <= 0 || n_items <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid persistent launch (%d workgroups, %d items)", grid, n_items);
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, d_counter);
    UFOMI_CHECK_LAUNCH();
    return by_batch_config(tile, achunk, dtype, [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, stream, d_reqs, d_items, n_items, d_counter, d_trig);
        UFOMI_CHECK_LAUNCH();
        return (int) UFOMI_OK;
    });
}
