// General backprojection (`general-backproject` task, parameter z) for gfx950.
//
// The reference generates an OpenCL kernel per geometry (make_kernel,
// ufo-general-backproject-task.c:1069-1180) that walks, per voxel and
// projection, the transformation chain of make_static_transformations and
// make_transformations (:807-1048): magnification, volume rotation,
// tomographic rotation, axis tilt, projection from the source onto the
// (possibly tilted, shifted) detector, then the detector frame.  Every step is
// affine in the voxel except the division by the projection depth, so per
// projection the chain is a projective map
//     u = U.X / D.X,  v = V.X / D.X,  coeff = 1 / C.X,   X = (x, y, z, 1)
// with U, V, D, C four affine rows.  The host evaluates the chain in FP64 at
// four points to obtain them exactly (up to FP64 rounding) and rounds once to
// float; the kernel then needs 12 (parallel beam) or 16 (cone) FMAs, one
// reciprocal and one bilinear fetch per voxel and projection, and keeps the
// whole projection sum of a voxel in a register (the reference re-reads and
// re-writes the volume every `burst` projections).
#include "ufomi_internal.h"

#include <cfloat>

#include <cmath>
#include <cstring>
#include <vector>

namespace {

struct Geo {
    double tomo, ax, ay, vx, vy, vz, dax, day, daz, cx, cz, sx, sy, sz, dx, dy, dz;
};

struct Flags {
    bool with_axis, with_volume, perpendicular, shifted_detector, shifted_source, parallel;
};

struct V3 {
    double x, y, z;
};

// general_bp_definitions.in:1-3 with trig = (sin, cos)
V3 rot_x(double s, double c, V3 p) { return {p.x, c * p.y - s * p.z, s * p.y + c * p.z}; }
V3 rot_y(double s, double c, V3 p) { return {c * p.x + s * p.z, p.y, -s * p.x + c * p.z}; }
V3 rot_z(double s, double c, V3 p) { return {c * p.x - s * p.y, s * p.x + c * p.y, p.z}; }

// Detector coordinates (u, v), projection depth d (u = U/d etc. with U, V affine)
// and 1/coeff of one voxel, FP64, following the generated kernel's statement order.
void chain(const Geo &g, const Flags &f, V3 X, double out[4])
{
    V3 v0 = X, n = {0.0, -1.0, 0.0};
    double det_off = 0.0, project_tmp = 0.0, tmp_tr = 0.0;
    if (!f.parallel) {
        const double m = -g.sy / (g.dy - g.sy);
        v0 = {v0.x * m, v0.y * m, v0.z * m};
    }
    if (!f.perpendicular) {
        n = rot_z(sin(g.daz), cos(g.daz), n);
        n = rot_y(sin(g.day), cos(g.day), n);
        n = rot_x(sin(g.dax), cos(g.dax), n);
        det_off = -(g.dx * n.x + g.dy * n.y + g.dz * n.z);
    } else if (!f.parallel) {
        project_tmp = g.dy - g.sy;
    }
    if (f.with_volume) {
        v0 = rot_z(sin(g.vz), cos(g.vz), v0);
        v0 = rot_y(sin(g.vy), cos(g.vy), v0);
        v0 = rot_x(sin(g.vx), cos(g.vx), v0);
    }
    if (!(f.perpendicular || f.parallel))
        tmp_tr = -(det_off + (g.sx * n.x + g.sy * n.y + g.sz * n.z));

    V3 v = rot_z(sin(g.tomo), cos(g.tomo), v0);
    if (f.with_axis) {
        v = rot_y(sin(g.ay), cos(g.ay), v);
        v = rot_x(sin(g.ax), cos(g.ax), v);
    }
    double inv_coeff = 1.0, depth = 1.0;  // depth = 1 / t of the projection step
    if (!f.parallel)
        inv_coeff = (g.sy - v.y) / (g.sy - g.dy);
    if (f.perpendicular) {
        if (!f.parallel) {
            depth = (v.y - g.sy) / project_tmp;
            const double t = 1.0 / depth;
            if (f.shifted_source)
                v = {t * (v.x - g.sx) + g.sx, t * (v.y - g.sy) + g.sy, t * (v.z - g.sz) + g.sz};
            else
                v = {t * v.x + g.sx, t * v.y + g.sy, t * v.z + g.sz};
        }
    } else {
        if (f.parallel) {
            v.y = -(v.z * n.z + (v.x * n.x + det_off)) / n.y;
        } else {
            v = {v.x - g.sx, v.y - g.sy, v.z - g.sz};
            depth = (v.x * n.x + v.y * n.y + v.z * n.z) / tmp_tr;
            const double t = 1.0 / depth;
            v = {t * v.x + g.sx, t * v.y + g.sy, t * v.z + g.sz};
        }
    }
    if (!f.perpendicular || f.shifted_detector) {
        v = {v.x - g.dx, v.y - g.dy, v.z - g.dz};
        if (!f.perpendicular) {
            v = rot_x(-sin(g.dax), cos(g.dax), v);
            v = rot_y(-sin(g.day), cos(g.day), v);
            v = rot_z(-sin(g.daz), cos(g.daz), v);
        }
    }
    out[0] = (v.x + g.cx) * depth;
    out[1] = (v.z + g.cz) * depth;
    out[2] = depth;
    out[3] = inv_coeff;
}

// Rows of the affine maps (U, V, D, C) from four evaluations of the chain.
void compose(const Geo &g, const Flags &f, double m[16])
{
    double o[4], ex[4], ey[4], ez[4];
    chain(g, f, {0, 0, 0}, o);
    chain(g, f, {1, 0, 0}, ex);
    chain(g, f, {0, 1, 0}, ey);
    chain(g, f, {0, 0, 1}, ez);
    for (int r = 0; r < 4; r++) {
        m[4 * r + 0] = ex[r] - o[r];
        m[4 * r + 1] = ey[r] - o[r];
        m[4 * r + 2] = ez[r] - o[r];
        m[4 * r + 3] = o[r];
    }
}

// Projection mode of a launch: affine (parallel beam), projective with only
// V depending on z (cone beam without tilts: U_z = D_z = C_z = 0, exactly, as
// composed), or general projective.
enum { MODE_PARALLEL = 0, MODE_CONE_ZSEP = 1, MODE_CONE = 2, MODE_PARALLEL_ZSEP = 3 };

// Horizontal half of a linear texture fetch: texel column i0 and weight a of
// read_imagef at u (unnormalised: sample u - 0.5), with the border rule.
template <int ADDR> struct Col {
    int i0, i1;
    float a;
    bool c0, c1;
    __device__ __forceinline__ void set(float u, int W)
    {
        const float fu = u - 0.5f, fl = floorf(fu);
        a = fu - fl;
        i0 = (int) fl;
        i1 = i0 + 1;
        if constexpr (ADDR == UFOMI_ADDR_CLAMP) {
            c0 = i0 >= 0 && i0 < W;
            c1 = i1 >= 0 && i1 < W;
        } else {
            i0 = min(max(i0, 0), W - 1);
            i1 = min(max(i1, 0), W - 1);
            c0 = c1 = true;
        }
    }
};

// Vertical half and the 4 taps; fv = v - 0.5.  Zero border (clamp) or clamp to edge.
template <int ADDR>
__device__ __forceinline__ float fetch(const float *__restrict__ img, int W, int H, const Col<ADDR> &c, float fv)
{
    const float fl = floorf(fv);
    const float b = fv - fl;
    int j0 = (int) fl, j1 = j0 + 1;
    float t00, t10, t01, t11;
    if constexpr (ADDR == UFOMI_ADDR_CLAMP) {
        const bool r0 = j0 >= 0 && j0 < H, r1 = j1 >= 0 && j1 < H;
        const float *p0 = img + (long) j0 * W, *p1 = img + (long) j1 * W;
        t00 = (r0 && c.c0) ? p0[c.i0] : 0.0f;
        t10 = (r0 && c.c1) ? p0[c.i1] : 0.0f;
        t01 = (r1 && c.c0) ? p1[c.i0] : 0.0f;
        t11 = (r1 && c.c1) ? p1[c.i1] : 0.0f;
    } else {
        j0 = min(max(j0, 0), H - 1);
        j1 = min(max(j1, 0), H - 1);
        const float *p0 = img + (long) j0 * W, *p1 = img + (long) j1 * W;
        t00 = p0[c.i0];
        t10 = p0[c.i1];
        t01 = p1[c.i0];
        t11 = p1[c.i1];
    }
    const float r0 = fmaf(c.a, t10 - t00, t00);
    const float r1 = fmaf(c.a, t11 - t01, t01);
    return fmaf(b, r1 - r0, r0);
}

template <typename ST>
__device__ __forceinline__ void store_voxel(ST *__restrict__ vol, long o, float r, float gmin, float gscale, float gmax)
{
    if constexpr (std::is_same<ST, float>::value) {
        vol[o] = r;
    } else if constexpr (std::is_same<ST, double>::value) {
        vol[o] = (double) r;
    } else if constexpr (std::is_same<ST, __half>::value) {
        vol[o] = __float2half_rn(r);
    } else {
        // integer stores: clamp(gray_limit.y * (norm * result - gray_limit.x), 0, max), truncated
        float g = gscale * (r - gmin);
        g = fminf(fmaxf(g, 0.0f), gmax);
        if constexpr (std::is_same<ST, uint32_t>::value)
            vol[o] = g >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t) g;
        else
            vol[o] = (ST) g;
    }
}

// ---------------------------------------------------------------------------
// LDS-staged kernel.  A workgroup owns a tile of TX x TY x TZ voxels (32 x 32
// x 8 for parallel beam; 16 x 16 x 16 for cone beam, whose magnification
// widens the footprint); thread (lx, ly) owns the columns (lx + 16 kx,
// ly + 16 ky) of TZ voxels.
//   * Per chunk of 32 projections, the workgroup projects the 8 tile corners
//     of each (the image of a box under a projective map with positive depth
//     is the hull of the corner images) and keeps the 32 detector footprints
//     in LDS; the next chunk's footprints are computed a chunk ahead.
//   * The footprint of projection p + 1 is loaded into registers (one texel
//     column and 12 rows per thread) while the voxels sample projection p
//     from LDS; the two boxes alternate, one barrier per projection.
//   * A bilinear sample reads its 4 texels with two ds_read2_b32.
// A footprint larger than the box (or a tile reaching behind the source)
// takes the global-memory path for that projection.
// ---------------------------------------------------------------------------
constexpr int kBoxW = 64, kBoxH = 48;  // texels per staged footprint
// Row pitch of a staged box: 72 floats, 8 banks (of 64) apart, so the rows a
// wave reads at once (each a window of ~9 banks) barely share banks (pitches
// of 70, 74 and 76 floats measured the same or slower, 80 much slower).
constexpr int kBoxS = kBoxW + 8;
constexpr int kChunk = 32;             // projections per footprint chunk
// Word 3 of a raw (untyped, unswizzled) buffer resource on gfx9-family GPUs: 32-bit data format,
// so loads past num_records return 0.
constexpr int kRawBufferFlags = 0x00020000;
constexpr int kRowsPerThread = kBoxH / 4;

struct BoxGeom {
    int i_lo, j_lo, w, h;  // w = 0: global path
};

// A box read back from LDS is the same for every lane: say so, so that the
// staging's row tests and row offsets are scalar.
__device__ __forceinline__ BoxGeom uniform(BoxGeom g)
{
    g.i_lo = __builtin_amdgcn_readfirstlane(g.i_lo);
    g.j_lo = __builtin_amdgcn_readfirstlane(g.j_lo);
    g.w = __builtin_amdgcn_readfirstlane(g.w);
    g.h = __builtin_amdgcn_readfirstlane(g.h);
    return g;
}

template <typename T> using Glb = __attribute__((address_space(1))) T;

template <int MODE> struct TileShape {
    static constexpr bool CONE = MODE == MODE_CONE_ZSEP || MODE == MODE_CONE;
    static constexpr int TX = CONE ? 16 : 32, TY = TX, TZ = CONE ? 16 : 8;
    static constexpr int KA = TX / 16, KB = TY / 16;
};

// Texel (row, col) of a staged box.  Rows and columns inside a box are small
// and non-negative, so a full-rate 24-bit multiply-add forms the index (a plain
// int multiply-add becomes a quarter-rate 64-bit one).
__device__ __forceinline__ const float *box_texel(const float *bx, int row, int col)
{
    return bx + __umul24((unsigned) row, (unsigned) kBoxS) + (unsigned) col;
}

// Keeps a wave-uniform value in a VGPR: a VALU operation takes one SGPR
// operand, so a product of two uniform values would need a move per use.
__device__ __forceinline__ float in_vgpr(float v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Projective (or affine) map of one projection at X = (x, y, z).
template <bool CONE>
__device__ __forceinline__ void map_point(const float4 &mu, const float4 &mv, const float4 &md, float x, float y,
                                          float z, float &u, float &v, float &d)
{
    u = fmaf(mu.x, x, fmaf(mu.y, y, fmaf(mu.z, z, mu.w)));
    v = fmaf(mv.x, x, fmaf(mv.y, y, fmaf(mv.z, z, mv.w)));
    d = 1.0f;
    if constexpr (CONE) {
        d = fmaf(md.x, x, fmaf(md.y, y, fmaf(md.z, z, md.w)));
        u = u / d;
        v = v / d;
    }
}

template <int MODE, int ADDR, typename ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void gbp_lds_kernel(
    const float *__restrict__ proj, int W, int H, int P, const float4 *__restrict__ mats, float x0, float xs,
    float y0, float ys, const float *__restrict__ zs, int nx, int ny, int nz, float norm, float gmin, float gscale,
    float gmax, ST *__restrict__ vol, int tiles_x, int tiles_y)
{
    using TS = TileShape<MODE>;
    constexpr bool CONE = TS::CONE;
    constexpr int TX = TS::TX, TY = TS::TY, TZ = TS::TZ, KA = TS::KA, KB = TS::KB;
    // U (and for cone D, C) independent of z: checked on the host for the ZSEP modes
    constexpr bool ZSEP = MODE == MODE_CONE_ZSEP || MODE == MODE_PARALLEL_ZSEP;
    __shared__ float box[2][kBoxH * kBoxS];
    __shared__ BoxGeom bgs[2][kChunk];

    const int tile = blockIdx.x;
    const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, tz = tile / (tiles_x * tiles_y);
    const int X0 = tx * TX, Y0 = ty * TY, Z0 = tz * TZ;
    // Wave w covers the 8 x 8 voxel block (w & 1, w >> 1) of the 16 x 16
    // thread grid: its reads spread over ~4 rows of ~8 columns at any angle,
    // which the box pitch puts in distinct banks.
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int lx = ((wv & 1) << 3) | (lane & 7), ly = ((wv >> 1) << 3) | (lane >> 3);

    float xv[KA], yv[KB], z[TZ];
#pragma unroll
    for (int k = 0; k < KA; k++) xv[k] = fmaf((float) (X0 + lx + 16 * k), xs, x0);
#pragma unroll
    for (int k = 0; k < KB; k++) yv[k] = fmaf((float) (Y0 + ly + 16 * k), ys, y0);
#pragma unroll
    for (int k = 0; k < TZ; k++) z[k] = zs[min(Z0 + k, nz - 1)];
    // tile corners (all TX x TY x TZ positions, also those beyond the volume)
    const float cxa = fmaf((float) X0, xs, x0), cxb = fmaf((float) (X0 + TX - 1), xs, x0);
    const float cya = fmaf((float) Y0, ys, y0), cyb = fmaf((float) (Y0 + TY - 1), ys, y0);
    const float cza = z[0], czb = z[TZ - 1];

    // Footprints of projections [c0, c0 + 32) into bgs[buf]: thread t maps
    // corner t % 8 of projection c0 + t / 8.
    auto footprints = [&](int c0, int buf) {
        const int pj = c0 + (tid >> 3), c = tid & 7;
        const bool live = pj < P;
        float u = 0.0f, v = 0.0f, d = 1.0f;
        if (live) {
            const float4 mu = mats[4 * pj + 0], mv = mats[4 * pj + 1], md = mats[4 * pj + 2];
            map_point<CONE>(mu, mv, md, (c & 1) ? cxb : cxa, (c & 2) ? cyb : cya, (c & 4) ? czb : cza, u, v, d);
        }
        float umin = u, umax = u, vmin = v, vmax = v, dmin = d;
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            umin = fminf(umin, __shfl_xor(umin, o, 8));
            umax = fmaxf(umax, __shfl_xor(umax, o, 8));
            vmin = fminf(vmin, __shfl_xor(vmin, o, 8));
            vmax = fmaxf(vmax, __shfl_xor(vmax, o, 8));
            dmin = fminf(dmin, __shfl_xor(dmin, o, 8));
        }
        if (c == 0) {
            BoxGeom g;
            const bool finite = umin > -1e6f && umax < 1e6f && vmin > -1e6f && vmax < 1e6f;
            g.i_lo = finite ? (int) floorf(umin - 0.5f) - 1 : 0;
            g.j_lo = finite ? (int) floorf(vmin - 0.5f) - 1 : 0;
            g.w = finite ? (int) floorf(umax - 0.5f) - g.i_lo + 2 : 0;
            g.h = finite ? (int) floorf(vmax - 0.5f) - g.j_lo + 2 : 0;
            if (!live || !(dmin > 0.0f) || g.w > kBoxW || g.h > kBoxH || g.w <= 0 || g.h <= 0)
                g.w = 0;
            bgs[buf][tid >> 3] = g;
        }
    };

    // Staging registers: texel column tid % 64, rows tid / 64 + 4 k.
    // (the wave index, hence every staged row of a wave, is uniform)
    const int sc = tid & 63, sr0 = __builtin_amdgcn_readfirstlane(tid >> 6);
    float st[kRowsPerThread];
    const long plane = (long) W * H;
    // A row of a box is one load per lane with no VALU work: each row is a buffer resource
    // over the detector row (scalar), and the lane adds its column offset.  A lane whose column
    // lies past the box's width, or off the detector with a zero border, reads at the range's
    // end, and a row off the detector with a zero border has an empty range: the hardware
    // returns 0 for those without a memory access.  Other addressing modes clamp rows and
    // columns.  Texels past the box's width or height are never sampled.
    auto stage_load = [&](int p, const BoxGeom &g) {
        const char *img = (const char *) (proj + (long) p * plane);
        int i = g.i_lo + sc;
        bool ci = sc < g.w;
        if constexpr (ADDR == UFOMI_ADDR_CLAMP)
            ci = ci && i >= 0 && i < W;
        else
            i = min(max(i, 0), W - 1);
        const unsigned ib = (unsigned) (ci ? i : W) * 4u;
#pragma unroll
        for (int k = 0; k < kRowsPerThread; k++) {
            const int r = sr0 + 4 * k;
            if (r >= g.h) continue;
            int j = g.j_lo + r;
            bool rin = true;
            if constexpr (ADDR == UFOMI_ADDR_CLAMP) {
                rin = j >= 0 && j < H;
                j = rin ? j : 0;
            } else {
                j = min(max(j, 0), H - 1);
            }
            // byte offsets inside one projection (< 4 GiB, checked on the host)
            const __amdgpu_buffer_rsrc_t row = __builtin_amdgcn_make_buffer_rsrc(
                (void *) (img + (unsigned) j * (unsigned) W * 4u), (short) 0, rin ? W * 4 : 0, kRawBufferFlags);
            st[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(row, ib, 0, 0));
        }
    };
    auto commit = [&](int buf, const BoxGeom &g) {
        // one address per projection; the rows are immediate offsets from it
        float *dst = &box[buf][sr0 * kBoxS + sc];
#pragma unroll
        for (int k = 0; k < kRowsPerThread; k++) {
            if (sr0 + 4 * k < g.h) dst[4 * k * kBoxS] = st[k];
        }
    };

    float acc[KA][KB][TZ];
#pragma unroll
    for (int a = 0; a < KA; a++)
#pragma unroll
        for (int b = 0; b < KB; b++)
#pragma unroll
            for (int k = 0; k < TZ; k++)
                acc[a][b][k] = 0.0f;

    footprints(0, 0);
    __syncthreads();
    {
        const BoxGeom g0 = uniform(bgs[0][0]);
        if (g0.w > 0) {
            stage_load(0, g0);
            commit(0, g0);
        }
    }
    __syncthreads();
    bool any_global = false;  // workgroup-uniform, like the boxes
    for (int p = 0; p < P; p++) {
        if (p % kChunk == 0 && p + kChunk < P + kChunk - 1)
            footprints(p + kChunk, ((p / kChunk) + 1) & 1);  // read a chunk later
        const BoxGeom g = uniform(bgs[(p / kChunk) & 1][p % kChunk]);
        const bool next = p + 1 < P;
        BoxGeom gn;
        gn.w = 0;
        if (next) {
            gn = uniform(bgs[((p + 1) / kChunk) & 1][(p + 1) % kChunk]);
            if (gn.w > 0) stage_load(p + 1, gn);
        }
        const float4 mu = mats[4 * p + 0], mv = mats[4 * p + 1], md = mats[4 * p + 2], mc = mats[4 * p + 3];
        if (g.w > 0) {
            const float *bx = box[p & 1];
            const float ilo = (float) g.i_lo + 0.5f, jlo = (float) g.j_lo + 0.5f;
#pragma unroll
            for (int a = 0; a < KA; a++) {
#pragma unroll
                for (int b = 0; b < KB; b++) {
                    const float x = xv[a], y = yv[b];
                    const float ub = fmaf(mu.x, x, fmaf(mu.y, y, mu.w));
                    const float vb = fmaf(mv.x, x, fmaf(mv.y, y, mv.w));
                    if constexpr (ZSEP) {
                        float rd = 1.0f, w = 1.0f;
                        if constexpr (CONE) {
                            rd = __builtin_amdgcn_rcpf(fmaf(md.x, x, fmaf(md.y, y, md.w)));
                            const float cf = __builtin_amdgcn_rcpf(fmaf(mc.x, x, fmaf(mc.y, y, mc.w)));
                            w = cf * cf;
                        }
                        const float uu = fmaf(ub, rd, -ilo);
                        const int ci = (int) uu;
                        const float fa = __builtin_amdgcn_fractf(uu);
                        // (z[k] is uniform: keep the other factor in a VGPR)
                        const float dv = CONE ? mv.z * rd : in_vgpr(mv.z), v0 = fmaf(vb, rd, -jlo);
#pragma unroll
                        for (int k = 0; k < TZ; k++) {
                            const float vv = fmaf(dv, z[k], v0);
                            const float *t = box_texel(bx, (int) vv, ci);
                            const float fb = __builtin_amdgcn_fractf(vv);
                            const float top = fmaf(fa, t[1] - t[0], t[0]);
                            const float bot = fmaf(fa, t[kBoxS + 1] - t[kBoxS], t[kBoxS]);
                            const float val = fmaf(fb, bot - top, top);
                            if constexpr (CONE)
                                acc[a][b][k] = fmaf(val, w, acc[a][b][k]);
                            else
                                acc[a][b][k] += val;
                        }
                    } else {
                        const float db = fmaf(md.x, x, fmaf(md.y, y, md.w));
                        const float cb = fmaf(mc.x, x, fmaf(mc.y, y, mc.w));
                        const float muz = in_vgpr(mu.z), mvz = in_vgpr(mv.z);
                        const float mdz = CONE ? in_vgpr(md.z) : 0.0f, mcz = CONE ? in_vgpr(mc.z) : 0.0f;
#pragma unroll
                        for (int k = 0; k < TZ; k++) {
                            float rd = 1.0f, w = 1.0f;
                            if constexpr (CONE) {
                                rd = __builtin_amdgcn_rcpf(fmaf(mdz, z[k], db));
                                const float cf = __builtin_amdgcn_rcpf(fmaf(mcz, z[k], cb));
                                w = cf * cf;
                            }
                            const float uu = fmaf(fmaf(muz, z[k], ub), rd, -ilo);
                            const float vv = fmaf(fmaf(mvz, z[k], vb), rd, -jlo);
                            const float *t = box_texel(bx, (int) vv, (int) uu);
                            const float fa = __builtin_amdgcn_fractf(uu), fb = __builtin_amdgcn_fractf(vv);
                            const float top = fmaf(fa, t[1] - t[0], t[0]);
                            const float bot = fmaf(fa, t[kBoxS + 1] - t[kBoxS], t[kBoxS]);
                            const float val = fmaf(fb, bot - top, top);
                            if constexpr (CONE)
                                acc[a][b][k] = fmaf(val, w, acc[a][b][k]);
                            else
                                acc[a][b][k] += val;
                        }
                    }
                }
            }
        } else {
            any_global = true;
        }
        if (next && gn.w > 0) commit((p + 1) & 1, gn);
        __syncthreads();  // box p + 1 staged; box p free for p + 2
    }

    // Projections whose footprint does not fit a box (too large, or the tile reaches behind
    // the source) take the global-memory path in a pass of their own: with it out of the main
    // loop, the accumulators stay in their registers across that loop's iterations (a branch
    // there that also updated them made the compiler copy them twice per projection).
    if (any_global) {
        for (int c0 = 0; c0 < P; c0 += kChunk) {
            footprints(c0, 0);  // the main loop's last barrier freed both footprint buffers
            __syncthreads();
            for (int p = c0; p < min(c0 + kChunk, P); p++) {
                if (uniform(bgs[0][p - c0]).w > 0) continue;
                const float4 mu = mats[4 * p + 0], mv = mats[4 * p + 1], md = mats[4 * p + 2], mc = mats[4 * p + 3];
                // One voxel per iteration of a rolled loop and a select chain into the
                // accumulators, so this rare path does not inflate the register budget.
                const float *img = proj + (long) p * plane;
#pragma unroll 1
                for (int q = 0; q < KA * KB * TZ; q++) {
                    const int a = q / (KB * TZ), b = (q / TZ) % KB, k = q % TZ;
                    const float x = a ? xv[KA - 1] : xv[0], y = b ? yv[KB - 1] : yv[0];
                    const float zz = zs[min(Z0 + k, nz - 1)];
                    float u, v, d;
                    map_point<CONE>(mu, mv, md, x, y, zz, u, v, d);
                    Col<ADDR> col;
                    col.set(u, W);
                    float val = fetch<ADDR>(img, W, H, col, v - 0.5f);
                    if constexpr (CONE) {
                        const float cf = 1.0f / fmaf(mc.x, x, fmaf(mc.y, y, fmaf(mc.z, zz, mc.w)));
                        val *= cf * cf;
                    }
#pragma unroll
                    for (int aa = 0; aa < KA; aa++)
#pragma unroll
                        for (int bb = 0; bb < KB; bb++)
#pragma unroll
                            for (int kk = 0; kk < TZ; kk++)
                                acc[aa][bb][kk] += (q == (aa * KB + bb) * TZ + kk) ? val : 0.0f;
                }
            }
            __syncthreads();  // bgs[0] read before the next chunk's footprints
        }
    }

#pragma unroll
    for (int a = 0; a < KA; a++) {
        const int ix = X0 + lx + 16 * a;
#pragma unroll
        for (int b = 0; b < KB; b++) {
            const int iy = Y0 + ly + 16 * b;
            if (ix < nx && iy < ny) {
#pragma unroll
                for (int k = 0; k < TZ; k++) {
                    const int iz = Z0 + k;
                    if (iz < nz)
                        store_voxel(vol, ((long) iz * ny + iy) * nx + ix, norm * acc[a][b][k], gmin, gscale, gmax);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Precise kernel: compute-type double, result-type half or double, and the
// repeat addressing modes.  One thread per voxel, global-memory gathers (these
// are options, not the hot path):
//   * CT = double keeps the composed FP64 rows and evaluates the projective map,
//     the cone weight and the bilinear weights in FP64 (the reference's double
//     kernel computes the chain in FP64 but rounds the sample coordinate to float
//     and interpolates in float, so this is at least as accurate);
//   * the running sum of each burst of `burst` projections is rounded to the
//     result type after every projection (`rtype result += ...`,
//     general_bp_body.in:7, 21) and norm * result is added to the volume in the
//     store type once per burst (:26-30), as the reference's burst launches do;
//     integer store types take the gray mapping of the whole sum once, as the
//     LDS kernel does;
//   * repeat and mirrored_repeat apply the OpenCL 1.2 normalised-coordinate
//     rules (section 8.2) to s = u / W, t = v / H; the reference's sampler uses
//     unnormalised coordinates, for which OpenCL leaves these modes undefined.
// ---------------------------------------------------------------------------
enum { RT_HALF = 0, RT_FLOAT = 1, RT_DOUBLE = 2 };

// double -> binary16 with one rounding: truncate to float, mark an inexact
// truncation in the last bit (round to odd), then round to nearest even.
__device__ __forceinline__ float round_half(double d)
{
    float f = __double2float_rz(d);
    if ((double) f != d)
        f = __uint_as_float(__float_as_uint(f) | 1u);
    return __half2float(__float2half_rn(f));
}
__device__ __forceinline__ float round_half(float f) { return __half2float(__float2half_rn(f)); }

template <typename CT> __device__ __forceinline__ void axis_taps(CT x, int w, int addr, int &i0, int &i1, CT &a,
                                                                 bool &c0, bool &c1)
{
    c0 = c1 = true;
    if (addr == UFOMI_ADDR_REPEAT || addr == UFOMI_ADDR_MIRRORED_REPEAT) {
        const CT s = x / (CT) w;
        CT u;
        if (addr == UFOMI_ADDR_REPEAT) {
            u = (s - floor(s)) * (CT) w;
        } else {
            const CT sp = (CT) 2 * rint((CT) 0.5 * s);
            u = fabs(s - sp) * (CT) w;
        }
        const CT f = floor(u - (CT) 0.5);
        i0 = (int) f;
        i1 = i0 + 1;
        a = (u - (CT) 0.5) - f;
        if (addr == UFOMI_ADDR_REPEAT) {
            if (i0 < 0) i0 += w;
            if (i1 > w - 1) i1 -= w;
        } else {
            i0 = max(i0, 0);
            i1 = min(i1, w - 1);
        }
        i0 = min(max(i0, 0), w - 1);  // u within [0, w]; keeps every index inside for any rounding
        i1 = min(max(i1, 0), w - 1);
        return;
    }
    // far coordinates stay far on their side (no overflow in the conversion)
    const CT f = fmin(fmax(floor(x - (CT) 0.5), (CT) -1e9), (CT) 1e9);
    a = (x - (CT) 0.5) - f;
    i0 = (int) f;
    i1 = i0 + 1;
    if (addr == UFOMI_ADDR_CLAMP) {
        c0 = i0 >= 0 && i0 < w;
        c1 = i1 >= 0 && i1 < w;
    } else {
        i0 = min(max(i0, 0), w - 1);
        i1 = min(max(i1, 0), w - 1);
    }
}

template <typename CT>
__device__ __forceinline__ CT sample_precise(const float *__restrict__ img, int W, int H, CT x, CT y, int addr)
{
    if (!isfinite(x) || !isfinite(y))
        return (CT) 0;
    int i0, i1, j0, j1;
    CT a, b;
    bool ci0, ci1, cj0, cj1;
    axis_taps(x, W, addr, i0, i1, a, ci0, ci1);
    axis_taps(y, H, addr, j0, j1, b, cj0, cj1);
    const float *r0 = img + (long) (cj0 ? j0 : 0) * W, *r1 = img + (long) (cj1 ? j1 : 0) * W;
    const CT t00 = (cj0 && ci0) ? (CT) r0[i0] : (CT) 0, t10 = (cj0 && ci1) ? (CT) r0[i1] : (CT) 0;
    const CT t01 = (cj1 && ci0) ? (CT) r1[i0] : (CT) 0, t11 = (cj1 && ci1) ? (CT) r1[i1] : (CT) 0;
    const CT top = t00 + a * (t10 - t00), bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

template <bool CONE, typename CT, int RT>
__global__ __launch_bounds__(256) void gbp_precise_kernel(
    const float *__restrict__ proj, int W, int H, int P, const CT *__restrict__ rows, CT x0, CT xs, CT y0, CT ys,
    const CT *__restrict__ zs, int nx, int ny, int nz, CT norm, int addr, int burst, int st, float gmin, float gscale,
    float gmax, void *__restrict__ vol)
{
    // the sum of a burst: in the wider of the compute and result types
    using SumT = typename std::conditional<std::is_same<CT, double>::value || RT == RT_DOUBLE, double, float>::type;
    const long n = (long) nx * ny * nz;
    const long plane = (long) W * H;
    for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < n; i += (long) gridDim.x * 256) {
        const int ix = (int) (i % nx), iy = (int) ((i / nx) % ny), iz = (int) (i / ((long) nx * ny));
        const CT x = x0 + (CT) ix * xs, y = y0 + (CT) iy * ys, z = zs[iz];
        SumT r = 0;
        float vf = 0.0f;   // volume in float (store float or half, the latter rounded to half)
        double vd = 0.0;   // volume in double (store double) or the whole sum (integer stores)
        for (int p = 0; p < P; p++) {
            const CT *m = rows + 16 * (long) p;
            CT u = m[0] * x + m[1] * y + m[2] * z + m[3];
            CT v = m[4] * x + m[5] * y + m[6] * z + m[7];
            CT w2 = 1;
            if constexpr (CONE) {
                const CT d = m[8] * x + m[9] * y + m[10] * z + m[11];
                u = u / d;
                v = v / d;
                const CT c = (CT) 1 / (m[12] * x + m[13] * y + m[14] * z + m[15]);
                w2 = c * c;
            }
            const CT contrib = sample_precise<CT>(proj + p * plane, W, H, u, v, addr) * w2;
            const SumT sum = (SumT) r + (SumT) contrib;
            if constexpr (RT == RT_HALF)
                r = (SumT) round_half(sum);
            else if constexpr (RT == RT_FLOAT)
                r = (SumT) (float) sum;
            else
                r = sum;
            if ((p + 1) % burst == 0 || p + 1 == P) {
                const SumT part = (SumT) norm * r;
                const bool first = p + 1 <= burst;
                if (st == UFOMI_ST_FLOAT)
                    vf = first ? (float) part : vf + (float) part;
                else if (st == UFOMI_ST_HALF)
                    vf = first ? round_half(part) : round_half(vf + round_half(part));
                else
                    vd = first ? (double) part : vd + (double) part;
                r = 0;
            }
        }
        switch (st) {
        case UFOMI_ST_FLOAT: ((float *) vol)[i] = vf; break;
        case UFOMI_ST_HALF: ((__half *) vol)[i] = __float2half_rn(vf); break;
        case UFOMI_ST_DOUBLE: ((double *) vol)[i] = vd; break;
        case UFOMI_ST_UCHAR: store_voxel((uint8_t *) vol, i, (float) vd, gmin, gscale, gmax); break;
        case UFOMI_ST_USHORT: store_voxel((uint16_t *) vol, i, (float) vd, gmin, gscale, gmax); break;
        default: store_voxel((uint32_t *) vol, i, (float) vd, gmin, gscale, gmax); break;
        }
    }
}

double at(const double *a, int n, int i, double dflt) { return (!a || n <= 0) ? dflt : (n == 1 ? a[0] : a[i]); }

bool almost_zero(const double *a, int n)
{
    if (!a || n <= 0)
        return true;
    for (int i = 0; i < n; i++)
        if (!(fabs(a[i]) < 1e-7))
            return false;
    return true;
}

int check_len(const char *name, int n, unsigned P)
{
    if (n > 1 && (unsigned) n < P)
        return ufomi_fail(UFOMI_ERR_ARG, "%s has %d values, needs 1 or %u", name, n, P);
    return UFOMI_OK;
}

int gbp_launch(ufomi_ctx *ctx, const float *d_proj, int W, int H, unsigned n, const std::vector<float> &mats,
               const Flags &f, int kaddr, const ufomi_gbp_params *prm, int nx, int ny, int nz, void *d_vol)
{
    int rc;
    const int st = prm->store_type;
    const bool integer_store = st == UFOMI_ST_UCHAR || st == UFOMI_ST_USHORT || st == UFOMI_ST_UINT;
    // compute_slice_region (:246-259): step 0 -> (-(float) length / 2, 1)
    const float x0 = prm->x_region[2] ? (float) prm->x_region[0] : -((float) nx / 2);
    const float xs = prm->x_region[2] ? (float) prm->x_region[2] : 1.0f;
    const float y0 = prm->y_region[2] ? (float) prm->y_region[0] : -((float) ny / 2);
    const float ys = prm->y_region[2] ? (float) prm->y_region[2] : 1.0f;
    // create_regions (:38-76): z_k = start + k * step in FP64, rounded.
    double zstart = 0.0, zstep = 1.0;
    if (!(fabs(prm->region[2]) < 1e-7)) {
        zstart = prm->region[0];
        zstep = prm->region[2];
    }
    std::vector<float> zs(nz);
    for (int k = 0; k < nz; k++)
        zs[k] = (float) (zstart + k * zstep);

    const float norm = (float) (fabs(prm->overall_angle) / n);
    double maxval = st == UFOMI_ST_UCHAR ? 255.0 : st == UFOMI_ST_USHORT ? 65535.0 : 4294967295.0;
    const float gmin = (float) prm->gray_map_min;
    const float gscale = integer_store ? (float) (maxval / (prm->gray_map_max - prm->gray_map_min)) : 0.0f;
    const float gmax = (float) maxval;

    const size_t mbytes = mats.size() * sizeof(float), zbytes = zs.size() * sizeof(float);
    char *scratch = nullptr;
    if ((rc = ufomi_scratch(ctx, mbytes + zbytes + 256, (void **) &scratch))) return rc;
    std::vector<char> host(mbytes + zbytes + 256);
    memcpy(host.data(), mats.data(), mbytes);
    const size_t zoff = (mbytes + 255) & ~(size_t) 255;
    memcpy(host.data() + zoff, zs.data(), zbytes);
    if ((rc = ufomi_upload(ctx, scratch, host.data(), zoff + zbytes))) return rc;

    // rows that do not depend on z (U; for cone also D and C) are shared along z
    bool zsep = true;
    for (size_t i = 0; i < n && zsep; i++)
        zsep = mats[16 * i + 2] == 0.0f && (f.parallel || (mats[16 * i + 10] == 0.0f && mats[16 * i + 14] == 0.0f));
    const int mode = f.parallel ? (zsep ? MODE_PARALLEL_ZSEP : MODE_PARALLEL) : (zsep ? MODE_CONE_ZSEP : MODE_CONE);
    const bool cone_tile = !f.parallel;
    const int TXl = cone_tile ? 16 : 32, TZl = cone_tile ? 16 : 8;
    const int tiles_x = (nx + TXl - 1) / TXl, tiles_y = (ny + TXl - 1) / TXl, tiles_z = (nz + TZl - 1) / TZl;
    const long n_tiles = (long) tiles_x * tiles_y * tiles_z;
    if (n_tiles > 0x7fffffffL)
        return ufomi_fail(UFOMI_ERR_ARG, "volume too large for one launch");
    // the staging loads index a projection with 24-bit factors and a 32-bit offset
    if (W >= (1 << 24) || H >= (1 << 24) || (long) W * H * 4 > 0xffffffffL)
        return ufomi_fail(UFOMI_ERR_ARG, "projections of %d x %d texels are too large", W, H);
    dim3 grid((unsigned) n_tiles);
    auto go = [&](auto mode_c, auto addr_c, auto *vol) {
        constexpr int C = decltype(mode_c)::value;
        constexpr int A = decltype(addr_c)::value;
        using ST = std::remove_pointer_t<decltype(vol)>;
        hipLaunchKernelGGL((gbp_lds_kernel<C, A, ST>), grid, dim3(256), 0, ctx->stream, d_proj, W, H, (int) n,
                           (const float4 *) scratch, x0, xs, y0, ys, (const float *) (scratch + zoff), nx, ny, nz,
                           norm, gmin, gscale, gmax, vol, tiles_x, tiles_y);
        UFOMI_CHECK_LAUNCH();
        return (int) UFOMI_OK;
    };
    auto by_store = [&](auto cone, auto addr_c) {
        switch (st) {
        case UFOMI_ST_HALF: return go(cone, addr_c, (__half *) d_vol);
        case UFOMI_ST_DOUBLE: return go(cone, addr_c, (double *) d_vol);
        case UFOMI_ST_UCHAR: return go(cone, addr_c, (uint8_t *) d_vol);
        case UFOMI_ST_USHORT: return go(cone, addr_c, (uint16_t *) d_vol);
        case UFOMI_ST_UINT: return go(cone, addr_c, (uint32_t *) d_vol);
        default: return go(cone, addr_c, (float *) d_vol);
        }
    };
    auto by_addr = [&](auto cone) {
        if (kaddr == UFOMI_ADDR_CLAMP)
            return by_store(cone, std::integral_constant<int, UFOMI_ADDR_CLAMP>());
        return by_store(cone, std::integral_constant<int, UFOMI_ADDR_CLAMP_TO_EDGE>());
    };
    switch (mode) {
    case MODE_PARALLEL: return by_addr(std::integral_constant<int, MODE_PARALLEL>());
    case MODE_PARALLEL_ZSEP: return by_addr(std::integral_constant<int, MODE_PARALLEL_ZSEP>());
    case MODE_CONE_ZSEP: return by_addr(std::integral_constant<int, MODE_CONE_ZSEP>());
    default: return by_addr(std::integral_constant<int, MODE_CONE>());
    }
}

// The precise kernel over FP64 rows (16 per projection) for one parameter-z call.
int gbp_precise_launch(ufomi_ctx *ctx, const float *d_proj, int W, int H, unsigned n, const std::vector<double> &rows,
                       bool parallel, int kaddr, const ufomi_gbp_params *prm, int nx, int ny, int nz, void *d_vol)
{
    int rc;
    const int st = prm->store_type;
    const bool integer_store = st == UFOMI_ST_UCHAR || st == UFOMI_ST_USHORT || st == UFOMI_ST_UINT;
    const bool f64 = prm->compute_type == UFOMI_CT_DOUBLE;
    const int burst = prm->burst ? (int) prm->burst : 8;
    double zstart = 0.0, zstep = 1.0;
    if (!(fabs(prm->region[2]) < 1e-7)) {
        zstart = prm->region[0];
        zstep = prm->region[2];
    }
    const double x0 = prm->x_region[2] ? prm->x_region[0] : -((double) nx / 2);
    const double xs = prm->x_region[2] ? prm->x_region[2] : 1.0;
    const double y0 = prm->y_region[2] ? prm->y_region[0] : -((double) ny / 2);
    const double ys = prm->y_region[2] ? prm->y_region[2] : 1.0;
    const double norm = fabs(prm->overall_angle) / n;
    const double maxval = st == UFOMI_ST_UCHAR ? 255.0 : st == UFOMI_ST_USHORT ? 65535.0 : 4294967295.0;
    const float gscale = integer_store ? (float) (maxval / (prm->gray_map_max - prm->gray_map_min)) : 0.0f;

    auto run = [&](auto ct_tag) {
        using CT = decltype(ct_tag);
        std::vector<CT> tab(rows.size() + (size_t) nz);
        for (size_t i = 0; i < rows.size(); i++) tab[i] = (CT) rows[i];
        for (int k = 0; k < nz; k++) tab[rows.size() + k] = (CT) (zstart + k * zstep);
        void *d_tab = nullptr;
        int rc2;
        if ((rc2 = ufomi_scratch(ctx, tab.size() * sizeof(CT), &d_tab))) return rc2;
        if ((rc2 = ufomi_upload(ctx, d_tab, tab.data(), tab.size() * sizeof(CT)))) return rc2;
        const CT *d_rows = (const CT *) d_tab, *d_zs = d_rows + rows.size();
        const long voxels = (long) nx * ny * nz;
        const unsigned grid = (unsigned) std::min<long>((voxels + 255) / 256, 16384);
        auto go = [&](auto cone_c, auto rt_c) {
            constexpr bool CONE = decltype(cone_c)::value;
            constexpr int RT = decltype(rt_c)::value;
            hipLaunchKernelGGL((gbp_precise_kernel<CONE, CT, RT>), dim3(grid), dim3(256), 0, ctx->stream, d_proj, W,
                               H, (int) n, d_rows, (CT) x0, (CT) xs, (CT) y0, (CT) ys, d_zs, nx, ny, nz, (CT) norm,
                               kaddr, burst, st, (float) prm->gray_map_min, gscale, (float) maxval, d_vol);
            UFOMI_CHECK_LAUNCH();
            return (int) UFOMI_OK;
        };
        auto by_rt = [&](auto cone_c) {
            switch (prm->result_type) {
            case UFOMI_RT_HALF: return go(cone_c, std::integral_constant<int, RT_HALF>());
            case UFOMI_RT_DOUBLE: return go(cone_c, std::integral_constant<int, RT_DOUBLE>());
            default: return go(cone_c, std::integral_constant<int, RT_FLOAT>());
            }
        };
        return parallel ? by_rt(std::false_type()) : by_rt(std::true_type());
    };
    rc = f64 ? run(0.0) : run(0.0f);
    return rc;
}

}  // namespace

extern "C" void ufomi_gbp_params_init(ufomi_gbp_params *p)
{
    memset(p, 0, sizeof(*p));
    p->overall_angle = 2.0 * M_PI;
    p->store_type = UFOMI_ST_FLOAT;
    p->addressing_mode = UFOMI_ADDR_CLAMP;
    p->compute_type = UFOMI_CT_FLOAT;
    p->result_type = UFOMI_RT_FLOAT;
    p->burst = 0;
}

extern "C" void ufomi_gbp_geometry_init(ufomi_gbp_geometry *g) { memset(g, 0, sizeof(*g)); }

extern "C" int ufomi_gbp_volume_shape(const ufomi_gbp_params *p, int W, int *nx, int *ny, int *nz)
{
    if (!p || W <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid volume shape arguments");
    // ufo-general-backproject-task.c:1494-1507 (REGION_SIZE) and :1516-1528
    *nx = p->x_region[2] == 0.0 ? W : (int) ceil((p->x_region[1] - p->x_region[0]) / p->x_region[2]);
    *ny = p->y_region[2] == 0.0 ? W : (int) ceil((p->y_region[1] - p->y_region[0]) / p->y_region[2]);
    if (fabs(p->region[2]) < 1e-7)
        *nz = 1;
    else
        *nz = (int) ceil((p->region[1] - p->region[0]) / p->region[2]);
    if (*nx <= 0 || *ny <= 0 || *nz <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "empty region");
    return UFOMI_OK;
}

extern "C" int ufomi_gbp(ufomi_ctx *ctx, const float *d_proj, int W, int H, int P, const ufomi_gbp_geometry *g,
                         const ufomi_gbp_params *prm, void *d_vol)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_proj || !d_vol || !g || !prm || W <= 0 || H <= 0 || P <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid general backprojection arguments");
    const unsigned n = prm->num_projections;
    if (!n)
        return ufomi_fail(UFOMI_ERR_NUM_PROJECTIONS, "Number of projections not set");
    if ((unsigned) P < n)
        return ufomi_fail(UFOMI_ERR_INCOMPLETE,
                          "general-backproject received only %d projections out of %u specified", P, n);
    if ((unsigned) P > n)
        return ufomi_fail(UFOMI_ERR_ARG, "%d projections given but num-projections is %u", P, n);
    const int st = prm->store_type;
    if (st < UFOMI_ST_HALF || st > UFOMI_ST_UINT)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown store type %d", st);
    const bool integer_store = st == UFOMI_ST_UCHAR || st == UFOMI_ST_USHORT || st == UFOMI_ST_UINT;
    if (integer_store && prm->gray_map_min >= prm->gray_map_max)
        return ufomi_fail(UFOMI_ERR_GRAY_MAP, "Gray mapping minimum must be less then the maximum");
    int addr = prm->addressing_mode;
    if (addr < UFOMI_ADDR_NONE || addr > UFOMI_ADDR_MIRRORED_REPEAT)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown addressing mode %d", addr);
    if (prm->compute_type != UFOMI_CT_FLOAT && prm->compute_type != UFOMI_CT_DOUBLE)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown compute type %d", prm->compute_type);
    if (prm->result_type < UFOMI_RT_HALF || prm->result_type > UFOMI_RT_DOUBLE)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown result type %d", prm->result_type);
    if (prm->burst > 128)
        return ufomi_fail(UFOMI_ERR_ARG, "burst %u above 128", prm->burst);
    // "none" leaves out-of-range reads undefined in OpenCL; clamp to the edge here.
    const int kaddr = addr == UFOMI_ADDR_NONE ? UFOMI_ADDR_CLAMP_TO_EDGE : addr;
    const bool precise = prm->compute_type == UFOMI_CT_DOUBLE || prm->result_type != UFOMI_RT_FLOAT ||
                         kaddr == UFOMI_ADDR_REPEAT || kaddr == UFOMI_ADDR_MIRRORED_REPEAT;

    std::vector<double> rows(16 * (size_t) n);
    int parallel = 0;
    if ((rc = ufomi_gbp_projection_rows(g, prm, rows.data(), &parallel))) return rc;
    if (precise) {
        int nx, ny, nz;
        if ((rc = ufomi_gbp_volume_shape(prm, W, &nx, &ny, &nz))) return rc;
        return gbp_precise_launch(ctx, d_proj, W, H, n, rows, parallel != 0, kaddr, prm, nx, ny, nz, d_vol);
    }
    std::vector<float> mats(rows.size());
    for (size_t i = 0; i < rows.size(); i++)
        mats[i] = (float) rows[i];
    const Flags f = {false, false, false, false, false, parallel != 0};

    int nx, ny, nz;
    if ((rc = ufomi_gbp_volume_shape(prm, W, &nx, &ny, &nz))) return rc;
    if (nz > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "at most 65535 slices per call");
    return gbp_launch(ctx, d_proj, W, H, n, mats, f, kaddr, prm, nx, ny, nz, d_vol);
}

// The geometry array a swept parameter replaces (the reference's parameter
// enum order, ufo-general-backproject-task.c:345-364).
static bool sweep_field(ufomi_gbp_geometry *g, int parameter, const double ***ptr, int **n)
{
    switch (parameter) {
    case 0: *ptr = &g->axis_angle_x; *n = &g->n_axis_angle_x; return true;
    case 1: *ptr = &g->axis_angle_y; *n = &g->n_axis_angle_y; return true;
    case 3: *ptr = &g->volume_angle_x; *n = &g->n_volume_angle_x; return true;
    case 4: *ptr = &g->volume_angle_y; *n = &g->n_volume_angle_y; return true;
    case 5: *ptr = &g->volume_angle_z; *n = &g->n_volume_angle_z; return true;
    case 6: *ptr = &g->detector_angle_x; *n = &g->n_detector_angle_x; return true;
    case 7: *ptr = &g->detector_angle_y; *n = &g->n_detector_angle_y; return true;
    case 8: *ptr = &g->detector_angle_z; *n = &g->n_detector_angle_z; return true;
    case 9: *ptr = &g->detector_position_x; *n = &g->n_detector_position_x; return true;
    case 10: *ptr = &g->detector_position_y; *n = &g->n_detector_position_y; return true;
    case 11: *ptr = &g->detector_position_z; *n = &g->n_detector_position_z; return true;
    case 12: *ptr = &g->source_position_x; *n = &g->n_source_position_x; return true;
    case 13: *ptr = &g->source_position_y; *n = &g->n_source_position_y; return true;
    case 14: *ptr = &g->source_position_z; *n = &g->n_source_position_z; return true;
    case 15: *ptr = &g->center_position_x; *n = &g->n_center_position_x; return true;
    case 16: *ptr = &g->center_position_z; *n = &g->n_center_position_z; return true;
    default: return false;
    }
}

extern "C" int ufomi_gbp_sweep(ufomi_ctx *ctx, const float *d_proj, int W, int H, int P, const ufomi_gbp_geometry *g,
                               const ufomi_gbp_params *prm, int parameter, void *d_vol)
{
    if (parameter == UFOMI_GBP_PARAMETER_Z)
        return ufomi_gbp(ctx, d_proj, W, H, P, g, prm, d_vol);
    if (!g || !prm || !d_vol)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid general backprojection arguments");
    if (parameter == 2) {
        // The generated kernel rotates by its tomographic-angle arguments and never
        // reads axis_angle_z.  Without per-projection geometry (the scalar kernel,
        // general_bp_header_scalar.in) the assignment `axis_angle_z = region[idz]`
        // (make_parameter_assignment, :762-764) names an undeclared variable and the
        // kernel fails to build; with it (:1103-1116, the vectorized kernel) the
        // variable is declared and unused, so every slice is the reconstruction at z.
        const unsigned np = prm->num_projections;
        const int counts[] = {g->n_axis_angle_x, g->n_axis_angle_y, g->n_volume_angle_x, g->n_volume_angle_y,
                              g->n_volume_angle_z, g->n_detector_angle_x, g->n_detector_angle_y,
                              g->n_detector_angle_z, g->n_detector_position_x, g->n_detector_position_y,
                              g->n_detector_position_z, g->n_source_position_x, g->n_source_position_y,
                              g->n_source_position_z, g->n_center_position_x, g->n_center_position_z};
        bool vectorized = false;
        for (int c : counts)
            vectorized = vectorized || (np > 0 && c == (int) np);
        if (!vectorized)
            return ufomi_fail(UFOMI_ERR_UNSUPPORTED,
                              "parameter axis-angle-z without per-projection geometry: the reference's kernel "
                              "assigns the undeclared axis_angle_z and does not build");
        int nx, ny, nz, rc;
        if ((rc = ufomi_gbp_volume_shape(prm, W, &nx, &ny, &nz))) return rc;
        static const size_t st_bytes[] = {2, 4, 8, 1, 2, 4};
        if (prm->store_type < UFOMI_ST_HALF || prm->store_type > UFOMI_ST_UINT)
            return ufomi_fail(UFOMI_ERR_ARG, "unknown store type %d", prm->store_type);
        const size_t slice = (size_t) nx * ny * st_bytes[prm->store_type];
        ufomi_gbp_params pk = *prm;
        pk.region[0] = prm->z;
        pk.region[1] = prm->z + 1.0;
        pk.region[2] = 1.0;
        if ((rc = ufomi_gbp(ctx, d_proj, W, H, P, g, &pk, d_vol))) return rc;
        for (int k = 1; k < nz; k++)
            UFOMI_HIP(hipMemcpyAsync((char *) d_vol + (size_t) k * slice, d_vol, slice, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        return UFOMI_OK;
    }
    ufomi_gbp_geometry gk = *g;
    const double **field;
    int *count;
    if (!sweep_field(&gk, parameter, &field, &count))
        return ufomi_fail(UFOMI_ERR_ARG, "unknown parameter %d", parameter);
    int nx, ny, nz, rc;
    if ((rc = ufomi_gbp_volume_shape(prm, W, &nx, &ny, &nz))) return rc;
    static const size_t st_size[] = {2, 4, 8, 1, 2, 4};
    if (prm->store_type < UFOMI_ST_HALF || prm->store_type > UFOMI_ST_UINT)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown store type %d", prm->store_type);
    const size_t slice = (size_t) nx * ny * st_size[prm->store_type];
    // slice k: the parameter takes region[0] + k region[2] for every
    // projection (make_parameter_assignment, :755-775); voxels lie at z
    ufomi_gbp_params pk = *prm;
    pk.region[0] = prm->z;
    pk.region[1] = prm->z + 1.0;
    pk.region[2] = 1.0;
    for (int k = 0; k < nz; k++) {
        const double value = prm->region[0] + (double) k * prm->region[2];
        *field = &value;
        *count = 1;
        if ((rc = ufomi_gbp(ctx, d_proj, W, H, P, &gk, &pk, (char *) d_vol + (size_t) k * slice))) return rc;
    }
    return UFOMI_OK;
}

extern "C" void ufomi_lamino_params_init(ufomi_lamino_params *p)
{
    memset(p, 0, sizeof(*p));
    p->overall_angle = M_PI;
    p->tomo_angle = -FLT_MAX;
    p->parameter = UFOMI_LAMINO_Z;
    p->addressing_mode = UFOMI_ADDR_CLAMP;
}

extern "C" int ufomi_lamino_volume_shape(const ufomi_lamino_params *p, int *nx, int *ny, int *nz)
{
    if (!p || !nx || !ny || !nz)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid arguments");
    // REGION_SIZE and get_requisition (ufo-lamino-backproject-task.c:41-43, 247-262)
    auto size = [](const int *r) { return r[2] == 0 ? 0 : (r[1] - r[0] - 1) / r[2] + 1; };
    *nx = size(p->x_region);
    *ny = size(p->y_region);
    const float start = (float) p->region[0], stop = (float) p->region[1], step = (float) p->region[2];
    *nz = step == 0.0f ? 0 : (int) ceil((double) ((stop - start) / step));
    if (*nx <= 0 || *ny <= 0 || *nz <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "lamino-backproject: x-region, y-region and region must each give at "
                                         "least one voxel");
    return UFOMI_OK;
}

// The map of lamino_template.in / z_template.in / center_template.in /
// roll_template.in and rotate() (definitions.in) for tomographic angle t, as
// rows over (x, y, z, 1):
//   a = x cos t + y sin t + xc - xr,  b = x cos l sin t - y cos l cos t + z sin l,
//   a' = a cos r + b sin r,  b' = -a' sin r + b cos r   (rotate() reuses the new a'),
//   u = a' + xr,  v = b' + yc
// with xc the (swept) x centre, xr the x centre rotate() pivots on (the first
// value of a sweep) and r = -roll.
static void lamino_rows(double t, double xc, double xr, double yc, double lam, double roll, double *m)
{
    const double ct = cos(t), st = sin(t), cl = cos(lam), sl = sin(lam), cr = cos(-roll), sr = sin(-roll);
    const double a[4] = {ct, st, 0.0, xc - xr}, b[4] = {cl * st, -cl * ct, sl, 0.0};
    for (int k = 0; k < 4; k++) {
        const double a2 = a[k] * cr + b[k] * sr;
        const double b2 = -a2 * sr + b[k] * cr;
        m[k] = a2 + (k == 3 ? xr : 0.0);
        m[4 + k] = b2 + (k == 3 ? yc : 0.0);
        m[8 + k] = k == 3 ? 1.0 : 0.0;
        m[12 + k] = k == 3 ? 1.0 : 0.0;
    }
}

extern "C" int ufomi_lamino_bp(ufomi_ctx *ctx, const float *d_proj, int W, int H, int P, const ufomi_lamino_params *p,
                               float *d_vol)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!d_proj || !d_vol || !p || W <= 0 || H <= 0 || P <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid lamino backprojection arguments");
    if (!p->num_projections)
        return ufomi_fail(UFOMI_ERR_NUM_PROJECTIONS, "Number of projections not set");
    if (p->parameter < UFOMI_LAMINO_Z || p->parameter > UFOMI_LAMINO_ROLL_ANGLE)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown parameter %d", p->parameter);
    const int addr = p->addressing_mode;
    if (addr == UFOMI_ADDR_REPEAT || addr == UFOMI_ADDR_MIRRORED_REPEAT)
        return ufomi_fail(UFOMI_ERR_UNSUPPORTED,
                          "repeat addressing needs normalized coordinates; unnormalized sampling is undefined");
    if (addr != UFOMI_ADDR_CLAMP && addr != UFOMI_ADDR_CLAMP_TO_EDGE && addr != UFOMI_ADDR_NONE)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown addressing mode %d", addr);
    const int kaddr = addr == UFOMI_ADDR_CLAMP ? UFOMI_ADDR_CLAMP : UFOMI_ADDR_CLAMP_TO_EDGE;
    int nx, ny, nz;
    if ((rc = ufomi_lamino_volume_shape(p, &nx, &ny, &nz))) return rc;
    if (nz > 65535)
        return ufomi_fail(UFOMI_ERR_ARG, "at most 65535 slices per call");

    // process (:336-381): tomographic angle tomo-angle if set, else overall * i / n in float
    std::vector<double> tomo(P);
    for (int i = 0; i < P; i++)
        tomo[i] = p->tomo_angle > -FLT_MAX ? (double) (float) p->tomo_angle
                                           : (double) ((float) p->overall_angle * (float) i / (float) p->num_projections);
    const double xc0 = p->center[0] - p->projection_offset[0], yc = p->center[1] - p->projection_offset[1];
    ufomi_gbp_params g;
    ufomi_gbp_params_init(&g);
    g.num_projections = (unsigned) P;
    g.overall_angle = fabs(p->overall_angle) * P / p->num_projections;  // norm |overall| / num-projections
    g.x_region[0] = p->x_region[0];
    g.x_region[1] = p->x_region[1];
    g.x_region[2] = p->x_region[2];
    g.y_region[0] = p->y_region[0];
    g.y_region[1] = p->y_region[1];
    g.y_region[2] = p->y_region[2];
    Flags f = {false, false, false, false, false, true};
    std::vector<float> mats(16 * (size_t) P);
    double m[16];
    if (p->parameter == UFOMI_LAMINO_Z) {
        for (int k = 0; k < 3; k++)
            g.region[k] = p->region[k];
        for (int i = 0; i < P; i++) {
            lamino_rows(tomo[i], xc0, xc0, yc, p->lamino_angle, p->roll_angle, m);
            for (int k = 0; k < 16; k++)
                mats[16 * (size_t) i + k] = (float) m[k];
        }
        return gbp_launch(ctx, d_proj, W, H, (unsigned) P, mats, f, kaddr, &g, nx, ny, nz, d_vol);
    }
    // a swept parameter: slice k at z with the parameter at region[0] + k region[2]
    g.region[0] = p->z;
    g.region[1] = p->z + 1.0;
    g.region[2] = 1.0;
    const double xr = p->parameter == UFOMI_LAMINO_X_CENTER ? p->region[0] - p->projection_offset[0] : xc0;
    for (int s = 0; s < nz; s++) {
        const double v = p->region[0] + (double) s * p->region[2];
        const double xc = p->parameter == UFOMI_LAMINO_X_CENTER ? v - p->projection_offset[0] : xc0;
        const double lam = p->parameter == UFOMI_LAMINO_LAMINO_ANGLE ? v : p->lamino_angle;
        const double roll = p->parameter == UFOMI_LAMINO_ROLL_ANGLE ? v : p->roll_angle;
        for (int i = 0; i < P; i++) {
            lamino_rows(tomo[i], xc, xr, yc, lam, roll, m);
            for (int k = 0; k < 16; k++)
                mats[16 * (size_t) i + k] = (float) m[k];
        }
        if ((rc = gbp_launch(ctx, d_proj, W, H, (unsigned) P, mats, f, kaddr, &g, nx, ny, 1,
                             d_vol + (size_t) s * nx * ny)))
            return rc;
    }
    return UFOMI_OK;
}

extern "C" int ufomi_gbp_projection_rows(const ufomi_gbp_geometry *g, const ufomi_gbp_params *prm, double *rows,
                                         int *parallel_beam)
{
    if (!g || !prm || !rows)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid arguments");
    const unsigned n = prm->num_projections;
    if (!n)
        return ufomi_fail(UFOMI_ERR_NUM_PROJECTIONS, "Number of projections not set");
    int rc;
    struct { const char *name; int n; } lens[] = {
        {"axis-angle-x", g->n_axis_angle_x}, {"axis-angle-y", g->n_axis_angle_y},
        {"volume-angle-x", g->n_volume_angle_x}, {"volume-angle-y", g->n_volume_angle_y},
        {"volume-angle-z", g->n_volume_angle_z}, {"detector-angle-x", g->n_detector_angle_x},
        {"detector-angle-y", g->n_detector_angle_y}, {"detector-angle-z", g->n_detector_angle_z},
        {"center-position-x", g->n_center_position_x}, {"center-position-z", g->n_center_position_z},
        {"source-position-x", g->n_source_position_x}, {"source-position-y", g->n_source_position_y},
        {"source-position-z", g->n_source_position_z}, {"detector-position-x", g->n_detector_position_x},
        {"detector-position-y", g->n_detector_position_y}, {"detector-position-z", g->n_detector_position_z}};
    for (auto &l : lens)
        if ((rc = check_len(l.name, l.n, n))) return rc;

    // Flags of node_setup (ufo-general-backproject-task.c:1305-1336).
    Flags f;
    f.with_axis = !(almost_zero(g->axis_angle_x, g->n_axis_angle_x) && almost_zero(g->axis_angle_y, g->n_axis_angle_y));
    f.with_volume = !(almost_zero(g->volume_angle_x, g->n_volume_angle_x) &&
                      almost_zero(g->volume_angle_y, g->n_volume_angle_y) &&
                      almost_zero(g->volume_angle_z, g->n_volume_angle_z));
    f.shifted_detector = !(almost_zero(g->detector_position_x, g->n_detector_position_x) &&
                           almost_zero(g->detector_position_z, g->n_detector_position_z));
    f.shifted_source = !(almost_zero(g->source_position_x, g->n_source_position_x) &&
                         almost_zero(g->source_position_z, g->n_source_position_z));
    f.perpendicular = almost_zero(g->detector_angle_x, g->n_detector_angle_x) &&
                      almost_zero(g->detector_angle_y, g->n_detector_angle_y) &&
                      almost_zero(g->detector_angle_z, g->n_detector_angle_z);
    f.parallel = true;
    for (unsigned i = 0; i < n; i++)
        f.parallel = f.parallel && std::isinf(at(g->source_position_y, g->n_source_position_y, i, -INFINITY));

    // Tomographic angles: axis-angle-z if it has n values, else i / n * overall (:1435-1446).
    const bool own_tomo = g->axis_angle_z && g->n_axis_angle_z == (int) n;
    for (unsigned i = 0; i < n; i++) {
        Geo q;
        q.tomo = own_tomo ? g->axis_angle_z[i] : ((double) i) / n * prm->overall_angle;
        q.ax = at(g->axis_angle_x, g->n_axis_angle_x, i, 0.0);
        q.ay = at(g->axis_angle_y, g->n_axis_angle_y, i, 0.0);
        q.vx = at(g->volume_angle_x, g->n_volume_angle_x, i, 0.0);
        q.vy = at(g->volume_angle_y, g->n_volume_angle_y, i, 0.0);
        q.vz = at(g->volume_angle_z, g->n_volume_angle_z, i, 0.0);
        q.dax = at(g->detector_angle_x, g->n_detector_angle_x, i, 0.0);
        q.day = at(g->detector_angle_y, g->n_detector_angle_y, i, 0.0);
        q.daz = at(g->detector_angle_z, g->n_detector_angle_z, i, 0.0);
        q.cx = at(g->center_position_x, g->n_center_position_x, i, 0.0);
        q.cz = at(g->center_position_z, g->n_center_position_z, i, 0.0);
        q.sx = at(g->source_position_x, g->n_source_position_x, i, 0.0);
        q.sy = at(g->source_position_y, g->n_source_position_y, i, -INFINITY);
        q.sz = at(g->source_position_z, g->n_source_position_z, i, 0.0);
        q.dx = at(g->detector_position_x, g->n_detector_position_x, i, 0.0);
        q.dy = at(g->detector_position_y, g->n_detector_position_y, i, 0.0);
        q.dz = at(g->detector_position_z, g->n_detector_position_z, i, 0.0);
        if (!f.parallel && !(std::isfinite(q.sy) && std::isfinite(q.dy) && q.dy != q.sy))
            return ufomi_fail(UFOMI_ERR_ARG, "cone beam needs finite source and detector y positions");
        compose(q, f, rows + 16 * (size_t) i);
    }
    if (parallel_beam)
        *parallel_beam = f.parallel ? 1 : 0;
    return UFOMI_OK;
}
