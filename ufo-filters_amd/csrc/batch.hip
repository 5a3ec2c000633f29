// Batched FBP of uneven single-slice requests (ufomi_fbp_batch_*).
//
// A batch owns, in HBM: the per-request descriptors, the concatenated FP64
// trig tables (rounded to float2), one buffer of filtered rows in the storage
// type (each request's rows at a 256-byte aligned offset), per FFT size the
// coefficients and the segment table, the (request, tile) work list and the
// persistent kernel's work counter.  One HIP graph replays:
//   ramp16_segments_kernel<N>  x (number of distinct FFT sizes)
//   reset_counter_kernel + bp_batch_kernel<TILE, A>  (one persistent launch)
// The graph holds kernel nodes only (the counter is cleared by a kernel, not
// a memset node).  The autotuner times the persistent kernel for every tile /
// angle-chunk pair on the batch itself and caches the winner per process
// under a signature of the batch's shapes.
#include "ufomi_internal.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

struct Group {
    int N = 0;
    int pair = 0;
    float *d_coef = nullptr;
    ufomi_ramp_seg *d_segs = nullptr;
    int nseg = 0;
    long nblocks = 0;
};

std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tune;

const int kCandidates[][2] = {{32, 8}, {32, 16}, {32, 32}, {64, 8}, {64, 16}};

}  // namespace

struct ufomi_fbp_batch {
    ufomi_ctx *ctx = nullptr;
    int n = 0;
    int dtype = UFOMI_F32;
    std::vector<ufomi_bp_batch_req> reqs;
    ufomi_bp_batch_req *d_reqs = nullptr;
    float2 *d_trig = nullptr;
    void *d_filtered = nullptr;
    size_t filtered_bytes = 0;
    int *d_counter = nullptr;
    int2 *d_items = nullptr;
    size_t items_cap = 0;
    int n_items = 0;
    std::vector<Group> groups;
    int tile = 0, achunk = 0, grid = 0;
    double updates = 0.0;
    std::string signature;
    hipGraphExec_t exec = nullptr;
};

namespace {

int free_batch(ufomi_fbp_batch *b)
{
    if (b->exec) (void) hipGraphExecDestroy(b->exec);
    for (Group &g : b->groups) {
        if (g.d_coef) (void) hipFree(g.d_coef);
        if (g.d_segs) (void) hipFree(g.d_segs);
    }
    if (b->d_reqs) (void) hipFree(b->d_reqs);
    if (b->d_trig) (void) hipFree(b->d_trig);
    if (b->d_filtered) (void) hipFree(b->d_filtered);
    if (b->d_counter) (void) hipFree(b->d_counter);
    if (b->d_items) (void) hipFree(b->d_items);
    delete b;
    return UFOMI_OK;
}

// Kernels of `stages` (1 filters, 2 backprojection) on `stream`.
int enqueue(ufomi_fbp_batch *b, hipStream_t stream, int stages)
{
    int rc;
    if (stages & 1)
        for (const Group &g : b->groups)
            if ((rc = ufomi_ramp_segments_launch(stream, b->ctx->twiddles, g.N, b->dtype, g.d_segs, g.nseg,
                                                 g.nblocks, g.d_coef, g.pair)))
                return rc;
    if (stages & 2)
        return ufomi_bp_batch_launch(stream, b->tile, b->achunk, b->dtype, b->grid, b->d_reqs, b->d_items,
                                     b->n_items, b->d_counter, b->d_trig);
    return UFOMI_OK;
}

// Work list for (tile, achunk): requests by decreasing angle count, each
// request's tiles in row-major order.
int set_items(ufomi_fbp_batch *b, int tile, int achunk)
{
    if (!ufomi_bp_batch_config_ok(tile, achunk))
        return ufomi_fail(UFOMI_ERR_ARG, "no persistent backprojection kernel for tile %d, angle chunk %d",
                          tile, achunk);
    std::vector<int> order(b->n);
    for (int i = 0; i < b->n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->reqs[a].P > b->reqs[c].P; });
    std::vector<int2> items;
    for (int i : order) {
        ufomi_bp_batch_req &r = b->reqs[i];
        r.tiles_x = (r.roi_w + tile - 1) / tile;
        const long tiles = (long) r.tiles_x * ((r.roi_h + tile - 1) / tile);
        if ((long) items.size() + tiles > INT_MAX)
            return ufomi_fail(UFOMI_ERR_SIZE, "more than 2^31 work items");
        for (long t = 0; t < tiles; t++)
            items.push_back(make_int2(i, (int) t));
    }
    UFOMI_HIP(hipDeviceSynchronize());  // no replay may still read the lists
    if (items.size() > b->items_cap) {
        if (b->d_items) UFOMI_HIP(hipFree(b->d_items));
        b->d_items = nullptr;
        UFOMI_HIP(hipMalloc(&b->d_items, items.size() * sizeof(int2)));
        b->items_cap = items.size();
    }
    UFOMI_HIP(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
    UFOMI_HIP(hipMemcpy(b->d_reqs, b->reqs.data(), b->reqs.size() * sizeof(ufomi_bp_batch_req),
                        hipMemcpyHostToDevice));
    int per_cu = 1, rc;
    if ((rc = ufomi_bp_batch_blocks_per_cu(tile, achunk, b->dtype, &per_cu))) return rc;
    hipDeviceProp_t prop;
    UFOMI_HIP(hipGetDeviceProperties(&prop, b->ctx->device));
    b->n_items = (int) items.size();
    b->grid = std::min(prop.multiProcessorCount * per_cu, b->n_items);
    b->tile = tile;
    b->achunk = achunk;
    return UFOMI_OK;
}

int build_graph(ufomi_fbp_batch *b)
{
    if (b->exec) {
        UFOMI_HIP(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    hipStream_t cs;
    UFOMI_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void) hipStreamDestroy(cs);
        return ufomi_fail(UFOMI_ERR_HIP, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
    }
    const int rc = enqueue(b, cs, 3);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cs, &graph);
    (void) hipStreamDestroy(cs);
    if (rc) {
        if (graph) (void) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess)
        return ufomi_fail(UFOMI_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void) hipGraphDestroy(graph);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return ufomi_fail(UFOMI_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    return UFOMI_OK;
}

int check_batch(ufomi_ctx *ctx, ufomi_fbp_batch *b)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!b) return ufomi_fail(UFOMI_ERR_ARG, "batch is NULL");
    if (b->ctx != ctx) return ufomi_fail(UFOMI_ERR_ARG, "batch belongs to another context");
    return UFOMI_OK;
}

template <typename T>
int upload(T **dst, const T *src, size_t n)
{
    UFOMI_HIP(hipMalloc(dst, n * sizeof(T)));
    UFOMI_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return UFOMI_OK;
}

}  // namespace

extern "C" int ufomi_fbp_batch_create(ufomi_ctx *ctx, const ufomi_fbp_request *requests, int n, int filter_type,
                                      float filter_scale, int storage_dtype, ufomi_fbp_batch **out)
{
    int rc = ufomi_enter(ctx);
    if (rc) return rc;
    if (!out || !requests || n <= 0)
        return ufomi_fail(UFOMI_ERR_ARG, "invalid batch arguments");
    *out = nullptr;
    if (storage_dtype != UFOMI_F32 && storage_dtype != UFOMI_F16 && storage_dtype != UFOMI_BF16)
        return ufomi_fail(UFOMI_ERR_ARG, "unknown storage dtype %d", storage_dtype);
    const size_t esize = storage_dtype == UFOMI_F32 ? 4 : 2;

    ufomi_fbp_batch *b = new ufomi_fbp_batch;
    b->ctx = ctx;
    b->n = n;
    b->dtype = storage_dtype;
    b->reqs.resize(n);
    std::vector<float2> trig;
    std::vector<int> fft_n(n);
    std::vector<size_t> offs(n);
    uint64_t sig = 1469598103934665603ull;  // FNV-1a over the shapes
    auto mix = [&](uint64_t v) { sig = (sig ^ v) * 1099511628211ull; };
    mix((uint64_t) storage_dtype);

    auto bad = [&](int code, const char *fmt, int i, auto... args) {
        char msg[400];
        snprintf(msg, sizeof msg, fmt, args...);
        free_batch(b);
        return ufomi_fail(code, "request %d: %s", i, msg);
    };
    for (int i = 0; i < n; i++) {
        const ufomi_fbp_request &q = requests[i];
        const ufomi_bp_params &prm = q.params;
        const int W = q.width, P = q.num_angles;
        if (!q.sinogram || !q.slice || W <= 0 || P <= 0)
            return bad(UFOMI_ERR_ARG, "invalid sinogram, slice or size (%d x %d)", i, P, W);
        if (prm.mode != UFOMI_BP_TEXTURE)
            return bad(UFOMI_ERR_UNSUPPORTED, "batched backprojection supports texture mode only (mode %d)", i,
                       prm.mode);
        int N = q.fft_size;
        if (N == 0) {
            N = 16;
            while (N < 2 * W) N *= 2;
        }
        if (!ufomi_is_pow2(N))
            return bad(UFOMI_ERR_NOT_POW2, "fft size %d is not a power of two", i, N);
        if (N < W)
            return bad(UFOMI_ERR_SIZE, "Specified size %d must be greater than or equal to input size %d", i, N, W);
        if (N < 16 || N > UFOMI_FFT_MAX_N)
            return bad(UFOMI_ERR_UNSUPPORTED, "fft size %d outside 16..%d", i, N, UFOMI_FFT_MAX_N);
        // ufo_backproject_task_get_requisition (ufo-backproject-task.c:212-267)
        const unsigned n_proj = prm.num_projections ? prm.num_projections : (unsigned) P;
        if ((unsigned) P > n_proj)
            return bad(UFOMI_ERR_PROJECTIONS,
                       "Total number of projections (%u) must be greater than or equal to sinogram height (%u)", i,
                       n_proj, (unsigned) P);
        if ((unsigned long) prm.offset + (unsigned) P > n_proj)
            return bad(UFOMI_ERR_ARG, "offset %u + sinogram height %d exceeds the %u LUT entries", i, prm.offset, P,
                       n_proj);
        const int roi_w = prm.roi_width ? (int) prm.roi_width : W;
        const int roi_h = prm.roi_height ? (int) prm.roi_height : W;
        const double step = prm.angle_step <= 0.0 ? M_PI / (double) n_proj : prm.angle_step;
        const float axis = prm.axis_pos <= 0.0 ? (float) W / 2.0f : (float) prm.axis_pos;
        if (!std::isfinite(axis) || !std::isfinite(step) || !std::isfinite(prm.angle_offset))
            return bad(UFOMI_ERR_ARG, "non-finite geometry", i, 0);
        const double lim = (double) (1 << 20);
        if ((double) roi_w + prm.roi_x > lim || (double) roi_h + prm.roi_y > lim || W > (1 << 20) ||
            fabs((double) axis) > lim)
            return bad(UFOMI_ERR_ARG, "region of interest or axis beyond 2^20 pixels", i, 0);
        // the staging loads address one sinogram with 32-bit byte offsets
        if ((double) P * W * 4.0 >= 4294967296.0)
            return bad(UFOMI_ERR_ARG, "sinogram exceeds 4 GiB", i, 0);

        ufomi_bp_batch_req &r = b->reqs[i];
        r.sino = nullptr;  // set once the filtered buffer exists
        r.out = q.slice;
        r.W = W;
        r.P = P;
        r.roi_w = roi_w;
        r.roi_h = roi_h;
        r.tiles_x = 0;
        r.trig0 = (int) trig.size();
        r.axis = axis;
        r.roi_x = (float) prm.roi_x;
        r.roi_y = (float) prm.roi_y;
        r.pad = 0.0f;
        // LUT of ufo-backproject-task.c:174-196: FP64 sin/cos rounded to float.
        for (int k = 0; k < P; k++) {
            const double a = prm.angle_offset + (double) (prm.offset + k) * step;
            trig.push_back(make_float2((float) sin(a), (float) cos(a)));
        }
        if (trig.size() > (size_t) INT_MAX)
            return bad(UFOMI_ERR_SIZE, "more than 2^31 angles in the batch", i, 0);
        fft_n[i] = N;
        offs[i] = b->filtered_bytes;
        b->filtered_bytes += ((size_t) P * W * esize + 255) & ~(size_t) 255;
        b->updates += (double) roi_w * roi_h * P;
        mix((uint64_t) W); mix((uint64_t) P); mix((uint64_t) N); mix((uint64_t) roi_w); mix((uint64_t) roi_h);
    }

    hipError_t e = hipMalloc(&b->d_filtered, b->filtered_bytes);
    if (e != hipSuccess) {
        free_batch(b);
        return ufomi_fail(UFOMI_ERR_MEMORY, "cannot allocate %zu bytes of filtered rows: %s", b->filtered_bytes,
                          hipGetErrorString(e));
    }
    for (int i = 0; i < n; i++)
        b->reqs[i].sino = (const char *) b->d_filtered + offs[i];

    // One segmented ramp launch per FFT size.
    std::map<int, std::vector<int>> by_n;
    for (int i = 0; i < n; i++) by_n[fft_n[i]].push_back(i);
    for (auto &kv : by_n) {
        const int N = kv.first;
        std::vector<float> coef(2 * (size_t) N);
        if ((rc = ufomi_filter_coefficients(filter_type, 2u * N, filter_scale, 0.5f, 4.0f, 0.1f, 1.0f, coef.data()))) {
            free_batch(b);
            return rc;
        }
        // Real even spectrum: two rows per complex transform (classify_kernel's test).
        int pair = 1;
        for (int k = 0; k < N && pair; k++)
            if (coef[2 * k] != coef[2 * k + 1] || (k && coef[2 * k] != coef[2 * (N - k)]))
                pair = 0;
        const long per_block = ufomi_ramp_rows_per_block(N, pair);
        if (per_block <= 0) {
            free_batch(b);
            return per_block < 0 ? (int) per_block : ufomi_fail(UFOMI_ERR_UNSUPPORTED, "no ramp kernel for N %d", N);
        }
        Group g;
        g.N = N;
        g.pair = pair;
        std::vector<ufomi_ramp_seg> segs;
        for (int i : kv.second) {
            ufomi_ramp_seg sg;
            sg.in = requests[i].sinogram;
            sg.out = (char *) b->d_filtered + offs[i];
            sg.W = b->reqs[i].W;
            sg.rows = b->reqs[i].P;
            sg.block0 = g.nblocks;
            g.nblocks += (sg.rows + per_block - 1) / per_block;
            segs.push_back(sg);
        }
        g.nseg = (int) segs.size();
        b->groups.push_back(g);
        Group &gg = b->groups.back();
        if ((rc = upload(&gg.d_coef, coef.data(), coef.size())) || (rc = upload(&gg.d_segs, segs.data(), segs.size()))) {
            free_batch(b);
            return rc;
        }
    }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, ctx->device)) != hipSuccess) {
        free_batch(b);
        return ufomi_fail(UFOMI_ERR_HIP, "hipGetDeviceProperties failed: %s", hipGetErrorString(e));
    }
    b->signature = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount) + ":" +
                   std::to_string(n) + ":" + std::to_string(sig);
    const int zero = 0;
    if ((rc = upload(&b->d_trig, trig.data(), trig.size())) ||
        (rc = upload(&b->d_reqs, b->reqs.data(), b->reqs.size())) || (rc = upload(&b->d_counter, &zero, 1)) ||
        (rc = set_items(b, 32, 16)) || (rc = build_graph(b))) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return UFOMI_OK;
}

extern "C" int ufomi_fbp_batch_configure(ufomi_ctx *ctx, ufomi_fbp_batch *b, int tile, int angle_chunk)
{
    int rc = check_batch(ctx, b);
    if (rc) return rc;
    if ((rc = set_items(b, tile, angle_chunk))) return rc;
    return build_graph(b);
}

extern "C" int ufomi_fbp_batch_autotune(ufomi_ctx *ctx, ufomi_fbp_batch *b, int reps, int force, int *tile,
                                        int *angle_chunk, float *ms)
{
    int rc = check_batch(ctx, b);
    if (rc) return rc;
    if (reps <= 0) reps = 2;
    if (!force) {
        std::pair<int, int> hit(0, 0);
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            auto it = g_tune.find(b->signature);
            if (it != g_tune.end()) hit = it->second;
        }
        if (hit.first) {
            if ((rc = ufomi_fbp_batch_configure(ctx, b, hit.first, hit.second))) return rc;
            if (tile) *tile = hit.first;
            if (angle_chunk) *angle_chunk = hit.second;
            if (ms) *ms = 0.0f;
            return UFOMI_OK;
        }
    }
    // The filtered rows the backprojection reads.
    if ((rc = enqueue(b, ctx->stream, 1))) return rc;
    hipEvent_t e0, e1;
    UFOMI_HIP(hipEventCreate(&e0));
    UFOMI_HIP(hipEventCreate(&e1));
    float best = INFINITY;
    int bt = 32, ba = 16;
    for (const auto &c : kCandidates) {
        if ((rc = set_items(b, c[0], c[1]))) break;
        if ((rc = enqueue(b, ctx->stream, 2))) break;  // warm-up
        (void) hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && !rc; r++)
            rc = enqueue(b, ctx->stream, 2);
        (void) hipEventRecord(e1, ctx->stream);
        if (rc) break;
        hipError_t e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) {
            rc = ufomi_fail(UFOMI_ERR_HIP, "timing the persistent kernel failed: %s", hipGetErrorString(e));
            break;
        }
        t /= (float) reps;
        if (t < best) {
            best = t;
            bt = c[0];
            ba = c[1];
        }
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    if (rc) return rc;
    if ((rc = ufomi_fbp_batch_configure(ctx, b, bt, ba))) return rc;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune[b->signature] = std::make_pair(bt, ba);
    }
    if (tile) *tile = bt;
    if (angle_chunk) *angle_chunk = ba;
    if (ms) *ms = best;
    return UFOMI_OK;
}

extern "C" int ufomi_fbp_batch_run(ufomi_ctx *ctx, ufomi_fbp_batch *b)
{
    int rc = check_batch(ctx, b);
    if (rc) return rc;
    UFOMI_HIP(hipGraphLaunch(b->exec, ctx->stream));
    return UFOMI_OK;
}

extern "C" int ufomi_fbp_batch_launch(ufomi_ctx *ctx, ufomi_fbp_batch *b, int stages)
{
    int rc = check_batch(ctx, b);
    if (rc) return rc;
    if (stages < 1 || stages > 3)
        return ufomi_fail(UFOMI_ERR_ARG, "stages must be 1 (filters), 2 (backprojection) or 3");
    return enqueue(b, ctx->stream, stages);
}

extern "C" int ufomi_fbp_batch_get_info(const ufomi_fbp_batch *b, ufomi_fbp_batch_info *info)
{
    if (!b || !info) return ufomi_fail(UFOMI_ERR_ARG, "batch or info is NULL");
    info->voxel_updates = b->updates;
    info->tile = b->tile;
    info->angle_chunk = b->achunk;
    info->workgroups = b->grid;
    info->items = b->n_items;
    info->filter_launches = (int) b->groups.size();
    info->filtered_bytes = b->filtered_bytes;
    return UFOMI_OK;
}

extern "C" int ufomi_fbp_batch_destroy(ufomi_fbp_batch *b)
{
    if (!b) return UFOMI_OK;
    if (b->ctx) {
        (void) hipSetDevice(b->ctx->device);
        (void) hipDeviceSynchronize();
    }
    return free_batch(b);
}
