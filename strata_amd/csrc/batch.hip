// batch.hip — host side of batches: many tiles with mixed ops run by one
// persistent launch (batch_kernel) over a device work list, optionally
// captured in a HIP graph, with a launch-shape tuner and the per-tile
// plain-path baseline.
//
// A task is exec.ProcessN (internal/exec/process.go:28) on its own rasters,
// then point ops on its output in place; the work list only changes how
// that work is scheduled on the device, never a result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "strata_hip.h"

using namespace strata;

struct strata_batch {
    std::vector<strata_task> tasks;           // as given, taps repointed at `taps`
    std::vector<std::vector<float>> taps;     // copies of every kernel's taps
    std::vector<BatchOp> ops;                 // distinct ops
    std::vector<int32_t> op_of;               // op index of each task
    std::vector<int32_t> order;               // tasks in work-list order (by source address)
    std::vector<BatchTask> list;              // tasks in work-list order (made by make_task)
    std::vector<char> image;                  // host image of the device work lists
    ListLayout layout{};
    char *dev = nullptr;  // the work lists (see ListLayout)
    size_t dev_bytes = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int rows = 64, blocks = 0, blocks_per_cu = 0, cus = 1;
    bool graph = false;
    hipGraph_t g = nullptr;
    hipGraphExec_t gx = nullptr;
    int64_t cells = 0;

    ~strata_batch()
    {
        if (gx) (void)hipGraphExecDestroy(gx);
        if (g) (void)hipGraphDestroy(g);
        if (dev) (void)hipFree(dev);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }
};

namespace strata {

int64_t cdiv64(int64_t a, int64_t b) { return (a + b - 1) / b; }

int batch_kind(const strata_kernel &k)
{
    switch (k.op) {
    case STRATA_OP_SLOPE:
    case STRATA_OP_ASPECT:
    case STRATA_OP_HILLSHADE:
    case STRATA_OP_CURVATURE:
        return (k.flags & STRATA_FLAG_FIT) ? -1 : BK_HORN;
    case STRATA_OP_MEAN: return k.radius <= 3 ? BK_MEAN : -1;
    case STRATA_OP_MIN: return k.radius <= 3 ? BK_MIN : -1;
    case STRATA_OP_MAX: return k.radius <= 3 ? BK_MAX : -1;
    }
    return -1;
}

int check_task(const strata_task &t, int i)
{
    int rc = check_kernel(&t.k, 1, 1);
    if (rc) return rc;
    if (batch_kind(t.k) < 0)
        return fail(STRATA_EINVAL,
                    "strata: task %d: op %d (radius %d, flags %#x) does not run in a batch: batches run Slope, "
                    "Aspect, Hillshade and Curvature at radius 1 and focal Mean, Min and Max at radius 1 to 3",
                    i, t.k.op, t.k.radius, t.k.flags);
    if (t.npost < 0 || t.npost > 2) return fail(STRATA_EINVAL, "strata: task %d: npost %d not in 0..2", i, t.npost);
    for (int j = 0; j < t.npost; ++j) {
        if (t.post[j].op != STRATA_OP_LOOKUP && t.post[j].op != STRATA_OP_RECLASS)
            return fail(STRATA_EINVAL, "strata: task %d: post op %d is not Lookup or Reclass", i, t.post[j].op);
        if ((rc = check_kernel(&t.post[j], 1, 1))) return rc;
        if (t.post[j].ntaps > kBatchPostTaps)
            return fail(STRATA_EINVAL, "strata: task %d: post op %d has %d taps; a batch holds at most %d", i, j,
                        t.post[j].ntaps, kBatchPostTaps);
    }
    if ((rc = check_raster("src", i, t.src)) || (rc = check_raster("dst", i, t.dst))) return rc;
    if (t.src.mem != STRATA_MEM_DEVICE || t.dst.mem != STRATA_MEM_DEVICE)
        return fail(STRATA_EINVAL, "strata: task %d: batch rasters must be in device memory", i);
    if (t.src.width != t.dst.width || t.src.height != t.dst.height)
        return fail(STRATA_EINVAL, "engine: src[0] is %lld×%lld, dst[0] is %lld×%lld", (long long)t.src.width,
                    (long long)t.src.height, (long long)t.dst.width, (long long)t.dst.height);
    if (t.src.valid && !t.dst.valid)
        return fail(STRATA_EINVAL,
                    "engine: an input has a validity mask but dst[0].Valid is nil; allocate it with "
                    "raster.NewFloat32Like, or set Valid on its root raster before windowing");
    return STRATA_OK;
}

BatchOp make_op(const strata_task &t)
{
    BatchOp op;
    memset(&op, 0, sizeof op);
    op.op = t.k.op;
    op.radius = t.k.radius;
    op.flags = t.k.flags;
    op.kind = batch_kind(t.k);
    memcpy(op.k, t.k.k, sizeof op.k);
    op.npost = t.npost;
    for (int j = 0; j < t.npost; ++j) {
        op.post_op[j] = t.post[j].op;
        op.post_n[j] = t.post[j].ntaps;
        memcpy(op.post_taps[j], t.post[j].taps, sizeof(float) * (size_t)t.post[j].ntaps);
    }
    return op;
}

BatchTask make_task(const strata_task &t, int32_t op)
{
    BatchTask b;
    memset(&b, 0, sizeof b);
    const int64_t w = t.src.width, h = t.src.height;
    b.rg = Region{w, h, 0, 0, w, h};
    b.s.p = t.src.data;
    b.s.stride = t.src.stride;
    b.s.sx = b.s.sy = 0;
    b.s.lx0 = 0;
    b.s.lx1 = w;
    b.s.ly0 = 0;
    b.s.ly1 = h;
    b.d.p = t.dst.data;
    b.d.stride = t.dst.stride;
    auto mask = [&](const strata_raster &r) {
        MaskView m;
        m.bits = r.valid;
        m.off = r.valid_offset;
        m.stride = r.stride;
        m.x0 = m.y0 = 0;
        m.lx0 = 0;
        m.lx1 = w;
        m.ly0 = 0;
        m.ly1 = h;
        return m;
    };
    b.sm = mask(t.src);
    b.dm = mask(t.dst);
    b.op = op;
    b.has_src = t.src.valid ? 1 : 0;
    auto al8 = [](const void *p) { return ((uintptr_t)p & 7) == 0; };
    b.aligned = al8(t.src.data) && al8(t.dst.data) && (t.src.stride & 1) == 0 && (t.dst.stride & 1) == 0;
    return b;
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t list_bytes(int nops, size_t ntasks)
{
    return 256 + align256(sizeof(BatchOp) * (size_t)nops) + 3 * align256(sizeof(BatchTask) * ntasks) + 256 * kBatchFamilies;
}

// Rows per radius-1 unit (128 columns wide) for a whole list: the smallest
// power of two in [4, 64] that splits the list's radius-1 tasks into at
// most kLatencyUnits wave units. A lone small window then runs as many
// short units at once instead of a few long row chains, and a list of many
// windows (a served batch at saturation) runs long units, whose two halo
// rows and prologue weigh little.
constexpr int64_t kLatencyUnits = 8192;
int adaptive_rows(const BatchOp *ops, const BatchTask *tasks, size_t n)
{
    int rows = 4;
    for (; rows < 64; rows *= 2) {
        int64_t units = 0;
        for (size_t i = 0; i < n && units <= kLatencyUnits; ++i)
            if (ops[tasks[i].op].kind == BK_HORN) units += cdiv64(tasks[i].rg.w, 128) * cdiv64(tasks[i].rg.h, rows);
        if (units <= kLatencyUnits) break;
    }
    return rows;
}

ListLayout write_lists(const BatchOp *ops, int nops, const BatchTask *tasks, const bool *masked, size_t n,
                       int rows, char *out)
{
    ListLayout L;
    L.ops = 256;
    size_t count[kBatchFamilies] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        count[ops[tasks[i].op].kind == BK_HORN ? BF_HORN : BF_FOCAL] += 1;
        if (masked[i]) count[BF_VALIDITY] += 1;
    }
    size_t p = L.ops + align256(sizeof(BatchOp) * (size_t)nops);
    for (int f = 0; f < kBatchFamilies; ++f) {
        L.tasks[f] = p;
        p += align256(sizeof(BatchTask) * count[f]);
    }
    L.counters = p;
    L.bytes = p + 256 * kBatchFamilies;  // one 256-byte line per counter
    memcpy(out + L.ops, ops, sizeof(BatchOp) * (size_t)nops);
    size_t k[kBatchFamilies] = {0, 0, 0};
    int64_t u[kBatchFamilies] = {0, 0, 0};
    if (rows <= 0) rows = adaptive_rows(ops, tasks, n);
    BatchTask *lt[kBatchFamilies];
    for (int f = 0; f < kBatchFamilies; ++f) lt[f] = reinterpret_cast<BatchTask *>(out + L.tasks[f]);
    for (size_t i = 0; i < n; ++i) {
        BatchTask t = tasks[i];
        batch_units(ops[t.op], t, rows);
        const int64_t data = (int64_t)t.blocks_x * cdiv64(t.rg.h, t.rows);
        const int64_t vrows =
            (t.vgroup ? kVgRows * (64 / (int64_t)t.vgroup) : kVsRows) * kVUnitChunks;  // rows per validity unit
        t.vunits = masked[i] ? (int32_t)((int64_t)t.vwaves_x * cdiv64(t.rg.h, vrows)) : 0;
        const int f = ops[t.op].kind == BK_HORN ? BF_HORN : BF_FOCAL;
        t.u0 = u[f];
        t.u1 = u[f] += data;
        lt[f][k[f]++] = t;
        if (masked[i]) {
            t.u0 = u[BF_VALIDITY];
            t.u1 = u[BF_VALIDITY] += t.vunits;
            lt[BF_VALIDITY][k[BF_VALIDITY]++] = t;
        }
    }
    for (int f = 0; f < kBatchFamilies; ++f) {
        L.h[f].ntasks = (int64_t)count[f];
        L.h[f].nunits = u[f];
    }
    // One focal kind and radius in the whole list: its own kernel.
    for (size_t i = 0; i < n; ++i) {
        const BatchOp &o = ops[tasks[i].op];
        if (o.kind == BK_HORN) continue;
        const int code = batch_focal_one_code(o.kind, o.radius);
        if (L.focal_one == -1) {
            L.focal_one = code;
        } else if (L.focal_one != code) {
            L.focal_one = -2;  // mixed
            break;
        }
    }
    if (L.focal_one < 0) L.focal_one = -1;
    memcpy(out, L.h, sizeof L.h);
    memset(out + L.counters, 0, 256 * kBatchFamilies);
    return L;
}

BatchLaunch launch_view(const ListLayout &L, char *dev, int max_wave_blocks, int max_focal_blocks, bool clear)
{
    BatchLaunch b;
    b.ops = reinterpret_cast<const BatchOp *>(dev + L.ops);
    b.hdr = reinterpret_cast<const BatchHeader *>(dev);
    for (int f = 0; f < kBatchFamilies; ++f) {
        b.tasks[f] = reinterpret_cast<const BatchTask *>(dev + L.tasks[f]);
        b.nunits[f] = L.h[f].nunits;
    }
    const int64_t waves = L.h[BF_HORN].nunits + L.h[BF_VALIDITY].nunits;
    b.blocks_wave = (int)std::max<int64_t>(1, std::min<int64_t>(max_wave_blocks, cdiv64(waves, 4)));
    b.blocks_focal = (int)std::max<int64_t>(1, std::min<int64_t>(max_focal_blocks, L.h[BF_FOCAL].nunits));
    b.counters = reinterpret_cast<unsigned *>(dev + L.counters);
    b.clear = clear;
    b.focal_one = L.focal_one;
    return b;
}

} // namespace strata

namespace {

// Lays the work lists out for `rows` rows per radius-1 unit and uploads them.
int build_list(strata_batch *b, int rows)
{
    b->rows = rows;
    std::vector<char> masked(b->list.size());
    for (size_t n = 0; n < b->list.size(); ++n) masked[n] = b->tasks[(size_t)b->order[n]].dst.valid != nullptr;
    b->image.assign(list_bytes((int)b->ops.size(), b->list.size()), 0);
    b->layout = write_lists(b->ops.data(), (int)b->ops.size(), b->list.data(),
                            reinterpret_cast<const bool *>(masked.data()), b->list.size(), rows, b->image.data());
    int64_t total = 0;
    for (int f = 0; f < kBatchFamilies; ++f) total = std::max<int64_t>(total, b->layout.h[f].nunits);
    if (total >= ((int64_t)1 << 31))
        return fail(STRATA_EINVAL, "strata: a batch list of %lld work units exceeds the limit of 2^31",
                    (long long)total);
    const size_t bytes = b->layout.bytes;
    if (bytes > b->dev_bytes) {
        if (b->dev) HIP_TRY(hipFree(b->dev), "freeing a work list");
        b->dev = nullptr;
        HIP_TRY(hipMalloc((void **)&b->dev, bytes), "allocating a work list");
        b->dev_bytes = bytes;
    }
    HIP_TRY(hipMemcpy(b->dev, b->image.data(), bytes, hipMemcpyHostToDevice), "uploading a work list");
    return STRATA_OK;
}

hipError_t enqueue(strata_batch *b)
{
    auto cap = [&](int family) {
        const int occ = batch_blocks_per_cu(family, b->layout.focal_one);
        return b->cus * (b->blocks_per_cu > 0 ? std::min(b->blocks_per_cu, occ) : occ);
    };
    // Replays reuse the uploaded lists: the counters are cleared by a kernel.
    const BatchLaunch L = launch_view(b->layout, b->dev, cap(BF_HORN), cap(BF_FOCAL), true);
    return launch_batch(L, b->stream);
}

int set_shape(strata_batch *b, int rows, int blocks_per_cu, bool graph)
{
    if (rows < 4 || rows > 4096) return fail(STRATA_EINVAL, "strata: rows %d not in 4..4096", rows);
    // Runs still in flight read the work lists about to be rewritten.
    HIP_TRY(hipStreamSynchronize(b->stream), "waiting for a batch");
    int rc = build_list(b, rows);
    if (rc) return rc;
    b->blocks_per_cu = blocks_per_cu > 0 ? blocks_per_cu : 0;
    b->blocks = b->cus * (blocks_per_cu > 0 ? std::min(blocks_per_cu, batch_blocks_per_cu(BF_HORN))
                                            : batch_blocks_per_cu(BF_HORN));
    if (b->gx) HIP_TRY(hipGraphExecDestroy(b->gx), "destroying a graph");
    if (b->g) HIP_TRY(hipGraphDestroy(b->g), "destroying a graph");
    b->gx = nullptr;
    b->g = nullptr;
    b->graph = graph;
    if (graph) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "capturing a batch");
        const hipError_t e = enqueue(b);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        HIP_TRY(e, "capturing a batch");
        HIP_TRY(e2, "capturing a batch");
        b->g = g;
        HIP_TRY(hipGraphInstantiate(&b->gx, b->g, nullptr, nullptr, 0), "instantiating a batch graph");
    }
    return STRATA_OK;
}

} // namespace

extern "C" {

int strata_batch_create(const strata_task *tasks, int ntasks, const strata_batch_options *opts,
                        strata_batch **out)
{
    if (!out) return fail(STRATA_EINVAL, "strata: nil batch result pointer");
    *out = nullptr;
    if (ntasks <= 0 || !tasks) return fail(STRATA_EINVAL, "strata: a batch needs at least one task");
    for (int i = 0; i < ntasks; ++i) {
        const int rc = check_task(tasks[i], i);
        if (rc) return rc;
    }
    strata_batch *b = new strata_batch;
    b->tasks.assign(tasks, tasks + ntasks);
    // Own copies of every table, so the caller's taps need not outlive
    // this call. A task has at most three tables; the outer vector holds
    // them all without growing, and an inner vector's data stays put when
    // the outer one moves it anyway.
    b->taps.reserve(3 * (size_t)ntasks);
    for (auto &t : b->tasks) {
        auto keep = [&](strata_kernel &k) {
            if (k.ntaps > 0) {
                b->taps.emplace_back(k.taps, k.taps + k.ntaps);
                k.taps = b->taps.back().data();
            }
        };
        keep(t.k);
        for (int j = 0; j < t.npost; ++j) keep(t.post[j]);
    }
    b->op_of.resize(ntasks);
    for (int i = 0; i < ntasks; ++i) {
        const BatchOp op = make_op(b->tasks[i]);
        int found = -1;
        for (size_t j = 0; j < b->ops.size(); ++j)
            if (memcmp(&b->ops[j], &op, sizeof op) == 0) found = (int)j;
        if (found < 0) {
            found = (int)b->ops.size();
            b->ops.push_back(op);
        }
        b->op_of[i] = found;
        b->cells += b->tasks[i].src.width * b->tasks[i].src.height;
    }
    // Tasks in the order of their source's address: windows of one raster
    // in row-major order of their corners. The waves running at any moment
    // then read nearby rows of the parent, a few pages of it, instead of
    // windows scattered over the whole raster (which thrashes the address
    // translation caches). Units are at most 128x128 cells, so the tail of
    // the list is short whatever the order.
    b->order.resize(ntasks);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
        return (uintptr_t)b->tasks[a].src.data < (uintptr_t)b->tasks[c].src.data;
    });
    b->list.resize(ntasks);
    for (int n = 0; n < ntasks; ++n) b->list[n] = make_task(b->tasks[b->order[n]], b->op_of[b->order[n]]);

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || b->cus < 1) {
        delete b;
        return fail(STRATA_EDEVICE, "strata: cannot query the device");
    }
    if (opts && opts->stream) {
        b->stream = (hipStream_t)opts->stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return fail(STRATA_EDEVICE, "strata: cannot create a stream");
        }
        b->own_stream = true;
    }
    const int rc = set_shape(b, opts && opts->rows ? opts->rows : 64, opts ? opts->blocks_per_cu : 0,
                             opts && opts->graph);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return STRATA_OK;
}

int strata_batch_run(strata_batch *b)
{
    if (!b) return fail(STRATA_EINVAL, "strata: nil batch");
    if (b->graph)
        HIP_TRY(hipGraphLaunch(b->gx, b->stream), "launching a batch graph");
    else
        HIP_TRY(enqueue(b), "launching a batch");
    return STRATA_OK;
}

int strata_batch_run_per_tile(strata_batch *b)
{
    if (!b) return fail(STRATA_EINVAL, "strata: nil batch");
    strata_options o;
    memset(&o, 0, sizeof o);
    o.stream = b->stream;
    o.async = 1;
    for (const strata_task &t : b->tasks) {
        int rc = strata_process(&t.k, &t.dst, 1, &t.src, 1, &o, nullptr);
        if (rc) return rc;
        for (int j = 0; j < t.npost; ++j)
            if ((rc = strata_process(&t.post[j], &t.dst, 1, &t.dst, 1, &o, nullptr))) return rc;
    }
    return STRATA_OK;
}

int strata_batch_tune(strata_batch *b, int reps, strata_batch_options *chosen)
{
    if (!b) return fail(STRATA_EINVAL, "strata: nil batch");
    if (reps < 1) reps = 1;
    const bool graph = b->graph;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "creating an event");
    HIP_TRY(hipEventCreate(&e1), "creating an event");
    float best = 0.0f;
    int best_rows = b->rows, best_bpc = b->blocks_per_cu;
    const int occ = batch_blocks_per_cu(BF_HORN);
    const int rows_c[] = {16, 32, 64, 128};
    std::vector<int> bpc_c = {occ};
    if (occ > 1) bpc_c.push_back(std::max(1, occ / 2));
    int rc = STRATA_OK;
    for (int rows : rows_c) {
        for (int bpc : bpc_c) {
            if ((rc = set_shape(b, rows, bpc, false))) goto done;
            if ((rc = strata_batch_run(b))) goto done;  // warm-up
            if (hipEventRecord(e0, b->stream) != hipSuccess) goto device;
            for (int r = 0; r < reps; ++r)
                if ((rc = strata_batch_run(b))) goto done;
            if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto device;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto device;
            if (best == 0.0f || ms < best) {
                best = ms;
                best_rows = rows;
                best_bpc = bpc;
            }
        }
    }
    rc = set_shape(b, best_rows, best_bpc, graph);
    if (!rc && chosen) {
        memset(chosen, 0, sizeof *chosen);
        chosen->rows = b->rows;
        chosen->blocks_per_cu = b->blocks_per_cu;
        chosen->graph = b->graph ? 1 : 0;
        chosen->stream = b->stream;
    }
    goto done;
device:
    rc = fail(STRATA_EDEVICE, "strata: timing a batch failed");
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int strata_batch_sync(strata_batch *b)
{
    if (!b) return fail(STRATA_EINVAL, "strata: nil batch");
    HIP_TRY(hipStreamSynchronize(b->stream), "waiting for a batch");
    return STRATA_OK;
}

int strata_batch_info_get(const strata_batch *b, strata_batch_info *out)
{
    if (!b || !out) return fail(STRATA_EINVAL, "strata: nil batch or info");
    memset(out, 0, sizeof *out);
    out->tasks = (int64_t)b->tasks.size();
    out->cells = b->cells;
    out->data_units = b->layout.h[BF_HORN].nunits + b->layout.h[BF_FOCAL].nunits;
    out->validity_units = b->layout.h[BF_VALIDITY].nunits;
    out->rows = b->rows;
    out->blocks = b->blocks;
    out->graph = b->graph ? 1 : 0;
    return STRATA_OK;
}

void strata_batch_destroy(strata_batch *b) { delete b; }

} // extern "C"
