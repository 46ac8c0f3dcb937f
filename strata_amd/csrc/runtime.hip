// runtime.hip — host side of the C ABI: checks, the plan, Stats, staging
// of host-memory operands, the cancellation units and the chunked tile
// queue. Device work is launched through kernels.hpp.
//
// Reference behaviour restated here:
//   checks           internal/exec/process.go:37-135, chunked.go:42-104
//   plan and bands   internal/exec/tile.go:298-359
//   band accounting  internal/exec/halo.go:21-57,97-105; worker.go:59-68
//   chunked tiles    internal/exec/chunked.go:159-374
//   cancellation     internal/exec/worker.go:112-181 (claims in order,
//                    a claimed unit always finishes: a prefix is written)
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <condition_variable>
#include <thread>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "strata_hip.h"

using namespace strata;

namespace strata {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(STRATA_EDEVICE, "strata: %s: %s", what, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Kernel descriptors
// ---------------------------------------------------------------------------

int n_outputs(const strata_kernel *k)
{
    if (k->op == STRATA_OP_GRADIENT) return 2;
    if (k->op == STRATA_OP_SURFACE) return __builtin_popcount((unsigned)k->flags & STRATA_SURF_ALL);
    return 1;
}

int check_kernel(const strata_kernel *k, int ndst, int nsrc)
{
    if (!k) return fail(STRATA_EINVAL, "engine: nil kernel");
    const bool fit = (k->flags & STRATA_FLAG_FIT) != 0;
    switch (k->op) {
    case STRATA_OP_SURFACE:
        if ((k->flags & STRATA_SURF_ALL) == 0)
            return fail(STRATA_EINVAL,
                        "terrain: Surface has no outputs; set at least one of Dx, Dy, Slope, Aspect and Hillshade");
        [[fallthrough]];
    case STRATA_OP_GRADIENT:
    case STRATA_OP_SLOPE:
    case STRATA_OP_ASPECT:
    case STRATA_OP_HILLSHADE:
    case STRATA_OP_CURVATURE:
        if (fit ? (k->radius < 1 || k->radius > kMaxRadius) : k->radius != 1)
            return fail(STRATA_EINVAL, "terrain: FitRadius must be 0 to %d, got radius %d", kMaxRadius, k->radius);
        if (k->op == STRATA_OP_CURVATURE && ((k->flags >> STRATA_CURV_SHIFT) & 3) > 2)
            return fail(STRATA_EINVAL, "terrain: unknown CurvatureType %d", (k->flags >> STRATA_CURV_SHIFT) & 3);
        break;
    case STRATA_OP_CORRELATE:
    case STRATA_OP_SEPARABLE:
    case STRATA_OP_MEAN:
    case STRATA_OP_MIN:
    case STRATA_OP_MAX: {
        if (k->radius < 1 || k->radius > kMaxRadius)
            return fail(STRATA_EINVAL, "focal: Radius must be 1 to %d, got %d", kMaxRadius, k->radius);
        const int n = 2 * k->radius + 1;
        if (k->op == STRATA_OP_CORRELATE && k->ntaps != n * n)
            return fail(STRATA_EINVAL, "focal: Radius %d needs %d×%d = %d Weights, got %d", k->radius, n, n, n * n,
                        k->ntaps);
        if (k->op == STRATA_OP_SEPARABLE && k->ntaps != 2 * n)
            return fail(STRATA_EINVAL, "focal: Radius %d needs %d Row and %d Col taps, got %d in all", k->radius, n,
                        n, k->ntaps);
        break;
    }
    case STRATA_OP_LOOKUP:
        if (k->radius != 0 || k->ntaps < 2 || (k->ntaps & 1))
            return fail(STRATA_EINVAL, "transfer.Lookup: xs and ys must have equal, non-zero length");
        break;
    case STRATA_OP_RECLASS:
        if (k->radius != 0 || k->ntaps < 1 || !(k->ntaps & 1))
            return fail(STRATA_EINVAL, "transfer.Reclass: values must hold one more element than breaks");
        break;
    default:
        return fail(STRATA_EINVAL, "engine: unknown kernel op %d", k->op);
    }
    if (k->ntaps < 0 || k->ntaps > kMaxTaps)
        return fail(STRATA_EINVAL, "engine: %d taps exceed the device limit of %d", k->ntaps, kMaxTaps);
    if (k->ntaps > 0 && !k->taps) return fail(STRATA_EINVAL, "engine: kernel has %d taps and a nil pointer", k->ntaps);
    if (k->op == STRATA_OP_LOOKUP || k->op == STRATA_OP_RECLASS) {
        // transfer/transfer.go:176-221: the x side holds no NaN and strictly
        // increases (the kernels count the entries at or below a value,
        // which is the scan's index only for such a table), and a Lookup's
        // xs are finite.
        const bool lookup = k->op == STRATA_OP_LOOKUP;
        const char *op = lookup ? "transfer.Lookup" : "transfer.Reclass", *name = lookup ? "xs" : "breaks";
        const int n = lookup ? k->ntaps / 2 : (k->ntaps - 1) / 2;
        const float *x = k->taps;
        for (int i = 0; i < n; ++i)
            if (x[i] != x[i]) return fail(STRATA_EINVAL, "%s: %s must not hold NaN, but %s[%d] is NaN", op, name, name, i);
        for (int i = 1; i < n; ++i)
            if (!(x[i - 1] < x[i]))
                return fail(STRATA_EINVAL, "%s: %s must be strictly increasing, but %s[%d] = %g is not above %s[%d] = %g",
                            op, name, name, i, (double)x[i], name, i - 1, (double)x[i - 1]);
        if (lookup)
            for (int i = 0; i < n; ++i)
                if (x[i] - x[i] != 0.0f)
                    return fail(STRATA_EINVAL, "%s: xs must be finite, but xs[%d] = %g", op, i, (double)x[i]);
    }
    if (nsrc != 1 || ndst != n_outputs(k))
        return fail(STRATA_EINVAL, "engine: kernel takes %d inputs and %d outputs, got %d and %d", 1, n_outputs(k),
                    nsrc, ndst);
    return STRATA_OK;
}

Params make_params(const strata_kernel *k)
{
    Params P;
    memset(&P, 0, sizeof P);
    P.op = k->op;
    P.radius = k->radius;
    P.flags = k->flags;
    P.ntaps = k->ntaps;
    memcpy(P.k, k->k, sizeof P.k);
    if (k->ntaps > 0) memcpy(P.taps, k->taps, sizeof(float) * (size_t)k->ntaps);
    return P;
}

// ---------------------------------------------------------------------------
// Jobs: the device launches of one call
// ---------------------------------------------------------------------------

constexpr int kMaxDst = 16;  // outputs of one call (the parts of a fan-out)

// One data launch writing outputs out[0..nout) of the call.
struct Group {
    Params P;
    int nout = 0;
    int out[kMaxOutputs] = {};
};

// A group of a chain's stages: a stencil stage (or, leading the chain, a
// point stage) and the point stages after it. A radius-1 Horn stage takes
// them into its store (P.post); otherwise they run over the group's values
// in place. grow is how far past the output later groups read the group's
// values.
struct ChainGroup {
    Params P;
    int radius = 0, grow = 0;
    std::vector<Params> inplace;
};

// A call's device work: its data launches, and each output's radius, which
// is the width of its edge ring and the erosion of its validity. r is the
// call's radius (how far its reads reach past an output cell) and rmin the
// narrowest ring (internal/exec/tile.go:136-184). A chain's groups run in
// order instead of the launches (one output).
struct Job {
    int nout = 0, r = 0, rmin = 0;
    int radius[kMaxDst] = {};
    std::vector<Group> groups;
    std::vector<ChainGroup> chain;
};

Job make_job(const strata_kernel *k)
{
    Job J;
    J.nout = n_outputs(k);
    J.r = J.rmin = k->radius;
    Group g;
    g.P = make_params(k);
    g.nout = J.nout;
    for (int o = 0; o < J.nout; ++o) {
        g.out[o] = o;
        J.radius[o] = k->radius;
    }
    J.groups.push_back(g);
    return J;
}

// Part i of a fan-out: a kernel of one input and one output.
int check_parts(const strata_kernel *parts, int nparts)
{
    if (!parts || nparts < 1) return fail(STRATA_EINVAL, "engine: pipeline has no stages");
    if (nparts > kMaxDst)
        return fail(STRATA_EINVAL, "engine: %d outputs exceed the library's limit of %d", nparts, kMaxDst);
    for (int i = 0; i < nparts; ++i) {
        if (n_outputs(&parts[i]) != 1)
            return fail(STRATA_EINVAL, "engine: stage %d has %d outputs; a fan-out stage has one", i,
                        n_outputs(&parts[i]));
        int rc = check_kernel(&parts[i], 1, 1);
        if (rc) return rc;
    }
    return STRATA_OK;
}

// exec.NewPipeline(1, stages, outs) with stage i reading input 0 and
// writing output i (terrain/features.go:103-117): the call's radius is the
// largest stage radius, and each output keeps its own (its reach,
// internal/exec/kernel.go:94-110). Radius-1 Horn ops share one sweep per
// kMaxOutputs of them; every other stage is a launch of its own.
Job make_fan_job(const strata_kernel *parts, int nparts)
{
    Job J;
    J.nout = nparts;
    J.r = 0;
    J.rmin = kMaxRadius;
    Group fan;
    auto flush = [&]() {
        if (fan.nout == 0) return;
        fan.P.fan.n = fan.nout;
        J.groups.push_back(fan);
        fan = Group();
    };
    for (int i = 0; i < nparts; ++i) {
        const strata_kernel &k = parts[i];
        J.radius[i] = k.radius;
        J.r = std::max(J.r, k.radius);
        J.rmin = std::min(J.rmin, k.radius);
        const bool horn = (k.op == STRATA_OP_SLOPE || k.op == STRATA_OP_ASPECT || k.op == STRATA_OP_HILLSHADE ||
                           k.op == STRATA_OP_CURVATURE) &&
                          !(k.flags & STRATA_FLAG_FIT);
        if (!horn) {
            Group g;
            g.P = make_params(&k);
            g.nout = 1;
            g.out[0] = i;
            J.groups.push_back(g);
            continue;
        }
        if (fan.nout == 0) {
            memset(&fan.P, 0, sizeof fan.P);
            fan.P.op = kOpHornFan;
            fan.P.radius = 1;
        }
        const int j = fan.nout++;
        fan.out[j] = i;
        fan.P.fan.op[j] = k.op;
        fan.P.fan.flags[j] = k.flags;
        memcpy(fan.P.fan.k[j], k.k, sizeof fan.P.fan.k[j]);
        if (fan.nout == kMaxOutputs) flush();
    }
    flush();
    return J;
}

constexpr int kMaxChain = 8;  // stages of one chain

bool is_point(const strata_kernel &k) { return k.op == STRATA_OP_LOOKUP || k.op == STRATA_OP_RECLASS; }

// exec.NewPipeline(1, stages, [last]) with stage i reading stage i-1's
// value (internal/exec/pipeline.go:166-187): one input, one output, every
// stage a kernel of one input and one output. The chain's reach is the sum
// of the stage radii (radii add along a chain), at most the library's 8.
int check_chain(const strata_kernel *stages, int n)
{
    if (!stages || n < 1) return fail(STRATA_EINVAL, "engine: pipeline has no stages");
    if (n > kMaxChain) return fail(STRATA_EINVAL, "engine: %d stages exceed the library's limit of %d", n, kMaxChain);
    int reach = 0;
    for (int i = 0; i < n; ++i) {
        if (n_outputs(&stages[i]) != 1)
            return fail(STRATA_EINVAL, "engine: stage %d has %d outputs; a chain stage has one", i,
                        n_outputs(&stages[i]));
        int rc = check_kernel(&stages[i], 1, 1);
        if (rc) return rc;
        reach += stages[i].radius;
    }
    if (reach > kMaxRadius)
        return fail(STRATA_EINVAL, "engine: the chain reaches %d cells; the library's limit is %d", reach, kMaxRadius);
    return STRATA_OK;
}

// The chain's job: its output's ring and validity erosion are the chain's
// reach, and so is its radius (internal/exec/pipeline.go:32-41, 83-100).
Job make_chain_job(const strata_kernel *stages, int n)
{
    Job J;
    J.nout = 1;
    for (int i = 0; i < n; ++i) {
        const strata_kernel &k = stages[i];
        J.r += k.radius;
        if (is_point(k) && !J.chain.empty()) {
            ChainGroup &g = J.chain.back();
            PostArgs &pa = g.P.post;
            const bool horn = (g.P.op == STRATA_OP_SLOPE || g.P.op == STRATA_OP_ASPECT ||
                               g.P.op == STRATA_OP_HILLSHADE || g.P.op == STRATA_OP_CURVATURE) &&
                              !(g.P.flags & STRATA_FLAG_FIT) && g.inplace.empty();
            const int used = pa.n ? pa.off[pa.n - 1] + pa.cnt[pa.n - 1] : 0;
            if (horn && pa.n < kMaxPost && used + k.ntaps <= kPostTaps) {
                pa.op[pa.n] = k.op;
                pa.off[pa.n] = used;
                pa.cnt[pa.n] = k.ntaps;
                memcpy(pa.taps + used, k.taps, sizeof(float) * (size_t)k.ntaps);
                ++pa.n;
            } else {
                g.inplace.push_back(make_params(&k));
            }
            continue;
        }
        ChainGroup g;
        g.P = make_params(&k);
        g.radius = k.radius;
        J.chain.push_back(g);
    }
    for (size_t g = J.chain.size() - 1; g > 0; --g) J.chain[g - 1].grow = J.chain[g].grow + J.chain[g].radius;
    J.rmin = J.r;
    J.radius[0] = J.r;
    return J;
}

// The tables of a chain's fused point stages, in device memory for one
// call: written once on the call's stream before its launches (every launch
// of the call is on that stream) and freed after them in stream order.
struct ChainTables {
    void *dev = nullptr;
    hipStream_t st = nullptr;
    ChainTables() = default;
    ChainTables(const ChainTables &) = delete;
    ChainTables &operator=(const ChainTables &) = delete;
    ~ChainTables()
    {
        if (dev) (void)hipFreeAsync(dev, st);
    }
    int stage(Job &J, hipStream_t s)
    {
        size_t n = 0;
        for (const ChainGroup &G : J.chain) n += G.P.post.n > 0;
        if (n == 0) return STRATA_OK;
        st = s;
        HIP_TRY(scratch_alloc_async(&dev, n * sizeof(PostArgs), st), "a chain's tables");
        PostArgs *p = static_cast<PostArgs *>(dev);
        for (ChainGroup &G : J.chain) {
            if (G.P.post.n == 0) continue;
            HIP_TRY(launch_stage_post(G.P.post, p, st), "staging a chain's tables");
            G.P.post_dev = p++;
        }
        return STRATA_OK;
    }
};

// ---------------------------------------------------------------------------
// Rasters
// ---------------------------------------------------------------------------

int64_t data_len(const strata_raster &r) { return (r.height - 1) * r.stride + r.width; }

// raster.Float32Raster.Validate (raster/raster.go:172-203), minus the
// slice lengths, which the caller owns.
int check_raster(const char *name, int i, const strata_raster &r)
{
    if (r.width <= 0 || r.height <= 0)
        return fail(STRATA_EINVAL, "engine: %s[%d]: raster: dimensions must be positive, got %lld×%lld", name, i,
                    (long long)r.width, (long long)r.height);
    if (r.stride < r.width)
        return fail(STRATA_EINVAL, "engine: %s[%d]: raster: stride %lld is less than width %lld", name, i,
                    (long long)r.stride, (long long)r.width);
    if (!r.data) return fail(STRATA_EINVAL, "engine: %s[%d]: nil data", name, i);
    if (r.valid && r.valid_offset < 0)
        return fail(STRATA_EINVAL, "engine: %s[%d]: raster: negative ValidOffset %lld", name, i,
                    (long long)r.valid_offset);
    if (r.mem != STRATA_MEM_DEVICE && r.mem != STRATA_MEM_HOST)
        return fail(STRATA_EINVAL, "engine: %s[%d]: unknown memory kind %d", name, i, r.mem);
    return STRATA_OK;
}

} // namespace strata

namespace {

SrcView srcview(const strata_raster &r, int64_t sx, int64_t sy, int64_t W, int64_t H)
{
    SrcView v;
    v.p = r.data;
    v.stride = r.stride;
    v.sx = sx;
    v.sy = sy;
    v.lx0 = std::max<int64_t>(0, sx);
    v.ly0 = std::max<int64_t>(0, sy);
    v.lx1 = std::min<int64_t>(W, sx + r.width);
    v.ly1 = std::min<int64_t>(H, sy + r.height);
    return v;
}

MaskView mask_view(const strata_raster &r, int64_t x0, int64_t y0, int64_t W, int64_t H)
{
    MaskView m;
    m.bits = r.valid;
    m.off = r.valid_offset;
    m.stride = r.stride;
    m.x0 = x0;
    m.y0 = y0;
    m.lx0 = std::max<int64_t>(0, x0);
    m.ly0 = std::max<int64_t>(0, y0);
    m.lx1 = std::min<int64_t>(W, x0 + r.width);
    m.ly1 = std::min<int64_t>(H, y0 + r.height);
    return m;
}

// ---------------------------------------------------------------------------
// Device execution of one rectangle of the output
// ---------------------------------------------------------------------------

// Device views of one call's operands. The source's data[0] is raster cell
// (sx, sy); output o's data[0] is raster cell (dx[o], dy[o]).
struct Operands {
    const strata_raster *src;
    int64_t sx, sy;
    const strata_raster *dst[kMaxDst];
    int64_t dx[kMaxDst], dy[kMaxDst];
    int nout;
    int64_t W, H;
};

// Kernel timing for measurement: when enabled, each data and validity
// launch is bracketed by HIP events on its own stream and the elapsed times
// are summed when the profile is read.
struct Timing {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> data, validity;
};
Timing g_timing;
std::mutex g_timing_mu;

hipEvent_t timing_event(hipStream_t st)
{
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, st);
    return e;
}

// The library's device scratch (defined with the scratch pool below).
int scratch_alloc(void **p, size_t bytes, hipStream_t st);
void scratch_free(void *p, hipStream_t st);

// A chain's groups over the output rect rc: every group but the last
// writes its values over rc grown by its `grow` (clipped to the raster)
// into a buffer of the library's scratch, which the next group reads as its
// input; the last writes the output. Each stage computes its own edge ring
// (NaN), which the stages after it carry through, so the cells within the
// chain's reach of the raster's edge hold NaN as the separate calls leave
// them (internal/exec/pipeline.go:102-114).
int run_chain_data(const Job &J, const Operands &op, const Rect &rc, hipStream_t st)
{
    const int64_t W = op.W, H = op.H;
    SrcView s = srcview(*op.src, op.sx, op.sy, W, H);
    void *prev = nullptr;
    struct Free {
        void *&p;
        hipStream_t st;
        ~Free() { scratch_free(p, st); }
    } guard{prev, st};
    for (size_t g = 0; g < J.chain.size(); ++g) {
        const ChainGroup &G = J.chain[g];
        const bool last = g + 1 == J.chain.size();
        Rect out = rc;
        DstView d;
        void *buf = nullptr;
        if (last) {
            d.p = op.dst[0]->data + (rc.y0 - op.dy[0]) * op.dst[0]->stride + (rc.x0 - op.dx[0]);
            d.stride = op.dst[0]->stride;
        } else {
            out.x0 = std::max<int64_t>(0, rc.x0 - G.grow);
            out.y0 = std::max<int64_t>(0, rc.y0 - G.grow);
            out.x1 = std::min<int64_t>(W, rc.x1 + G.grow);
            out.y1 = std::min<int64_t>(H, rc.y1 + G.grow);
            const int64_t bs = (out.x1 - out.x0 + 63) & ~(int64_t)63;
            int rc = scratch_alloc(&buf, (size_t)((out.y1 - out.y0) * bs) * 4, st);
            if (rc) return rc;
            d.p = (float *)buf;
            d.stride = bs;
        }
        const Region rg{W, H, out.x0, out.y0, out.x1 - out.x0, out.y1 - out.y0};
        // A group with fused point stages reads their tables at G.P.post_dev,
        // staged once for the call (ChainTables).
        HIP_TRY(launch_data(G.P, rg, s, &d, 1, st), "launching a chain stage");
        for (const Params &Q : G.inplace) {
            const SrcView v{d.p, d.stride, out.x0, out.y0, out.x0, out.x1, out.y0, out.y1};
            HIP_TRY(launch_data(Q, rg, v, &d, 1, st), "launching a chain's point stage");
        }
        scratch_free(prev, st);
        prev = buf;
        if (!last) s = SrcView{(const float *)buf, d.stride, out.x0, out.y0, out.x0, out.x1, out.y0, out.y1};
    }
    return STRATA_OK;
}

int run_rect(const Job &J, const Operands &op, const Rect &rc, hipStream_t st)
{
    Region rg{op.W, op.H, rc.x0, rc.y0, rc.x1 - rc.x0, rc.y1 - rc.y0};
    if (rg.w <= 0 || rg.h <= 0) return STRATA_OK;
    const SrcView s = srcview(*op.src, op.sx, op.sy, op.W, op.H);
    const bool timed = g_timing.on;
    if (!J.chain.empty()) {
        hipEvent_t e0 = timed ? timing_event(st) : nullptr;
        const int e = run_chain_data(J, op, rc, st);
        if (e) return e;
        if (timed) {
            std::lock_guard<std::mutex> lk(g_timing_mu);
            g_timing.data.emplace_back(e0, timing_event(st));
        }
    }
    bool valid_done[kMaxDst] = {};
    const bool has_src = op.src->valid != nullptr;
    for (const Group &g : J.groups) {
        DstView d[kMaxOutputs];
        for (int i = 0; i < g.nout; ++i) {
            const int o = g.out[i];
            d[i].p = op.dst[o]->data + (rc.y0 - op.dy[o]) * op.dst[o]->stride + (rc.x0 - op.dx[o]);
            d[i].stride = op.dst[o]->stride;
        }
        hipEvent_t e0 = timed ? timing_event(st) : nullptr;
        bool fused = false;
        if (g.nout == 1 && op.dst[g.out[0]]->valid) {
            // A radius-1 Horn op writes its one output's validity in the
            // same launch as its data.
            const int o = g.out[0];
            const MaskView dm = mask_view(*op.dst[o], op.dx[o], op.dy[o], op.W, op.H);
            const MaskView sm = has_src ? mask_view(*op.src, op.sx, op.sy, op.W, op.H) : dm;
            HIP_TRY(launch_data_valid(g.P, rg, s, d[0], sm, has_src ? 1 : 0, dm, st, &fused),
                    "launching the data and validity kernel");
            valid_done[o] = fused && J.radius[o] == 1;
        }
        if (!fused) HIP_TRY(launch_data(g.P, rg, s, d, g.nout, st), "launching the data kernel");
        if (timed) {
            std::lock_guard<std::mutex> lk(g_timing_mu);
            g_timing.data.emplace_back(e0, timing_event(st));
        }
    }
    for (int o = 0; o < op.nout; ++o) {
        if (!op.dst[o]->valid || valid_done[o]) continue;
        const MaskView dm = mask_view(*op.dst[o], op.dx[o], op.dy[o], op.W, op.H);
        const MaskView sm = has_src ? mask_view(*op.src, op.sx, op.sy, op.W, op.H) : dm;
        hipEvent_t v0 = timed ? timing_event(st) : nullptr;
        HIP_TRY(launch_validity(J.radius[o], rg, sm, has_src ? 1 : 0, dm, st), "launching the validity kernel");
        if (timed) {
            std::lock_guard<std::mutex> lk(g_timing_mu);
            g_timing.validity.emplace_back(v0, timing_event(st));
        }
    }
    return STRATA_OK;
}

bool cancelled(const strata_options *o) { return o && o->cancel && *o->cancel != 0; }

// Device work to claim between cancellation checks: whole bands of the
// plan, within one tile, at least this many cells (about 1.6 ms of a
// radius-1 sweep). Each unit is a launch of its own, which drains the
// device at its end (about 10 us), so units stay large: a call of one unit
// runs as the plain call does.
constexpr int64_t kUnitCells = (int64_t)1 << 30;

int check_operands(const strata_raster *dst, int ndst, const strata_raster *src, int nsrc,
                   const strata_options *opts)
{
    int rc = STRATA_OK;
    if (!dst || !src) return fail(STRATA_EINVAL, "engine: nil operand list");
    if (opts && (opts->tile_width < 0 || opts->tile_height < 0 || opts->workers < 0))
        return fail(STRATA_EINVAL, "engine: negative Options {TileWidth:%lld TileHeight:%lld Workers:%lld}",
                    (long long)opts->tile_width, (long long)opts->tile_height, (long long)opts->workers);
    for (int i = 0; i < ndst; ++i)
        if ((rc = check_raster("dst", i, dst[i]))) return rc;
    for (int i = 0; i < nsrc; ++i)
        if ((rc = check_raster("src", i, src[i]))) return rc;
    for (int i = 0; i < nsrc; ++i)
        if (src[i].width != dst[0].width || src[i].height != dst[0].height)
            return fail(STRATA_EINVAL, "engine: src[%d] is %lld×%lld, dst[0] is %lld×%lld", i,
                        (long long)src[i].width, (long long)src[i].height, (long long)dst[0].width,
                        (long long)dst[0].height);
    for (int i = 1; i < ndst; ++i)
        if (dst[i].width != dst[0].width || dst[i].height != dst[0].height)
            return fail(STRATA_EINVAL, "engine: dst[%d] is %lld×%lld, dst[0] is %lld×%lld", i,
                        (long long)dst[i].width, (long long)dst[i].height, (long long)dst[0].width,
                        (long long)dst[0].height);
    bool masked = false;
    for (int i = 0; i < nsrc; ++i) masked = masked || src[i].valid;
    for (int i = 0; i < ndst; ++i)
        if (masked && !dst[i].valid)
            return fail(STRATA_EINVAL,
                        "engine: an input has a validity mask but dst[%d].Valid is nil; allocate it with "
                        "raster.NewFloat32Like, or set Valid on its root raster before windowing",
                        i);
    return STRATA_OK;
}

int ensure_device()
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    return STRATA_OK;
}

// Stream-ordered scratch of the chunked passes (the buffers of outputs
// that are only folded) comes from a pool of the library's own per device
// that keeps what calls free: a later call of the same shape reuses the
// memory instead of mapping it again. The pool holds at most the largest
// scratch one call used, until strata_trim.
constexpr int kMaxDevices = 64;
std::mutex g_pool_mu;
hipMemPool_t g_pools[kMaxDevices] = {};

int scratch_pool(hipMemPool_t *out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(STRATA_EDEVICE, "strata: device %d past the pool table", dev);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!g_pools[dev]) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        HIP_TRY(hipMemPoolCreate(&pool, &props), "creating the scratch pool");
        uint64_t keep = UINT64_MAX;
        const hipError_t e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        if (e != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            return hip_fail(e, "configuring the scratch pool");
        }
        g_pools[dev] = pool;
    }
    *out = g_pools[dev];
    return STRATA_OK;
}

// Large scratch (64 MiB and up: rows of outputs that are only folded,
// mirrors of host rasters, a chain's intermediate values) comes from blocks
// of the library's own per device, allocated once and reused: taken from
// the stream-ordered pool, each such buffer was mapped again on every call
// (a fold-only pass over 1.6 GB ran 0.45 ms slower than one writing the
// caller's raster). A block goes back to the free list when the call's
// stream reaches the point where the call gives it up (a host function on
// the stream), so a later call, on any stream, only takes it once every
// use is done. strata_trim returns the free blocks.
constexpr size_t kBigScratch = (size_t)64 << 20;
struct BigBlock {
    int dev;
    size_t bytes;
};
std::mutex g_big_mu;
std::multimap<std::pair<int, size_t>, void *> g_big_free;  // (device, bytes) -> block
std::map<void *, BigBlock> g_big_used;

int scratch_alloc(void **p, size_t bytes, hipStream_t st)
{
    *p = nullptr;
    if (bytes < kBigScratch) {
        hipMemPool_t pool = nullptr;
        int rc = scratch_pool(&pool);
        if (rc) return rc;
        HIP_TRY(hipMallocFromPoolAsync(p, bytes, pool, st), "scratch");
        return STRATA_OK;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    {
        // The smallest free block that holds the request and is at most
        // twice its size.
        std::lock_guard<std::mutex> lk(g_big_mu);
        auto it = g_big_free.lower_bound({dev, bytes});
        if (it != g_big_free.end() && it->first.first == dev && it->first.second <= 2 * bytes) {
            *p = it->second;
            g_big_used[*p] = BigBlock{dev, it->first.second};
            g_big_free.erase(it);
            return STRATA_OK;
        }
    }
    const size_t rounded = (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
    HIP_TRY(hipMalloc(p, rounded), "scratch block");
    std::lock_guard<std::mutex> lk(g_big_mu);
    g_big_used[*p] = BigBlock{dev, rounded};
    return STRATA_OK;
}

void big_release(void *p)
{
    std::lock_guard<std::mutex> lk(g_big_mu);
    auto it = g_big_used.find(p);
    if (it == g_big_used.end()) return;
    g_big_free.emplace(std::make_pair(it->second.dev, it->second.bytes), p);
    g_big_used.erase(it);
}

// Gives p back once st reaches this point.
void scratch_free(void *p, hipStream_t st)
{
    if (!p) return;
    bool big = false;
    {
        std::lock_guard<std::mutex> lk(g_big_mu);
        big = g_big_used.count(p) != 0;
    }
    if (!big) {
        (void)hipFreeAsync(p, st);
        return;
    }
    if (hipLaunchHostFunc(st, [](void *q) { big_release(q); }, p) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        big_release(p);
    }
}

// Frees the free blocks (those of finished calls).
int big_trim()
{
    std::vector<void *> blocks;
    {
        std::lock_guard<std::mutex> lk(g_big_mu);
        for (auto &b : g_big_free) blocks.push_back(b.second);
        g_big_free.clear();
    }
    for (void *b : blocks) HIP_TRY(hipFree(b), "freeing a scratch block");
    return STRATA_OK;
}

// A host-memory raster mirrored in device memory for a call: the
// raster's whole span (cells and padding) and the mask words its bits
// touch, in the library's scratch (reused from call to call, given back
// when the call's stream reaches the end of the call).
struct Staged {
    strata_raster dev{};
    void *dbuf = nullptr;
    uint64_t *dmask = nullptr;
    int64_t w0 = 0, nwords = 0;
    hipStream_t st = nullptr;
    ~Staged()
    {
        scratch_free(dbuf, st);
        scratch_free(dmask, st);
    }
};

// Pageable host memory to and from device memory through pinned buffers,
// overlapping the host copies with the DMA (defined with the IO pool).
int host_to_device(void *dev, const void *host, size_t bytes, hipStream_t st);
int device_to_host_2d(void *host, size_t hpitch, const void *dev, size_t dpitch, size_t width, size_t rows,
                      hipStream_t st);

int stage_in(const strata_raster &h, Staged &s, bool copy_data, hipStream_t st)
{
    int rc = STRATA_OK;
    s.st = st;
    s.dev = h;
    s.dev.mem = STRATA_MEM_DEVICE;
    const int64_t n = data_len(h);
    if ((rc = scratch_alloc(&s.dbuf, (size_t)n * sizeof(float), st))) return rc;
    s.dev.data = (float *)s.dbuf;
    if (copy_data && (rc = host_to_device(s.dbuf, h.data, (size_t)n * sizeof(float), st))) return rc;
    if (h.valid) {
        s.w0 = h.valid_offset >> 6;
        const int64_t w1 = (h.valid_offset + n - 1) >> 6;
        s.nwords = w1 - s.w0 + 1;
        if ((rc = scratch_alloc((void **)&s.dmask, (size_t)s.nwords * 8, st))) return rc;
        if ((rc = host_to_device(s.dmask, h.valid + s.w0, (size_t)s.nwords * 8, st))) return rc;
        s.dev.valid = s.dmask;
        s.dev.valid_offset = h.valid_offset - 64 * s.w0;
    }
    return STRATA_OK;
}

// The staged output's cells (rows of width cells, not the padding) and the
// mask words its bits touch back into the caller's raster. The mask words
// were copied in whole, so bits of other cells go back as they came.
int stage_out(const strata_raster &h, Staged &s, hipStream_t st)
{
    int rc = device_to_host_2d(h.data, (size_t)h.stride * 4, s.dbuf, (size_t)h.stride * 4, (size_t)h.width * 4,
                               (size_t)h.height, st);
    if (rc) return rc;
    if (h.valid && (rc = device_to_host_2d(h.valid + s.w0, (size_t)s.nwords * 8, s.dmask, (size_t)s.nwords * 8,
                                           (size_t)s.nwords * 8, 1, st)))
        return rc;
    return STRATA_OK;
}

// Device memory of strata_reduce_noalloc: the fold's bins, its running
// total and one partial per workgroup of the fold launch. Each GPU the
// library runs on holds its own copy from the moment the code object is
// loaded until the process ends. Every call on one GPU uses that GPU's copy.
constexpr size_t kNoAllocScratchBytes = 128 << 10;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The largest raster strata_reduce_noalloc folds: the int64 bins hold the
// exact sums of at most 2^40 values.
constexpr int64_t kNoAllocMaxCells = (int64_t)1 << 40;

// Frees the pinned host buffers no call holds (defined with their cache).
void pinned_trim();

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current. In host code the array's name is only a stand-in for
// the symbol; device work must be given this address.
int scratch_address(void **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(STRATA_EDEVICE, "strata: device %d past the scratch table", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "locating the no-allocation scratch");
        if (!p) return fail(STRATA_EDEVICE, "strata: the no-allocation scratch has no device address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return STRATA_OK;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int strata_init(int device)
{
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    HIP_TRY(hipFree(nullptr), "device init");
    void *p = nullptr;
    const int rc = scratch_address(&p);
    if (rc) return rc;
    g_err.clear();
    return STRATA_OK;
}

const char *strata_last_error(void) { return g_err.c_str(); }

const char *strata_version(void) { return "strata-hip 0.1 gfx950"; }

int strata_trim(void)
{
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        for (hipMemPool_t p : g_pools)
            if (p) HIP_TRY(hipMemPoolTrimTo(p, 0), "trimming the scratch pool");
    }
    int rc = big_trim();
    if (rc) return rc;
    pinned_trim();
    g_err.clear();
    return STRATA_OK;
}

int strata_selfcheck(int kind, uint64_t lo, uint64_t n, uint64_t *mismatches, uint64_t *first, uint64_t *checked)
{
    if (kind < 0 || kind > 3) return fail(STRATA_EINVAL, "strata: unknown self-check %d", kind);
    unsigned long long *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, 3 * sizeof(unsigned long long)), "self-check buffer");
    unsigned long long h[3] = {0, ~0ull, 0};
    hipError_t e = hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_selfcheck(kind, lo, n, d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "self-check");
    if (mismatches) *mismatches = h[0];
    if (first) *first = h[1];
    if (checked) *checked = h[2];
    g_err.clear();
    return STRATA_OK;
}

int strata_correlate_error(const strata_kernel *k, const strata_raster *src, const strata_raster *out,
                           double *max_abs, double *max_rel, double *max_scaled, int64_t *cells)
{
    int rc = check_kernel(k, 1, 1);
    if (rc) return rc;
    if (k->op != STRATA_OP_CORRELATE) return fail(STRATA_EINVAL, "strata: the error report is for Correlate kernels");
    if (!src || !out) return fail(STRATA_EINVAL, "strata: nil raster");
    if ((rc = check_raster("src", 0, *src)) || (rc = check_raster("dst", 0, *out))) return rc;
    if (src->mem != STRATA_MEM_DEVICE || out->mem != STRATA_MEM_DEVICE)
        return fail(STRATA_EINVAL, "strata: the error report reads device rasters");
    if (src->width != out->width || src->height != out->height)
        return fail(STRATA_EINVAL, "engine: src[0] is %lld×%lld, dst[0] is %lld×%lld", (long long)src->width,
                    (long long)src->height, (long long)out->width, (long long)out->height);
    const Params P = make_params(k);
    const Region rg{src->width, src->height, 0, 0, src->width, src->height};
    const SrcView sv = srcview(*src, 0, 0, src->width, src->height);
    unsigned long long *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, 4 * sizeof(unsigned long long)), "error report buffer");
    unsigned long long h[4] = {0, 0, 0, 0};
    hipError_t e = hipMemcpy(d, h, sizeof h, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_correlate_error(P, rg, sv, out->data, out->stride, d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "error report");
    double v[3];
    memcpy(v, h, sizeof v);
    if (max_abs) *max_abs = v[0];
    if (max_rel) *max_rel = v[1];
    if (max_scaled) *max_scaled = v[2];
    if (cells) *cells = (int64_t)h[3];
    g_err.clear();
    return STRATA_OK;
}

int strata_get_tuning(strata_tuning *t)
{
    if (!t) return fail(STRATA_EINVAL, "strata: nil tuning");
    const Tuning g = get_tuning();
    memset(t, 0, sizeof *t);
    t->sweep_rows = g.sweep_rows;
    t->fold_stripe_mb = g.fold_stripe_mb;
    t->unit_kcells = g.unit_kcells;
    t->stop_after_units = g.stop_after_units;
    return STRATA_OK;
}

int strata_set_tuning(const strata_tuning *t)
{
    if (!t) return fail(STRATA_EINVAL, "strata: nil tuning");
    Tuning g = get_tuning();
    g.sweep_rows = t->sweep_rows;
    g.fold_stripe_mb = t->fold_stripe_mb;
    g.unit_kcells = t->unit_kcells;
    g.stop_after_units = t->stop_after_units;
    if (!set_tuning(g))
        return fail(STRATA_EINVAL,
                    "strata: sweep_rows %d outside [4, 4096], fold_stripe_mb %d outside [1, 4096], unit_kcells %d "
                    "outside [0, 1048576] or stop_after_units %d negative",
                    t->sweep_rows, t->fold_stripe_mb, t->unit_kcells, t->stop_after_units);
    g_err.clear();
    return STRATA_OK;
}

int strata_timing_begin(void)
{
    std::lock_guard<std::mutex> lk(g_timing_mu);
    for (auto &p : g_timing.data) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
    for (auto &p : g_timing.validity) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
    g_timing.data.clear();
    g_timing.validity.clear();
    g_timing.on = true;
    return STRATA_OK;
}

int strata_timing_end(double *data_ms, int64_t *data_launches, double *validity_ms, int64_t *validity_launches)
{
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing.on = false;
    double dms = 0, vms = 0;
    for (auto &p : g_timing.data) {
        float ms = 0;
        HIP_TRY(hipEventSynchronize(p.second), "timing");
        HIP_TRY(hipEventElapsedTime(&ms, p.first, p.second), "timing");
        dms += ms;
    }
    for (auto &p : g_timing.validity) {
        float ms = 0;
        HIP_TRY(hipEventSynchronize(p.second), "timing");
        HIP_TRY(hipEventElapsedTime(&ms, p.first, p.second), "timing");
        vms += ms;
    }
    if (data_ms) *data_ms = dms;
    if (data_launches) *data_launches = (int64_t)g_timing.data.size();
    if (validity_ms) *validity_ms = vms;
    if (validity_launches) *validity_launches = (int64_t)g_timing.validity.size();
    return STRATA_OK;
}

} // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Raw files (engine/raw.go)
// ---------------------------------------------------------------------------

// Host threads for raw file IO: a pool created on first use and kept for
// the process (never destroyed, so no thread is joined at exit). run(n, f)
// calls f(0..n-1) on the pool and the calling thread and returns when all
// are done; calls from several threads take turns.
class IoPool {
  public:
    static IoPool &get()
    {
        static IoPool *p = new IoPool();
        return *p;
    }
    void run(int n, const std::function<void(int)> &f)
    {
        std::lock_guard<std::mutex> turn(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            n_ = n;
            next_ = 0;
            left_ = n;
            ++gen_;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return left_ == 0; });
        job_ = nullptr;
    }
    int threads() const { return (int)th_.size() + 1; }

  private:
    IoPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const int n = (int)std::min(7u, hw > 1 ? hw - 1 : 1u);
        for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
    }
    // Claims parts of the current job until none is left.
    void work()
    {
        for (;;) {
            int k;
            const std::function<void(int)> *f;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (!job_ || next_ >= n_) return;
                k = next_++;
                f = job_;
            }
            (*f)(k);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_.notify_all();
        }
    }
    void loop()
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            work();
        }
    }
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> th_;
    const std::function<void(int)> *job_ = nullptr;
    int n_ = 0, next_ = 0, left_ = 0;
    uint64_t gen_ = 0;
};

// Rows [r0, r1) of a W-wide raw file at (x, y) to or from buf (row stride
// `stride`): one pread/pwrite per row, or one for all of them when they are
// contiguous both in the file and in buf. Returns -1, or the first row (of
// the file) that failed, with *eof telling a short read from an errno.
template <bool WRITE>
int64_t raw_rows(const strata_raw &r, int64_t W, char *buf, int64_t w, int64_t stride, int64_t x, int64_t y,
                 int64_t r0, int64_t r1, bool *eof, int *err)
{
    const bool contiguous = x == 0 && w == W && stride == W;
    const int64_t step = contiguous ? r1 - r0 : 1;
    for (int64_t row = r0; row < r1; row += step) {
        const int64_t nrows = std::min(step, r1 - row);
        char *p = buf + 4 * row * stride;
        const size_t total = (size_t)(4 * w * nrows);
        size_t done = 0;
        off_t off = (off_t)(r.offset + 4 * ((y + row) * W + x));
        while (done < total) {
            const ssize_t n = WRITE ? pwrite(r.fd, p + done, total - done, off) : pread(r.fd, p + done, total - done, off);
            if (n < 0 && errno == EINTR) continue;
            if (n <= 0) {
                *eof = n == 0;
                *err = n == 0 ? 0 : errno;
                return y + row + (int64_t)(done / (size_t)(4 * w));
            }
            done += (size_t)n;
            off += n;
        }
    }
    return -1;
}

// Rows of a tile split into parts of at least 2 MiB for the IO pool; the
// lowest failing row is the one reported, as a row-by-row loop reports it.
// Rows of a part past a failure in another part may still have moved.
template <bool WRITE>
int64_t raw_io(const strata_raw &r, int64_t W, char *buf, int64_t w, int64_t h, int64_t stride, int64_t x,
               int64_t y, bool *eof, int *err)
{
    const int64_t bytes = 4 * w * h;
    IoPool &pool = IoPool::get();
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)pool.threads(), h, bytes >> 21}));
    if (parts == 1) return raw_rows<WRITE>(r, W, buf, w, stride, x, y, 0, h, eof, err);
    std::vector<int64_t> bad(parts, -1);
    std::vector<char> beof(parts, 0);
    std::vector<int> berr(parts, 0);
    pool.run(parts, [&](int k) {
        bool e = false;
        int en = 0;
        bad[k] = raw_rows<WRITE>(r, W, buf, w, stride, x, y, h * k / parts, h * (k + 1) / parts, &e, &en);
        beof[k] = e;
        berr[k] = en;
    });
    for (int k = 0; k < parts; ++k)
        if (bad[k] >= 0) {
            *eof = beof[k] != 0;
            *err = berr[k];
            return bad[k];
        }
    return -1;
}

// Reads rows y..y+h of columns x..x+w of a W-wide raw file into buf (rows
// `stride` floats apart). Errors as the reference's wrap of a raw source's
// error (internal/exec/chunked.go:379-393, engine/raw.go:86-92).
int raw_read(const strata_raw &r, int64_t W, float *buf, int64_t w, int64_t h, int64_t stride, int64_t x, int64_t y,
             const char *what)
{
    bool eof = false;
    int err = 0;
    const int64_t row = raw_io<false>(r, W, (char *)buf, w, h, stride, x, y, &eof, &err);
    if (row < 0) return STRATA_OK;
    return fail(STRATA_EIO, "engine: reading %s at (%lld, %lld): engine: raw source: reading row %lld: %s", what,
                (long long)x, (long long)y, (long long)row, eof ? "unexpected EOF" : strerror(err));
}

// Writes rows of buf into the W-wide raw file at (x, y).
int raw_write(const strata_raw &r, int64_t W, const float *buf, int64_t w, int64_t h, int64_t stride, int64_t x,
              int64_t y, int o)
{
    bool eof = false;
    int err = 0;
    const int64_t row = raw_io<true>(r, W, (char *)buf, w, h, stride, x, y, &eof, &err);
    if (row < 0) return STRATA_OK;
    return fail(STRATA_EIO, "engine: writing dst[%d] at (%lld, %lld): engine: raw sink: writing row %lld: %s", o,
                (long long)x, (long long)y, (long long)row, eof ? "short write" : strerror(err));
}

int check_raw(const strata_raw *r, int32_t masked, const char *name, int i)
{
    if (r->fd < 0) return fail(STRATA_EINVAL, "engine: %s[%d]: nil file", name, i);
    if ((r->has_fill != 0) != (masked != 0))
        return fail(STRATA_EINVAL, "engine: %s[%d]: a raw file is Masked exactly when it has a fill value", name, i);
    if (r->offset < 0) return fail(STRATA_EINVAL, "engine: %s[%d]: negative file offset", name, i);
    return STRATA_OK;
}

int plan_stats(const Job &J, int nsrc, int64_t width, int64_t height, const strata_options *opts, int chunked,
               strata_stats *stats)
{
    const int ndst = J.nout;
    if (width <= 0 || height <= 0) return fail(STRATA_EINVAL, "engine: sizes must be positive");
    strata_stats s{};
    if (chunked) {
        const TileGrid g = tile_grid(width, height, opts);
        for (int64_t t = 0; t < g.tiles; ++t) tile_stats(s, J.r, J.rmin, g, nsrc, ndst, tile_rect(g, t));
    } else {
        const Plan p = new_plan(width, height, opts ? opts->tile_width : 0, opts ? opts->tile_height : 0);
        for (int64_t b = 0; b < p.bands; ++b)
            band_stats(s, J.r, J.rmin, width, height, nsrc, ndst, plan_band(p, b));
    }
    add_stats(stats, s, nsrc, ndst);
    g_err.clear();
    return STRATA_OK;
}

// ---------------------------------------------------------------------------
// Reused host memory and streams of staged and cancellable calls
// ---------------------------------------------------------------------------

// Pinned host buffers kept for reuse: a call takes the smallest free buffer
// that holds what it needs, or maps a new one (rounded up to 2 MiB), and
// gives it back when it ends, so calls of a repeated shape map nothing.
// strata_trim frees the buffers not in use.
class PinnedCache {
  public:
    static PinnedCache &get()
    {
        static PinnedCache *c = new PinnedCache();
        return *c;
    }
    int acquire(size_t bytes, void **out)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(bytes);
            if (it != free_.end()) {
                *out = it->second;
                free_.erase(it);
                return STRATA_OK;
            }
        }
        const size_t cap = (bytes + ((size_t)2 << 20) - 1) & ~(((size_t)2 << 20) - 1);
        void *p = nullptr;
        HIP_TRY(hipHostMalloc(&p, cap, hipHostMallocDefault), "pinned buffer");
        std::lock_guard<std::mutex> lk(mu_);
        size_[p] = cap;
        *out = p;
        return STRATA_OK;
    }
    void release(void *p)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(size_[p], p);
    }
    void trim()
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto &e : free_) {
            (void)hipHostFree(e.second);
            size_.erase(e.second);
        }
        free_.clear();
    }

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> size_;
};

void pinned_trim() { PinnedCache::get().trim(); }

// The pieces pageable copies move through pinned memory in, and the size
// below which the runtime's own copy is used.
constexpr size_t kHostPiece = (size_t)32 << 20, kHostDirect = (size_t)4 << 20;

// A copy between pageable and pinned memory split over the IO pool's
// threads, in parts of at least 2 MiB.
void pool_memcpy(void *dst, const void *src, size_t n)
{
    IoPool &pool = IoPool::get();
    const int parts = (int)std::max<size_t>(1, std::min<size_t>((size_t)pool.threads(), n >> 21));
    if (parts == 1) {
        memcpy(dst, src, n);
        return;
    }
    pool.run(parts, [&](int k) {
        const size_t a = n * (size_t)k / (size_t)parts, b = n * (size_t)(k + 1) / (size_t)parts;
        memcpy(static_cast<char *>(dst) + a, static_cast<const char *>(src) + a, b - a);
    });
}

// Two pinned pieces and their events, given back to the cache when done.
struct PinnedPair {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~PinnedPair()
    {
        for (int k = 0; k < 2; ++k) {
            if (ev[k]) {
                (void)hipEventSynchronize(ev[k]);
                (void)hipEventDestroy(ev[k]);
            }
            PinnedCache::get().release(buf[k]);
        }
    }
    int init()
    {
        for (int k = 0; k < 2; ++k) {
            int rc = PinnedCache::get().acquire(kHostPiece, &buf[k]);
            if (rc) return rc;
            HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming), "staging event");
        }
        return STRATA_OK;
    }
};

// Host to device: the IO pool copies piece i into one pinned buffer while
// the DMA of piece i-1 runs from the other (the runtime's copy from
// pageable memory runs the two steps one after the other).
int host_to_device(void *dev, const void *host, size_t bytes, hipStream_t st)
{
    if (bytes < kHostDirect) {
        HIP_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st), "copy in");
        return STRATA_OK;
    }
    PinnedPair pp;
    int rc = pp.init();
    if (rc) return rc;
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kHostPiece, ++i) {
        const int k = (int)(i & 1);
        const size_t n = std::min(kHostPiece, bytes - off);
        if (i >= 2) HIP_TRY(hipEventSynchronize(pp.ev[k]), "waiting for a staging buffer");
        pool_memcpy(pp.buf[k], static_cast<const char *>(host) + off, n);
        HIP_TRY(hipMemcpyAsync(static_cast<char *>(dev) + off, pp.buf[k], n, hipMemcpyHostToDevice, st), "copy in");
        HIP_TRY(hipEventRecord(pp.ev[k], st), "staging event");
    }
    return STRATA_OK;
}

// Copies n bits from src at bit sb to dst at bit db (LSB-first 64-bit
// words); dst's other bits stay as they were.
void copy_bits(uint64_t *dst, int64_t db, const uint64_t *src, int64_t sb, int64_t n)
{
    while (n > 0) {
        const int64_t dw = db >> 6, sw = sb >> 6;
        const int dbit = (int)(db & 63), sbit = (int)(sb & 63);
        const int m = (int)std::min<int64_t>(n, 64 - dbit);
        uint64_t v = src[sw] >> sbit;
        if (sbit + m > 64) v |= src[sw + 1] << (64 - sbit);
        const uint64_t mask = m == 64 ? ~0ull : ((1ull << m) - 1);
        dst[dw] = (dst[dw] & ~(mask << dbit)) | ((v & mask) << dbit);
        db += m;
        sb += m;
        n -= m;
    }
}

// A window of a host-memory raster to and from a tile buffer, as
// engine.MemorySource.ReadWindow and MemorySink.WriteWindow do: cells of w x h
// at (x, y), rows `stride` cells apart in the buffer, the buffer's bit of its
// cell (i, j) at voff + j*stride + i. Rows are copied by the IO pool; bits
// one row after another (rows of the raster may share mask words).
void host_window(const strata_raster &r, bool write, float *buf, uint64_t *bits, int64_t w, int64_t h,
                 int64_t stride, int64_t voff, int64_t x, int64_t y)
{
    IoPool &pool = IoPool::get();
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)pool.threads(), h, (4 * w * h) >> 21}));
    pool.run(parts, [&](int k) {
        for (int64_t j = h * k / parts; j < h * (k + 1) / parts; ++j) {
            float *rp = r.data + (y + j) * r.stride + x, *bp = buf + j * stride;
            if (write)
                memcpy(rp, bp, (size_t)w * 4);
            else
                memcpy(bp, rp, (size_t)w * 4);
        }
    });
    if (!bits || !r.valid) return;
    for (int64_t j = 0; j < h; ++j) {
        const int64_t rb = r.valid_offset + (y + j) * r.stride + x, bb = voff + j * stride;
        if (write)
            copy_bits(r.valid, rb, bits, bb, w);
        else
            copy_bits(bits, bb, r.valid, rb, w);
    }
}

// Device to host, rows of `width` bytes (dpitch apart on the device, hpitch
// in host memory): whole device rows go to a pinned buffer by DMA, two
// pieces in flight, and the IO pool copies each piece's rows out while the
// next piece's DMA runs.
int device_to_host_2d(void *host, size_t hpitch, const void *dev, size_t dpitch, size_t width, size_t rows,
                      hipStream_t st)
{
    if (rows == 0 || width == 0) return STRATA_OK;
    const size_t piece = kHostPiece / 4;
    if (rows == 1 && width > piece) {
        // One long row: rows of `piece` bytes, then the rest (each call
        // below has rows of at most `piece` bytes, so it does not split again).
        const size_t full = width / piece;
        int rc = full ? device_to_host_2d(host, piece, dev, piece, piece, full, st) : STRATA_OK;
        if (rc || width == full * piece) return rc;
        return device_to_host_2d(static_cast<char *>(host) + full * piece, width - full * piece,
                                 static_cast<const char *>(dev) + full * piece, width - full * piece,
                                 width - full * piece, 1, st);
    }
    if (dpitch * rows < kHostDirect || dpitch > kHostPiece) {
        HIP_TRY(hipMemcpy2DAsync(host, hpitch, dev, dpitch, width, rows, hipMemcpyDeviceToHost, st), "copy out");
        HIP_TRY(hipStreamSynchronize(st), "copy out");
        return STRATA_OK;
    }
    PinnedPair pp;
    int rc = pp.init();
    if (rc) return rc;
    const size_t per = kHostPiece / dpitch;  // rows per piece
    const size_t pieces = (rows + per - 1) / per;
    auto issue = [&](size_t i) -> int {
        const int k = (int)(i & 1);
        const size_t r0 = i * per, nr = std::min(per, rows - r0);
        // The last row ends at its width: the device buffer may end there.
        HIP_TRY(hipMemcpyAsync(pp.buf[k], static_cast<const char *>(dev) + r0 * dpitch, (nr - 1) * dpitch + width,
                               hipMemcpyDeviceToHost, st),
                "copy out");
        HIP_TRY(hipEventRecord(pp.ev[k], st), "staging event");
        return STRATA_OK;
    };
    for (size_t i = 0; i < std::min<size_t>(2, pieces); ++i)
        if ((rc = issue(i))) return rc;
    for (size_t i = 0; i < pieces; ++i) {
        const int k = (int)(i & 1);
        HIP_TRY(hipEventSynchronize(pp.ev[k]), "waiting for a staging buffer");
        const size_t r0 = i * per, nr = std::min(per, rows - r0);
        const char *src = static_cast<const char *>(pp.buf[k]);
        char *dst = static_cast<char *>(host) + r0 * hpitch;
        if (hpitch == dpitch && width == dpitch)
            pool_memcpy(dst, src, nr * width);
        else {
            IoPool &pool = IoPool::get();
            const int parts = (int)std::max<size_t>(1, std::min<size_t>((size_t)pool.threads(), nr));
            pool.run(parts, [&](int q) {
                for (size_t r = nr * (size_t)q / (size_t)parts; r < nr * (size_t)(q + 1) / (size_t)parts; ++r)
                    memcpy(dst + r * hpitch, src + r * dpitch, width);
            });
        }
        if (i + 2 < pieces && (rc = issue(i + 2))) return rc;
    }
    return STRATA_OK;
}

// Device memory for a call's buffers from the library's stream-ordered pool.
int pool_alloc(void **p, size_t bytes, hipStream_t st, const char *what)
{
    hipMemPool_t pool = nullptr;
    int rc = scratch_pool(&pool);
    if (rc) return rc;
    HIP_TRY(hipMallocFromPoolAsync(p, bytes, pool, st), what);
    return STRATA_OK;
}

// Cells per unit of a cancellable call: a sixteenth of the raster, at
// least kUnitCells, unless tuned.
int64_t unit_target(int64_t W, int64_t H)
{
    const int64_t tuned = (int64_t)get_tuning().unit_kcells * 1024;
    if (tuned > 0) return tuned;
    return std::max<int64_t>(kUnitCells, cdiv(W * H, 16));
}

// A cancellable call: units of the plan (plan_unit) claimed in order on the
// call's stream, kUnitsAhead in flight. Before claiming unit k the call
// waits for unit k-kUnitsAhead, so while the host waits the device holds
// the next units. The flag is checked before every claim and a claimed
// unit always finishes, so the finished bands are a prefix of the plan
// (internal/exec/worker.go:144-181, internal/exec/doc.go:165-183). A plan
// of one unit is one launch. Everything runs on the call's stream: a
// second stream overlapped the ends of small units, but waits between two
// hardware queues cost about 50 us a call, and the runtime maps streams
// onto few hardware queues, so another stream of the caller's could queue
// behind the library's. Adds the finished units' band Stats to s; *any
// tells whether a unit ran.
constexpr int kUnitsAhead = 4;
int run_units(const Job &J, const Operands &op, const Plan &p, const strata_options *opts, hipStream_t st,
              int nsrc, strata_stats &s, bool *any)
{
    const int64_t W = op.W, H = op.H;
    const int64_t target = unit_target(W, H);
    const int32_t limit = get_tuning().stop_after_units;
    int rc = STRATA_OK;
    const Unit first = plan_unit(p, 0, target);
    if (first.b1 >= p.bands) {
        if (cancelled(opts)) return STRATA_CANCELLED;
        if ((rc = run_rect(J, op, first.rc, st))) return rc;
        *any = true;
        for (int64_t i = first.b0; i < first.b1; ++i)
            band_stats(s, J.r, J.rmin, W, H, nsrc, op.nout, plan_band(p, i));
        return STRATA_OK;
    }
    hipEvent_t ev[kUnitsAhead] = {};  // one per unit in flight
    struct Guard {
        hipEvent_t *e;
        ~Guard()
        {
            for (int k = 0; k < kUnitsAhead; ++k)
                if (e[k]) (void)hipEventDestroy(e[k]);
        }
    } guard{ev};
    for (int k = 0; k < kUnitsAhead; ++k)
        HIP_TRY(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming), "unit event");
    Unit pend[kUnitsAhead];
    bool live[kUnitsAhead] = {};
    auto stats_of = [&](int k) {
        live[k] = false;
        for (int64_t i = pend[k].b0; i < pend[k].b1; ++i)
            band_stats(s, J.r, J.rmin, W, H, nsrc, op.nout, plan_band(p, i));
    };
    int status = STRATA_OK;
    int64_t n = 0;
    for (int64_t i = 0; i < p.bands; ++n) {
        const int k = (int)(n % kUnitsAhead);
        if (live[k]) {
            HIP_TRY(hipEventSynchronize(ev[k]), "waiting for a unit");
            stats_of(k);
        }
        if (cancelled(opts) || (limit > 0 && n >= limit)) {
            status = STRATA_CANCELLED;
            break;
        }
        const Unit u = plan_unit(p, i, target);
        if ((rc = run_rect(J, op, u.rc, st))) return rc;
        HIP_TRY(hipEventRecord(ev[k], st), "unit event");
        pend[k] = u;
        live[k] = true;
        *any = true;
        i = u.b1;
    }
    // Claimed units always finish, cancelled or not: one wait, then their
    // Stats in claim order.
    HIP_TRY(hipStreamSynchronize(st), "waiting for the units");
    for (int64_t j = 0; j < kUnitsAhead; ++j) {
        const int k = (int)((n + j) % kUnitsAhead);
        if (live[k]) stats_of(k);
    }
    return status;
}

int process_chunked(const Job &J, const strata_sink *dst, int ndst, const strata_source *src, int nsrc,
                    const strata_options *opts, strata_stats *stats);

// Rows of a tile when a plain call's host operands stream through the tile
// pipeline: about 16 MiB of cells per tile, at least 64 rows.
int64_t host_tile_rows(int64_t W) { return std::max<int64_t>(64, ((int64_t)16 << 20) / (4 * W)); }

int process(const Job &J, const strata_raster *dst, int ndst, const strata_raster *src, int nsrc,
            const strata_options *opts, strata_stats *stats)
{
    int rc = check_operands(dst, ndst, src, nsrc, opts);
    if (rc) return rc;
    if ((rc = ensure_device())) return rc;
    hipStream_t st = opts ? (hipStream_t)opts->stream : nullptr;
    const int64_t W = dst[0].width, H = dst[0].height;
    const Plan p = new_plan(W, H, opts ? opts->tile_width : 0, opts ? opts->tile_height : 0);
    const bool can_cancel = opts && opts->cancel;
    // A context done before the call writes nothing (internal/exec/doc.go:
    // 165-178).
    if (can_cancel && cancelled(opts)) {
        g_err.clear();
        return STRATA_CANCELLED;
    }

    // Host operands and nothing to poll: the call streams through the
    // chunked path's tile pipeline (a tile's rows copied in while the device
    // computes the previous tile and the one before is copied out), which
    // computes the same cells; its Stats are the plain plan's.
    bool any_host = false;
    for (int i = 0; i < nsrc; ++i) any_host = any_host || src[i].mem == STRATA_MEM_HOST;
    for (int o = 0; o < ndst; ++o) any_host = any_host || dst[o].mem == STRATA_MEM_HOST;
    if (any_host && !can_cancel && nsrc == 1) {
        strata_source cs[1] = {};  // one input, as every job here has
        strata_sink cd[kMaxDst] = {};
        for (int i = 0; i < nsrc; ++i) {
            cs[i].width = W;
            cs[i].height = H;
            cs[i].masked = src[i].valid != nullptr;
            cs[i].resident = &src[i];
        }
        for (int o = 0; o < ndst; ++o) {
            cd[o].width = W;
            cd[o].height = H;
            cd[o].masked = dst[o].valid != nullptr;
            cd[o].resident = &dst[o];
        }
        strata_options co{};
        if (opts) co = *opts;
        co.tile_width = 0;
        co.tile_height = host_tile_rows(W);
        co.cancel = nullptr;
        co.async = 0;
        strata_stats ignored{};
        if ((rc = process_chunked(J, cd, ndst, cs, nsrc, &co, &ignored))) return rc;
        strata_stats s{};
        for (int64_t b = 0; b < p.bands; ++b) band_stats(s, J.r, J.rmin, W, H, nsrc, ndst, plan_band(p, b));
        add_stats(stats, s, nsrc, ndst);
        g_err.clear();
        return STRATA_OK;
    }

    // With a cancel flag, host operands are mirrored in device memory for
    // the call: a host output's mirror starts as the caller's raster, so the
    // bands the call never claims go back unchanged.
    Staged ssrc, sdst[kMaxDst];
    const strata_raster *dsrc = &src[0];
    Operands op{dsrc, 0, 0, {}, {}, {}, ndst, W, H};
    if (src[0].mem == STRATA_MEM_HOST) {
        if ((rc = stage_in(src[0], ssrc, true, st))) return rc;
        op.src = &ssrc.dev;
    }
    for (int o = 0; o < ndst; ++o) {
        op.dst[o] = &dst[o];
        if (dst[o].mem == STRATA_MEM_HOST) {
            any_host = true;
            if ((rc = stage_in(dst[o], sdst[o], can_cancel, st))) return rc;
            op.dst[o] = &sdst[o].dev;
        }
    }

    strata_stats s{};
    int status = STRATA_OK;
    bool any_unit = false;
    if (!can_cancel) {
        // Nothing to poll: the plan never changes a result, so the whole
        // raster is one launch and its Stats follow in closed form.
        if ((rc = run_rect(J, op, Rect{0, 0, W, H}, st))) return rc;
        for (int64_t b = 0; b < p.bands; ++b) band_stats(s, J.r, J.rmin, W, H, nsrc, ndst, plan_band(p, b));
        any_unit = true;
    } else {
        status = run_units(J, op, p, opts, st, nsrc, s, &any_unit);
        if (status < 0) return status;
    }
    if (any_unit)
        for (int o = 0; o < ndst; ++o)
            if (dst[o].mem == STRATA_MEM_HOST && (rc = stage_out(dst[o], sdst[o], st))) return rc;
    if (any_host || !(opts && opts->async) || can_cancel) HIP_TRY(hipStreamSynchronize(st), "synchronizing");
    add_stats(stats, s, nsrc, ndst);
    g_err.clear();
    return status;
}

// ---------------------------------------------------------------------------
// Reductions (internal/exec/reduce.go, reducechunked.go)
// ---------------------------------------------------------------------------

// Adds v * 2^k to the n-limb two's complement integer at limbs.
void add_shifted(uint64_t *limbs, int n, long long v, int k)
{
    const int w = k >> 6, sh = k & 63;
    const unsigned __int128 wide = (unsigned __int128)(__int128)v << sh;
    const uint64_t lo = (uint64_t)wide, hi = (uint64_t)(wide >> 64);
    const uint64_t ext = v < 0 ? ~0ull : 0ull;
    unsigned __int128 carry = 0;
    for (int j = w; j < n; ++j) {
        const uint64_t add = j == w ? lo : (j == w + 1 ? hi : ext);
        const unsigned __int128 t = (unsigned __int128)limbs[j] + add + carry;
        limbs[j] = (uint64_t)t;
        carry = t >> 64;
    }
}

// The device's fold of one call, read back: the exact integers carried
// out of the bins (internal/accum/result.go exact) and the tallies.
void fold_result(const FoldBins &b, const FoldPart &t, int what, strata_fold *out)
{
    memset(out, 0, sizeof *out);
    out->count = (int64_t)t.count;
    out->nan = (int64_t)t.nan;
    out->pos_inf = (int64_t)t.pinf;
    out->neg_inf = (int64_t)t.ninf;
    out->not_neg_zero = t.nnz != 0;
    out->min = t.mn;
    out->max = t.mx;
    if (what & (STRATA_REDUCE_SUM | STRATA_REDUCE_MOMENTS))
        for (int k = 0; k < kFoldSumBins; ++k)
            if (b.sum[k]) add_shifted(out->sum, 6, b.sum[k], k);
    if (what & STRATA_REDUCE_MOMENTS)
        for (int j = 0; j < kFoldSqBins; ++j)
            if (b.sq[j]) add_shifted(out->sumsq, 10, b.sq[j], j);
}

// Device scratch of one reduction: the bins, the total and the partials.
struct FoldScratch {
    char *dev = nullptr;
    hipStream_t st = nullptr;
    FoldBins *bins() const { return (FoldBins *)dev; }
    FoldPart *total() const { return (FoldPart *)(dev + sizeof(FoldBins)); }
    FoldPart *parts() const { return (FoldPart *)(dev + sizeof(FoldBins) + 256); }
    static size_t bytes() { return sizeof(FoldBins) + 256 + sizeof(FoldPart) * kFoldMaxBlocks; }
    ~FoldScratch()
    {
        if (dev) (void)hipFreeAsync(dev, st);
    }
};

int check_reduce(int what, const strata_options *opts)
{
    if (what == 0 || (what & ~(STRATA_REDUCE_COUNT | STRATA_REDUCE_MINMAX | STRATA_REDUCE_SUM |
                               STRATA_REDUCE_MOMENTS)))
        return fail(STRATA_EINVAL, "engine: unknown reduction %#x", what);
    if (opts && (opts->tile_width < 0 || opts->tile_height < 0 || opts->workers < 0))
        return fail(STRATA_EINVAL, "engine: negative Options {TileWidth:%lld TileHeight:%lld Workers:%lld}",
                    (long long)opts->tile_width, (long long)opts->tile_height, (long long)opts->workers);
    return STRATA_OK;
}

FoldArgs fold_args(const strata_raster &r, int what, bool masked)
{
    FoldArgs a{};
    a.p = r.data;
    a.stride = r.stride;
    a.w = r.width;
    a.h = r.height;
    a.bits = masked ? r.valid : nullptr;
    a.off = r.valid_offset;
    a.mstride = r.stride;
    a.what = what;
    return a;
}

// Reads the scratch back into out (waits for the stream).
int fold_read(const FoldScratch &fs, int what, strata_fold *out)
{
    FoldBins b;
    FoldPart t;
    HIP_TRY(hipMemcpyAsync(&b, fs.bins(), sizeof b, hipMemcpyDeviceToHost, fs.st), "reading the bins");
    HIP_TRY(hipMemcpyAsync(&t, fs.total(), sizeof t, hipMemcpyDeviceToHost, fs.st), "reading the fold");
    HIP_TRY(hipStreamSynchronize(fs.st), "synchronizing");
    if (out) fold_result(b, t, what, out);
    return STRATA_OK;
}

// A band's counters of a fold (reduce.go:242-257): no outputs, no halo.
void fold_band_stats(strata_stats &s, const Rect &b)
{
    const int64_t cells = (b.x1 - b.x0) * (b.y1 - b.y0);
    s.cells += cells;
    s.bands += 1;
    s.kernel_read += cells * kBytesPerCell;
}


int process_chunked(const Job &J, const strata_sink *dst, int ndst, const strata_source *src, int nsrc,
                    const strata_options *opts, strata_stats *stats)
{
    int rc = STRATA_OK;
    if (!dst || !src) return fail(STRATA_EINVAL, "engine: nil operand list");
    if (opts && (opts->tile_width < 0 || opts->tile_height < 0 || opts->workers < 0))
        return fail(STRATA_EINVAL, "engine: negative Options");
    const int64_t W = dst[0].width, H = dst[0].height;
    if (W <= 0 || H <= 0)
        return fail(STRATA_EINVAL, "engine: dst[0] is %lld×%lld; sizes must be positive", (long long)W, (long long)H);
    for (int i = 0; i < nsrc; ++i) {
        if (src[i].width != W || src[i].height != H)
            return fail(STRATA_EINVAL, "engine: src[%d] is %lld×%lld, dst[0] is %lld×%lld", i,
                        (long long)src[i].width, (long long)src[i].height, (long long)W, (long long)H);
        if (!src[i].resident && !src[i].raw && !src[i].read)
            return fail(STRATA_EINVAL, "engine: src[%d] is a nil source", i);
        if (!src[i].resident && src[i].raw && (rc = check_raw(src[i].raw, src[i].masked, "src", i))) return rc;
        if (src[i].resident) {
            if ((rc = check_raster("src", i, *src[i].resident))) return rc;
            if (src[i].resident->width != W || src[i].resident->height != H)
                return fail(STRATA_EINVAL, "engine: src[%d]'s raster does not match its size", i);
        }
    }
    bool masked = false;
    for (int i = 0; i < nsrc; ++i) masked = masked || src[i].masked;
    for (int i = 0; i < ndst; ++i) {
        if (dst[i].width != W || dst[i].height != H)
            return fail(STRATA_EINVAL, "engine: dst[%d] is %lld×%lld, dst[0] is %lld×%lld", i,
                        (long long)dst[i].width, (long long)dst[i].height, (long long)W, (long long)H);
        if (!dst[i].resident && !dst[i].raw && !dst[i].write && !dst[i].fold)
            return fail(STRATA_EINVAL, "engine: dst[%d] is a nil sink", i);
        if (dst[i].fold && (rc = check_reduce(dst[i].fold_what, nullptr))) return rc;
        if (!dst[i].resident && dst[i].raw && (rc = check_raw(dst[i].raw, dst[i].masked, "dst", i))) return rc;
        if (dst[i].resident && (rc = check_raster("dst", i, *dst[i].resident))) return rc;
        if (masked && !dst[i].masked)
            return fail(STRATA_EINVAL,
                        "engine: a source is Masked but dst[%d] is not; its validity would be lost", i);
    }
    for (int o = 0; o < ndst; ++o)
        if (dst[o].fold) memset(dst[o].fold, 0, sizeof *dst[o].fold);
    if ((rc = ensure_device())) return rc;
    hipStream_t st = opts ? (hipStream_t)opts->stream : nullptr;
    const int r = J.r, rmin = J.rmin;
    const TileGrid g = tile_grid(W, H, opts);
    const bool opts_cancel = opts && opts->cancel;

    // The outputs' folds (graph/sinks.go:17-71): one set of bins per
    // folded output, accumulated over every tile its sink is handed. A
    // fold-only sink writes nothing: its tiles live in device buffers.
    FoldScratch folds[kMaxDst];
    bool fold_first[kMaxDst];
    auto fold_only = [&](int o) { return dst[o].fold && !dst[o].resident && !dst[o].raw && !dst[o].write; };
    // Resident operands in device memory are read and written in place;
    // those in host memory go through the tile buffers, copied by the
    // library itself (as the callbacks of engine.MemorySource/MemorySink).
    auto dev_src = [&](int i) { return src[i].resident && src[i].resident->mem != STRATA_MEM_HOST; };
    auto dev_dst = [&](int o) { return dst[o].resident && dst[o].resident->mem != STRATA_MEM_HOST; };
    for (int o = 0; o < ndst; ++o) {
        fold_first[o] = true;
        folds[o].st = st;
        if (dst[o].fold && (rc = pool_alloc((void **)&folds[o].dev, FoldScratch::bytes(), st, "fold scratch")))
            return rc;
    }
    // Folds output o's cells of rect rc (raster coordinates), held in d
    // whose cell (0, 0) is raster cell (dx, dy): exec.Reduce over the
    // window the sink is handed, valid cells only when the sink is masked.
    auto fold_tile = [&](int o, const strata_raster &d, int64_t dx, int64_t dy, const Rect &rc) -> int {
        strata_raster v = d;
        const int64_t off = (rc.y0 - dy) * d.stride + (rc.x0 - dx);
        v.data = d.data + off;
        v.valid_offset = d.valid_offset + off;
        v.width = rc.x1 - rc.x0;
        v.height = rc.y1 - rc.y0;
        HIP_TRY(launch_fold(fold_args(v, dst[o].fold_what, dst[o].masked && d.valid != nullptr), folds[o].bins(),
                            folds[o].parts(), folds[o].total(), fold_first[o], st),
                "launching an output's fold");
        fold_first[o] = false;
        return STRATA_OK;
    };
    // The combined partials into the callers' structs (waits for the stream).
    auto fold_store = [&]() -> int {
        for (int o = 0; o < ndst; ++o) {
            if (!dst[o].fold) continue;
            if (fold_first[o]) {  // no tile: the empty fold
                dst[o].fold->min = INFINITY;
                dst[o].fold->max = -INFINITY;
                continue;
            }
            int e = fold_read(folds[o], dst[o].fold_what, dst[o].fold);
            if (e) return e;
        }
        return STRATA_OK;
    };

    // Buffers for callback operands: inputs hold a tile grown by the
    // radius, outputs a tile; a masked buffer's stride is a multiple of 64
    // so every row's bits start a word (chunked.go:184-200).
    const int64_t inW = std::min(W, g.tileW + 2 * r), inH = std::min(H, g.tileH + 2 * r);
    auto round64 = [](int64_t n) { return (n + 63) & ~(int64_t)63; };
    struct Buf {
        int64_t stride = 0, cells = 0, words = 0;
        float *hdata = nullptr;
        uint64_t *hmask = nullptr;
        float *ddata = nullptr;
        uint64_t *dmask = nullptr;
    };
    // Two sets of buffers: while the device computes one tile, the host
    // hands the previous tile's outputs to their sinks and reads the next
    // tile (the write-behind of internal/exec/behind.go:54-192).
    constexpr int kSlots = 2;
    const int nbuf = nsrc + ndst;
    std::vector<Buf> bufs(kSlots * nbuf);
    hipEvent_t done[kSlots] = {nullptr, nullptr};
    auto buf = [&](int slot, int i) -> Buf & { return bufs[slot * nbuf + i]; };
    auto free_bufs = [&]() {
        // Pinned buffers go back to the cache only once no copy uses them;
        // device buffers are freed in stream order. A call that took no
        // pinned buffer (resident operands) does not wait.
        bool pinned = false;
        for (auto &b : bufs) pinned = pinned || b.hdata || b.hmask;
        if (pinned) (void)hipStreamSynchronize(st);
        for (auto &b : bufs) {
            PinnedCache::get().release(b.hdata);
            PinnedCache::get().release(b.hmask);
            if (b.ddata) (void)hipFreeAsync(b.ddata, st);
            if (b.dmask) (void)hipFreeAsync(b.dmask, st);
        }
        for (auto &e : done)
            if (e) (void)hipEventDestroy(e);
    };
    struct Guard {
        std::function<void()> f;
        ~Guard() { f(); }
    };
    auto alloc_buf = [&](Buf &b, int64_t w, int64_t h, bool m, bool host = true) -> int {
        b.stride = m ? round64(w) : w;
        b.cells = (h - 1) * b.stride + w;
        b.words = m ? cdiv(b.cells, 64) : 0;
        int e = STRATA_OK;
        if (host && (e = PinnedCache::get().acquire((size_t)b.cells * 4, (void **)&b.hdata))) return e;
        if ((e = pool_alloc((void **)&b.ddata, (size_t)b.cells * 4, st, "tile buffer"))) return e;
        if (m) {
            if (host && (e = PinnedCache::get().acquire((size_t)b.words * 8, (void **)&b.hmask))) return e;
            if ((e = pool_alloc((void **)&b.dmask, (size_t)b.words * 8, st, "tile mask"))) return e;
        }
        return STRATA_OK;
    };
    int status = STRATA_OK;
    {
        Guard guard{free_bufs};
        strata_stats s{};
        bool all_resident = true, any_fold_only = false;
        for (int i = 0; i < nsrc; ++i) all_resident = all_resident && dev_src(i);
        for (int o = 0; o < ndst; ++o) {
            all_resident = all_resident && (dev_dst(o) || fold_only(o));
            any_fold_only = any_fold_only || fold_only(o);
        }
        if (all_resident && !opts_cancel) {
            // Device tile queue over resident operands with nothing to poll:
            // every tile reads its halo straight from HBM and writes in place,
            // so the queue is one launch over the raster (tiling never changes
            // a result) and the tiles' Stats follow in closed form. Outputs
            // that are only folded go through a device buffer of a few
            // hundred rows instead, one step of rows at a time, each step
            // folded while the cache still holds it.
            strata_raster sv = *src[0].resident, dv[kMaxDst]{};
            if (sv.mem == STRATA_MEM_HOST) return fail(STRATA_EINVAL, "engine: resident sources must be in device memory");
            if (!src[0].masked) sv.valid = nullptr;
            int64_t step = H;
            if (any_fold_only) {
                const int64_t bytes = (int64_t)get_tuning().fold_stripe_mb << 20;
                step = std::min(H, std::max<int64_t>(32, bytes / (4 * W) / 32 * 32));
            }
            // Stream-ordered buffers: the pool reuses them from call to call.
            struct AsyncBufs {
                hipStream_t st;
                std::vector<void *> p;
                ~AsyncBufs()
                {
                    for (void *q : p) scratch_free(q, st);
                }
            } abufs{st, {}};
            for (int o = 0; o < ndst; ++o) {
                if (fold_only(o)) {
                    const int64_t cells = step * W;
                    float *data = nullptr;
                    uint64_t *mask = nullptr;
                    if ((rc = scratch_alloc((void **)&data, (size_t)cells * 4, st))) return rc;
                    abufs.p.push_back(data);
                    if (dst[o].masked) {
                        if ((rc = scratch_alloc((void **)&mask, (size_t)cdiv(cells, 64) * 8, st))) return rc;
                        abufs.p.push_back(mask);
                    }
                    dv[o].data = data;
                    dv[o].width = W;
                    dv[o].height = step;
                    dv[o].stride = W;
                    dv[o].valid = mask;
                    dv[o].mem = STRATA_MEM_DEVICE;
                    continue;
                }
                dv[o] = *dst[o].resident;
                if (dv[o].mem == STRATA_MEM_HOST)
                    return fail(STRATA_EINVAL, "engine: resident sinks must be in device memory");
                if (!dst[o].masked) dv[o].valid = nullptr;
            }
            for (int64_t y0 = 0; y0 < H; y0 += step) {
                const Rect sr{0, y0, W, std::min(H, y0 + step)};
                Operands op{&sv, 0, 0, {}, {}, {}, ndst, W, H};
                for (int o = 0; o < ndst; ++o) {
                    op.dst[o] = &dv[o];
                    if (fold_only(o)) op.dy[o] = y0;
                }
                if ((rc = run_rect(J, op, sr, st))) return rc;
                for (int o = 0; o < ndst; ++o)
                    if (dst[o].fold && (rc = fold_tile(o, dv[o], op.dx[o], op.dy[o], sr))) return rc;
            }
            for (int64_t t = 0; t < g.tiles; ++t) tile_stats(s, r, rmin, g, nsrc, ndst, tile_rect(g, t));
            bool any_fold = false;
            for (int o = 0; o < ndst; ++o) any_fold = any_fold || dst[o].fold;
            if (any_fold) {
                if ((rc = fold_store())) return rc;
            } else if (!(opts && opts->async)) {
                HIP_TRY(hipStreamSynchronize(st), "synchronizing");
            }
            add_stats(stats, s, nsrc, ndst);
            g_err.clear();
            return STRATA_OK;
        }
        for (int slot = 0; slot < kSlots; ++slot) {
            for (int i = 0; i < nsrc; ++i)
                if (!dev_src(i) && (rc = alloc_buf(buf(slot, i), inW, inH, src[i].masked))) return rc;
            for (int o = 0; o < ndst; ++o)
                if (!dev_dst(o) &&
                    (rc = alloc_buf(buf(slot, nsrc + o), g.tileW, g.tileH, dst[o].masked, !fold_only(o))))
                    return rc;
            HIP_TRY(hipEventCreateWithFlags(&done[slot], hipEventDisableTiming), "tile event");
        }

        // The tile whose outputs wait to be handed to the sinks.
        struct Pending {
            bool live = false;
            int slot = 0;
            Rect tr{};
        } pend;
        auto finish = [&](Pending &p) -> int {
            if (!p.live) return STRATA_OK;
            p.live = false;
            HIP_TRY(hipEventSynchronize(done[p.slot]), "waiting for a tile");
            const int64_t tw = p.tr.x1 - p.tr.x0, th = p.tr.y1 - p.tr.y0;
            for (int o = 0; o < ndst; ++o) {
                if (dev_dst(o) || fold_only(o)) continue;
                Buf &b = buf(p.slot, nsrc + o);
                if (dst[o].resident) {
                    host_window(*dst[o].resident, true, b.hdata, dst[o].masked ? b.hmask : nullptr, tw, th, b.stride,
                                0, p.tr.x0, p.tr.y0);
                } else if (dst[o].raw) {
                    int e = raw_write(*dst[o].raw, W, b.hdata, tw, th, b.stride, p.tr.x0, p.tr.y0, o);
                    if (e) return e;
                } else if (dst[o].write(dst[o].user, b.hdata, tw, th, b.stride, b.hmask, 0, p.tr.x0, p.tr.y0) != 0) {
                    return fail(STRATA_EIO, "engine: writing dst[%d] at (%lld, %lld): sink error", o,
                                (long long)p.tr.x0, (long long)p.tr.y0);
                }
            }
            tile_stats(s, r, rmin, g, nsrc, ndst, p.tr);
            return STRATA_OK;
        };

        for (int64_t t = 0; t < g.tiles; ++t) {
            if (cancelled(opts)) {
                status = STRATA_CANCELLED;
                break;
            }
            const int slot = (int)(t % kSlots);
            const Rect tr = tile_rect(g, t);
            const int64_t rx0 = std::max<int64_t>(0, tr.x0 - r), ry0 = std::max<int64_t>(0, tr.y0 - r);
            const int64_t rx1 = std::min<int64_t>(W, tr.x1 + r), ry1 = std::min<int64_t>(H, tr.y1 + r);
            const int64_t tw = tr.x1 - tr.x0, th = tr.y1 - tr.y0;

            // The source window: resident (read in place) or read into the
            // slot's pinned buffer and copied to the device. The slot's last
            // tile was handed off by the previous iteration's finish().
            strata_raster sv{};
            int64_t sx = 0, sy = 0;
            if (dev_src(0)) {
                sv = *src[0].resident;
                if (!src[0].masked) sv.valid = nullptr;
            } else {
                Buf &b = buf(slot, 0);
                const int64_t ww = rx1 - rx0, hh = ry1 - ry0;
                const int64_t cells = (hh - 1) * b.stride + ww;
                if (src[0].raw) {
                    // A raw file: rows read straight into the pinned buffer,
                    // the validity derived from the fill on the device.
                    if ((status = raw_read(*src[0].raw, W, b.hdata, ww, hh, b.stride, rx0, ry0, "src[0]"))) break;
                    HIP_TRY(hipMemcpyAsync(b.ddata, b.hdata, (size_t)cells * 4, hipMemcpyHostToDevice, st),
                            "tile in");
                    if (b.dmask)
                        HIP_TRY(launch_raw_validity(RawTile{b.ddata, b.stride, ww, hh, b.dmask}, src[0].raw->fill, st),
                                "launching the fill validity kernel");
                } else if (src[0].resident) {
                    if (b.hmask) memset(b.hmask, 0, (size_t)b.words * 8);
                    host_window(*src[0].resident, false, b.hdata, b.hmask, ww, hh, b.stride, 0, rx0, ry0);
                    HIP_TRY(hipMemcpyAsync(b.ddata, b.hdata, (size_t)cells * 4, hipMemcpyHostToDevice, st),
                            "tile in");
                    if (b.hmask)
                        HIP_TRY(hipMemcpyAsync(b.dmask, b.hmask, (size_t)cdiv(cells, 64) * 8, hipMemcpyHostToDevice,
                                               st),
                                "tile mask in");
                } else {
                    if (b.hmask) memset(b.hmask, 0, (size_t)b.words * 8);
                    if (src[0].read(src[0].user, b.hdata, ww, hh, b.stride, b.hmask, 0, rx0, ry0) != 0) {
                        status = fail(STRATA_EIO, "engine: reading src[0] at (%lld, %lld): source error",
                                      (long long)rx0, (long long)ry0);
                        break;
                    }
                    HIP_TRY(hipMemcpyAsync(b.ddata, b.hdata, (size_t)cells * 4, hipMemcpyHostToDevice, st),
                            "tile in");
                    if (b.hmask)
                        HIP_TRY(hipMemcpyAsync(b.dmask, b.hmask, (size_t)cdiv(cells, 64) * 8, hipMemcpyHostToDevice,
                                               st),
                                "tile mask in");
                }
                sv.data = b.ddata;
                sv.width = ww;
                sv.height = hh;
                sv.stride = b.stride;
                sv.valid = b.hmask ? b.dmask : nullptr;
                sv.valid_offset = 0;
                sx = rx0;
                sy = ry0;
            }
            strata_raster dv[kMaxDst]{};
            Operands op{&sv, sx, sy, {}, {}, {}, ndst, W, H};
            for (int o = 0; o < ndst; ++o) op.dst[o] = &dv[o];
            for (int o = 0; o < ndst; ++o) {
                if (dev_dst(o)) {
                    dv[o] = *dst[o].resident;
                    if (!dst[o].masked) dv[o].valid = nullptr;
                } else {
                    Buf &b = buf(slot, nsrc + o);
                    dv[o].data = b.ddata;
                    dv[o].width = tw;
                    dv[o].height = th;
                    dv[o].stride = b.stride;
                    dv[o].valid = b.dmask;
                    dv[o].valid_offset = 0;
                    op.dx[o] = tr.x0;
                    op.dy[o] = tr.y0;
                }
            }
            if ((rc = run_rect(J, op, tr, st))) return rc;
            // Folds see the tile before its sink does (graph/sinks.go:32-38).
            for (int o = 0; o < ndst; ++o)
                if (dst[o].fold && (rc = fold_tile(o, dv[o], op.dx[o], op.dy[o], tr))) return rc;
            // Buffered outputs back to the host: a raw sink's fill first.
            for (int o = 0; o < ndst; ++o) {
                if (dev_dst(o) || fold_only(o)) continue;
                Buf &b = buf(slot, nsrc + o);
                const int64_t cells = (th - 1) * b.stride + tw;
                if (dst[o].raw && b.dmask)
                    HIP_TRY(launch_raw_fill(RawTile{b.ddata, b.stride, tw, th, b.dmask}, dst[o].raw->fill, st),
                            "launching the fill kernel");
                HIP_TRY(hipMemcpyAsync(b.hdata, b.ddata, (size_t)cells * 4, hipMemcpyDeviceToHost, st), "tile out");
                if (b.hmask && !dst[o].raw)
                    HIP_TRY(hipMemcpyAsync(b.hmask, b.dmask, (size_t)cdiv(cells, 64) * 8, hipMemcpyDeviceToHost, st),
                            "tile mask out");
            }
            HIP_TRY(hipEventRecord(done[slot], st), "tile event");
            // While the device runs this tile, the previous one goes out.
            if ((status = finish(pend))) break;
            pend.live = true;
            pend.slot = slot;
            pend.tr = tr;
            // A cancellable call observes its flag once per finished tile.
            if (opts_cancel && (status = finish(pend))) break;
        }
        // The claimed tile always completes (internal/exec/worker.go:112-181).
        const int last = finish(pend);
        if (status == STRATA_OK || status == STRATA_CANCELLED) status = last ? last : status;
        HIP_TRY(hipStreamSynchronize(st), "synchronizing");
        add_stats(stats, s, nsrc, ndst);
        if (status == STRATA_OK) {
            status = fold_store();
        }
        if (status != STRATA_OK)
            for (int o = 0; o < ndst; ++o)
                if (dst[o].fold) memset(dst[o].fold, 0, sizeof *dst[o].fold);
    }
    if (status == STRATA_OK) g_err.clear();
    return status;
}

} // namespace

extern "C" {

int strata_reduce(int what, const strata_raster *src, const strata_options *opts, strata_fold *out,
                  strata_stats *stats)
{
    int rc = check_reduce(what, opts);
    if (rc) return rc;
    if (!src) return fail(STRATA_EINVAL, "engine: reducer takes 1 inputs, got 0");
    if ((rc = check_raster("src", 0, *src))) return rc;
    if (out) memset(out, 0, sizeof *out);
    // runWorkers checks the context before every band (internal/exec/
    // worker.go:144-181); the device folds the whole raster in one launch,
    // so the check comes once, before it.
    if (cancelled(opts)) return STRATA_CANCELLED;
    if ((rc = ensure_device())) return rc;
    hipStream_t st = opts ? (hipStream_t)opts->stream : nullptr;
    Staged staged;
    const strata_raster *d = src;
    if (src->mem == STRATA_MEM_HOST) {
        if ((rc = stage_in(*src, staged, true, st))) return rc;
        d = &staged.dev;
    }
    FoldScratch fs;
    fs.st = st;
    if ((rc = pool_alloc((void **)&fs.dev, FoldScratch::bytes(), st, "reduction scratch"))) return rc;
    HIP_TRY(launch_fold(fold_args(*d, what, d->valid != nullptr), fs.bins(), fs.parts(), fs.total(), true, st),
            "launching the fold");
    if ((rc = fold_read(fs, what, out))) return rc;
    if (stats) {
        strata_stats s{};
        const Plan p = new_plan(src->width, src->height, opts ? opts->tile_width : 0, opts ? opts->tile_height : 0);
        for (int64_t b = 0; b < p.bands; ++b) fold_band_stats(s, plan_band(p, b));
        add_stats(stats, s, 1, 0);
    }
    g_err.clear();
    return STRATA_OK;
}

// strata_reduce without any allocation: the fold runs in the library's
// static device scratch (the bins, the running total and the workgroups'
// partials) instead of memory taken from the stream-ordered allocator.
// Device rasters only. Calls that overlap on one GPU share that scratch
// and race; the fold's first kernel clears the bins and the total, so
// nothing of an earlier call carries over.
int strata_reduce_noalloc(int what, const strata_raster *src, const strata_options *opts, strata_fold *out,
                          strata_stats *stats)
{
    static_assert(sizeof(FoldBins) + 256 + sizeof(FoldPart) * kFoldMaxBlocks <= kNoAllocScratchBytes,
                  "the fold's scratch outgrew the static array");
    int rc = check_reduce(what, opts);
    if (rc) return rc;
    if (!src) return fail(STRATA_EINVAL, "engine: reducer takes 1 inputs, got 0");
    if ((rc = check_raster("src", 0, *src))) return rc;
    if (src->mem != STRATA_MEM_DEVICE)
        return fail(STRATA_EINVAL,
                    "strata: strata_reduce_noalloc reads device rasters only (a host raster needs staging memory)");
    if (src->width > kNoAllocMaxCells / src->height)
        return fail(STRATA_EINVAL, "strata: a %lld×%lld raster exceeds strata_reduce_noalloc's limit of 2^40 cells",
                    (long long)src->width, (long long)src->height);
    if (out) memset(out, 0, sizeof *out);
    if (cancelled(opts)) return STRATA_CANCELLED;
    void *base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    hipStream_t st = opts ? (hipStream_t)opts->stream : nullptr;
    char *dev = (char *)base;
    FoldBins *bins = (FoldBins *)dev;
    FoldPart *total = (FoldPart *)(dev + sizeof(FoldBins));
    FoldPart *parts = (FoldPart *)(dev + sizeof(FoldBins) + 256);
    HIP_TRY(launch_fold(fold_args(*src, what, src->valid != nullptr), bins, parts, total, true, st),
            "launching the fold");
    FoldBins b;
    FoldPart t;
    HIP_TRY(hipMemcpyAsync(&b, bins, sizeof b, hipMemcpyDeviceToHost, st), "reading the bins");
    HIP_TRY(hipMemcpyAsync(&t, total, sizeof t, hipMemcpyDeviceToHost, st), "reading the fold");
    HIP_TRY(hipStreamSynchronize(st), "synchronizing");
    if (out) fold_result(b, t, what, out);
    if (stats) {
        strata_stats s{};
        const Plan p = new_plan(src->width, src->height, opts ? opts->tile_width : 0, opts ? opts->tile_height : 0);
        for (int64_t i = 0; i < p.bands; ++i) fold_band_stats(s, plan_band(p, i));
        add_stats(stats, s, 1, 0);
    }
    g_err.clear();
    return STRATA_OK;
}

int strata_reduce_chunked(int what, const strata_source *src, const strata_options *opts, strata_fold *out,
                          strata_stats *stats)
{
    int rc = check_reduce(what, opts);
    if (rc) return rc;
    if (!src || (!src->resident && !src->raw && !src->read))
        return fail(STRATA_EINVAL, "engine: src[0] is a nil source");
    if (!src->resident && src->raw && (rc = check_raw(src->raw, src->masked, "src", 0))) return rc;
    const int64_t W = src->width, H = src->height;
    if (W <= 0 || H <= 0)
        return fail(STRATA_EINVAL, "engine: src[0] is %lld×%lld; sizes must be positive", (long long)W,
                    (long long)H);
    if (src->resident) {
        if ((rc = check_raster("src", 0, *src->resident))) return rc;
        if (src->resident->width != W || src->resident->height != H)
            return fail(STRATA_EINVAL, "engine: src[0]'s raster does not match its size");
        if (src->resident->mem == STRATA_MEM_HOST)
            return fail(STRATA_EINVAL, "engine: resident sources must be in device memory");
    }
    if (out) memset(out, 0, sizeof *out);
    if ((rc = ensure_device())) return rc;
    hipStream_t st = opts ? (hipStream_t)opts->stream : nullptr;
    const TileGrid g = tile_grid(W, H, opts);
    strata_stats s{};
    auto tile_counts = [&](const Rect &t) {
        s.tiles += 1;
        s.source_read += (t.x1 - t.x0) * (t.y1 - t.y0) * kBytesPerCell;
    };
    auto tile_bands = [&](const Rect &t) {
        const Plan p = new_plan(t.x1 - t.x0, t.y1 - t.y0, 0, 0);
        for (int64_t b = 0; b < p.bands; ++b) {
            Rect band = plan_band(p, b);
            band.x0 += t.x0;
            band.x1 += t.x0;
            band.y0 += t.y0;
            band.y1 += t.y0;
            fold_band_stats(s, band);
        }
    };
    FoldScratch fs;
    fs.st = st;
    int status = STRATA_OK;
    if (src->resident) {
        // Read in place: one launch over the raster, the tiles' counters in
        // closed form.
        if (cancelled(opts)) return STRATA_CANCELLED;
        if ((rc = pool_alloc((void **)&fs.dev, FoldScratch::bytes(), st, "reduction scratch"))) return rc;
        HIP_TRY(launch_fold(fold_args(*src->resident, what, src->masked && src->resident->valid), fs.bins(),
                            fs.parts(), fs.total(), true, st),
                "launching the fold");
        if ((rc = fold_read(fs, what, out))) return rc;
        for (int64_t t = 0; t < g.tiles; ++t) {
            const Rect tr = tile_rect(g, t);
            tile_counts(tr);
            tile_bands(tr);
        }
    } else {
        // Tiles read by the callback into a pinned buffer, copied to the
        // device and folded into the same bins (reducechunked.go:186-205).
        const bool masked = src->masked != 0;
        const int64_t stride = masked ? ((g.tileW + 63) & ~(int64_t)63) : g.tileW;
        const int64_t cells = (g.tileH - 1) * stride + g.tileW;
        const int64_t words = masked ? cdiv(cells, 64) : 0;
        float *hdata = nullptr, *ddata = nullptr;
        uint64_t *hmask = nullptr, *dmask = nullptr;
        struct Bufs {
            float *&hd, *&dd;
            uint64_t *&hm, *&dm;
            hipStream_t st;
            ~Bufs()
            {
                (void)hipStreamSynchronize(st);
                PinnedCache::get().release(hd);
                PinnedCache::get().release(hm);
                if (dd) (void)hipFreeAsync(dd, st);
                if (dm) (void)hipFreeAsync(dm, st);
            }
        } bufs{hdata, ddata, hmask, dmask, st};
        if ((rc = PinnedCache::get().acquire((size_t)cells * 4, (void **)&hdata))) return rc;
        if ((rc = pool_alloc((void **)&ddata, (size_t)cells * 4, st, "tile buffer"))) return rc;
        if (masked) {
            if ((rc = PinnedCache::get().acquire((size_t)words * 8, (void **)&hmask))) return rc;
            if ((rc = pool_alloc((void **)&dmask, (size_t)words * 8, st, "tile mask"))) return rc;
        }
        if ((rc = pool_alloc((void **)&fs.dev, FoldScratch::bytes(), st, "reduction scratch"))) return rc;
        bool first = true;
        for (int64_t t = 0; t < g.tiles; ++t) {
            if (cancelled(opts)) {
                status = STRATA_CANCELLED;
                break;
            }
            const Rect tr = tile_rect(g, t);
            const int64_t tw = tr.x1 - tr.x0, th = tr.y1 - tr.y0;
            tile_counts(tr);
            const int64_t n = (th - 1) * stride + tw;
            if (src->raw) {
                if ((status = raw_read(*src->raw, W, hdata, tw, th, stride, tr.x0, tr.y0, "src[0]"))) break;
                HIP_TRY(hipMemcpyAsync(ddata, hdata, (size_t)n * 4, hipMemcpyHostToDevice, st), "tile in");
                if (dmask)
                    HIP_TRY(launch_raw_validity(RawTile{ddata, stride, tw, th, dmask}, src->raw->fill, st),
                            "launching the fill validity kernel");
            } else {
                if (hmask) memset(hmask, 0, (size_t)words * 8);
                if (src->read(src->user, hdata, tw, th, stride, hmask, 0, tr.x0, tr.y0) != 0) {
                    status = fail(STRATA_EIO, "engine: reading src[0] at (%lld, %lld): source error",
                                  (long long)tr.x0, (long long)tr.y0);
                    break;
                }
                HIP_TRY(hipMemcpyAsync(ddata, hdata, (size_t)n * 4, hipMemcpyHostToDevice, st), "tile in");
                if (hmask)
                    HIP_TRY(hipMemcpyAsync(dmask, hmask, (size_t)cdiv(n, 64) * 8, hipMemcpyHostToDevice, st),
                            "tile mask in");
            }
            strata_raster tv{};
            tv.data = ddata;
            tv.width = tw;
            tv.height = th;
            tv.stride = stride;
            tv.valid = dmask;
            tv.valid_offset = 0;
            HIP_TRY(launch_fold(fold_args(tv, what, masked), fs.bins(), fs.parts(), fs.total(), first, st),
                    "launching the fold");
            first = false;
            // The pinned buffer is reused by the next tile.
            HIP_TRY(hipStreamSynchronize(st), "waiting for a tile");
            tile_bands(tr);
        }
        if (status == STRATA_OK && (rc = fold_read(fs, what, out))) return rc;
    }
    add_stats(stats, s, 1, 0);
    if (status != STRATA_OK) {
        if (out) memset(out, 0, sizeof *out);
        return status;
    }
    g_err.clear();
    return STRATA_OK;
}

int strata_plan_stats(const strata_kernel *k, int ndst, int nsrc, int64_t width, int64_t height,
                      const strata_options *opts, int chunked, strata_stats *stats)
{
    int rc = check_kernel(k, ndst, nsrc);
    if (rc) return rc;
    return plan_stats(make_job(k), nsrc, width, height, opts, chunked, stats);
}

int strata_process(const strata_kernel *k, const strata_raster *dst, int ndst, const strata_raster *src, int nsrc,
                   const strata_options *opts, strata_stats *stats)
{
    int rc = check_kernel(k, ndst, nsrc);
    if (rc) return rc;
    return process(make_job(k), dst, ndst, src, nsrc, opts, stats);
}

int strata_process_chunked(const strata_kernel *k, const strata_sink *dst, int ndst, const strata_source *src,
                           int nsrc, const strata_options *opts, strata_stats *stats)
{
    int rc = check_kernel(k, ndst, nsrc);
    if (rc) return rc;
    return process_chunked(make_job(k), dst, ndst, src, nsrc, opts, stats);
}

int strata_plan_stats_fanout(const strata_kernel *parts, int nparts, int64_t width, int64_t height,
                             const strata_options *opts, int chunked, strata_stats *stats)
{
    int rc = check_parts(parts, nparts);
    if (rc) return rc;
    return plan_stats(make_fan_job(parts, nparts), 1, width, height, opts, chunked, stats);
}

int strata_process_fanout(const strata_kernel *parts, int nparts, const strata_raster *dst,
                          const strata_raster *src, const strata_options *opts, strata_stats *stats)
{
    int rc = check_parts(parts, nparts);
    if (rc) return rc;
    return process(make_fan_job(parts, nparts), dst, nparts, src, 1, opts, stats);
}

int strata_process_fanout_chunked(const strata_kernel *parts, int nparts, const strata_sink *dst,
                                  const strata_source *src, const strata_options *opts, strata_stats *stats)
{
    int rc = check_parts(parts, nparts);
    if (rc) return rc;
    return process_chunked(make_fan_job(parts, nparts), dst, nparts, src, 1, opts, stats);
}

int strata_plan_stats_chain(const strata_kernel *stages, int nstages, int64_t width, int64_t height,
                            const strata_options *opts, int chunked, strata_stats *stats)
{
    int rc = check_chain(stages, nstages);
    if (rc) return rc;
    return plan_stats(make_chain_job(stages, nstages), 1, width, height, opts, chunked, stats);
}

int strata_process_chain(const strata_kernel *stages, int nstages, const strata_raster *dst, const strata_raster *src,
                         const strata_options *opts, strata_stats *stats)
{
    int rc = check_chain(stages, nstages);
    if (rc) return rc;
    if ((rc = ensure_device())) return rc;
    Job J = make_chain_job(stages, nstages);
    ChainTables T;
    if ((rc = T.stage(J, opts ? (hipStream_t)opts->stream : nullptr))) return rc;
    return process(J, dst, 1, src, 1, opts, stats);
}

int strata_process_chain_chunked(const strata_kernel *stages, int nstages, const strata_sink *dst,
                                 const strata_source *src, const strata_options *opts, strata_stats *stats)
{
    int rc = check_chain(stages, nstages);
    if (rc) return rc;
    if ((rc = ensure_device())) return rc;
    Job J = make_chain_job(stages, nstages);
    ChainTables T;
    if ((rc = T.stage(J, opts ? (hipStream_t)opts->stream : nullptr))) return rc;
    return process_chunked(J, dst, 1, src, 1, opts, stats);
}

} // extern "C"

namespace strata {

hipError_t scratch_alloc_async(void **p, size_t bytes, hipStream_t st)
{
    hipMemPool_t pool = nullptr;
    if (scratch_pool(&pool) != STRATA_OK) return hipErrorOutOfMemory;
    return hipMallocFromPoolAsync(p, bytes, pool, st);
}

} // namespace strata
