// api.cpp -- the C ABI of include/vlb_bake.h over the HIP kernels.
//
// Host-side replacement of the reference's Vulkan plumbing: Application
// (src/application.cpp:540-556) becomes one HIP device + one stream;
// LightBaker's per-probe submit/fence loop (src/baker/light_baker.cpp:287-328)
// becomes one asynchronous plan execution (bake kernel + FP64 merge).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "vlb_internal.hpp"

using namespace vlb;

namespace vlb {

DevBuf dev_alloc(Context& ctx, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    const uint64_t now = ctx.in_use.load();
    if (ctx.budget && now + bytes > ctx.budget)
        throw Error{VLB_ERR_BUDGET, "request of " + std::to_string(bytes) +
                                        " device bytes exceeds the HBM budget (" +
                                        std::to_string(now) + " of " + std::to_string(ctx.budget) +
                                        " in use)"};
    DevBuf b;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error{VLB_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e)};
    }
    b.bytes = bytes;
    b.acct = &ctx.in_use;
    ctx.in_use.fetch_add(bytes);
    return b;
}

Plan::~Plan()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
}

// Reads the event triples of the executes recorded since the last drain.
static void drain_timing(Plan& pl)
{
    for (uint32_t i = 0; i < pl.timed_pending; ++i) {
        hipEvent_t* e = &pl.ev[Plan::kEv * (size_t)i];
        float a = 0.f, b = 0.f, c = 0.f, d = 0.f;
        VLB_HIP(hipEventElapsedTime(&d, e[0], e[3]));
        if (!pl.use_graph) {
            VLB_HIP(hipEventElapsedTime(&a, e[0], e[1]));
            VLB_HIP(hipEventElapsedTime(&b, e[1], e[2]));
            VLB_HIP(hipEventElapsedTime(&c, e[2], e[3]));
        }
        pl.sum_trace += a;
        pl.sum_project += b;
        pl.sum_merge += c;
        pl.sum_total += d;
    }
    pl.timed_done += pl.timed_pending;
    pl.timed_pending = 0;
    if (pl.timed_done) {
        pl.stats.trace_ms = pl.sum_trace / pl.timed_done;
        pl.stats.project_ms = pl.sum_project / pl.timed_done;
        pl.stats.merge_ms = pl.sum_merge / pl.timed_done;
        pl.stats.total_ms = pl.sum_total / pl.timed_done;
    }
    pl.stats.timed_runs = pl.timed_done;
}

}  // namespace vlb

namespace {

// trace_variant: 1..12 fused k_trace (1 + map + 4 * occupancy index),
// 13..24 wavefront k_rays/k_shade (13 + map + 4 * refill index), 25..36 as
// 1..12 with far-first shadow queries
constexpr uint32_t kFusedDefault = 12;  // fused, lane map 3, 8 waves/SIMD

// trace_variant of a plan that leaves it to the library: the fused kernel,
// best or within ~10% of the best on rooms, terrains and a city block of
// 11k-2M triangles at 256-65,536 probes, and with an indirect term too (its
// eight corner visibility rays per hit: 411k probes/s against 388k for the
// wavefront kernels, variant 20, on the 2048-probe room).  Kept per term so
// that the two can differ again.
uint32_t default_variant(bool /*indirect*/)
{
    return kFusedDefault;
}
constexpr uint32_t kMaxVariant = 36;
constexpr int kOccTab[3] = {0, 6, 8};
constexpr int kRefillTab[3] = {16, 32, 8};

struct Variant {
    bool wave, far;
    int map, occ, refill;
};
Variant decode_variant(uint32_t v)
{
    Variant d{};
    d.far = v >= 25;
    if (d.far) v -= 24;
    if (v >= 13) {
        d.wave = true;
        d.map = (int)((v - 13) % 4);
        d.refill = kRefillTab[((v - 13) / 4) % 3];
    } else {
        d.wave = false;
        d.map = (int)((v - 1) % 4);
        d.occ = kOccTab[((v - 1) / 4) % 3];
        d.refill = 16;
    }
    return d;
}

// Allocates the wavefront buffers of a plan on first use.
void ensure_wave_buffers(Context& ctx, Plan& pl)
{
    if (pl.hits.p) return;
    const size_t nslots = (size_t)pl.args.ntrace * 512;
    pl.hits = dev_alloc(ctx, nslots * sizeof(int4));
    pl.vis = dev_alloc(ctx, std::max<size_t>(16, nslots * (size_t)pl.args.nvis));
    pl.args.hits = pl.hits.as<int4>();
    pl.args.vis = pl.vis.as<uint8_t>();
}

// Applies a trace_variant to a plan (launch shapes and lane map).
void apply_variant(Context& ctx, Plan& pl, uint32_t v)
{
    const Variant d = decode_variant(v);
    pl.variant = v;
    pl.wave = d.wave;
    pl.far = d.far;
    pl.occ = d.occ;
    pl.args.map = d.map;
    pl.args.refill = d.refill;
    if (d.wave) {
        ensure_wave_buffers(ctx, pl);
        pl.args.spill_stride = (uint32_t)pl.wgrid * 256u;
    } else {
        pl.grid = trace_grid_size(ctx, d.occ, d.far);
        pl.args.spill_stride = (uint32_t)pl.grid * 256u;
    }
    pl.stats.grid = (uint32_t)(d.wave ? pl.wgrid : pl.grid);
    if (pl.graph_exec) {  // launch shape changed: recapture on the next execute
        (void)hipGraphExecDestroy(pl.graph_exec);
        (void)hipGraphDestroy(pl.graph);
        pl.graph_exec = nullptr;
        pl.graph = nullptr;
    }
}

// The trace + shade stage of the plan's variant.
void launch_trace_stage(Plan& pl, bool record, bool count, hipStream_t st)
{
    if (pl.wave) launch_wave(pl.args, record, count, pl.wgrid, st);
    else launch_trace(pl.args, record, count, record || count ? 0 : pl.occ, pl.far, pl.grid, st);
}



void build_plan(Context& ctx, Scene& sc, const vlb_bake_params& p, Plan& pl)
{
    validate(p);
    pl.ctx = &ctx;
    pl.scene = &sc;
    pl.params = p;
    pl.lights.assign(p.lights, p.lights + p.num_lights);
    pl.params.lights = pl.lights.data();
    pl.params.probe_xyz = nullptr;
    pl.K = p.sh_order == 2 ? 9 : 16;
    pl.precision = p.precision;
    const int K = pl.K;
    const uint32_t P = p.num_probes;

    // ---- tuning: trace_variant 0 = the autotuner's earlier choice for this
    // scene and ray count, else the default
    const uint32_t R = rays_of(p);
    uint32_t tv = p.trace_variant;
    if (tv == 0) {
        auto it = ctx.tuned.find({sc.id, R});
        pl.default_variant = it == ctx.tuned.end();
        tv = pl.default_variant ? default_variant(false) : it->second;
    }
    VLB_REQUIRE(tv >= 1 && tv <= kMaxVariant, VLB_ERR_INVALID, "trace_variant out of range");
    pl.grid = trace_grid_size(ctx, 0, false);
    pl.wgrid = wave_grid_size(ctx);
    // 8 waves/SIMD is the most any trace kernel runs at
    const int max_grid = std::max(std::max(pl.grid, pl.wgrid), trace_grid_size(ctx, 8, false));
    pl.pgrid = ctx.num_cus * 8;
    const int npg = (int)((P + 15) / 16);

    // ---- direction + basis tables (sorted directions), probes (Morton order)
    build_ray_table(ctx, p, K, slice_rays(p, R, npg, pl.pgrid), pl.rt);
    build_probe_set(ctx, p.probe_xyz, P, pl.ps);
    const int Ppad = pl.ps.Ppad;
    const int nslices = pl.rt.nslices;
    const int Rpad = (int)pl.rt.Rpad;
    const double wref = pl.rt.wref;
    VLB_REQUIRE((uint64_t)npg * nslices < (1ull << 31), VLB_ERR_INVALID, "too many tiles");
    VLB_REQUIRE((uint64_t)npg * (Rpad / 32) * 512 < (1ull << 40), VLB_ERR_INVALID, "bake too large");

    // ---- device buffers
    pl.rad = dev_alloc(ctx, (size_t)npg * (Rpad / 32) * 1536 * sizeof(float));
    pl.partials = dev_alloc(ctx, (size_t)nslices * Ppad * 48 * sizeof(float2));
    // spill area sized for the largest candidate grid of the autotuner
    const uint32_t spill_stride = (uint32_t)pl.grid * 256u;
    pl.spill = dev_alloc(ctx, scene_spill_bytes(sc, (uint32_t)max_grid * 256u));
    pl.out = dev_alloc(ctx, (size_t)P * K * 3 * sizeof(float));
    pl.out64 = dev_alloc(ctx, (size_t)P * K * 3 * sizeof(double));
    pl.counters = dev_alloc(ctx, 4 * sizeof(unsigned long long));
    // work counters of every variant: 8 per-XCD counters 128 B apart (fused),
    // or 16 packed ones (wavefront)
    pl.wctr = dev_alloc(ctx, 32 * kCtrStride * sizeof(unsigned long long));

    // ---- kernel arguments
    BakeArgs& a = pl.args;
    memset(&a, 0, sizeof a);
    a.sc = sc.view();
    a.sp = shade_params(p);
    a.probes = pl.ps.probes.as<float4>();
    a.P = (int)P;
    a.Ppad = Ppad;
    a.dirs = pl.rt.dirs.as<float4>();
    a.ytab = pl.rt.ytab.p;
    a.R = (int)R;
    a.Rpad = Rpad;
    a.RS = (int)pl.rt.RS;
    a.nslices = nslices;
    a.npg = npg;
    a.ntiles = npg * nslices;
    a.nchunk = Rpad / 32;
    a.ntrace = npg * a.nchunk;
    a.rad = pl.rad.as<float>();
    a.nvis = p.shade_mode == VLB_SHADE_REFERENCE ? 1 : (int)p.num_lights;
    a.tmin = p.t_min;
    a.tmax = p.t_max;
    a.partials = pl.partials.as<float2>();
    a.spill = pl.spill.as<uint32_t>();
    a.spill_stride = spill_stride;
    a.probe_perm = pl.ps.probe_perm.as<int>();
    a.dir_perm = pl.rt.dir_perm.as<int>();
    a.counters = pl.counters.as<unsigned long long>();
    a.wctr = pl.wctr.as<unsigned long long>();

    MergeArgs& m = pl.margs;
    m.partials = pl.partials.as<float2>();
    m.nslices = nslices;
    m.P = (int)P;
    m.Ppad = Ppad;
    m.K = K;
    m.probe_perm = pl.ps.probe_perm.as<int>();
    // FP16 forms carry radiance / 16 (range headroom); undo it here.
    const bool fp16 = p.precision == VLB_PREC_FP16 || p.precision == VLB_PREC_FP16_FAST;
    m.scale = fp16 ? wref * 16.0 : wref;
    m.out = pl.out.as<float>();
    m.out64 = pl.out64.as<double>();

    pl.stats.primary_rays = (uint64_t)P * R;
    pl.stats.tiles = (uint32_t)a.ntiles;
    apply_variant(ctx, pl, tv);
    pl.stats.rays_per_slice = pl.rt.RS;
    pl.stats.slices = (uint32_t)nslices;
    VLB_HIP(hipStreamSynchronize(ctx.stream));
}

void plan_sync(Plan& pl);

// Multi-bounce: the result just merged becomes the next execute's source.
void refresh_indirect(Plan& pl, hipStream_t st)
{
    if (pl.indirect)
        launch_copy_f32(pl.ind.as<float>(), pl.out.as<float>(), (size_t)pl.args.P * pl.K * 3, st);
}

// Drops a captured graph (kernel arguments changed).
void drop_graph(Plan& pl)
{
    if (pl.graph_exec) {
        (void)hipGraphExecDestroy(pl.graph_exec);
        (void)hipGraphDestroy(pl.graph);
        pl.graph_exec = nullptr;
        pl.graph = nullptr;
    }
}

void plan_launch(Plan& pl)
{
    hipStream_t st = pl.ctx->stream;
    if (pl.timing && pl.timed_pending == (uint32_t)Plan::kTimingRing) plan_sync(pl);
    hipEvent_t* ev = pl.timing ? &pl.ev[Plan::kEv * (size_t)pl.timed_pending] : nullptr;
    if (pl.use_graph) {
        if (!pl.graph_exec) {
            VLB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                launch_trace_stage(pl, false, false, st);
                launch_project(pl.args, pl.precision, pl.pgrid, st);
                launch_merge(pl.margs, st);
                refresh_indirect(pl, st);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            VLB_HIP(hipStreamEndCapture(st, &pl.graph));
            VLB_HIP(hipGraphInstantiate(&pl.graph_exec, pl.graph, nullptr, nullptr, 0));
        }
        if (ev) VLB_HIP(hipEventRecord(ev[0], st));
        VLB_HIP(hipGraphLaunch(pl.graph_exec, st));
        if (ev) VLB_HIP(hipEventRecord(ev[3], st));
    } else {
        if (ev) VLB_HIP(hipEventRecord(ev[0], st));
        launch_trace_stage(pl, false, false, st);
        if (ev) VLB_HIP(hipEventRecord(ev[1], st));
        launch_project(pl.args, pl.precision, pl.pgrid, st);
        if (ev) VLB_HIP(hipEventRecord(ev[2], st));
        launch_merge(pl.margs, st);
        refresh_indirect(pl, st);
        if (ev) VLB_HIP(hipEventRecord(ev[3], st));
    }
    if (ev) ++pl.timed_pending;
    pl.ran = true;
}

void plan_sync(Plan& pl)
{
    VLB_HIP(hipStreamSynchronize(pl.ctx->stream));
    if (pl.timed_pending) drain_timing(pl);
}

std::vector<float> solid_angle_table(int N)
{
    auto A = [](double x, double y) { return atan2(x * y, sqrt(x * x + y * y + 1.0)); };
    std::vector<float> w((size_t)N * N);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            double x0 = 2.0 * i / N - 1.0, x1 = 2.0 * (i + 1) / N - 1.0;
            double y0 = 2.0 * j / N - 1.0, y1 = 2.0 * (j + 1) / N - 1.0;
            w[(size_t)j * N + i] = (float)(A(x0, y0) - A(x0, y1) - A(x1, y0) + A(x1, y1));
        }
    return w;
}

}  // namespace

namespace vlb {
// Validates a probe grid and lookup parameters; returns the probe count.
uint64_t check_grid(const vlb_probe_grid_desc& g, const vlb_lookup_params& lp)
{
    VLB_REQUIRE(g.num_coefs == 9 || g.num_coefs == 16, VLB_ERR_INVALID, "num_coefs must be 9 or 16");
    VLB_REQUIRE(g.order == VLB_ORDER_LINEAR || g.order == VLB_ORDER_REFERENCE_BFS, VLB_ERR_INVALID,
                "bad probe order");
    VLB_REQUIRE(lp.mode == VLB_LOOKUP_REFERENCE || lp.mode == VLB_LOOKUP_IRRADIANCE, VLB_ERR_INVALID,
                "bad lookup mode");
    VLB_REQUIRE(std::isfinite(lp.bias), VLB_ERR_INVALID, "lookup bias must be finite");
    uint64_t P = 1;
    for (int a = 0; a < 3; ++a) {
        VLB_REQUIRE(g.count[a] >= 1, VLB_ERR_INVALID, "grid counts must be >= 1");
        VLB_REQUIRE(g.step[a] != 0.0f && std::isfinite(g.step[a]) && std::isfinite(g.origin[a]),
                    VLB_ERR_INVALID, "grid origin and step must be finite, step non-zero");
        P *= g.count[a];
    }
    VLB_REQUIRE(P < (1ull << 28), VLB_ERR_INVALID, "grid too large");
    return P;
}

ProbeGridView grid_view(const vlb_probe_grid_desc& g, const vlb_lookup_params& lp,
                        const float* dcoeffs, float scale)
{
    ProbeGridView v;
    memset(&v, 0, sizeof v);
    v.coeffs = dcoeffs;
    for (int ax = 0; ax < 3; ++ax) {
        v.origin[ax] = g.origin[ax];
        v.step[ax] = g.step[ax];
        v.count[ax] = g.count[ax];
    }
    v.order = g.order;
    v.K = (int)g.num_coefs;
    v.exact = g.sh_constants == VLB_SH_EXACT;
    v.mode = lp.mode;
    v.vis = lp.visibility ? 1 : 0;
    v.bias = lp.bias;
    v.wmax = sqrtf(g.step[0] * g.step[0] + g.step[1] * g.step[1] + g.step[2] * g.step[2]);
    v.scale = scale;
    return v;
}
}  // namespace vlb

extern "C" {

int32_t vlb_abi_version(void) { return VLB_ABI_VERSION; }

vlb_status vlb_ctx_create(int32_t device, uint64_t hbm_budget_bytes, vlb_ctx** out)
{
    if (!out) return VLB_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return VLB_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) return VLB_ERR_INVALID;
    Context* c = new Context();
    vlb_status s = guard(c, [&] {
        c->device = device;
        VLB_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        VLB_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount;
        c->name = prop.gcnArchName;
        size_t freeb = 0, total = 0;
        VLB_HIP(hipMemGetInfo(&freeb, &total));
        c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)freeb;
        VLB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto& e : c->ev) VLB_HIP(hipEventCreate(&e));
        c->sky_scratch = skybox_scratch_init(device);
    });
    if (s != VLB_OK) {
        delete c;
        return s;
    }
    *out = reinterpret_cast<vlb_ctx*>(c);
    return VLB_OK;
}

vlb_status vlb_ctx_destroy(vlb_ctx* h)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    {
        std::lock_guard<std::recursive_mutex> lk(c->mtx);
        if (c->children > 0) {
            set_error(c, "context still has " + std::to_string(c->children) +
                             " live plans, batches or servers: destroy them first");
            return VLB_ERR_INVALID;
        }
    }
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->scenes.clear();
    c->user_bufs.clear();
    c->wtab.release();
    c->sky_tab.release();
    c->sky_partial.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VLB_OK;
}

vlb_status vlb_debug_inject(vlb_ctx* h, int32_t site, int32_t count)
{
    if (!h || site < 0 || site >= VLB_INJECT_SITES || count < 0) return VLB_ERR_INVALID;
    reinterpret_cast<Context*>(h)->inject[site].store(count);
    return VLB_OK;
}

const char* vlb_last_error(const vlb_ctx* h)
{
    if (!h) return "null context";
    // a copy per calling thread: another thread's failure may replace the
    // context's message at any time
    thread_local std::string msg;
    msg = get_error(reinterpret_cast<const Context*>(h));
    return msg.c_str();
}

vlb_status vlb_device_info(vlb_ctx* h, int32_t* num_cus, char* name, size_t name_len)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    if (num_cus) *num_cus = c->num_cus;
    if (name && name_len) {
        strncpy(name, c->name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    return VLB_OK;
}

// ------------------------------------------------------------------ scenes
vlb_status vlb_scene_upload(vlb_ctx* h, const float* xyz, uint32_t nverts, const uint32_t* idx,
                            uint32_t ntris, const float* normals, const float* albedo,
                            uint32_t* out_scene)
{
    if (!h || !out_scene) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(ntris == 0 || (xyz && idx && nverts > 0), VLB_ERR_INVALID, "missing geometry");
        VLB_REQUIRE(ntris < (1u << 30), VLB_ERR_INVALID, "too many triangles");
        for (uint64_t i = 0; i < (uint64_t)ntris * 3; ++i)
            VLB_REQUIRE(idx[i] < nverts, VLB_ERR_INVALID, "index out of range at " + std::to_string(i));
        // the BVH4 node grids need finite boxes of bounded magnitude
        for (uint64_t i = 0; i < (uint64_t)nverts * 3 && ntris > 0; ++i)
            VLB_REQUIRE(std::isfinite(xyz[i]) && std::fabs(xyz[i]) <= 0x1p100f, VLB_ERR_INVALID,
                        "vertex coordinate not finite or beyond 2^100 at " + std::to_string(i));
        VLB_HIP(hipSetDevice(c->device));
        auto s = std::make_unique<Scene>();
        build_scene_bvh(*c, *s, xyz, nverts, idx, ntris, normals, albedo);
        VLB_REQUIRE(s->depth <= 4096, VLB_ERR_INVALID, "BVH build produced a malformed tree");
        s->id = c->next_scene++;
        *out_scene = s->id;
        c->scenes[s->id] = std::move(s);
    });
}

vlb_status vlb_scene_release(vlb_ctx* h, uint32_t scene)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        Scene& s = find_scene(*c, scene);
        const int users = s.users.load();
        VLB_REQUIRE(users == 0, VLB_ERR_INVALID,
                    "scene " + std::to_string(scene) + " is in use by " + std::to_string(users) +
                        " plans, batch jobs or unfinished server requests: destroy or finish them first");
        VLB_HIP(hipStreamSynchronize(c->stream));
        c->scenes.erase(scene);
    });
}

vlb_status vlb_scene_info_get(vlb_ctx* h, uint32_t scene, vlb_scene_info* out)
{
    if (!h || !out) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        Scene& s = find_scene(*c, scene);
        memset(out, 0, sizeof *out);
        out->num_tris = (uint32_t)s.ntris;
        out->num_nodes = (uint32_t)s.nnodes;
        out->depth = s.depth;
        out->device_bytes = s.bytes();
        out->build_ms = s.build_ms;
        for (int a = 0; a < 3; ++a) {
            out->bounds_min[a] = s.bmin[a];
            out->bounds_max[a] = s.bmax[a];
        }
    });
}

vlb_status vlb_scene_export_bvh(vlb_ctx* h, uint32_t scene, float* nodes, size_t nodes_len,
                                int32_t* leaf_prims, size_t leaf_len)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        Scene& s = find_scene(*c, scene);
        VLB_HIP(hipSetDevice(c->device));
        if (nodes) {
            VLB_REQUIRE(nodes_len >= (size_t)s.nnodes * 16, VLB_ERR_INVALID, "nodes buffer too small");
            VLB_HIP(hipMemcpyAsync(nodes, s.nodes.p, (size_t)s.nnodes * sizeof(BNode),
                                   hipMemcpyDeviceToHost, c->stream));
        }
        if (leaf_prims) {
            VLB_REQUIRE(leaf_len >= (size_t)s.ntris, VLB_ERR_INVALID, "leaf buffer too small");
            std::vector<float4> t((size_t)s.ntris * 3);
            VLB_HIP(hipMemcpyAsync(t.data(), s.tris.p, t.size() * sizeof(float4), hipMemcpyDeviceToHost,
                                   c->stream));
            VLB_HIP(hipStreamSynchronize(c->stream));
            for (int i = 0; i < s.ntris; ++i) {
                int32_t v;
                memcpy(&v, &t[3 * (size_t)i].w, 4);
                leaf_prims[i] = v;
            }
        }
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

// ------------------------------------------------------------- ray queries
vlb_status vlb_trace_closest(vlb_ctx* h, uint32_t scene, const float* org, const float* dir,
                             uint32_t n, float t_min, float t_max, int32_t* prim, float* t,
                             float* u, float* v)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(org && dir && prim && t && u && v, VLB_ERR_INVALID, "null buffer");
        Scene& s = find_scene(*c, scene);
        VLB_HIP(hipSetDevice(c->device));
        if (n == 0) return;
        const int grid = std::max(8, c->num_cus * 4);
        const uint32_t stride = (uint32_t)grid * 256u;
        DevBuf d_o = dev_alloc(*c, (size_t)n * 12), d_d = dev_alloc(*c, (size_t)n * 12);
        DevBuf d_p = dev_alloc(*c, (size_t)n * 4), d_t = dev_alloc(*c, (size_t)n * 4);
        DevBuf d_u = dev_alloc(*c, (size_t)n * 4), d_v = dev_alloc(*c, (size_t)n * 4);
        DevBuf spill = dev_alloc(*c, scene_spill_bytes(s, stride));
        VLB_HIP(hipMemcpyAsync(d_o.p, org, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VLB_HIP(hipMemcpyAsync(d_d.p, dir, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        launch_trace_closest(s.view(), d_o.as<float>(), d_d.as<float>(), (int)n, t_min, t_max,
                             d_p.as<int>(), d_t.as<float>(), d_u.as<float>(), d_v.as<float>(),
                             spill.as<uint32_t>(), stride, grid, c->stream);
        VLB_HIP(hipMemcpyAsync(prim, d_p.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipMemcpyAsync(t, d_t.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipMemcpyAsync(u, d_u.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipMemcpyAsync(v, d_v.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

vlb_status vlb_trace_occluded(vlb_ctx* h, uint32_t scene, const float* org, const float* dir,
                              const float* t_max, uint32_t n, float t_min, uint8_t* occluded)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(org && dir && t_max && occluded, VLB_ERR_INVALID, "null buffer");
        Scene& s = find_scene(*c, scene);
        VLB_HIP(hipSetDevice(c->device));
        if (n == 0) return;
        const int grid = std::max(8, c->num_cus * 4);
        const uint32_t stride = (uint32_t)grid * 256u;
        DevBuf d_o = dev_alloc(*c, (size_t)n * 12), d_d = dev_alloc(*c, (size_t)n * 12);
        DevBuf d_m = dev_alloc(*c, (size_t)n * 4), d_r = dev_alloc(*c, (size_t)n);
        DevBuf spill = dev_alloc(*c, scene_spill_bytes(s, stride));
        VLB_HIP(hipMemcpyAsync(d_o.p, org, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VLB_HIP(hipMemcpyAsync(d_d.p, dir, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VLB_HIP(hipMemcpyAsync(d_m.p, t_max, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        launch_trace_occluded(s.view(), d_o.as<float>(), d_d.as<float>(), d_m.as<float>(), (int)n,
                              t_min, d_r.as<uint8_t>(), spill.as<uint32_t>(), stride, grid, c->stream);
        VLB_HIP(hipMemcpyAsync(occluded, d_r.p, (size_t)n, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

// ------------------------------------------------------------------ plans
vlb_status vlb_plan_create(vlb_ctx* h, uint32_t scene, const vlb_bake_params* p, vlb_plan** out)
{
    if (!h || !p || !out) return VLB_ERR_INVALID;
    *out = nullptr;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        Scene& s = find_scene(*c, scene);
        VLB_HIP(hipSetDevice(c->device));
        auto pl = std::make_unique<Plan>();
        build_plan(*c, s, *p, *pl);
        *out = reinterpret_cast<vlb_plan*>(pl.release());
        ++s.users;
        ++c->children;
    });
}

vlb_status vlb_plan_execute(vlb_plan* h)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] {
        VLB_REQUIRE(!take_injected(*pl->ctx, VLB_INJECT_PLAN_EXECUTE), VLB_ERR_INJECTED,
                    "vlb_plan_execute: injected failure");
        plan_launch(*pl);
    });
}

vlb_status vlb_plan_sync(vlb_plan* h)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] { plan_sync(*pl); });
}

vlb_status vlb_plan_download(vlb_plan* h, float* out, size_t out_len)
{
    if (!h || !out) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] {
        const size_t n = (size_t)pl->args.P * pl->K * 3;
        VLB_REQUIRE(out_len >= n, VLB_ERR_INVALID, "output buffer too small");
        VLB_REQUIRE(pl->ran, VLB_ERR_INVALID, "plan has not been executed");
        VLB_HIP(hipMemcpyAsync(out, pl->out.p, n * sizeof(float), hipMemcpyDeviceToHost, pl->ctx->stream));
        plan_sync(*pl);
    });
}

vlb_status vlb_plan_download_f64(vlb_plan* h, double* out, size_t out_len)
{
    if (!h || !out) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] {
        const size_t n = (size_t)pl->args.P * pl->K * 3;
        VLB_REQUIRE(out_len >= n, VLB_ERR_INVALID, "output buffer too small");
        VLB_REQUIRE(pl->ran, VLB_ERR_INVALID, "plan has not been executed");
        VLB_HIP(hipMemcpyAsync(out, pl->out64.p, n * sizeof(double), hipMemcpyDeviceToHost,
                               pl->ctx->stream));
        plan_sync(*pl);
    });
}

vlb_status vlb_plan_set_timing(vlb_plan* h, int32_t on)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] {
        plan_sync(*pl);
        if (on && pl->ev.empty()) {
            pl->ev.assign(Plan::kEv * (size_t)Plan::kTimingRing, nullptr);
            for (auto& e : pl->ev) VLB_HIP(hipEventCreate(&e));
        }
        pl->timing = on != 0;
        if (on) {
            pl->timed_done = 0;
            pl->sum_trace = pl->sum_project = pl->sum_merge = pl->sum_total = 0.0;
            pl->stats.trace_ms = pl->stats.project_ms = pl->stats.merge_ms = pl->stats.total_ms = 0.0;
            pl->stats.timed_runs = 0;
        }
    });
}

vlb_status vlb_plan_use_graph(vlb_plan* h, int32_t on)
{
    if (!h) return VLB_ERR_INVALID;
    reinterpret_cast<Plan*>(h)->use_graph = on != 0;
    return VLB_OK;
}

vlb_status vlb_plan_set_indirect(vlb_plan* h, const vlb_indirect_params* ip, const float* coeffs)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] {
        Context& ctx = *pl->ctx;
        BakeArgs& a = pl->args;
        const int nvis0 = (int)pl->params.num_lights;
        if (!ip) {
            memset(&a.sp.ind, 0, sizeof a.sp.ind);
            if (pl->params.shade_mode == VLB_SHADE_LAMBERT) a.nvis = nvis0;
            pl->indirect = false;
            if (pl->default_variant) apply_variant(ctx, *pl, default_variant(false));
            drop_graph(*pl);
            return;
        }
        VLB_REQUIRE(pl->params.shade_mode == VLB_SHADE_LAMBERT, VLB_ERR_UNSUPPORTED,
                    "indirect light needs LAMBERT shading");
        const uint64_t P = check_grid(ip->grid, ip->lookup);
        VLB_REQUIRE(P == (uint64_t)a.P, VLB_ERR_INVALID, "grid probe count differs from the plan's");
        VLB_REQUIRE((int)ip->grid.num_coefs == pl->K, VLB_ERR_INVALID, "grid num_coefs differs from the plan's K");
        VLB_REQUIRE(std::isfinite(ip->scale), VLB_ERR_INVALID, "scale must be finite");
        VLB_REQUIRE(coeffs || pl->ran, VLB_ERR_INVALID,
                    "no source grid: pass coefficients or execute the plan first");
        VLB_HIP(hipSetDevice(ctx.device));
        const size_t n = (size_t)P * pl->K * 3;
        if (!pl->ind.p) pl->ind = dev_alloc(ctx, n * sizeof(float));
        if (coeffs) VLB_HIP(hipMemcpyAsync(pl->ind.p, coeffs, n * sizeof(float), hipMemcpyHostToDevice, ctx.stream));
        else launch_copy_f32(pl->ind.as<float>(), pl->out.as<float>(), n, ctx.stream);
        const int nvis = nvis0 + (ip->lookup.visibility ? 8 : 0);
        a.nvis = nvis;
        if (pl->default_variant) apply_variant(ctx, *pl, default_variant(true));
        if (pl->hits.p && pl->vis.bytes < (size_t)a.ntrace * 512 * (size_t)nvis) {
            pl->vis = dev_alloc(ctx, (size_t)a.ntrace * 512 * (size_t)nvis);
            a.vis = pl->vis.as<uint8_t>();
        }
        a.nvis = nvis;
        a.sp.ind = grid_view(ip->grid, ip->lookup, pl->ind.as<float>(), ip->scale);
        pl->indirect = true;
        drop_graph(*pl);
        // a host source buffer must stay valid until the copy is done
        VLB_HIP(hipStreamSynchronize(ctx.stream));
    });
}

vlb_status vlb_plan_count(vlb_plan* h)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    std::lock_guard<std::recursive_mutex> lk(pl->ctx->mtx);
    return guard(pl->ctx, [&] {
        hipStream_t st = pl->ctx->stream;
        VLB_HIP(hipMemsetAsync(pl->counters.p, 0, 4 * sizeof(unsigned long long), st));
        launch_trace_stage(*pl, false, true, st);
        unsigned long long cnt[4] = {0, 0, 0, 0};
        VLB_HIP(hipMemcpyAsync(cnt, pl->counters.p, sizeof cnt, hipMemcpyDeviceToHost, st));
        VLB_HIP(hipStreamSynchronize(st));
        pl->stats.node_visits = cnt[0];
        pl->stats.tri_tests = cnt[1];
        pl->stats.shadow_rays = cnt[2];
    });
}

vlb_status vlb_plan_autotune(vlb_plan* h, uint32_t iters, uint32_t* chosen)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    Context& c = *pl->ctx;
    std::lock_guard<std::recursive_mutex> lk(c.mtx);
    return guard(&c, [&] {
        hipStream_t st = c.stream;
        iters = std::max(1u, std::min(iters, 50u));
        hipEvent_t e0, e1;
        VLB_HIP(hipEventCreate(&e0));
        VLB_HIP(hipEventCreate(&e1));
        // fused: lane maps 0..3 unbounded + maps 2, 3 at 6 and 8 waves/SIMD,
        // each with slot-order and far-first shadows; wavefront: maps 0..3
        // with refill 16, maps 2, 3 with refill 32 / 8
        const uint32_t cand[] = {1,  2,  3,  4,  7,  8,  11, 12, 25, 26, 27, 28,
                                 31, 32, 35, 36, 13, 14, 15, 16, 19, 20, 23, 24};
        double best = 1e30;
        uint32_t best_v = pl->variant;
        for (uint32_t v : cand) {
            apply_variant(c, *pl, v);
            launch_trace_stage(*pl, false, false, st);  // warm-up
            VLB_HIP(hipEventRecord(e0, st));
            for (uint32_t i = 0; i < iters; ++i) launch_trace_stage(*pl, false, false, st);
            VLB_HIP(hipEventRecord(e1, st));
            VLB_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            VLB_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_v = v;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        apply_variant(c, *pl, best_v);
        pl->default_variant = false;
        // a pick made with the indirect term on does not carry over to
        // direct-only plans of the scene
        if (!pl->indirect) c.tuned[{pl->scene->id, (uint32_t)pl->args.R}] = best_v;
        if (chosen) *chosen = best_v;
    });
}

vlb_status vlb_plan_stats(vlb_plan* h, vlb_bake_stats* out)
{
    if (!h || !out) return VLB_ERR_INVALID;
    *out = reinterpret_cast<Plan*>(h)->stats;
    return VLB_OK;
}

vlb_status vlb_plan_destroy(vlb_plan* h)
{
    if (!h) return VLB_ERR_INVALID;
    Plan* pl = reinterpret_cast<Plan*>(h);
    Context* c = pl->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    (void)hipStreamSynchronize(c->stream);
    --pl->scene->users;
    delete pl;
    --c->children;
    return VLB_OK;
}

vlb_status vlb_bake(vlb_ctx* h, uint32_t scene, const vlb_bake_params* p, float* out, size_t out_len)
{
    if (!h || !p || !out) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        Scene& s = find_scene(*c, scene);
        VLB_REQUIRE(!take_injected(*c, VLB_INJECT_PLAN_EXECUTE), VLB_ERR_INJECTED,
                    "vlb_bake: injected failure");
        VLB_HIP(hipSetDevice(c->device));
        Plan pl;
        build_plan(*c, s, *p, pl);
        const size_t n = (size_t)p->num_probes * pl.K * 3;
        VLB_REQUIRE(out_len >= n, VLB_ERR_INVALID, "output buffer too small");
        plan_launch(pl);
        VLB_HIP(hipMemcpyAsync(out, pl.out.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        plan_sync(pl);
    });
}

vlb_status vlb_bake_radiance(vlb_ctx* h, uint32_t scene, const vlb_bake_params* p, float* radiance,
                             size_t radiance_len, int32_t* hit_prim, size_t hit_len)
{
    if (!h || !p) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        Scene& s = find_scene(*c, scene);
        VLB_HIP(hipSetDevice(c->device));
        Plan pl;
        build_plan(*c, s, *p, pl);
        const size_t nr = (size_t)pl.args.P * pl.args.R;
        VLB_REQUIRE(!radiance || radiance_len >= nr * 3, VLB_ERR_INVALID, "radiance buffer too small");
        VLB_REQUIRE(!hit_prim || hit_len >= nr, VLB_ERR_INVALID, "hit buffer too small");
        DevBuf d_rad, d_hit;
        if (radiance) {
            d_rad = dev_alloc(*c, nr * 3 * sizeof(float));
            pl.args.rad_out = d_rad.as<float>();
        }
        if (hit_prim) {
            d_hit = dev_alloc(*c, nr * sizeof(int));
            pl.args.hit_out = d_hit.as<int>();
        }
        launch_trace_stage(pl, true, false, c->stream);
        if (radiance)
            VLB_HIP(hipMemcpyAsync(radiance, d_rad.p, nr * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                   c->stream));
        if (hit_prim)
            VLB_HIP(hipMemcpyAsync(hit_prim, d_hit.p, nr * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------- skybox / images
// Solid-angle table of one face and, for faces whose per-texel direction
// table fits in 2 MiB (N <= 147, so it stays in every XCD's L2 next to the
// streamed texels), that table: the projection then reads each texel's
// direction instead of computing it (the same values, from the same code).
static void ensure_wtab(Context& c, int N)
{
    if (c.wtab_n == N && c.wtab.p) return;
    std::vector<float> w = solid_angle_table(N);
    c.wtab = upload_vec(c, w);
    c.sky_tab.release();
    const size_t tab_bytes = (size_t)6 * N * N * sizeof(float4);
    if (tab_bytes <= (2u << 20)) {
        c.sky_tab = dev_alloc(c, tab_bytes);
        launch_skybox_table(N, c.wtab.as<float>(), c.sky_tab.as<float4>(), c.stream);
    }
    c.wtab_n = N;
}

static int skybox_blocks_per_cube(int N, int batch)
{
    const int T = 6 * N * N;
    // ~2048+ workgroups in flight, 256..32768 texels each (larger ranges
    // amortise the end-of-workgroup reduction: 1024 cubemaps of 6 x 128^2
    // run 5.4 TB/s at 32768 against 5.2 at 8192 and 4.9 at 49152)
    long long per = (long long)T * batch / 2048;
    per = std::max(256LL, std::min(32768LL, per));
    per = (per + 255) / 256 * 256;
    return (int)((T + per - 1) / per);
}

vlb_status vlb_skybox_sh_d(vlb_ctx* h, const void* d_cubes, uint32_t face_size, uint32_t batch,
                           uint32_t sh_order, void* d_out)
{
    if (!h || !d_cubes || !d_out) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(face_size >= 1 && face_size <= 8192, VLB_ERR_INVALID, "face_size out of range");
        VLB_REQUIRE(batch >= 1 && (uint64_t)batch * 6 * face_size * face_size < (1ull << 31),
                    VLB_ERR_INVALID, "batch out of range");
        VLB_REQUIRE(sh_order == 2 || sh_order == 3, VLB_ERR_INVALID, "sh_order must be 2 or 3");
        VLB_HIP(hipSetDevice(c->device));
        const int N = (int)face_size, K = sh_order == 2 ? 9 : 16;
        ensure_wtab(*c, N);
        const int bpc = skybox_blocks_per_cube(N, (int)batch);
        const size_t pbytes = (size_t)batch * bpc * K * 3 * sizeof(float);
        if (c->sky_partial.bytes < pbytes) c->sky_partial = dev_alloc(*c, pbytes);
        launch_skybox_sh(static_cast<const float4*>(d_cubes), N, (int)batch, K, c->wtab.as<float>(),
                         c->sky_tab.p ? c->sky_tab.as<float4>() : nullptr, c->sky_partial.as<float>(), bpc, static_cast<float*>(d_out), c->stream,
                         c->ev[0], c->ev[1], c->ev[2]);
        c->sky_pending = true;
    });
}

vlb_status vlb_skybox_sh_noalloc(vlb_ctx* h, const void* d_cubes, uint32_t face_size, uint32_t batch,
                                 uint32_t sh_order, void* d_out)
{
    if (!h || !d_cubes || !d_out) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(face_size >= 1 && face_size <= 8192, VLB_ERR_INVALID, "face_size out of range");
        VLB_REQUIRE(batch >= 1 && (uint64_t)batch * 6 * face_size * face_size < (1ull << 31),
                    VLB_ERR_INVALID, "batch out of range");
        VLB_REQUIRE(sh_order == 2 || sh_order == 3, VLB_ERR_INVALID, "sh_order must be 2 or 3");
        VLB_REQUIRE(c->sky_scratch, VLB_ERR_DEVICE, "skybox scratch not resolved for this device");
        const int N = (int)face_size, K = sh_order == 2 ? 9 : 16;
        const int bpc = skybox_blocks_per_cube(N, (int)batch);
        // scratch layout: solid angles | directions (faces <= 147 texels) | partials
        auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
        const size_t wbytes = a256((size_t)N * N * sizeof(float));
        const size_t tab_bytes = (size_t)6 * N * N * sizeof(float4);
        const bool use_tab = tab_bytes <= (2u << 20);
        const size_t part_off = wbytes + (use_tab ? a256(tab_bytes) : 0);
        const size_t need = part_off + (size_t)batch * bpc * K * 3 * sizeof(float);
        VLB_REQUIRE(need <= kSkyScratchBytes, VLB_ERR_INVALID,
                    "vlb_skybox_sh_noalloc: " + std::to_string(batch) + " cubemaps of " +
                        std::to_string(N) + "^2 texels need " + std::to_string(need) +
                        " bytes of scratch, more than its " + std::to_string(kSkyScratchBytes) +
                        "; use vlb_skybox_sh_d");
        VLB_HIP(hipSetDevice(c->device));
        launch_skybox_sh_noalloc(static_cast<const float4*>(d_cubes), N, (int)batch, K, c->sky_scratch,
                                 wbytes, part_off, use_tab, bpc, static_cast<float*>(d_out), c->stream);
    });
}

vlb_status vlb_skybox_last_ms(vlb_ctx* h, double* project_ms, double* merge_ms)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_HIP(hipStreamSynchronize(c->stream));
        if (c->sky_pending) {
            float a = 0.f, b = 0.f;
            VLB_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
            VLB_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
            c->sky_project_ms = a;
            c->sky_merge_ms = b;
            c->sky_pending = false;
        }
        if (project_ms) *project_ms = c->sky_project_ms;
        if (merge_ms) *merge_ms = c->sky_merge_ms;
    });
}

vlb_status vlb_skybox_sh(vlb_ctx* h, const float* cubes, uint32_t face_size, uint32_t batch,
                         uint32_t sh_order, float* out, size_t out_len)
{
    if (!h || !cubes || !out) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(sh_order == 2 || sh_order == 3, VLB_ERR_INVALID, "sh_order must be 2 or 3");
        VLB_REQUIRE(face_size >= 1 && face_size <= 8192 && batch >= 1, VLB_ERR_INVALID, "bad size");
        const int K = sh_order == 2 ? 9 : 16;
        const size_t n_out = (size_t)batch * K * 3;
        VLB_REQUIRE(out_len >= n_out, VLB_ERR_INVALID, "output buffer too small");
        VLB_HIP(hipSetDevice(c->device));
        const size_t in_bytes = (size_t)batch * 6 * face_size * face_size * 16;
        DevBuf d_in = dev_alloc(*c, in_bytes);
        DevBuf d_out = dev_alloc(*c, n_out * sizeof(float));
        VLB_HIP(hipMemcpyAsync(d_in.p, cubes, in_bytes, hipMemcpyHostToDevice, c->stream));
        vlb_status s = vlb_skybox_sh_d(h, d_in.p, face_size, batch, sh_order, d_out.p);
        if (s != VLB_OK) throw Error{s, get_error(c)};
        VLB_HIP(hipMemcpyAsync(out, d_out.p, n_out * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

vlb_status vlb_project_equirect(vlb_ctx* h, const uint8_t* rgba8, uint32_t width, uint32_t height,
                                int32_t variant, float* out16x3)
{
    if (!h || !rgba8 || !out16x3) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(width >= 1 && height >= 1 && (uint64_t)width * height < (1ull << 30),
                    VLB_ERR_INVALID, "bad image size");
        VLB_REQUIRE(variant == VLB_EQUIRECT_SH_COMP || variant == VLB_EQUIRECT_SKYBOX_COMP,
                    VLB_ERR_INVALID, "bad variant");
        VLB_HIP(hipSetDevice(c->device));
        const size_t n = (size_t)width * height;
        const int nblocks = (int)std::max<size_t>(1, (n + 4095) / 4096);
        DevBuf d_img = dev_alloc(*c, n * 4);
        DevBuf d_part = dev_alloc(*c, (size_t)nblocks * 48 * sizeof(float));
        DevBuf d_out = dev_alloc(*c, 48 * sizeof(double));
        VLB_HIP(hipMemcpyAsync(d_img.p, rgba8, n * 4, hipMemcpyHostToDevice, c->stream));
        launch_equirect_project(d_img.as<uint8_t>(), (int)width, (int)height, variant,
                                d_out.as<double>(), d_part.as<float>(), nblocks, c->stream);
        double o[48];
        VLB_HIP(hipMemcpyAsync(o, d_out.p, sizeof o, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
        for (int i = 0; i < 48; ++i) out16x3[i] = (float)o[i];
    });
}

vlb_status vlb_sh_reconstruct(vlb_ctx* h, const float* coeffs, uint32_t width, uint32_t height,
                              uint8_t* out)
{
    if (!h || !coeffs || !out) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(width >= 1 && height >= 1 && (uint64_t)width * height < (1ull << 30),
                    VLB_ERR_INVALID, "bad image size");
        VLB_HIP(hipSetDevice(c->device));
        const size_t n = (size_t)width * height;
        DevBuf d_c = dev_alloc(*c, 48 * sizeof(float));
        DevBuf d_o = dev_alloc(*c, n * 4);
        VLB_HIP(hipMemcpyAsync(d_c.p, coeffs, 48 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        launch_sh_reconstruct(d_c.as<float>(), (int)width, (int)height, d_o.as<uint8_t>(), c->stream);
        VLB_HIP(hipMemcpyAsync(out, d_o.p, n * 4, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

// ------------------------------------------------------- probe-grid lookup
vlb_status vlb_probe_lookup(vlb_ctx* h, uint32_t scene, const vlb_probe_grid_desc* g,
                            const float* coeffs, const vlb_lookup_params* lp, const float* points,
                            const float* normals, uint32_t n, float* out)
{
    if (!h || !g || !coeffs || !lp || (n && (!points || !normals || !out))) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        const uint64_t P = check_grid(*g, *lp);
        VLB_REQUIRE(n < (1u << 28), VLB_ERR_INVALID, "too many lookup points");
        if (n == 0) return;
        Scene& s = find_scene(*c, scene);
        VLB_HIP(hipSetDevice(c->device));
        const size_t nc = (size_t)P * g->num_coefs * 3;
        DevBuf d_c = dev_alloc(*c, nc * sizeof(float));
        DevBuf d_p = dev_alloc(*c, (size_t)n * 3 * sizeof(float));
        DevBuf d_n = dev_alloc(*c, (size_t)n * 3 * sizeof(float));
        DevBuf d_o = dev_alloc(*c, (size_t)n * 3 * sizeof(float));
        VLB_HIP(hipMemcpyAsync(d_c.p, coeffs, d_c.bytes, hipMemcpyHostToDevice, c->stream));
        VLB_HIP(hipMemcpyAsync(d_p.p, points, d_p.bytes, hipMemcpyHostToDevice, c->stream));
        VLB_HIP(hipMemcpyAsync(d_n.p, normals, d_n.bytes, hipMemcpyHostToDevice, c->stream));
        const int grid = lookup_grid_size(*c);
        const uint32_t stride = (uint32_t)grid * 256u;
        DevBuf spill = dev_alloc(*c, scene_spill_bytes(s, stride));
        LookupArgs a;
        memset(&a, 0, sizeof a);
        a.sc = s.view();
        a.g = grid_view(*g, *lp, d_c.as<float>(), 1.0f);
        a.pts = d_p.as<float>();
        a.nrm = d_n.as<float>();
        a.n = (int)n;
        a.out = d_o.as<float>();
        a.spill = spill.as<uint32_t>();
        a.spill_stride = stride;
        launch_lookup(a, grid, c->stream);
        VLB_HIP(hipMemcpyAsync(out, d_o.p, d_o.bytes, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------- device buffers
vlb_status vlb_device_alloc(vlb_ctx* h, size_t bytes, void** dptr)
{
    if (!h || !dptr) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_HIP(hipSetDevice(c->device));
        DevBuf b = dev_alloc(*c, bytes);
        *dptr = b.p;
        c->user_bufs[b.p] = std::move(b);
    });
}

vlb_status vlb_device_free(vlb_ctx* h, void* dptr)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        auto it = c->user_bufs.find(dptr);
        VLB_REQUIRE(it != c->user_bufs.end(), VLB_ERR_NOT_FOUND, "pointer not allocated by vlb_device_alloc");
        VLB_HIP(hipStreamSynchronize(c->stream));
        c->user_bufs.erase(it);
    });
}

vlb_status vlb_upload(vlb_ctx* h, void* dptr, const void* src, size_t bytes)
{
    if (!h || !dptr || !src) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_HIP(hipMemcpyAsync(dptr, src, bytes, hipMemcpyHostToDevice, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

vlb_status vlb_download(vlb_ctx* h, void* dst, const void* dptr, size_t bytes)
{
    if (!h || !dptr || !dst) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_HIP(hipMemcpyAsync(dst, dptr, bytes, hipMemcpyDeviceToHost, c->stream));
        VLB_HIP(hipStreamSynchronize(c->stream));
    });
}

vlb_status vlb_device_copy(vlb_ctx* h, void* dst, const void* src, size_t bytes, double* ms)
{
    if (!h || !dst || !src) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(bytes % 16 == 0, VLB_ERR_INVALID, "copy size must be a multiple of 16");
        VLB_HIP(hipSetDevice(c->device));
        VLB_HIP(hipEventRecord(c->ev[0], c->stream));
        launch_copy(dst, src, bytes, c->stream);
        VLB_HIP(hipEventRecord(c->ev[1], c->stream));
        if (ms) {
            VLB_HIP(hipStreamSynchronize(c->stream));
            float f = 0.f;
            VLB_HIP(hipEventElapsedTime(&f, c->ev[0], c->ev[1]));
            *ms = f;
        }
    });
}

vlb_status vlb_device_read(vlb_ctx* h, const void* src, size_t bytes, uint32_t grid, double* ms)
{
    if (!h || !src) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    return guard(c, [&] {
        VLB_REQUIRE(bytes % 16 == 0, VLB_ERR_INVALID, "read size must be a multiple of 16");
        VLB_REQUIRE(grid <= (1u << 20), VLB_ERR_INVALID, "grid above 2^20 workgroups");
        VLB_HIP(hipSetDevice(c->device));
        const int g = grid ? (int)grid : 16 * c->num_cus;
        VLB_HIP(hipEventRecord(c->ev[0], c->stream));
        launch_read(src, bytes, g, c->stream);
        VLB_HIP(hipEventRecord(c->ev[1], c->stream));
        if (ms) {
            VLB_HIP(hipStreamSynchronize(c->stream));
            float f = 0.f;
            VLB_HIP(hipEventElapsedTime(&f, c->ev[0], c->ev[1]));
            *ms = f;
        }
    });
}

vlb_status vlb_ctx_sync(vlb_ctx* h)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = reinterpret_cast<Context*>(h);
    return guard(c, [&] { VLB_HIP(hipStreamSynchronize(c->stream)); });
}

}  // extern "C"
