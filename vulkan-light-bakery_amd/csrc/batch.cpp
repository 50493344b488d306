// batch.cpp -- many bake jobs over several resident scenes as one unit.
//
// The reference bakes one scene per process, one probe at a time
// (LightBaker::bake, src/baker/light_baker.cpp:287-328); a batch runs any
// number of such bakes, each with the semantics of vlb_bake, through the
// multi-job kernels of multi.hip: rounds of at most max_round_rays primary
// rays, four persistent launches per round whatever the number of jobs, and
// the whole sequence optionally captured as one hipGraph.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

#include "multi_host.hpp"

namespace vlb {

// ------------------------------------------------------------- ray tables
std::shared_ptr<RayTable> RayCache::get(Context& ctx, const vlb_bake_params& p, int K, uint32_t RS,
                                        bool* built)
{
    if (built) *built = false;
    const bool fib = p.dir_mode == VLB_DIRS_FIBONACCI;
    const Key key{p.dir_mode, fib ? p.rays_per_probe : p.equirect_w, fib ? 0u : p.equirect_h, K,
                  p.sh_constants, p.precision, RS};
    auto it = tabs_.find(key);
    if (it != tabs_.end()) return it->second;
    auto t = std::make_shared<RayTable>();
    build_ray_table(ctx, p, K, RS, *t);
    tabs_[key] = t;
    if (built) *built = true;
    return t;
}

size_t RayCache::bytes() const
{
    size_t b = 0;
    for (const auto& kv : tabs_) b += kv.second->dirs.bytes + kv.second->dir_perm.bytes + kv.second->ytab.bytes;
    return b;
}

// ------------------------------------------------------------------- jobs
uint32_t job_slice_rays(const vlb_bake_params& p, uint32_t R)
{
    if (p.rays_per_slice) return (p.rays_per_slice + 63) & ~63u;
    return std::min<uint32_t>((R + 63) & ~63u, 16384u);
}

size_t job_state_bytes(const vlb_bake_params& p, bool out64)
{
    const int K = p.sh_order == 2 ? 9 : 16;
    const uint32_t R = rays_of(p);
    const uint32_t RS = job_slice_rays(p, R);
    const size_t nslices = (R + RS - 1) / RS;
    const size_t Ppad = ((size_t)p.num_probes + 15) / 16 * 16;
    size_t b = Ppad * (sizeof(float4) + sizeof(int));
    b += nslices * Ppad * 48 * sizeof(float2);
    b += (size_t)p.num_probes * K * 3 * (sizeof(float) + (out64 ? sizeof(double) : 0));
    return b;
}

void init_job(Context& ctx, RayCache& cache, uint32_t scene, const vlb_bake_params& p, bool out64,
              JobState& js)
{
    validate(p);
    Scene& sc = find_scene(ctx, scene);
    js.scene_id = scene;
    js.params = p;
    js.lights.assign(p.lights, p.lights + p.num_lights);
    js.params.lights = js.lights.data();
    js.params.probe_xyz = nullptr;
    js.K = p.sh_order == 2 ? 9 : 16;
    const uint32_t R = rays_of(p);
    js.rt = cache.get(ctx, p, js.K, job_slice_rays(p, R));
    build_probe_set(ctx, p.probe_xyz, p.num_probes, js.ps);
    const RayTable& rt = *js.rt;
    js.npg = js.ps.Ppad / 16;
    VLB_REQUIRE((uint64_t)js.npg * rt.nslices < (1ull << 30), VLB_ERR_INVALID, "job too large");
    js.ntiles = js.npg * rt.nslices;
    js.nch = (int)(rt.RS / 32);
    js.rays = (size_t)p.num_probes * R;
    js.partials = dev_alloc(ctx, (size_t)rt.nslices * js.ps.Ppad * 48 * sizeof(float2));
    js.out = dev_alloc(ctx, (size_t)p.num_probes * js.K * 3 * sizeof(float));
    if (out64) js.out64 = dev_alloc(ctx, (size_t)p.num_probes * js.K * 3 * sizeof(double));

    JobDesc& d = js.desc;
    memset(&d, 0, sizeof d);
    d.sc = sc.view();
    d.sp = shade_params(js.params);
    d.probes = js.ps.probes.as<float4>();
    d.probe_perm = js.ps.probe_perm.as<int>();
    d.dirs = rt.dirs.as<float4>();
    d.ytab = rt.ytab.p;
    d.partials = js.partials.as<float2>();
    d.out = js.out.as<float>();
    d.out64 = out64 ? js.out64.as<double>() : nullptr;
    // FP16 forms carry radiance / 16 (range headroom); the merge undoes it
    const bool fp16 = p.precision == VLB_PREC_FP16 || p.precision == VLB_PREC_FP16_FAST;
    d.scale = fp16 ? rt.wref * 16.0 : rt.wref;
    d.P = (int)p.num_probes;
    d.Ppad = js.ps.Ppad;
    d.R = (int)R;
    d.Rpad = (int)rt.Rpad;
    d.RS = (int)rt.RS;
    d.nslices = rt.nslices;
    d.nchunk = (int)(rt.Rpad / 32);
    d.K = js.K;
    d.tmin = p.t_min;
    d.tmax = p.t_max;
    d.nvis = p.shade_mode == VLB_SHADE_REFERENCE ? 1 : (int)p.num_lights;
}

// ------------------------------------------------------------------ rounds
void round_add(RoundTable& r, int job, const JobState& js, int pt0, int npt)
{
    Segment s{};
    s.job = job;
    s.pt0 = pt0;
    s.npt = npt;
    s.tb = r.hdr.ntrace;
    s.pb = r.hdr.nproj;
    r.seg.push_back(s);
    r.hdr.nseg = (int)r.seg.size();
    r.hdr.ntrace += npt * js.nch;
    r.hdr.nproj += npt;
    r.hdr.nvis = std::max(r.hdr.nvis, js.desc.nvis);
    // primary rays of these tiles (padding probes and rays excluded)
    const int nsl = js.rt->nslices;
    for (int pt = pt0; pt < pt0 + npt; ++pt) {
        const int pg = pt / nsl, sl = pt - pg * nsl;
        const size_t np = std::min<size_t>(16, (size_t)js.desc.P - (size_t)pg * 16);
        const size_t nr = std::min<size_t>(js.rt->RS, (size_t)js.desc.R - (size_t)sl * js.rt->RS);
        r.rays += np * nr;
    }
}

void round_merge(RoundTable& r, int job, const JobState& js)
{
    MergeItem m{};
    m.job = job;
    m.base = r.hdr.merge_threads;
    r.merge.push_back(m);
    r.hdr.nmerge = (int)r.merge.size();
    r.hdr.merge_threads += (js.ps.Ppad * 48 + 255) / 256 * 256;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t round_bytes(const RoundTable& r)
{
    return align256(sizeof(WorkHeader)) + align256(std::max<size_t>(1, r.seg.size()) * sizeof(Segment)) +
           align256(std::max<size_t>(1, r.merge.size()) * sizeof(MergeItem));
}

void round_pack(const RoundTable& r, uint8_t* dst, size_t off[3])
{
    off[0] = 0;
    off[1] = align256(sizeof(WorkHeader));
    off[2] = off[1] + align256(std::max<size_t>(1, r.seg.size()) * sizeof(Segment));
    memcpy(dst + off[0], &r.hdr, sizeof(WorkHeader));
    if (!r.seg.empty()) memcpy(dst + off[1], r.seg.data(), r.seg.size() * sizeof(Segment));
    if (!r.merge.empty()) memcpy(dst + off[2], r.merge.data(), r.merge.size() * sizeof(MergeItem));
}

size_t round_bytes_cap(size_t seg_cap, size_t merge_cap)
{
    return align256(sizeof(WorkHeader)) + align256(std::max<size_t>(1, seg_cap) * sizeof(Segment)) +
           align256(std::max<size_t>(1, merge_cap) * sizeof(MergeItem));
}

void round_pack_cap(const RoundTable& r, uint8_t* dst, size_t seg_cap, size_t off[3])
{
    VLB_REQUIRE(r.seg.size() <= seg_cap, VLB_ERR_INVALID, "round has too many segments");
    off[0] = 0;
    off[1] = align256(sizeof(WorkHeader));
    off[2] = off[1] + align256(std::max<size_t>(1, seg_cap) * sizeof(Segment));
    memcpy(dst + off[0], &r.hdr, sizeof(WorkHeader));
    if (!r.seg.empty()) memcpy(dst + off[1], r.seg.data(), r.seg.size() * sizeof(Segment));
    if (!r.merge.empty()) memcpy(dst + off[2], r.merge.data(), r.merge.size() * sizeof(MergeItem));
}

void RoundBufs::alloc(Context& ctx, size_t cap, int nv, uint32_t dep)
{
    cap_tiles = cap;
    nvis = nv;
    depth = dep;
    grid = multi_grid_size(ctx);
    int max_grid = grid;
    for (int k = 0; k < 6; ++k) {
        fgrid[k] = multi_fused_grid_size(ctx, k % 3 == 0 ? 0 : (k % 3 == 1 ? 6 : 8), k >= 3);
        max_grid = std::max(max_grid, fgrid[k]);
    }
    shade_grid = (int)std::max<size_t>(8, std::min<size_t>(cap * 2, 8192));
    pgrid = ctx.num_cus * 8;
    mgrid = 1024;
    rad = dev_alloc(ctx, cap * 1536 * sizeof(float));
    hits = dev_alloc(ctx, cap * 512 * sizeof(int4));
    vis = dev_alloc(ctx, std::max<size_t>(16, cap * 512 * (size_t)nv));
    wctr = dev_alloc(ctx, 32 * kCtrStride * sizeof(unsigned long long));
    counters = dev_alloc(ctx, 4 * sizeof(unsigned long long));
    const uint32_t sb = stack_bound(dep);
    const uint32_t e = sb > (uint32_t)kLdsDepthMin ? sb - kLdsDepthMin : 0;
    spill = dev_alloc(ctx, std::max<size_t>(16, (size_t)e * max_grid * 256 * 4));
}

MultiArgs RoundBufs::args(const JobDesc* jobs, const uint8_t* table, const size_t off[3], int map,
                          int refill, int chunk, int focc, bool far) const
{
    MultiArgs a;
    memset(&a, 0, sizeof a);
    a.jobs = jobs;
    a.hdr = reinterpret_cast<const WorkHeader*>(table + off[0]);
    a.seg = reinterpret_cast<const Segment*>(table + off[1]);
    a.merge = reinterpret_cast<const MergeItem*>(table + off[2]);
    a.rad = rad.as<float>();
    a.hits = hits.as<int4>();
    a.vis = vis.as<uint8_t>();
    a.nvis = nvis;
    a.map = map;
    a.refill = refill;
    a.chunk = chunk;
    a.wctr = wctr.as<unsigned long long>();
    a.spill = spill.as<uint32_t>();
    a.focc = focc;
    a.far = focc >= 0 && far ? 1 : 0;
    a.fgrid = focc < 0 ? 0 : fgrid[(focc == 8 ? 2 : (focc == 6 ? 1 : 0)) + 3 * a.far];
    a.spill_stride = (uint32_t)(focc < 0 ? grid : a.fgrid) * 256u;
    a.counters = counters.as<unsigned long long>();
    return a;
}

void multi_variant(uint32_t v, int& map, int& refill, int& focc, bool& far)
{
    VLB_REQUIRE(v >= 1 && v <= 36, VLB_ERR_INVALID, "multi-job trace_variant must be 1..36");
    far = v >= 25;
    if (far) v -= 24;
    static const int kRefill[3] = {16, 32, 8};
    static const int kOcc[3] = {0, 6, 8};
    if (v >= 13) {
        map = (int)((v - 13) % 4);
        refill = kRefill[(v - 13) / 4];
        focc = -1;
    } else {
        map = (int)((v - 1) % 4);
        refill = 32;
        focc = kOcc[(v - 1) / 4];
    }
}

// ------------------------------------------------------------------ batch
namespace {

constexpr uint32_t kBatchDefaultVariant = 12;  // fused, lane map 3, 8 waves/SIMD
constexpr uint64_t kDefaultRoundRays = 1ull << 28;

struct Batch {
    Context* ctx = nullptr;
    RayCache cache;
    std::vector<std::unique_ptr<JobState>> jobs;
    int precision = 0;
    bool keep64 = false;
    DevBuf jobtab;                 // JobDesc[njobs]
    std::vector<RoundTable> rounds;
    std::vector<size_t> round_off; // byte offset of each round's table
    std::vector<std::array<size_t, 3>> parts;
    DevBuf table;
    RoundBufs rb;
    uint32_t variant = kBatchDefaultVariant;
    int map = 3, refill = 32, chunk = 256, focc = -1;
    bool far = false;
    std::vector<Scene*> scenes;    // one per job, each holding a use of its scene
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    static constexpr int kRing = 256;
    std::vector<hipEvent_t> ev;    // 2 per execute
    uint32_t pending = 0;
    double sum_ms = 0.0;
    uint32_t runs = 0;
    // per-kernel timing of eager executes: 4 events per round (before the
    // trace launch, then after the trace, projection and merge launches), a
    // ring of kKRing executes.  A captured graph is timed as a whole: events
    // recorded into a graph under capture cannot be timed on this runtime
    // (hipEventElapsedTime fails on them).
    static constexpr int kKRing = 32;
    bool ktiming = false;
    std::vector<hipEvent_t> kev;
    uint32_t kpending = 0;
    double ksum[3] = {0.0, 0.0, 0.0};
    uint32_t kruns = 0;
    vlb_batch_info info{};
    ~Batch()
    {
        drop_graph();
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : kev)
            if (e) (void)hipEventDestroy(e);
    }
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    size_t kset() const { return rounds.size() * 4; }
    MultiArgs args(size_t r) const
    {
        const size_t off[3] = {round_off[r] + parts[r][0], round_off[r] + parts[r][1],
                               round_off[r] + parts[r][2]};
        return rb.args(jobtab.as<JobDesc>(), table.as<uint8_t>(), off, map, refill, chunk, focc, far);
    }
    // E: kset() events to record around the kernels, or null
    void launch_all(bool count, hipStream_t st, hipEvent_t* E = nullptr)
    {
        for (size_t r = 0; r < rounds.size(); ++r) {
            const MultiArgs a = args(r);
            if (E) VLB_HIP(hipEventRecord(E[4 * r + 0], st));
            launch_multi_trace(a, count, rb.grid, rb.shade_grid, st);
            if (E) VLB_HIP(hipEventRecord(E[4 * r + 1], st));
            launch_multi_project(a, precision, rb.pgrid, st);
            if (E) VLB_HIP(hipEventRecord(E[4 * r + 2], st));
            launch_multi_merge(a, rb.mgrid, st);
            if (E) VLB_HIP(hipEventRecord(E[4 * r + 3], st));
        }
    }
    // Adds one execute's kernel times (its event set E, already complete).
    void read_kernels(const hipEvent_t* E)
    {
        for (size_t r = 0; r < rounds.size(); ++r)
            for (int k = 0; k < 3; ++k) {
                float ms = 0.f;
                VLB_HIP(hipEventElapsedTime(&ms, E[4 * r + k], E[4 * r + k + 1]));
                ksum[k] += ms;
            }
        ++kruns;
    }
    void drain()
    {
        for (uint32_t i = 0; i < pending; ++i) {
            float ms = 0.f;
            VLB_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            sum_ms += ms;
            ++runs;
        }
        pending = 0;
        for (uint32_t i = 0; i < kpending; ++i) read_kernels(&kev[kset() * i]);
        kpending = 0;
    }
    void sync()
    {
        VLB_HIP(hipStreamSynchronize(ctx->stream));
        drain();
    }
    void capture(hipStream_t st, hipGraph_t& g, hipGraphExec_t& x)
    {
        VLB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            launch_all(false, st);
        } catch (...) {
            hipGraph_t gg = nullptr;
            (void)hipStreamEndCapture(st, &gg);
            if (gg) (void)hipGraphDestroy(gg);
            throw;
        }
        VLB_HIP(hipStreamEndCapture(st, &g));
        VLB_HIP(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    }
    void set_timing(bool on)
    {
        sync();
        ktiming = on;
        ksum[0] = ksum[1] = ksum[2] = 0.0;
        kruns = 0;
        if (on && kev.size() < kset() * kKRing) {
            for (auto& e : kev)
                if (e) (void)hipEventDestroy(e);
            kev.assign(kset() * kKRing, nullptr);
            for (auto& e : kev) VLB_HIP(hipEventCreate(&e));
        }
    }
    void execute()
    {
        hipStream_t st = ctx->stream;
        if (pending == (uint32_t)kRing) sync();
        VLB_HIP(hipEventRecord(ev[2 * pending], st));
        if (use_graph) {
            if (!gexec) capture(st, graph, gexec);
            VLB_HIP(hipGraphLaunch(gexec, st));
        } else {
            hipEvent_t* E = nullptr;
            if (ktiming) {
                if (kpending == (uint32_t)kKRing) sync();
                E = &kev[kset() * kpending];
                ++kpending;
            }
            launch_all(false, st, E);
        }
        VLB_HIP(hipEventRecord(ev[2 * pending + 1], st));
        ++pending;
    }
};

void build_batch(Context& ctx, const vlb_batch_job* jobs, uint32_t njobs, const vlb_batch_options* opt,
                 Batch& b)
{
    VLB_REQUIRE(njobs >= 1 && jobs, VLB_ERR_INVALID, "a batch needs at least one job");
    b.ctx = &ctx;
    b.keep64 = opt && opt->keep_f64;
    uint64_t max_rays = opt && opt->max_round_rays ? opt->max_round_rays : kDefaultRoundRays;
    max_rays = std::min<uint64_t>(max_rays, 1ull << 30);
    b.variant = opt && opt->trace_variant ? opt->trace_variant : kBatchDefaultVariant;
    multi_variant(b.variant, b.map, b.refill, b.focc, b.far);
    b.chunk = opt && opt->chunk ? (int)opt->chunk : 256;
    VLB_REQUIRE(b.chunk >= 64 && b.chunk <= 1024 && 1024 % b.chunk == 0, VLB_ERR_INVALID,
                "chunk must divide 1024 and be at least 64");

    int nvis = 0;
    uint32_t depth = 0;
    int max_nch = 2;
    for (uint32_t j = 0; j < njobs; ++j) {
        VLB_REQUIRE(jobs[j].params, VLB_ERR_INVALID, "job params are NULL");
        if (j == 0) b.precision = jobs[j].params->precision;
        VLB_REQUIRE(jobs[j].params->precision == b.precision, VLB_ERR_INVALID,
                    "all jobs of a batch must use the same precision");
        auto js = std::make_unique<JobState>();
        init_job(ctx, b.cache, jobs[j].scene, *jobs[j].params, b.keep64, *js);
        nvis = std::max(nvis, js->desc.nvis);
        Scene& sc = find_scene(ctx, jobs[j].scene);
        depth = std::max(depth, sc.depth);
        b.scenes.push_back(&sc);
        max_nch = std::max(max_nch, js->nch);
        b.info.primary_rays += js->rays;
        b.info.projection_tiles += (uint64_t)js->ntiles;
        b.jobs.push_back(std::move(js));
    }
    // rounds: whole projection tiles, at most cap trace tiles each
    const size_t cap = std::max<size_t>((size_t)max_nch, max_rays / 512);
    RoundTable cur;
    size_t peak = 0;
    auto flush = [&] {
        if (cur.hdr.nseg == 0 && cur.hdr.nmerge == 0) return;
        peak = std::max(peak, (size_t)cur.hdr.ntrace);
        b.rounds.push_back(std::move(cur));
        cur = RoundTable();
    };
    for (uint32_t j = 0; j < njobs; ++j) {
        const JobState& js = *b.jobs[j];
        int pt = 0;
        while (pt < js.ntiles) {
            const size_t room = cap - (size_t)cur.hdr.ntrace;
            const int can = (int)std::min<size_t>(room / js.nch, (size_t)(js.ntiles - pt));
            if (can == 0) {
                flush();
                continue;
            }
            round_add(cur, (int)j, js, pt, can);
            pt += can;
        }
        round_merge(cur, (int)j, js);
    }
    flush();
    // device tables
    std::vector<JobDesc> descs;
    for (auto& js : b.jobs) descs.push_back(js->desc);
    b.jobtab = upload_vec(ctx, descs);
    size_t total = 0;
    for (const auto& r : b.rounds) {
        b.round_off.push_back(total);
        total += round_bytes(r);
    }
    std::vector<uint8_t> host(total, 0);
    for (size_t r = 0; r < b.rounds.size(); ++r) {
        size_t off[3];
        round_pack(b.rounds[r], host.data() + b.round_off[r], off);
        b.parts.push_back({off[0], off[1], off[2]});
    }
    b.table = upload_vec(ctx, host);
    b.rb.alloc(ctx, std::max<size_t>(peak, 2), nvis, depth);
    b.ev.assign(2 * Batch::kRing, nullptr);
    for (auto& e : b.ev) VLB_HIP(hipEventCreate(&e));
    b.info.jobs = njobs;
    b.info.rounds = (uint32_t)b.rounds.size();
    VLB_HIP(hipStreamSynchronize(ctx.stream));
    for (Scene* sc : b.scenes) ++sc->users;  // built: the jobs hold their scenes
}

// A sample of the batch for the autotuner: every k-th projection tile of
// every job, about `target` trace tiles in all.
RoundTable sample_round(const Batch& b, size_t target)
{
    size_t total = 0;
    for (const auto& js : b.jobs) total += (size_t)js->ntiles * js->nch;
    const size_t step = std::max<size_t>(1, total / std::max<size_t>(1, target));
    RoundTable r;
    size_t t = 0;
    for (size_t j = 0; j < b.jobs.size(); ++j) {
        const JobState& js = *b.jobs[j];
        for (int pt = 0; pt < js.ntiles; ++pt, ++t)
            if (t % step == 0 && (size_t)r.hdr.ntrace + js.nch <= b.rb.cap_tiles) round_add(r, (int)j, js, pt, 1);
    }
    return r;
}

Batch* as_batch(vlb_batch* h) { return reinterpret_cast<Batch*>(h); }

}  // namespace
}  // namespace vlb

using namespace vlb;

extern "C" {

vlb_status vlb_batch_create(vlb_ctx* h, const vlb_batch_job* jobs, uint32_t njobs,
                            const vlb_batch_options* opt, vlb_batch** out)
{
    if (!h || !out) return VLB_ERR_INVALID;
    *out = nullptr;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    Batch* b = new Batch();
    const vlb_status s = guard(c, [&] {
        VLB_HIP(hipSetDevice(c->device));
        build_batch(*c, jobs, njobs, opt, *b);
    });
    if (s != VLB_OK) {
        delete b;
        return s;
    }
    *out = reinterpret_cast<vlb_batch*>(b);
    ++c->children;
    return VLB_OK;
}

vlb_status vlb_batch_execute(vlb_batch* h)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] {
        VLB_REQUIRE(!take_injected(*b->ctx, VLB_INJECT_BATCH_EXECUTE), VLB_ERR_INJECTED,
                    "vlb_batch_execute: injected failure");
        VLB_HIP(hipSetDevice(b->ctx->device));
        b->execute();
    });
}

vlb_status vlb_batch_sync(vlb_batch* h)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] { b->sync(); });
}

vlb_status vlb_batch_use_graph(vlb_batch* h, int32_t on)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] {
        b->use_graph = on != 0;
        if (!b->use_graph) {
            b->sync();  // reads the kernel events of the graph's last replays
            b->drop_graph();
        }
    });
}

vlb_status vlb_batch_set_timing(vlb_batch* h, int32_t on)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] {
        VLB_HIP(hipSetDevice(b->ctx->device));
        b->set_timing(on != 0);
    });
}

static vlb_status batch_download(vlb_batch* h, uint32_t job, void* out, size_t out_len, bool f64)
{
    if (!h || !out) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] {
        VLB_REQUIRE(job < b->jobs.size(), VLB_ERR_NOT_FOUND, "no such job in the batch");
        const JobState& js = *b->jobs[job];
        VLB_REQUIRE(!f64 || b->keep64, VLB_ERR_INVALID, "batch was created without keep_f64");
        const size_t n = (size_t)js.desc.P * js.K * 3;
        VLB_REQUIRE(out_len >= n, VLB_ERR_INVALID, "output buffer too small");
        b->sync();
        const DevBuf& src = f64 ? js.out64 : js.out;
        VLB_HIP(hipMemcpy(out, src.p, n * (f64 ? sizeof(double) : sizeof(float)), hipMemcpyDeviceToHost));
    });
}

vlb_status vlb_batch_download(vlb_batch* h, uint32_t job, float* out, size_t out_len)
{
    return batch_download(h, job, out, out_len, false);
}

vlb_status vlb_batch_download_f64(vlb_batch* h, uint32_t job, double* out, size_t out_len)
{
    return batch_download(h, job, out, out_len, true);
}

vlb_status vlb_batch_autotune(vlb_batch* h, uint32_t iters, uint32_t* variant, uint32_t* chunk)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    Context& c = *b->ctx;
    std::lock_guard<std::recursive_mutex> lk(c.mtx);
    return guard(&c, [&] {
        VLB_HIP(hipSetDevice(c.device));
        b->sync();
        hipStream_t st = c.stream;
        iters = std::max(1u, std::min(iters, 20u));
        // the sample: ~16M primary rays spread over every job
        const RoundTable smp = sample_round(*b, (size_t)1 << 15);
        std::vector<uint8_t> host(round_bytes(smp), 0);
        size_t off[3];
        round_pack(smp, host.data(), off);
        DevBuf tab = upload_vec(c, host);
        hipEvent_t e0, e1;
        VLB_HIP(hipEventCreate(&e0));
        VLB_HIP(hipEventCreate(&e1));
        double best = 1e30;
        uint32_t best_v = b->variant;
        int best_c = b->chunk;
        // wavefront: every lane map x refill x work chunk {256, 1024}; fused
        // (one chunk size: it fetches 256 rays per workgroup): lane maps 2, 3
        // at 6 and 8 waves/SIMD, slot-order and far-first shadows
        std::vector<std::pair<uint32_t, int>> cand;
        for (uint32_t v = 13; v <= 24; ++v)
            for (int ch : {256, 1024}) cand.push_back({v, ch});
        for (uint32_t v : {7u, 8u, 11u, 12u, 31u, 32u, 35u, 36u}) cand.push_back({v, 256});
        try {
            for (const auto& vc : cand) {
                const uint32_t v = vc.first;
                const int ch = vc.second;
                int map, refill, focc;
                bool far;
                multi_variant(v, map, refill, focc, far);
                const MultiArgs a = b->rb.args(b->jobtab.as<JobDesc>(), tab.as<uint8_t>(), off, map,
                                               refill, ch, focc, far);
                launch_multi_trace(a, false, b->rb.grid, b->rb.shade_grid, st);  // warm-up
                VLB_HIP(hipEventRecord(e0, st));
                for (uint32_t i = 0; i < iters; ++i)
                    launch_multi_trace(a, false, b->rb.grid, b->rb.shade_grid, st);
                VLB_HIP(hipEventRecord(e1, st));
                VLB_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                VLB_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    best_v = v;
                    best_c = ch;
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        VLB_HIP(hipStreamSynchronize(st));
        b->variant = best_v;
        multi_variant(best_v, b->map, b->refill, b->focc, b->far);
        b->chunk = best_c;
        b->drop_graph();
        if (variant) *variant = best_v;
        if (chunk) *chunk = (uint32_t)best_c;
    });
}

vlb_status vlb_batch_count(vlb_batch* h)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] {
        VLB_HIP(hipSetDevice(b->ctx->device));
        b->sync();
        hipStream_t st = b->ctx->stream;
        VLB_HIP(hipMemsetAsync(b->rb.counters.p, 0, 4 * sizeof(unsigned long long), st));
        b->launch_all(true, st);
        unsigned long long cnt[4] = {0, 0, 0, 0};
        VLB_HIP(hipMemcpyAsync(cnt, b->rb.counters.p, sizeof cnt, hipMemcpyDeviceToHost, st));
        VLB_HIP(hipStreamSynchronize(st));
        b->info.node_visits = cnt[0];
        b->info.tri_tests = cnt[1];
        b->info.shadow_rays = cnt[2];
    });
}

vlb_status vlb_batch_info_get(vlb_batch* h, vlb_batch_info* out)
{
    if (!h || !out) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mtx);
    return guard(b->ctx, [&] {
        b->sync();
        *out = b->info;
        out->total_ms = b->runs ? b->sum_ms / b->runs : 0.0;
        out->timed_runs = b->runs;
        out->grid = (uint32_t)b->rb.grid;
        out->trace_variant = b->variant;
        out->chunk = (uint32_t)b->chunk;
        out->kernel_timed_runs = b->kruns;
        if (b->kruns) {
            out->trace_ms = b->ksum[0] / b->kruns;
            out->project_ms = b->ksum[1] / b->kruns;
            out->merge_ms = b->ksum[2] / b->kruns;
        }
        size_t bytes = b->jobtab.bytes + b->table.bytes + b->cache.bytes();
        for (const auto& js : b->jobs) bytes += js->state_bytes();
        bytes += b->rb.rad.bytes + b->rb.hits.bytes + b->rb.vis.bytes + b->rb.spill.bytes;
        out->device_bytes = bytes;
    });
}

vlb_status vlb_batch_destroy(vlb_batch* h)
{
    if (!h) return VLB_ERR_INVALID;
    Batch* b = as_batch(h);
    Context* c = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    (void)hipStreamSynchronize(c->stream);
    for (Scene* sc : b->scenes) --sc->users;
    delete b;
    --c->children;
    return VLB_OK;
}

}  // extern "C"
