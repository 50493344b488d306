// server.cpp -- rebake requests served by continuous batching.
//
// The reference rebakes nothing: LightBaker::bake (src/baker/light_baker.cpp:
// 287-328) bakes every probe of one scene once, synchronously.  The server
// keeps scenes resident and takes bake requests (e.g. "a light moved:
// rebake this region") at any time.  A scheduler thread runs the multi-job
// kernels (multi.hip) in steps of at most step_rays primary rays:
//  * admission: queued requests are admitted in arrival order while a
//    request slot is free and their probe state (probes, partial sums,
//    outputs) fits the device state arena (HBM budget) and their result fits
//    the pinned result arena; nothing is allocated or freed with hipMalloc /
//    hipFree while serving, so no step waits for the device to go idle;
//  * continuous batching: every step is filled with the next projection tiles
//    of the admitted requests -- least work left first (SRPT, the default;
//    with an optional age bound past which a request goes first),
//    equal shares, or arrival order -- so a request that arrives while others
//    are on the GPU starts at the next step and its work overlaps theirs;
//  * two steps in flight: while the device runs step n, the host admits and
//    stages step n+1 (its tables go through pinned staging buffers), and
//    retires step n-1 (finished requests release their state);
//  * failure handling: a request can be cancelled (dropped from the queue, or
//    given no further steps once admitted), waited for with a deadline, and a
//    step that fails before launch (an injected failure) fails only the
//    requests whose work it held.  A request holds a use of its scene from
//    submit until it is done, so the scene cannot be released under it.
// One hipGraph, captured once, runs every step: the step's counts live in
// device memory.  A request's coefficients are a function of the request
// alone (fixed tile -> partial-sum slot map, fixed merge order), so they
// are the same alone and in any mix.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

#include "multi_host.hpp"

namespace vlb {
namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator of aligned byte ranges in a fixed arena.
class Arena {
public:
    void init(size_t bytes)
    {
        free_.clear();
        if (bytes) free_[0] = bytes;
        size_ = bytes;
        used_ = 0;
        peak_ = 0;
    }
    bool alloc(size_t bytes, size_t& off)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(std::max<size_t>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0, peak_ = 0;
};

enum ReqState { QUEUED, ACTIVE, DONE };

struct Request {
    uint64_t id = 0;
    uint32_t scene = 0;
    Scene* sc = nullptr;                  // held scene (users raised) until done
    vlb_bake_params params{};
    std::vector<float> probes;
    std::vector<vlb_light> lights;
    ReqState state = QUEUED;
    vlb_status status = VLB_OK;
    double submit_ms = 0.0, admit_ms = 0.0, start_ms = -1.0, done_ms = 0.0;
    uint32_t steps = 0;
    // admitted state
    int slot = -1;
    size_t dev_off = 0, dev_bytes = 0;   // device state arena block
    size_t res_off = 0, res_bytes = 0;   // pinned result arena block
    JobState js;                          // desc points into the arena block
    int next_tile = 0;                    // next projection tile to schedule
    // failure handling
    bool cancel = false;                  // vlb_cancel (or destroy) asked for it
    bool abandoned = false;               // admitted, then dropped: no more steps
    vlb_status fail = VLB_OK;             // why it was abandoned
    int inflight = 0;                     // launched steps holding its work
};

struct InFlight {
    bool used = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<Request*> finishing;      // requests whose merge is in the step
    std::vector<Request*> members;        // every request with work in the step
};

struct Upload {
    size_t src;                           // offset in the staging buffer
    void* dst;
    size_t bytes;
};

struct Server {
    Context* ctx = nullptr;
    vlb_server_options opt{};
    hipStream_t stream = nullptr;
    std::thread thr;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    bool stop = false;
    bool hold = false;                    // vlb_server_hold: no admission
    std::deque<Request*> queue;
    std::vector<Request*> active;         // admission order
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    std::chrono::steady_clock::time_point t0;
    // resources
    Arena dev_arena, res_arena;
    DevBuf dev_state;                     // state arena
    uint8_t* res_host = nullptr;          // pinned result arena
    uint8_t* staging[2] = {nullptr, nullptr};
    size_t staging_bytes = 0, upload_cap = 0;
    DevBuf dtab;                          // [job table][round table]
    size_t jobtab_bytes = 0, round_cap_bytes = 0;
    std::vector<JobDesc> jobtab;          // host mirror
    std::vector<int> free_slots;
    RayCache cache;
    RoundBufs rb;
    int map = 3, refill = 32, chunk = 256, focc = -1;
    bool far = false;
    size_t cap_tiles = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    InFlight fl[2];
    uint64_t step_no = 0;
    // statistics
    uint64_t submitted = 0, completed = 0, steps = 0, rays = 0, probes = 0;
    double busy_ms = 0.0;
    uint32_t active_peak = 0;
    std::string error;
    vlb_status err_status = VLB_OK;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    bool admit_one(Request* r, size_t& upload_used, std::vector<Upload>& ups, uint8_t* stg);
    void retire(InFlight& f);
    void launch_step(const RoundTable& rt, const std::vector<Upload>& ups, uint8_t* stg,
                     InFlight& f, const std::vector<Request*>& finishing,
                     const std::vector<Request*>& members);
    void fail_all(vlb_status st, const std::string& msg);
    // with mu held:
    void finish(Request* r, vlb_status st);
    void abandon(Request* r, vlb_status st);
    void drop_cancelled();
};

// Ends a request (with mu held): releases its device state, slot and scene
// use, and, unless it succeeded, its result block; wakes its waiters.
void Server::finish(Request* r, vlb_status st)
{
    if (r->slot >= 0) {
        dev_arena.release(r->dev_off, r->dev_bytes);
        free_slots.push_back(r->slot);
        r->slot = -1;
    }
    if (st != VLB_OK && r->res_bytes) {
        res_arena.release(r->res_off, r->res_bytes);
        r->res_bytes = 0;
    }
    r->state = DONE;
    r->status = st;
    r->done_ms = now_ms();
    r->js.rt.reset();
    if (r->sc) {
        --r->sc->users;
        r->sc = nullptr;
    }
    if (st == VLB_OK) {
        ++completed;
        rays += r->js.rays;
        probes += r->params.num_probes;
    }
    cv_done.notify_all();
}

// An admitted request gets no further steps (with mu held); it ends once
// the steps already launched with its work have retired.
void Server::abandon(Request* r, vlb_status st)
{
    if (r->abandoned || r->state == DONE) return;
    auto it = std::find(active.begin(), active.end(), r);
    if (it != active.end()) active.erase(it);
    r->abandoned = true;
    r->fail = st;
    if (r->inflight == 0) finish(r, st);
}

// Cancellations asked for since the last step (with mu held).
void Server::drop_cancelled()
{
    if (stop) {
        while (!queue.empty()) {
            Request* r = queue.front();
            queue.pop_front();
            finish(r, VLB_ERR_CANCELLED);
        }
        for (Request* r : active) r->cancel = true;
    }
    std::vector<Request*> drop;
    for (Request* r : active)
        if (r->cancel) drop.push_back(r);
    for (Request* r : drop) abandon(r, VLB_ERR_CANCELLED);
}

// Device block of one request: probes | perm | partials | out.
struct BlockLayout {
    size_t probes, perm, partials, out, total;
};
BlockLayout block_layout(size_t Ppad, size_t nslices, size_t P, int K)
{
    BlockLayout L;
    L.probes = 0;
    L.perm = align_up(Ppad * sizeof(float4));
    L.partials = L.perm + align_up(Ppad * sizeof(int));
    L.out = L.partials + align_up(nslices * Ppad * 48 * sizeof(float2));
    L.total = L.out + align_up(P * (size_t)K * 3 * sizeof(float));
    return L;
}

bool Server::admit_one(Request* r, size_t& upload_used, std::vector<Upload>& ups, uint8_t* stg)
{
    const vlb_bake_params& p = r->params;
    const uint32_t R = rays_of(p);
    const uint32_t RS = job_slice_rays(p, R);
    const size_t nsl = (R + RS - 1) / RS;
    const size_t P = p.num_probes, Ppad = (P + 15) / 16 * 16;
    const int K = p.sh_order == 2 ? 9 : 16;
    const BlockLayout L = block_layout(Ppad, nsl, P, K);
    const size_t up = align_up(Ppad * sizeof(float4)) + align_up(Ppad * sizeof(int));
    if (free_slots.empty() || upload_used + up > upload_cap) return false;
    JobState& js = r->js;
    {
        // scene lookup and (first use of a direction set) the ray table;
        // lock order: the server's mutex, then the context's
        std::lock_guard<std::recursive_mutex> lk(ctx->mtx);
        Scene& sc = find_scene(*ctx, r->scene);
        bool built = false;
        js.rt = cache.get(*ctx, p, K, RS, &built);
        if (built) VLB_HIP(hipStreamSynchronize(ctx->stream));
        js.desc.sc = sc.view();
    }
    size_t doff = 0, roff = 0;
    if (!dev_arena.alloc(L.total, doff)) return false;
    const size_t rbytes = P * (size_t)K * 3 * sizeof(float);
    if (!res_arena.alloc(rbytes, roff)) {
        dev_arena.release(doff, L.total);
        return false;
    }
    r->dev_off = doff;
    r->dev_bytes = L.total;
    r->res_off = roff;
    r->res_bytes = rbytes;
    r->slot = free_slots.back();
    free_slots.pop_back();

    const RayTable& rt = *js.rt;
    js.scene_id = r->scene;
    js.K = K;
    js.ps.P = (int)P;
    js.ps.Ppad = (int)Ppad;
    js.npg = (int)(Ppad / 16);
    js.ntiles = js.npg * rt.nslices;
    js.nch = (int)(rt.RS / 32);
    js.rays = P * R;
    // probes in Morton order, padded with the last one, staged for upload
    const std::vector<uint32_t> order = morton_order(r->probes.data(), P);
    float4* p4 = reinterpret_cast<float4*>(stg + upload_used);
    int* perm = reinterpret_cast<int*>(stg + upload_used + align_up(Ppad * sizeof(float4)));
    for (size_t i = 0; i < Ppad; ++i) {
        const size_t src = order[std::min(i, P - 1)];
        const float* q = &r->probes[3 * src];
        p4[i] = make_float4(q[0], q[1], q[2], 1.0f);
        perm[i] = i < P ? (int)src : -1;
    }
    uint8_t* base = dev_state.as<uint8_t>() + doff;
    ups.push_back({upload_used, base + L.probes, Ppad * sizeof(float4)});
    ups.push_back({upload_used + align_up(Ppad * sizeof(float4)), base + L.perm, Ppad * sizeof(int)});
    upload_used += up;

    JobDesc& d = js.desc;
    d.sp = shade_params(p);
    d.probes = reinterpret_cast<const float4*>(base + L.probes);
    d.probe_perm = reinterpret_cast<const int*>(base + L.perm);
    d.dirs = rt.dirs.as<float4>();
    d.ytab = rt.ytab.p;
    d.partials = reinterpret_cast<float2*>(base + L.partials);
    d.out = reinterpret_cast<float*>(base + L.out);
    d.out64 = nullptr;
    const bool fp16 = p.precision == VLB_PREC_FP16 || p.precision == VLB_PREC_FP16_FAST;
    d.scale = fp16 ? rt.wref * 16.0 : rt.wref;
    d.P = (int)P;
    d.Ppad = (int)Ppad;
    d.R = (int)R;
    d.Rpad = (int)rt.Rpad;
    d.RS = (int)rt.RS;
    d.nslices = rt.nslices;
    d.nchunk = (int)(rt.Rpad / 32);
    d.K = K;
    d.tmin = p.t_min;
    d.tmax = p.t_max;
    d.nvis = p.shade_mode == VLB_SHADE_REFERENCE ? 1 : (int)p.num_lights;
    jobtab[r->slot] = d;
    r->state = ACTIVE;
    r->admit_ms = now_ms();
    r->next_tile = 0;
    return true;
}

void Server::retire(InFlight& f)
{
    VLB_HIP(hipEventSynchronize(f.e1));
    float ms = 0.f;
    VLB_HIP(hipEventElapsedTime(&ms, f.e0, f.e1));
    std::lock_guard<std::mutex> lk(mu);
    busy_ms += ms;
    for (Request* r : f.members) --r->inflight;
    // a request cancelled after its last step was launched does not finish
    for (Request* r : f.finishing) finish(r, r->cancel ? VLB_ERR_CANCELLED : VLB_OK);
    for (Request* r : f.members)
        if (r->abandoned && r->inflight == 0 && r->state != DONE) finish(r, r->fail);
    f.finishing.clear();
    f.members.clear();
    f.used = false;
    cv_done.notify_all();
}

void Server::launch_step(const RoundTable& rt, const std::vector<Upload>& ups, uint8_t* stg,
                         InFlight& f, const std::vector<Request*>& finishing,
                         const std::vector<Request*>& members)
{
    hipStream_t st = stream;
    memcpy(stg, jobtab.data(), jobtab_bytes);
    size_t off[3];
    round_pack_cap(rt, stg + jobtab_bytes, (size_t)opt.max_active, off);
    VLB_HIP(hipEventRecord(f.e0, st));
    VLB_HIP(hipMemcpyAsync(dtab.p, stg, jobtab_bytes + round_cap_bytes, hipMemcpyHostToDevice, st));
    for (const Upload& u : ups)
        VLB_HIP(hipMemcpyAsync(u.dst, stg + u.src, u.bytes, hipMemcpyHostToDevice, st));
    if (!gexec) {
        const MultiArgs a = rb.args(dtab.as<JobDesc>(), dtab.as<uint8_t>() + jobtab_bytes, off, map, refill, chunk,
                                    focc, far);
        VLB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            launch_multi_trace(a, false, rb.grid, rb.shade_grid, st);
            launch_multi_project(a, opt.precision, rb.pgrid, st);
            launch_multi_merge(a, rb.mgrid, st);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        VLB_HIP(hipStreamEndCapture(st, &graph));
        VLB_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    VLB_HIP(hipGraphLaunch(gexec, st));
    for (Request* r : finishing)
        VLB_HIP(hipMemcpyAsync(res_host + r->res_off, r->js.desc.out, r->res_bytes, hipMemcpyDeviceToHost, st));
    VLB_HIP(hipEventRecord(f.e1, st));
    f.finishing = finishing;
    f.members = members;
    f.used = true;
    ++steps;
    std::lock_guard<std::mutex> lk(mu);
    for (Request* r : members) ++r->inflight;
}

void Server::fail_all(vlb_status st, const std::string& msg)
{
    // work already on the device may still read the requests' scenes: let it
    // end before their uses are given back
    (void)hipStreamSynchronize(stream);
    (void)hipGetLastError();
    std::lock_guard<std::mutex> lk(mu);
    err_status = st;
    error = msg;
    for (auto& kv : reqs) {
        Request* r = kv.second.get();
        if (r->state != DONE) finish(r, st);
    }
    queue.clear();
    active.clear();
    cv_done.notify_all();
}

void Server::run()
{
    try {
        VLB_HIP(hipSetDevice(ctx->device));
        for (;;) {
            // steps in flight: fl[(step_no - 1) & 1] (newest), fl[step_no & 1]
            InFlight& next = fl[step_no & 1];
            InFlight& last = fl[(step_no + 1) & 1];
            bool work;
            {
                std::unique_lock<std::mutex> lk(mu);
                drop_cancelled();
                work = !hold && !stop && !queue.empty();
                for (Request* r : active) work = work || r->next_tile < r->js.ntiles;
                if (!work && !next.used && !last.used) {
                    if (stop) break;
                    cv_sched.wait(lk);
                    continue;
                }
            }
            if (next.used) {  // at most two steps in flight
                retire(next);
                continue;
            }
            if (!work) {      // drain
                retire(last);
                continue;
            }
            // ---- admit and build the next step
            uint8_t* stg = staging[step_no & 1];
            const size_t upl0 = jobtab_bytes + round_cap_bytes;
            std::vector<Upload> ups;
            RoundTable rt;
            std::vector<Request*> finishing, members, admitted;
            {
                std::lock_guard<std::mutex> lk(mu);
                size_t used = 0;
                while (!hold && !stop && !queue.empty()) {
                    Request* r = queue.front();
                    bool ok = false;
                    try {
                        ok = admit_one(r, used, ups, stg + upl0);
                    } catch (const Error& e) {
                        // this request only (e.g. its ray table did not fit)
                        queue.pop_front();
                        finish(r, e.code);
                        continue;
                    }
                    if (!ok) break;
                    queue.pop_front();
                    active.push_back(r);
                    admitted.push_back(r);
                }
                for (Upload& u : ups) u.src += upl0;
                active_peak = std::max(active_peak, (uint32_t)active.size());
                // the step's trace tiles: SRPT (least work left first), FIFO
                // (arrival order) or SHARE (equal shares, the rest to the
                // oldest)
                std::vector<Request*> runnable;
                for (Request* r : active)
                    if (r->next_tile < r->js.ntiles) runnable.push_back(r);
                auto left_tiles = [](const Request* r) {
                    return (size_t)(r->js.ntiles - r->next_tile) * (size_t)r->js.nch;
                };
                if (opt.policy == VLB_SCHED_SRPT) {
                    // requests past the age bound first, oldest first; then
                    // least work left
                    const double old_before = opt.age_ms ? now_ms() - (double)opt.age_ms : -1e300;
                    std::stable_sort(runnable.begin(), runnable.end(), [&](const Request* a, const Request* b) {
                        const bool oa = a->submit_ms < old_before, ob = b->submit_ms < old_before;
                        if (oa != ob) return oa;
                        if (oa) return a->submit_ms < b->submit_ms;
                        return left_tiles(a) < left_tiles(b);
                    });
                }
                std::vector<size_t> give(runnable.size(), 0);
                size_t room = cap_tiles;
                if (opt.policy == VLB_SCHED_SHARE && !runnable.empty()) {
                    const size_t share = std::max<size_t>(1, cap_tiles / runnable.size());
                    for (size_t i = 0; i < runnable.size(); ++i) {
                        const Request* r = runnable[i];
                        const size_t nch = (size_t)r->js.nch;
                        const size_t left = (size_t)(r->js.ntiles - r->next_tile);
                        const size_t n = std::min({left, std::max<size_t>(1, share / nch), room / nch});
                        give[i] = n;
                        room -= n * nch;
                    }
                }
                for (size_t i = 0; i < runnable.size() && room > 0; ++i) {
                    const Request* r = runnable[i];
                    const size_t nch = (size_t)r->js.nch;
                    const size_t left = (size_t)(r->js.ntiles - r->next_tile) - give[i];
                    const size_t n = std::min(left, room / nch);
                    give[i] += n;
                    room -= n * nch;
                }
                const double t = now_ms();
                for (size_t i = 0; i < runnable.size(); ++i) {
                    Request* r = runnable[i];
                    if (give[i] == 0) continue;
                    round_add(rt, r->slot, r->js, r->next_tile, (int)give[i]);
                    r->next_tile += (int)give[i];
                    if (r->start_ms < 0) r->start_ms = t;
                    ++r->steps;
                    members.push_back(r);
                    if (r->next_tile == r->js.ntiles) {
                        round_merge(rt, r->slot, r->js);
                        finishing.push_back(r);
                    }
                }
                for (Request* r : finishing) active.erase(std::find(active.begin(), active.end(), r));
                if (rt.hdr.nseg > 0 && take_injected(*ctx, VLB_INJECT_SERVER_STEP)) {
                    // the step fails before launch: so do the requests it held
                    // and those admitted with it (their probes were never
                    // uploaded)
                    for (Request* r : members) abandon(r, VLB_ERR_INJECTED);
                    for (Request* r : admitted) abandon(r, VLB_ERR_INJECTED);
                    continue;
                }
            }
            if (rt.hdr.nseg == 0) {
                // nothing admitted or runnable: the state or result arena is
                // full; retire a step, or wait for results to be collected
                if (last.used) {
                    retire(last);
                } else {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_sched.wait_for(lk, std::chrono::milliseconds(1));
                }
                continue;
            }
            launch_step(rt, ups, stg, next, finishing, members);
            ++step_no;
        }
    } catch (const Error& e) {
        fail_all(e.code, e.msg);
    } catch (const std::exception& e) {
        fail_all(VLB_ERR_DEVICE, e.what());
    }
}

Server* as_server(vlb_server* h) { return reinterpret_cast<Server*>(h); }

void destroy_server(Server* s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    if (s->thr.joinable()) s->thr.join();
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (InFlight& f : s->fl) {
        if (f.e0) (void)hipEventDestroy(f.e0);
        if (f.e1) (void)hipEventDestroy(f.e1);
    }
    for (uint8_t* p : s->staging)
        if (p) (void)hipHostFree(p);
    if (s->res_host) (void)hipHostFree(s->res_host);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // namespace
}  // namespace vlb

using namespace vlb;

extern "C" {

vlb_status vlb_server_create(vlb_ctx* h, const vlb_server_options* o, vlb_server** out)
{
    if (!h || !out) return VLB_ERR_INVALID;
    *out = nullptr;
    Context* c = reinterpret_cast<Context*>(h);
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    Server* s = new Server();
    const vlb_status st = guard(c, [&] {
        VLB_HIP(hipSetDevice(c->device));
        s->ctx = c;
        if (o) s->opt = *o;
        vlb_server_options& op = s->opt;
        if (!op.state_budget) op.state_budget = 2ull << 30;
        if (!op.step_rays) op.step_rays = 1ull << 25;
        if (!op.max_active) op.max_active = 64;
        if (!op.max_lights) op.max_lights = 4;
        // fused, lane map 2 (2 probes x 32 directions per wave), 8 waves/SIMD,
        // far-first shadow queries: on the request mix of 256-65,536 probes x
        // 4096 rays over rooms, terrains and a city with moving point lights,
        // 2.2% less device time than lane map 3 with slot-order shadows
        // (variant 12: 2,039-2,042 against 2,087-2,089 ms busy for 200
        // requests; variant 11, slot order, 2,048-2,054) and a p99 latency of
        // 132-133 against 143-144 ms
        if (!op.trace_variant) op.trace_variant = 35;
        if (!op.chunk) op.chunk = 256;
        if (!op.max_depth) op.max_depth = 64;
        if (!op.result_bytes) op.result_bytes = 256ull << 20;
        if (!op.upload_bytes) op.upload_bytes = 32ull << 20;
        VLB_REQUIRE(op.max_active <= 1024, VLB_ERR_INVALID, "max_active must be <= 1024");
        VLB_REQUIRE(op.max_lights <= VLB_MAX_LIGHTS, VLB_ERR_INVALID, "max_lights too large");
        VLB_REQUIRE(op.precision >= VLB_PREC_FP32 && op.precision <= VLB_PREC_FP16_FAST, VLB_ERR_INVALID,
                    "bad precision");
        VLB_REQUIRE(op.step_rays <= (1ull << 30), VLB_ERR_INVALID, "step_rays must be <= 2^30");
        VLB_REQUIRE(op.policy <= VLB_SCHED_FIFO, VLB_ERR_INVALID, "bad scheduling policy");
        VLB_REQUIRE(op.age_ms == 0 || op.policy == VLB_SCHED_SRPT, VLB_ERR_INVALID,
                    "age_ms applies to the SRPT policy only");
        multi_variant(op.trace_variant, s->map, s->refill, s->focc, s->far);
        s->chunk = (int)op.chunk;
        VLB_REQUIRE(s->chunk >= 64 && s->chunk <= 1024 && 1024 % s->chunk == 0, VLB_ERR_INVALID,
                    "chunk must divide 1024 and be at least 64");
        // largest projection tile: 16 probes x 16384 rays (job_slice_rays)
        s->cap_tiles = std::max<size_t>(512, op.step_rays / 512);
        VLB_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (InFlight& f : s->fl) {
            VLB_HIP(hipEventCreate(&f.e0));
            VLB_HIP(hipEventCreate(&f.e1));
        }
        s->dev_state = dev_alloc(*c, op.state_budget);
        s->dev_arena.init(op.state_budget);
        VLB_HIP(hipHostMalloc((void**)&s->res_host, op.result_bytes, 0));
        s->res_arena.init(op.result_bytes);
        s->jobtab_bytes = align_up(sizeof(JobDesc) * op.max_active);
        s->round_cap_bytes = round_bytes_cap(op.max_active, op.max_active);
        s->upload_cap = op.upload_bytes;
        s->staging_bytes = s->jobtab_bytes + s->round_cap_bytes + s->upload_cap;
        for (auto& p : s->staging) VLB_HIP(hipHostMalloc((void**)&p, s->staging_bytes, 0));
        s->dtab = dev_alloc(*c, s->jobtab_bytes + s->round_cap_bytes);
        s->jobtab.assign(op.max_active, JobDesc{});
        for (int i = (int)op.max_active - 1; i >= 0; --i) s->free_slots.push_back(i);
        s->rb.alloc(*c, s->cap_tiles, (int)op.max_lights, op.max_depth);
        s->t0 = std::chrono::steady_clock::now();
        s->thr = std::thread([s] { s->run(); });
    });
    if (st != VLB_OK) {
        destroy_server(s);
        return st;
    }
    *out = reinterpret_cast<vlb_server*>(s);
    ++c->children;
    return VLB_OK;
}

vlb_status vlb_submit(vlb_server* h, uint32_t scene, const vlb_bake_params* p, uint64_t* request)
{
    if (!h || !p) return VLB_ERR_INVALID;
    Server* s = as_server(h);
    return guard(s->ctx, [&] {
        validate(*p);
        const vlb_server_options& op = s->opt;
        VLB_REQUIRE(p->precision == op.precision, VLB_ERR_INVALID,
                    "request precision differs from the server's");
        const uint32_t nvis = p->shade_mode == VLB_SHADE_REFERENCE ? 1 : p->num_lights;
        VLB_REQUIRE(nvis <= op.max_lights, VLB_ERR_INVALID, "request has more lights than max_lights");
        {
            std::lock_guard<std::recursive_mutex> lk(s->ctx->mtx);
            const Scene& sc = find_scene(*s->ctx, scene);
            VLB_REQUIRE(sc.depth <= op.max_depth, VLB_ERR_INVALID, "scene BVH deeper than max_depth");
        }
        const uint32_t R = rays_of(*p);
        const uint32_t RS = job_slice_rays(*p, R);
        const size_t P = p->num_probes, Ppad = (P + 15) / 16 * 16;
        const int K = p->sh_order == 2 ? 9 : 16;
        VLB_REQUIRE((size_t)RS / 32 <= s->cap_tiles, VLB_ERR_INVALID, "rays per slice exceed a step");
        VLB_REQUIRE(block_layout(Ppad, (R + RS - 1) / RS, P, K).total <= op.state_budget, VLB_ERR_BUDGET,
                    "request state exceeds the server's state budget");
        VLB_REQUIRE(P * K * 3 * sizeof(float) <= op.result_bytes, VLB_ERR_BUDGET,
                    "request result exceeds the result arena");
        VLB_REQUIRE(align_up(Ppad * sizeof(float4)) + align_up(Ppad * sizeof(int)) <= s->upload_cap,
                    VLB_ERR_BUDGET, "request probes exceed the upload staging");
        auto r = std::make_unique<Request>();
        r->scene = scene;
        r->params = *p;
        r->probes.assign(p->probe_xyz, p->probe_xyz + 3 * P);
        r->lights.assign(p->lights, p->lights + p->num_lights);
        r->params.probe_xyz = r->probes.data();
        r->params.lights = r->lights.data();
        std::lock_guard<std::mutex> lk(s->mu);
        VLB_REQUIRE(s->err_status == VLB_OK, s->err_status, "server failed: " + s->error);
        VLB_REQUIRE(!s->stop, VLB_ERR_INVALID, "server is being destroyed");
        {
            // the request holds its scene until it is done; raised under the
            // context's mutex so that a release cannot slip in between
            std::lock_guard<std::recursive_mutex> lk2(s->ctx->mtx);
            Scene& sc = find_scene(*s->ctx, scene);
            ++sc.users;
            r->sc = &sc;
        }
        r->id = s->next_id++;
        r->submit_ms = s->now_ms();
        if (request) *request = r->id;
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        ++s->submitted;
        s->cv_sched.notify_one();
    });
}

vlb_status vlb_poll(vlb_server* h, uint64_t request, int32_t* done)
{
    if (!h || !done) return VLB_ERR_INVALID;
    Server* s = as_server(h);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) return VLB_ERR_NOT_FOUND;
    *done = it->second->state == DONE ? 1 : 0;
    return VLB_OK;
}

// timeout_ms < 0: no deadline
static vlb_status wait_request(vlb_server* h, uint64_t request, double timeout_ms, float* out,
                               size_t out_len, vlb_request_info* info)
{
    if (!h) return VLB_ERR_INVALID;
    Server* s = as_server(h);
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) return VLB_ERR_NOT_FOUND;
    Request* r = it->second.get();
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, [&] { return r->state == DONE; });
    } else {
        const auto deadline = std::chrono::steady_clock::now() +
                              std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                  std::chrono::duration<double, std::milli>(std::min(timeout_ms, 1e12)));
        if (!s->cv_done.wait_until(lk, deadline, [&] { return r->state == DONE; })) return VLB_ERR_TIMEOUT;
    }
    vlb_status st = r->status;
    const size_t n = (size_t)r->params.num_probes * (r->params.sh_order == 2 ? 9 : 16) * 3;
    if (st == VLB_OK && out) {
        if (out_len < n) return VLB_ERR_INVALID;  // the result stays for a retry
        memcpy(out, s->res_host + r->res_off, n * sizeof(float));
    }
    if (info) {
        info->submit_ms = r->submit_ms;
        info->admit_ms = r->admit_ms;
        info->start_ms = r->start_ms;
        info->done_ms = r->done_ms;
        info->primary_rays = (uint64_t)r->params.num_probes * rays_of(r->params);
        info->steps = r->steps;
        info->status = st;
    }
    if (r->res_bytes) s->res_arena.release(r->res_off, r->res_bytes);
    s->reqs.erase(it);
    s->cv_sched.notify_one();
    return st;
}

vlb_status vlb_wait(vlb_server* h, uint64_t request, float* out, size_t out_len, vlb_request_info* info)
{
    return wait_request(h, request, -1.0, out, out_len, info);
}

vlb_status vlb_wait_for(vlb_server* h, uint64_t request, double timeout_ms, float* out, size_t out_len,
                        vlb_request_info* info)
{
    if (!(timeout_ms >= 0.0)) return VLB_ERR_INVALID;  // negative or NaN
    return wait_request(h, request, timeout_ms, out, out_len, info);
}

vlb_status vlb_cancel(vlb_server* h, uint64_t request)
{
    if (!h) return VLB_ERR_INVALID;
    Server* s = as_server(h);
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) return VLB_ERR_NOT_FOUND;
    Request* r = it->second.get();
    if (r->state == QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r));
        s->finish(r, VLB_ERR_CANCELLED);
    } else if (r->state == ACTIVE) {
        r->cancel = true;  // the scheduler drops it before its next step
        s->cv_sched.notify_one();
    }
    return VLB_OK;
}

vlb_status vlb_server_hold(vlb_server* h, int32_t on)
{
    if (!h) return VLB_ERR_INVALID;
    Server* s = as_server(h);
    std::lock_guard<std::mutex> lk(s->mu);
    s->hold = on != 0;
    s->cv_sched.notify_one();
    return VLB_OK;
}

vlb_status vlb_server_stats_get(vlb_server* h, vlb_server_stats* out)
{
    if (!h || !out) return VLB_ERR_INVALID;
    Server* s = as_server(h);
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->steps = s->steps;
    out->primary_rays = s->rays;
    out->probes = s->probes;
    out->busy_ms = s->busy_ms;
    out->clock_ms = s->now_ms();
    out->state_bytes_peak = s->dev_arena.peak();
    out->active_peak = s->active_peak;
    out->queued = (uint32_t)s->queue.size();
    out->status = s->err_status;
    out->trace_variant = s->opt.trace_variant;
    return VLB_OK;
}

vlb_status vlb_server_destroy(vlb_server* h)
{
    if (!h) return VLB_ERR_INVALID;
    Context* c = as_server(h)->ctx;
    destroy_server(as_server(h));
    std::lock_guard<std::recursive_mutex> lk(c->mtx);
    --c->children;
    return VLB_OK;
}

}  // extern "C"
