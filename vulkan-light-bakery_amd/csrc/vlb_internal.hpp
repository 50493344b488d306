// vlb_internal.hpp -- host-side objects behind the C ABI: context, scene,
// bake plan, and the launch functions each .hip translation unit exports.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vlb_bake.h"
#include "device_common.hpp"

namespace vlb {

// ------------------------------------------------------------------ errors
struct Error {
    vlb_status code;
    std::string msg;
};

#define VLB_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw ::vlb::Error{e_ == hipErrorOutOfMemory ? VLB_ERR_OOM : VLB_ERR_DEVICE, \
                               std::string(#call) + ": " + hipGetErrorString(e_)};     \
    } while (0)

#define VLB_REQUIRE(cond, code, text)                           \
    do {                                                        \
        if (!(cond)) throw ::vlb::Error{(code), (text)};        \
    } while (0)

// --------------------------------------------------------- device buffers
// Device allocation owned by RAII; `acct` is the owning context's byte
// counter (HBM budget admission).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    std::atomic<uint64_t>* acct = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), acct(o.acct)
    {
        o.p = nullptr;
        o.bytes = 0;
        o.acct = nullptr;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            acct = o.acct;
            o.p = nullptr;
            o.bytes = 0;
            o.acct = nullptr;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            if (acct) acct->fetch_sub(bytes);
        }
        p = nullptr;
        bytes = 0;
        acct = nullptr;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Context;
DevBuf dev_alloc(Context& ctx, size_t bytes);

// ------------------------------------------------------------------ scene
struct Scene {
    uint32_t id = 0;
    int ntris = 0;
    int nnodes = 0;           // BVH2 inner nodes (build product, exported)
    int nnodes4 = 0;          // BVH4 nodes (traversed)
    DevBuf nodes, nodes4, qnodes4, tris, shade;
    float maxabs = 0.0f;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    uint32_t depth = 0;
    double build_ms = 0.0;
    // plans, batch jobs and server requests that read this scene's device
    // buffers; vlb_scene_release refuses while it is non-zero.  Raised under
    // the context's mutex (so a release cannot slip in between the lookup
    // and the raise), lowered from any thread.
    std::atomic<int> users{0};
    size_t bytes() const { return nodes.bytes + nodes4.bytes + qnodes4.bytes + tris.bytes + shade.bytes; }
    SceneView view() const
    {
        SceneView v;
        v.nodes = nodes4.as<BNode4>();
        v.qnodes = qnodes4.as<QNode4>();
        v.tris = tris.as<float4>();
        v.shade = shade.as<float4>();
        v.ntris = ntris;
        v.maxabs = maxabs;
        return v;
    }
};

// ---------------------------------------------------------------- context
struct Context {
    int device = 0;
    int num_cus = 0;
    std::string name;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;                  // bytes
    std::atomic<uint64_t> in_use{0};      // bytes allocated through dev_alloc
    // message of the last failure on this context, from any thread (the
    // serving entry points may fail concurrently): written and copied out
    // under err_mu only (set_error / get_error)
    mutable std::mutex err_mu;
    std::string last_error;
    std::map<uint32_t, std::unique_ptr<Scene>> scenes;
    uint32_t next_scene = 1;
    // skybox state: events, cached solid-angle table and partial buffer
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int wtab_n = 0;
    DevBuf wtab;
    DevBuf sky_tab;  // per-texel (direction, solid angle) for small faces, else empty
    DevBuf sky_partial;
    double sky_project_ms = 0.0, sky_merge_ms = 0.0;
    bool sky_pending = false;
    // user allocations (vlb_device_alloc)
    std::map<void*, DevBuf> user_bufs;
    // autotuner decisions: (scene id, rays per probe) -> trace_variant
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> tuned;
    // live plans, batches and servers: they use the stream and the memory
    // accounting, so the context outlives them (guarded by mtx)
    int children = 0;
    std::recursive_mutex mtx;
    // failures injected by vlb_debug_inject: calls left to fail per site
    std::atomic<int> inject[VLB_INJECT_SITES] = {};
    // this device's address of the allocation-free skybox scratch
    void* sky_scratch = nullptr;
};

// Takes one injected failure of `site` if any is armed: the caller then
// fails with VLB_ERR_INJECTED before it launches or allocates anything.
inline bool take_injected(Context& ctx, int site)
{
    int n = ctx.inject[site].load();
    while (n > 0)
        if (ctx.inject[site].compare_exchange_weak(n, n - 1)) return true;
    return false;
}

// ------------------------------------------------------- shared tables
// Direction set of a bake in traversal order (Hilbert curve over the
// octahedral map) with its weighted SH basis table: [Rpad/8][16][8] FP32, or
// 16-bit [hi][lo] halves for the MFMA forms.  Rpad = nslices * RS.
struct RayTable {
    uint32_t R = 0, Rpad = 0, RS = 0;
    int nslices = 0;
    int K = 0, precision = 0;
    double wref = 0.0;        // 4 pi / R
    DevBuf dirs;              // float4[Rpad]: xyz, w = 1 valid / 0 padding
    DevBuf dir_perm;          // int[Rpad]: sorted ray -> canonical ray index
    DevBuf ytab;
};

// Probes in Morton order, padded to a multiple of 16 with copies of the last.
struct ProbeSet {
    int P = 0, Ppad = 0;
    DevBuf probes;            // float4[Ppad]
    DevBuf probe_perm;        // int[Ppad]: sorted -> original index, -1 padding
};

template <class T>
DevBuf upload_vec(Context& ctx, const std::vector<T>& v)
{
    DevBuf b = dev_alloc(ctx, v.size() * sizeof(T));
    VLB_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx.stream));
    return b;
}

inline void set_error(Context* ctx, const std::string& msg)
{
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    ctx->last_error = msg;
}

inline std::string get_error(const Context* ctx)
{
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    return ctx->last_error;
}

// Runs f, turning exceptions into a status code and the context's last error.
template <class F>
inline vlb_status guard(Context* ctx, F&& f)
{
    try {
        f();
        return VLB_OK;
    } catch (const Error& e) {
        if (ctx) set_error(ctx, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) set_error(ctx, "host allocation failed");
        return VLB_ERR_OOM;
    } catch (const std::exception& e) {
        if (ctx) set_error(ctx, e.what());
        return VLB_ERR_INVALID;
    } catch (...) {
        if (ctx) set_error(ctx, "unknown error");
        return VLB_ERR_INVALID;
    }
}

// Scratch of vlb_skybox_sh_noalloc (sh_image.hip): solid-angle table,
// direction table and partial sums of one call, in a device array each GPU
// holds until the process ends.
constexpr size_t kSkyScratchBytes = VLB_SKYBOX_NOALLOC_SCRATCH;

// tables.cpp
Scene& find_scene(Context& ctx, uint32_t id);
void validate(const vlb_bake_params& p);
uint32_t rays_of(const vlb_bake_params& p);
uint32_t slice_rays(const vlb_bake_params& p, uint32_t R, int npg, int pgrid);
void build_ray_table(Context& ctx, const vlb_bake_params& p, int K, uint32_t RS, RayTable& t);
void build_probe_set(Context& ctx, const float* xyz, uint32_t P, ProbeSet& s);
ShadeParams shade_params(const vlb_bake_params& p);
size_t scene_spill_bytes(const Scene& s, uint32_t stride);
std::vector<uint32_t> morton_order(const float* xyz, size_t n);
std::vector<uint32_t> direction_order(const float* xyz, size_t n);

// ------------------------------------------------------------ bake plan
struct BakeArgs {
    SceneView sc;
    ShadeParams sp;
    const float4* probes;     // Ppad, w unused
    int P, Ppad;
    const float4* dirs;       // Rpad: xyz, w = 1 valid / 0 padding
    const void* ytab;         // [Rpad/8][16][8] fp32 | bf16 | fp16 (+ lo table)
    int R, Rpad, RS, nslices, npg, ntiles;
    int nchunk;               // 32-ray chunks per probe group (Rpad / 32)
    int ntrace;               // trace tiles = npg * nchunk
    float* rad;               // [ntrace][3][64][8] radiance, MFMA A order
    int map;                  // lane -> ray map of k_trace / k_rays (0..3)
    // wavefront pipeline (wave.hip)
    int4* hits;               // [slot] prim, t, u, v (bits); -1 miss, -2 no ray
    uint8_t* vis;             // [slot][nvis] shadow-ray occlusion
    int nvis;                 // shadow rays per primary ray (lights)
    int refill;               // idle lanes that trigger a refill (1..64)
    unsigned long long* wctr; // 32 work counters, kCtrStride apart
    float tmin, tmax;
    float2* partials;         // [nslices][Ppad][3][16] (sum, compensation)
    uint32_t* spill;
    uint32_t spill_stride;    // grid * 256
    // record / count mode
    float* rad_out;           // [P][R][3] canonical order (probe_perm, dir_perm)
    int* hit_out;             // [P][R]
    const int* probe_perm;    // sorted probe -> original index (-1 padding)
    const int* dir_perm;      // sorted ray -> canonical index
    unsigned long long* counters;  // node visits, tri tests, shadow rays
};

struct MergeArgs {
    const float2* partials;
    int nslices, P, Ppad, K;
    const int* probe_perm;
    double scale;
    float* out;               // [P][K][3]
    double* out64;            // [P][K][3]
};

struct Plan {
    Context* ctx = nullptr;
    Scene* scene = nullptr;
    vlb_bake_params params{};
    std::vector<vlb_light> lights;
    int K = 9;
    int precision = VLB_PREC_FP32;
    BakeArgs args{};
    MergeArgs margs{};
    int grid = 0;             // trace grid (persistent)
    int occ = 0;              // k_trace waves-per-SIMD variant (0, 6, 8)
    uint32_t variant = 0;     // trace_variant in use
    bool wave = false;        // wavefront pipeline (variants 13..24)
    bool far = false;         // far-first shadow queries (variants 25..36)
    int wgrid = 0;            // k_rays grid
    DevBuf hits, vis, wctr;
    int pgrid = 0;            // projection grid
    RayTable rt;              // sorted directions + basis table
    ProbeSet ps;              // probes in Morton order
    DevBuf rad, partials, spill, out, out64, counters;
    bool timing = false;
    bool use_graph = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    // timing: kEv events per execute, kTimingRing executes deep
    static constexpr int kTimingRing = 512;
    static constexpr int kEv = 4;
    std::vector<hipEvent_t> ev;   // kEv * kTimingRing, created on set_timing
    uint32_t timed_pending = 0;   // executes recorded but not yet read
    uint32_t timed_done = 0;
    double sum_trace = 0.0, sum_project = 0.0, sum_merge = 0.0, sum_total = 0.0;
    bool ran = false;
    vlb_bake_stats stats{};
    // multi-bounce: source grid of the indirect term (a copy of the last
    // result, refreshed at the end of every execute)
    DevBuf ind;
    bool indirect = false;
    // trace_variant left to the library (0 at creation, no autotuner pick):
    // the default follows the indirect term on and off
    bool default_variant = false;
    ~Plan();
};

// ------------------------------------------------------- multi-job work
// One bake job of a batch (or one request of the server) as the multi-job
// kernels see it.  Probe groups and slices tile the job exactly as in a
// single bake: projection tile pt = pg * nslices + sl covers probes
// 16pg..16pg+15 and rays sl*RS..sl*RS+RS-1 (nchunk 32-ray chunks per group).
struct JobDesc {
    SceneView sc;
    ShadeParams sp;
    const float4* probes;     // Ppad, Morton order
    const int* probe_perm;    // sorted -> original probe (-1 padding)
    const float4* dirs;       // Rpad, Hilbert order
    const void* ytab;         // basis table of the job's ray table
    float2* partials;         // [nslices][Ppad][3][16]
    float* out;               // [P][K][3]
    double* out64;            // [P][K][3] or null
    double scale;             // 4 pi / R (x16 for the FP16 forms)
    int P, Ppad, R, Rpad, RS, nslices, nchunk, K;
    float tmin, tmax;
    int nvis;                 // shadow rays per primary ray
    int pad_;
};

// A run of projection tiles pt0 .. pt0+npt-1 of one job inside a round.  Its
// trace tiles are tb .. tb + npt*(RS/32) - 1 of the round (tb even, so a
// pair of trace tiles never straddles two runs), its projection tiles
// pb .. pb+npt-1.
struct Segment {
    int job, pt0, npt, tb, pb;
    int pad_[3];
};
// Jobs merged by a round: merge threads base .. base + Ppad*48 - 1 (base a
// multiple of 256, so a workgroup never straddles two jobs).
struct MergeItem {
    int job, base;
};
// Counts of a round, read by the kernels (device memory, so one captured
// graph serves rounds of any size up to the buffers' capacity).
struct WorkHeader {
    int nseg, ntrace, nproj, nmerge, merge_threads;
    int nvis;                 // shadow rays per primary ray in this round (<= MultiArgs::nvis)
    int pad_[2];
};

struct MultiArgs {
    const JobDesc* jobs;
    const WorkHeader* hdr;
    const Segment* seg;
    const MergeItem* merge;
    float* rad;               // [ntrace][3][512] radiance, MFMA A order
    int4* hits;               // [ntrace*512]
    uint8_t* vis;             // [ntrace*512][nvis]
    int nvis;                 // max shadow rays per primary ray over the jobs
    int map;                  // lane -> ray map (0..3)
    int refill;               // idle lanes that trigger a refill
    int chunk;                // items per work fetch (64..1024, divides 1024)
    unsigned long long* wctr; // 32 work counters, kCtrStride apart
    uint32_t* spill;
    uint32_t spill_stride;
    unsigned long long* counters;  // node visits, tri tests, shadow rays (count mode)
    int focc;                 // -1: wavefront kernels; else the fused kernel at this
                              // occupancy (0 = compiler's choice, 6, 8 waves/SIMD)
    int fgrid;                // grid of the fused kernel
    int far;                  // fused kernel: far-first shadow queries
};

// ------------------------------------------------------ probe-grid lookup
// Validates a probe grid and lookup parameters; returns the probe count.
uint64_t check_grid(const vlb_probe_grid_desc& g, const vlb_lookup_params& lp);
ProbeGridView grid_view(const vlb_probe_grid_desc& g, const vlb_lookup_params& lp,
                        const float* dcoeffs, float scale);

struct LookupArgs {
    SceneView sc;
    ProbeGridView g;
    const float* pts;         // n x 3
    const float* nrm;         // n x 3
    int n;
    float* out;               // n x 3
    uint32_t* spill;
    uint32_t spill_stride;
};

// ----------------------------------------------------------- launchers
// bvh.hip
void build_scene_bvh(Context& ctx, Scene& s, const float* xyz, uint32_t nverts,
                     const uint32_t* idx, uint32_t ntris, const float* normals,
                     const float* albedo);
// bake.hip
int  trace_grid_size(Context& ctx, int occ, bool far);
void launch_trace(const BakeArgs& a, bool record, bool count, int occ, bool far, int grid,
                  hipStream_t stream);
void launch_project(const BakeArgs& a, int precision, int grid, hipStream_t stream);
void launch_copy_f32(float* dst, const float* src, size_t n, hipStream_t stream);
// wave.hip
int  wave_grid_size(Context& ctx);
void launch_wave(const BakeArgs& a, bool record, bool count, int grid, hipStream_t stream);
void launch_merge(const MergeArgs& m, hipStream_t stream);
void launch_trace_closest(const SceneView& sc, const float* org, const float* dir, int n,
                          float tmin, float tmax, int* prim, float* t, float* u, float* v,
                          uint32_t* spill, uint32_t spill_stride, int grid, hipStream_t st);
void launch_trace_occluded(const SceneView& sc, const float* org, const float* dir,
                           const float* tmax, int n, float tmin, uint8_t* occ,
                           uint32_t* spill, uint32_t spill_stride, int grid, hipStream_t st);
// multi.hip
int  multi_grid_size(Context& ctx);
int  multi_fused_grid_size(Context& ctx, int occ, bool far);
void launch_multi_trace(const MultiArgs& a, bool count, int grid, int shade_grid, hipStream_t st);
void launch_multi_project(const MultiArgs& a, int precision, int grid, hipStream_t st);
void launch_multi_merge(const MultiArgs& a, int grid, hipStream_t st);
// lookup.hip
int  lookup_grid_size(Context& ctx);
void launch_lookup(const LookupArgs& a, int grid, hipStream_t st);
// sh_image.hip
void launch_skybox_table(int N, const float* wtab, float4* tab, hipStream_t st);
void launch_skybox_sh(const float4* cubes, int N, int batch, int K, const float* wtab,
                      const float4* tab,
                      float* partial, int blocks_per_cube, float* out, hipStream_t st,
                      hipEvent_t e0, hipEvent_t e1, hipEvent_t e2);
void* skybox_scratch_init(int device);
void launch_skybox_sh_noalloc(const float4* cubes, int N, int batch, int K, void* scratch,
                              size_t tab_off, size_t part_off, bool use_tab, int blocks_per_cube,
                              float* out, hipStream_t st);
void launch_equirect_project(const uint8_t* img, int W, int H, int variant, double* out,
                             float* partial, int nblocks, hipStream_t st);
void launch_sh_reconstruct(const float* coeffs, int W, int H, uint8_t* out, hipStream_t st);
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
void launch_read(const void* src, size_t bytes, int grid, hipStream_t st);

// host_util.cpp
void fibonacci_dirs(uint32_t R, float* xyz);
void equirect_dirs(uint32_t W, uint32_t H, float* xyz, float* weight);
void sh_eval_double(int K, bool exact, const double d[3], double* out);
void geometric_normal(const float* v0, const float* v1, const float* v2, float g[3]);

}  // namespace vlb
