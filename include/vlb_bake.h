/*
 * vlb_bake.h -- C ABI of the MI355X-native probe baker.
 *
 * This is the drop-in boundary for vulkan-light-bakery's probe-baking hot
 * path: a triangle scene and a probe grid go in, a per-probe spherical-
 * harmonics (SH) coefficient buffer comes out.  The reference has no FFI;
 * its seam is the C++ class vlb::LightBaker (src/baker/light_baker.hpp:16-54)
 * driving Vulkan RT shaders.  Each entry point below names the reference
 * function or shader it replaces.  The C++ facade in vlb/light_baker.hpp
 * keeps LightBaker's ctor/bake/serialize calls on top of this ABI.
 *
 * Conventions
 *  - Every function returns vlb_status (0 == VLB_OK); nothing throws across
 *    the ABI.  vlb_last_error(ctx) returns a message for the last failure on
 *    that context, from any thread; the string is the calling thread's copy,
 *    valid until that thread's next vlb_last_error call.
 *  - Host pointers unless the name says "device" / "_d".  The caller owns all
 *    buffers it passes in.
 *  - Coefficient layout everywhere: out[probe][coef][rgb], float32, with
 *    coef index l*(l+1)+m as in shaders/sh.comp:35-41.
 *  - SH frame: the real SH basis of shaders/sh_common.h:26-152 is evaluated
 *    at the world direction swizzled to (x, z, y), i.e. SH "z" is world up
 *    (+y).  This is what sh.comp does: env_map.rgen:21 traces toVector(..).xzy
 *    while sh.comp:30 evaluates SH at the un-swizzled toVector(..).
 *  - A context is thread-compatible: calls on one context are serialised by
 *    the caller (the serving entry points are the exception, they are
 *    internally synchronised).
 */
#ifndef VLB_BAKE_H
#define VLB_BAKE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VLB_ABI_VERSION 1

/* ------------------------------------------------------------------ status */
typedef int32_t vlb_status;
enum {
    VLB_OK              = 0,
    VLB_ERR_INVALID     = 1,  /* bad argument / size                          */
    VLB_ERR_DEVICE      = 2,  /* HIP runtime error                            */
    VLB_ERR_OOM         = 3,  /* device allocation failed                     */
    VLB_ERR_BUDGET      = 4,  /* request exceeds the context's HBM budget     */
    VLB_ERR_NOT_FOUND   = 5,  /* unknown scene / ticket / plan                */
    VLB_ERR_BUSY        = 6,  /* ticket not finished yet (poll)               */
    VLB_ERR_UNSUPPORTED = 7,  /* option not supported by this library         */
    VLB_ERR_TIMEOUT     = 8,  /* vlb_wait_for deadline passed                 */
    VLB_ERR_CANCELLED   = 9,  /* request cancelled (vlb_cancel, or pending at */
                              /* vlb_server_destroy) before it finished      */
    VLB_ERR_INJECTED    = 10, /* failure injected through vlb_debug_inject    */
    VLB_ERR_NO_DEVICE   = 11  /* no HIP device visible                        */
};

/* Failure-injection sites of vlb_debug_inject. */
enum {
    VLB_INJECT_PLAN_EXECUTE  = 0, /* vlb_plan_execute, vlb_bake               */
    VLB_INJECT_BATCH_EXECUTE = 1, /* vlb_batch_execute                        */
    VLB_INJECT_SERVER_STEP   = 2, /* one server step: its requests fail       */
    VLB_INJECT_SITES         = 3
};

/* ----------------------------------------------------------------- options */
/* Arithmetic of the SH projection (hit search and shading are always FP32).
 *  FP32      : v_mfma_f32_16x16x4_f32, FP32 operands.
 *  BF16/FP16 : v_mfma_f32_16x16x32_{bf16,f16}, operands split into a high
 *              and a low 16-bit part, three products (hi*hi + hi*lo + lo*hi)
 *              accumulated in FP32: ~16-bit mantissa operands.
 *  *_FAST    : one 16-bit MFMA product per chunk (8 / 11-bit mantissas);
 *              systematic error of a few 1e-3 on flat-lit scenes, kept for
 *              the precision study.
 * FP16 forms scale radiance by 1/16 (and clamp to the FP16 range) before
 * conversion; the FP64 merge undoes the scale.
 * All forms accumulate chunks with FP32 Kahan and merge slices in FP64. */
enum { VLB_PREC_FP32 = 0, VLB_PREC_BF16 = 1, VLB_PREC_FP16 = 2,
       VLB_PREC_BF16_FAST = 3, VLB_PREC_FP16_FAST = 4 };

/* Shading of a primary hit.
 *  LAMBERT   : albedo/pi * sum_lights E * max(N.L,0) * visibility (linear HDR),
 *              N flipped to face the ray, procedural sky on miss.
 *  REFERENCE : shaders/env_map.rchit:51-102 / main.rchit:95-116 -- point light,
 *              Phong diffuse + specular, optional sRGB encode and UNORM8
 *              quantisation of the stored radiance (the RGBA8 storage image
 *              of env_map_generator.cpp:39). */
enum { VLB_SHADE_LAMBERT = 0, VLB_SHADE_REFERENCE = 1 };

/* Ray direction set per probe.
 *  FIBONACCI : R directions on the Fibonacci sphere, weight 4*pi/R.
 *  EQUIRECT  : W x H pixel centres of shaders/env_map.rgen:21 (toVector(
 *              x2phi, y2theta).xzy), weight (2pi/W)(pi/H) sin(theta) of
 *              shaders/sh.comp:32-33. */
enum { VLB_DIRS_FIBONACCI = 0, VLB_DIRS_EQUIRECT = 1 };

/* SH constants: LITERAL = the 6-significant-figure values of sh_common.h,
 * EXACT = closed-form normalisation constants. */
enum { VLB_SH_LITERAL = 0, VLB_SH_EXACT = 1 };

/* Probe ordering of vlb_probe_grid. */
enum { VLB_ORDER_LINEAR = 0, VLB_ORDER_REFERENCE_BFS = 1 };

enum { VLB_LIGHT_DIRECTIONAL = 0, VLB_LIGHT_POINT = 1 };

/* Reference-shading flags. */
enum { VLB_REF_SRGB = 1u, VLB_REF_UNORM8 = 2u };

/* Equirect projection variants (vlb_project_equirect). */
enum {
    VLB_EQUIRECT_SH_COMP     = 0, /* shaders/sh.comp:23-42        */
    VLB_EQUIRECT_SKYBOX_COMP = 1  /* shaders/skybox_sh.comp:23-42 */
};

#define VLB_MAX_LIGHTS 8
#define VLB_MAX_K      16

/* ------------------------------------------------------------------ types */
typedef struct vlb_ctx  vlb_ctx;
typedef struct vlb_plan vlb_plan;

typedef struct vlb_light {
    int32_t type;    /* VLB_LIGHT_*                                           */
    float   vec[3];  /* DIRECTIONAL: direction the light travels (unit);      */
                     /* POINT: position                                       */
    float   rgb[3];  /* DIRECTIONAL: irradiance; POINT: intensity (/r^2)      */
} vlb_light;

/* Procedural sky, replacing main.rmiss:37-41 whose skybox set is unbound in
 * the baker (env_map_generator.cpp:218).  radiance(d) =
 *   lerp(horizon, zenith, max(d.y, 0)) + sun_rgb * exp(sun_sharpness*(d.s-1)) */
typedef struct vlb_sky {
    float zenith_rgb[3];
    float horizon_rgb[3];
    float sun_dir[3];      /* unit vector towards the sun */
    float sun_rgb[3];
    float sun_sharpness;
} vlb_sky;

/* Parameters of shaders/main.rchit:37-47 (defaults of env_map.rchit:25,75,
 * 81,92,95-96: light (1,10,1), bias 0.005, ambient 0, Cd 0.5, Cs 0.5, 16). */
typedef struct vlb_ref_shading {
    float    light_position[3];
    float    shadow_bias;
    float    ambient;
    float    c_diffuse;
    float    c_specular;
    float    glossiness;
    uint32_t flags;        /* VLB_REF_SRGB | VLB_REF_UNORM8 */
} vlb_ref_shading;

typedef struct vlb_bake_params {
    /* probes (host array, num_probes * 3 floats) */
    uint32_t         num_probes;
    const float*     probe_xyz;
    /* ray set */
    int32_t          dir_mode;        /* VLB_DIRS_*                         */
    uint32_t         rays_per_probe;  /* FIBONACCI: R (EQUIRECT: ignored)   */
    uint32_t         equirect_w;      /* EQUIRECT only                      */
    uint32_t         equirect_h;
    float            t_min;           /* primary interval, env_map.rgen:22-23 */
    float            t_max;
    /* projection */
    uint32_t         sh_order;        /* 2 (9 coefs) or 3 (16 coefs)        */
    int32_t          sh_constants;    /* VLB_SH_*                           */
    int32_t          precision;       /* VLB_PREC_*                         */
    /* shading */
    int32_t          shade_mode;      /* VLB_SHADE_*                        */
    uint32_t         num_lights;      /* LAMBERT only, <= VLB_MAX_LIGHTS    */
    const vlb_light* lights;
    float            shadow_bias;     /* LAMBERT only                       */
    vlb_sky          sky;
    vlb_ref_shading  ref;
    /* tuning (0 = automatic) */
    uint32_t         rays_per_slice;  /* projection slice length (rays)     */
    uint32_t         trace_variant;   /* 1..12 one fused trace+shade kernel: */
                                      /*   1 + map + 4*occ, lane->ray map    */
                                      /*   0..3 x waves/SIMD {none, 6, 8};   */
                                      /* 13..24 wavefront kernels (closest,  */
                                      /*   shadow, shade): 13 + map +        */
                                      /*   4*refill, refill {16, 32, 8}      */
                                      /*   idle lanes; 25..36 as 1..12 with  */
                                      /*   shadow rays visiting BVH children */
                                      /*   far first (same results); 0 =     */
                                      /*   autotuned choice, else 12         */
    uint32_t         reserved[2];
} vlb_bake_params;

typedef struct vlb_scene_info {
    uint32_t num_tris;
    uint32_t num_nodes;
    uint32_t depth;            /* inner-node levels, root = 1             */
    uint32_t reserved;
    uint64_t device_bytes;
    double   build_ms;         /* device time of the BVH build            */
    float    bounds_min[3];
    float    bounds_max[3];
} vlb_scene_info;

typedef struct vlb_bake_stats {
    uint64_t primary_rays;     /* probes * rays (padding excluded)        */
    uint64_t shadow_rays;      /* shadow rays actually traced             */
    uint64_t node_visits;      /* inner nodes fetched (stats pass only)   */
    uint64_t tri_tests;        /* triangles tested (stats pass only)      */
    double   trace_ms;         /* trace + shade kernel  \                  */
    double   project_ms;       /* MFMA projection kernel > mean over the  */
    double   merge_ms;         /* FP64 merge kernel     /  timed executes */
    double   total_ms;         /* whole execute on the device             */
    uint32_t tiles;
    uint32_t grid;
    uint32_t rays_per_slice;
    uint32_t slices;
    uint32_t timed_runs;       /* executes averaged (since set_timing(1)) */
    uint32_t reserved;
} vlb_bake_stats;

/* ------------------------------------------------------------ context API */
int32_t     vlb_abi_version(void);
/* Replaces Application::Application (src/application.cpp:540-556): device
 * selection, queues and pools become one HIP device + one stream.
 * hbm_budget_bytes == 0 means "all free memory". */
vlb_status  vlb_ctx_create(int32_t device, uint64_t hbm_budget_bytes, vlb_ctx** out);
/* Frees the context's scenes and device buffers.  VLB_ERR_INVALID (and
 * nothing freed) while plans, batches or servers made from it are alive:
 * destroy those first. */
vlb_status  vlb_ctx_destroy(vlb_ctx* ctx);
const char* vlb_last_error(const vlb_ctx* ctx);
/* Number of compute units and device name (for reports). */
vlb_status  vlb_device_info(vlb_ctx* ctx, int32_t* num_cus, char* name, size_t name_len);
/* Failure injection for tests of callers' error handling (no counterpart in
 * the reference, whose only failure path is an exception caught in main,
 * src/baker/main.cpp:26-40): the next `count` calls at `site` (VLB_INJECT_*)
 * fail with VLB_ERR_INJECTED before they launch or allocate anything; count 0
 * disarms the site.  A server step that takes an injected failure fails the
 * requests whose work it held (their vlb_wait returns VLB_ERR_INJECTED) and
 * the server keeps serving the others. */
vlb_status  vlb_debug_inject(vlb_ctx* ctx, int32_t site, int32_t count);

/* ------------------------------------------------------------- scene API */
/* Replaces SceneManager::pushScene + Scene_t::buildAccelerationStructures
 * (src/scene_manager.cpp:1013-1034, 339-443): world-space triangles are
 * uploaded and a BVH is built on the GPU (PLOC, Meister & Bittner 2018,
 * collapsed to a 4-wide BVH kept as float and as 8-bit quantized nodes).
 * xyz: nverts*3, finite with |x| <= 2^100 (else VLB_ERR_INVALID),
 * idx: ntris*3, normals: nverts*3 or NULL (geometric normals),
 * albedo_rgb: ntris*3 (base colour, env_map.rchit:36-49). */
vlb_status  vlb_scene_upload(vlb_ctx* ctx, const float* xyz, uint32_t nverts,
                             const uint32_t* idx, uint32_t ntris,
                             const float* normals, const float* albedo_rgb,
                             uint32_t* out_scene);
/* Frees the scene's device buffers.  VLB_ERR_INVALID (and nothing freed)
 * while plans, batch jobs or unfinished server requests use the scene:
 * destroy or finish those first. */
vlb_status  vlb_scene_release(vlb_ctx* ctx, uint32_t scene);
vlb_status  vlb_scene_info_get(vlb_ctx* ctx, uint32_t scene, vlb_scene_info* out);
/* Copies the BVH out: nodes as 16 floats each (child boxes + child ids, see
 * DESIGN), leaf-order triangle ids.  For tests and the traversal byte count. */
vlb_status  vlb_scene_export_bvh(vlb_ctx* ctx, uint32_t scene, float* nodes,
                                 size_t nodes_len, int32_t* leaf_prims,
                                 size_t leaf_len);

/* ------------------------------------------------------------- ray queries */
/* traceRayEXT closest hit (env_map.rgen:26): prim = -1 on miss; t, u, v are
 * gl_HitTEXT and hitAttributeEXT (u = weight of v1, v = weight of v2). */
vlb_status  vlb_trace_closest(vlb_ctx* ctx, uint32_t scene, const float* org,
                              const float* dir, uint32_t n, float t_min,
                              float t_max, int32_t* prim, float* t, float* u,
                              float* v);
/* Shadow query of env_map.rchit:84-88 (TerminateOnFirstHit | SkipClosestHit):
 * occluded[i] = 1 if any triangle is hit in [t_min, t_max[i]]. */
vlb_status  vlb_trace_occluded(vlb_ctx* ctx, uint32_t scene, const float* org,
                               const float* dir, const float* t_max,
                               uint32_t n, float t_min, uint8_t* occluded);

/* ---------------------------------------------------------------- baking */
/* One-shot bake: LightBaker::bake (src/baker/light_baker.cpp:287-328) with
 * per-probe reset ("intended sum").  out_coeffs: num_probes*K*3 floats. */
vlb_status  vlb_bake(vlb_ctx* ctx, uint32_t scene, const vlb_bake_params* p,
                     float* out_coeffs, size_t out_len);

/* Plan API: a prepared, device-resident bake (probe positions, direction and
 * SH tables, partial buffers).  execute() is asynchronous on the context
 * stream; sync() waits; download() copies out[P][K][3]. */
vlb_status  vlb_plan_create(vlb_ctx* ctx, uint32_t scene, const vlb_bake_params* p,
                            vlb_plan** out);
vlb_status  vlb_plan_execute(vlb_plan* plan);
vlb_status  vlb_plan_sync(vlb_plan* plan);
vlb_status  vlb_plan_download(vlb_plan* plan, float* out_coeffs, size_t out_len);
vlb_status  vlb_plan_download_f64(vlb_plan* plan, double* out_coeffs, size_t out_len);
/* Enable per-kernel hipEvent timing in execute: every execute records event
 * pairs on the context stream around the bake and merge kernels; stats()
 * reports the means over the executes since the last set_timing(plan, 1). */
vlb_status  vlb_plan_set_timing(vlb_plan* plan, int32_t on);
/* Capture execute() as a hipGraph and replay it from then on. */
vlb_status  vlb_plan_use_graph(vlb_plan* plan, int32_t on);
/* Runs a counting variant of the trace kernel once (node visits, triangle
 * tests, shadow rays) and fills the counters of vlb_bake_stats. */
vlb_status  vlb_plan_count(vlb_plan* plan);
/* Autotuner: times the trace-kernel variants (lane->ray maps x occupancy
 * bounds x shadow order, and the wavefront kernels) on this plan's real
 * workload, `iters` runs each, keeps the fastest
 * for the plan and remembers it for later plans on the same scene and ray
 * count.  *chosen (may be NULL) receives the trace_variant picked. */
vlb_status  vlb_plan_autotune(vlb_plan* plan, uint32_t iters, uint32_t* chosen);
vlb_status  vlb_plan_stats(vlb_plan* plan, vlb_bake_stats* out);
vlb_status  vlb_plan_destroy(vlb_plan* plan);

/* Per-ray debug output of a bake (EnvMapGenerator::getMap's storage image,
 * env_map_generator.cpp:325-359): radiance[P][R][3] as fed to the
 * projection, and the primary hit prim id (-1 = sky) per ray.  Either output
 * may be NULL. */
vlb_status  vlb_bake_radiance(vlb_ctx* ctx, uint32_t scene, const vlb_bake_params* p,
                              float* radiance, size_t radiance_len,
                              int32_t* hit_prim, size_t hit_len);

/* ---------------------------------------------------------------- batches */
/* Many bakes over any of the context's resident scenes, run as one unit.
 * Each job has the semantics of vlb_bake (LightBaker::bake,
 * src/baker/light_baker.cpp:287-328, with per-probe reset); the reference
 * itself bakes one scene per process.  The batch is cut into rounds of at
 * most max_round_rays primary rays; a round is one persistent launch per
 * pass (closest hit, shadow rays, shading, MFMA projection) over every job
 * it holds, and vlb_batch_use_graph captures all rounds as one hipGraph.
 * A job's coefficients do not depend on the other jobs of the batch.  All
 * jobs must use the same precision; params->trace_variant is ignored (the
 * batch has its own, below). */
typedef struct vlb_batch vlb_batch;
typedef struct vlb_batch_job {
    uint32_t               scene;
    uint32_t               reserved;
    const vlb_bake_params* params;   /* rays_per_slice 0: whole direction */
                                     /* set up to 16384 rays per slice    */
} vlb_batch_job;
typedef struct vlb_batch_options {
    uint64_t max_round_rays;  /* 0 = 2^28; at most 2^30                   */
    uint32_t trace_variant;   /* 1..12 fused kernel (map x waves/SIMD as */
                              /* in vlb_bake_params), 13..24 wavefront  */
                              /* lane map x refill, 25..36 fused with   */
                              /* far-first shadow rays; 0 = 12          */
    uint32_t chunk;           /* rays per work fetch: 64..1024, divides   */
                              /* 1024; 0 = 256                            */
    uint32_t keep_f64;        /* also keep FP64 coefficients              */
    uint32_t reserved[3];
} vlb_batch_options;
typedef struct vlb_batch_info {
    uint64_t primary_rays;
    uint64_t shadow_rays;     /* after vlb_batch_count                    */
    uint64_t node_visits;     /* after vlb_batch_count                    */
    uint64_t tri_tests;       /* after vlb_batch_count                    */
    uint64_t projection_tiles;
    uint64_t device_bytes;    /* job state + tables + round buffers       */
    double   total_ms;        /* mean device time of the timed executes   */
    uint32_t jobs, rounds, grid, trace_variant, chunk, timed_runs;
    /* with vlb_batch_set_timing(b, 1): mean device time per execute of the */
    /* trace, projection and merge launches (summed over the rounds)        */
    double   trace_ms, project_ms, merge_ms;
    uint32_t kernel_timed_runs;
    uint32_t reserved;
} vlb_batch_info;
vlb_status  vlb_batch_create(vlb_ctx* ctx, const vlb_batch_job* jobs, uint32_t njobs,
                             const vlb_batch_options* opt /* may be NULL */,
                             vlb_batch** out);
/* Enqueues every round (asynchronous; each execute is timed on the device). */
vlb_status  vlb_batch_execute(vlb_batch* b);
vlb_status  vlb_batch_sync(vlb_batch* b);
/* Capture execute() as one hipGraph and replay it from then on. */
vlb_status  vlb_batch_use_graph(vlb_batch* b, int32_t on);
/* Per-kernel timing: from now on every eager execute (use_graph off) also
 * records hipEvents around each round's trace, projection and merge launches
 * on the context stream; vlb_batch_info reports the means per execute (summed
 * over the rounds).  A graph replay is timed as a whole only: events recorded
 * into a graph under capture cannot be timed on this runtime.  Each call
 * resets the means. */
vlb_status  vlb_batch_set_timing(vlb_batch* b, int32_t on);
/* Coefficients of job `job`: [P][K][3] in the job's probe order. */
vlb_status  vlb_batch_download(vlb_batch* b, uint32_t job, float* out, size_t out_len);
vlb_status  vlb_batch_download_f64(vlb_batch* b, uint32_t job, double* out, size_t out_len);
/* Autotuner: times the wavefront lane maps x refill thresholds
 * (trace_variant 13..24) x work-fetch chunks {256, 1024} and the fused
 * kernel's lane maps 2, 3 at 6 and 8 waves/SIMD with either shadow order
 * (7, 8, 11, 12, 31, 32, 35, 36) on a sample of
 * every job's tiles (about 16M primary rays), `iters` runs each, and keeps
 * the fastest. */
vlb_status  vlb_batch_autotune(vlb_batch* b, uint32_t iters, uint32_t* variant,
                               uint32_t* chunk);
/* One counting pass (node visits, triangle tests, shadow rays). */
vlb_status  vlb_batch_count(vlb_batch* b);
vlb_status  vlb_batch_info_get(vlb_batch* b, vlb_batch_info* out);
vlb_status  vlb_batch_destroy(vlb_batch* b);

/* ---------------------------------------------------------------- serving */
/* Rebake requests over resident scenes, served by continuous batching (no
 * counterpart in the reference, whose baker runs once per process).  A
 * scheduler thread admits requests in arrival order while a request slot is
 * free and their probe state (probes, partial sums, outputs) fits the
 * state_budget arena, and runs steps of at most step_rays primary rays that
 * hold a share of every admitted request's remaining work, two steps in
 * flight (see VLB_SCHED_*).  Each request has the semantics of vlb_bake and its coefficients
 * do not depend on what else is served.  A request holds its scene from
 * vlb_submit until it is done (vlb_scene_release refuses meanwhile).
 * vlb_submit copies everything it needs. */
typedef struct vlb_server vlb_server;
typedef struct vlb_server_options {
    uint64_t state_budget;    /* device bytes of admitted request state;  */
                              /* 0 = 2 GiB (allocated once, up front)     */
    uint64_t step_rays;       /* primary rays per step; 0 = 2^25          */
    uint64_t result_bytes;    /* pinned host arena for finished results   */
                              /* until vlb_wait; 0 = 256 MiB              */
    uint64_t upload_bytes;    /* pinned staging for new requests' probes  */
                              /* per step; 0 = 32 MiB                     */
    uint32_t max_active;      /* requests on the device at once; 0 = 64   */
    uint32_t max_lights;      /* shadow rays per primary ray; 0 = 4       */
    int32_t  precision;       /* VLB_PREC_* of every request              */
    uint32_t trace_variant;   /* 1..36 as in vlb_batch_options; 0 = 35    */
    uint32_t chunk;           /* rays per work fetch; 0 = 256             */
    uint32_t max_depth;       /* deepest scene BVH served; 0 = 64         */
    uint32_t policy;          /* VLB_SCHED_*: how a step's rays are shared */
    uint32_t age_ms;          /* SRPT: requests submitted more than age_ms */
                              /* ago go first, oldest first; 0 = no bound */
} vlb_server_options;
/* Step scheduling: SRPT fills each step with the admitted requests that have
 * the least work left first (best median latency), except that requests
 * older than age_ms (when set) go first, oldest first, which bounds how long
 * a large request can be passed over; SHARE gives every admitted request an
 * equal share (processor sharing); FIFO serves them in arrival order. */
enum { VLB_SCHED_SRPT = 0, VLB_SCHED_SHARE = 1, VLB_SCHED_FIFO = 2 };
typedef struct vlb_request_info {
    double   submit_ms;       /* server clock: ms since vlb_server_create */
    double   admit_ms;        /* state allocated on the device            */
    double   start_ms;        /* first step holding its work launched     */
    double   done_ms;         /* coefficients on the host                 */
    uint64_t primary_rays;
    uint32_t steps;           /* steps that traced part of the request    */
    int32_t  status;
} vlb_request_info;
typedef struct vlb_server_stats {
    uint64_t submitted, completed, steps, primary_rays, probes;
    uint64_t state_bytes_peak;
    double   busy_ms;         /* device time of all retired steps         */
    double   clock_ms;        /* server clock now                         */
    uint32_t active_peak;     /* most requests admitted at once           */
    uint32_t queued;          /* waiting for admission now                */
    int32_t  status;          /* VLB_OK unless the scheduler failed       */
    uint32_t trace_variant;   /* the trace variant the server runs        */
} vlb_server_stats;
vlb_status  vlb_server_create(vlb_ctx* ctx, const vlb_server_options* opt /* may be NULL */,
                              vlb_server** out);
/* Queues a request (non-blocking); *request receives its id. */
vlb_status  vlb_submit(vlb_server* s, uint32_t scene, const vlb_bake_params* p,
                       uint64_t* request);
/* *done = 1 once the request's coefficients are on the host. */
vlb_status  vlb_poll(vlb_server* s, uint64_t request, int32_t* done);
/* Blocks until the request is done, copies its [P][K][3] coefficients to
 * out (may be NULL), fills *info (may be NULL) and forgets the request.
 * Returns the request's status: VLB_OK, or why it did not finish
 * (VLB_ERR_CANCELLED, VLB_ERR_INJECTED, a device error). */
vlb_status  vlb_wait(vlb_server* s, uint64_t request, float* out, size_t out_len,
                     vlb_request_info* info);
/* vlb_wait with a deadline: VLB_ERR_TIMEOUT if the request is not done
 * within timeout_ms milliseconds of the call (0: just look); the request
 * then stays and can be waited for again.  The reference waits on its fences
 * without bound and only prints a failed wait (src/application.cpp:263-276). */
vlb_status  vlb_wait_for(vlb_server* s, uint64_t request, double timeout_ms, float* out,
                         size_t out_len, vlb_request_info* info);
/* Cancels a request that has not finished: a queued one is dropped at once;
 * an admitted one gets no further steps and releases its device state when
 * the steps already launched with its work retire.  Its vlb_wait returns
 * VLB_ERR_CANCELLED.  A request that finished first keeps its result (the
 * call returns VLB_OK either way; vlb_wait tells which happened). */
vlb_status  vlb_cancel(vlb_server* s, uint64_t request);
/* Admission hold: while on, queued requests are not admitted (admitted ones
 * keep running); off resumes admission.  For draining a server and for
 * tests that need requests to stay queued. */
vlb_status  vlb_server_hold(vlb_server* s, int32_t on);
vlb_status  vlb_server_stats_get(vlb_server* s, vlb_server_stats* out);
/* Stops the scheduler: requests still queued or running are cancelled (their
 * work in flight on the device is waited for, not abandoned), results not yet
 * collected are dropped, and every resource of the server is freed.  Collect
 * the results you need with vlb_wait first. */
vlb_status  vlb_server_destroy(vlb_server* s);

/* --------------------------------------------------- SH of images / skies */
/* Skybox SH of float cubemaps: this library's form of skybox_sh.comp
 * (shaders/skybox_sh.comp:23-42, host src/skybox_manager.cpp:107-130) for a
 * 6 x N x N RGBA32F cubemap with exact texel solid angles.  Faces +X,-X,+Y,
 * -Y,+Z,-Z in Vulkan cube-map orientation.  The context keeps the solid-angle
 * table of the last face size used and, for N <= 147, a per-texel direction
 * table (at most 2 MiB) until the next size or vlb_ctx_destroy.  Host in/out. */
vlb_status  vlb_skybox_sh(vlb_ctx* ctx, const float* cubemaps, uint32_t face_size,
                          uint32_t batch, uint32_t sh_order, float* out, size_t out_len);
/* Device-resident variant: d_cubemaps and d_out are device pointers from
 * vlb_device_alloc; asynchronous on the context stream.  *kernel_ms (may be
 * NULL) receives the projection kernel's event-timed duration after a
 * vlb_ctx_sync. */
vlb_status  vlb_skybox_sh_d(vlb_ctx* ctx, const void* d_cubemaps, uint32_t face_size,
                            uint32_t batch, uint32_t sh_order, void* d_out);
vlb_status  vlb_skybox_last_ms(vlb_ctx* ctx, double* project_ms, double* merge_ms);
/* Allocation-free variant of vlb_skybox_sh_d.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc/hipFree or other allocating routine: its scratch
 * (the face's solid-angle and direction tables and the per-workgroup partial
 * sums) is a fixed device array of VLB_SKYBOX_NOALLOC_SCRATCH bytes that
 * each GPU holds until the process ends, shared by every call on that GPU:
 * calls that overlap on one GPU race.  Requests whose scratch would exceed it
 * fail with VLB_ERR_INVALID and launch nothing (1024 cubemaps of 6 x 128^2
 * need about 2 MiB; one of 6 x 1024^2 about 4.4 MiB).  d_cubemaps, d_out:
 * device pointers as for vlb_skybox_sh_d; asynchronous on the context
 * stream; no event timing. */
#define VLB_SKYBOX_NOALLOC_SCRATCH (8u << 20)
vlb_status  vlb_skybox_sh_noalloc(vlb_ctx* ctx, const void* d_cubemaps, uint32_t face_size,
                                  uint32_t batch, uint32_t sh_order, void* d_out);

/* Projection of an equirect RGBA8 image to 16 SH coefficients x RGB:
 * variant VLB_EQUIRECT_SH_COMP is shaders/sh.comp:23-42 (host
 * light_baker.cpp:269-285), VLB_EQUIRECT_SKYBOX_COMP is
 * shaders/skybox_sh.comp:23-42 (phi - pi/2, SH at dir.xzy).  Intended-sum
 * semantics (no race, fresh zero).  out: 16*3 floats. */
vlb_status  vlb_project_equirect(vlb_ctx* ctx, const uint8_t* rgba8, uint32_t width,
                                 uint32_t height, int32_t variant, float* out16x3);

/* shaders/sh_sum.comp:35-52: per pixel sum_k c_k Y_k(dir) * 1250 stored as
 * RGBA8 UNORM (alpha 255).  coeffs: 16*3 floats, out: width*height*4. */
vlb_status  vlb_sh_reconstruct(vlb_ctx* ctx, const float* coeffs16x3, uint32_t width,
                               uint32_t height, uint8_t* out_rgba8);

/* ------------------------------------------------------ probe-grid lookup */
/* Consumer side of a baked grid: shaders/main.rchit:124-168 with
 * shaders/sh.rmiss:23-39 -- for each surface point, the 8 probes of its grid
 * cell, each visibility-tested with a shadow ray from the point (offset by
 * `bias` along the normal) and weighted, and the SH of each evaluated at the
 * point's normal.  The reference's lookup disagrees with its own producer in
 * stride ((lmax+1)^2 = 289 vs 16 written), probe order (x + 7y + 49z vs the
 * BFS order), grid (hard-coded 7^3, origin 0) and SH frame (normal unswizzled
 * vs the bake's (x, z, y)); this form uses the grid as baked: `order`,
 * `origin`, `count`, `num_coefs`, and the bake's SH frame.
 *  REFERENCE  : weight = max(|step| - distance, 0) (main.rchit:141,155),
 *               value = sum_{k<K} c_k Y_k(n) (sh.rmiss:30-37, SH radiance at
 *               the normal).
 *  IRRADIANCE : trilinear weights, value = sum_l (A_l / pi) sum_m c_lm Y_lm(n)
 *               with A_0 = pi, A_1 = 2pi/3, A_2 = pi/4, A_3 = 0 (irradiance
 *               / pi of the probe's radiance field).
 * Occluded probes are dropped and the remaining weights renormalised; a point
 * with no visible probe gets 0. */
typedef struct vlb_probe_grid_desc {
    float    origin[3];     /* position of probe (0, 0, 0)                   */
    float    step[3];       /* gridStep                                      */
    uint32_t count[3];      /* probes per axis                               */
    int32_t  order;         /* VLB_ORDER_* of the coefficient buffer         */
    uint32_t num_coefs;     /* K: 9 or 16                                    */
    int32_t  sh_constants;  /* VLB_SH_* used to evaluate the basis           */
} vlb_probe_grid_desc;
enum { VLB_LOOKUP_REFERENCE = 0, VLB_LOOKUP_IRRADIANCE = 1 };
typedef struct vlb_lookup_params {
    int32_t  mode;          /* VLB_LOOKUP_*                                  */
    uint32_t visibility;    /* 1: shadow ray to every probe                  */
    float    bias;          /* ray origin offset along the normal            */
    uint32_t reserved;
} vlb_lookup_params;
/* coeffs: [P][K][3] in the grid's order (host); points, normals: n x 3 host
 * floats (normals unit length); out_rgb: n x 3. */
vlb_status  vlb_probe_lookup(vlb_ctx* ctx, uint32_t scene, const vlb_probe_grid_desc* grid,
                             const float* coeffs, const vlb_lookup_params* lp,
                             const float* points, const float* normals, uint32_t n,
                             float* out_rgb);

/* Multi-bounce indirect light: the step after direct-only baking that the
 * reference names as its goal (README.md:9-11), in the form its renderer
 * shades surfaces with a baked grid (shaders/main.rchit:118-172: albedo x
 * (lookup + direct)).  After this call every execute of a LAMBERT plan adds,
 * at each primary hit, albedo x scale x the grid lookup of vlb_probe_lookup
 * at the hit point and the two-sided shading normal (IRRADIANCE mode with
 * scale 1 is the Lambertian bounce: albedo x E / pi), reading a source grid
 * kept on the device; the execute then copies its own result into that
 * source, so N executes are N further bounces.  The grid describes the
 * plan's own probes: count product = num_probes, num_coefs = K, order and
 * origin/step as they were laid out.  Visibility rays to the corner probes
 * are traced like shadow rays (and counted as such).
 * coeffs: [P][K][3] host floats to start from, or NULL to start from the
 * plan's last result (the plan must have executed).  ip NULL: back to
 * direct light only.  A plan created with trace_variant 0 and no autotuner
 * pick runs variant 12 with and without the indirect term.
 * No counterpart call in the reference. */
typedef struct vlb_indirect_params {
    vlb_probe_grid_desc grid;
    vlb_lookup_params   lookup;
    float               scale;     /* factor on the looked-up value          */
    uint32_t            reserved[3];
} vlb_indirect_params;
vlb_status  vlb_plan_set_indirect(vlb_plan* plan, const vlb_indirect_params* ip,
                                  const float* coeffs);

/* ------------------------------------------------------- device buffers */
vlb_status  vlb_device_alloc(vlb_ctx* ctx, size_t bytes, void** dptr);
vlb_status  vlb_device_free(vlb_ctx* ctx, void* dptr);
vlb_status  vlb_upload(vlb_ctx* ctx, void* dptr, const void* src, size_t bytes);
vlb_status  vlb_download(vlb_ctx* ctx, void* dst, const void* dptr, size_t bytes);
vlb_status  vlb_ctx_sync(vlb_ctx* ctx);
/* Streaming copy kernel (16-byte loads/stores) of `bytes` (multiple of 16)
 * between device buffers, on the context stream.  No counterpart in the
 * reference: it calibrates HBM byte counters and measures attainable
 * bandwidth.  *ms (may be NULL) gets the event-timed duration. */
vlb_status  vlb_device_copy(vlb_ctx* ctx, void* dst, const void* src, size_t bytes,
                            double* ms);
/* Read-only streaming pass over `bytes` (multiple of 16) of device memory,
 * split into `grid` contiguous workgroup ranges (0: 16 per compute unit),
 * on the context stream.  No counterpart in the reference: it measures the
 * attainable read bandwidth that read-bound kernels (the skybox projection)
 * are compared with.  *ms (may be NULL) gets the event-timed duration. */
vlb_status  vlb_device_read(vlb_ctx* ctx, const void* src, size_t bytes, uint32_t grid,
                            double* ms);

/* ------------------------------------------------------------ host helpers */
/* LightBaker::probePositionsFromBoudingBox (src/baker/light_baker.cpp:80-101):
 * counts[0]*counts[1]*counts[2] probes between bmin and bmax, step =
 * (bmax-bmin)/(counts-1) in FP32.  REFERENCE_BFS reproduces the reference's
 * breadth-first order and repeated "+= step"; LINEAR is x-fastest with
 * position bmin + i*step.  out_xyz: P*3 floats, out_step: 3 floats. */
vlb_status  vlb_probe_grid(const float bmin[3], const float bmax[3],
                           const uint32_t counts[3], int32_t order,
                           float* out_xyz, float* out_step);
/* Direction table used by the bake (float x,y,z per ray). */
vlb_status  vlb_directions(int32_t dir_mode, uint32_t rays, uint32_t w, uint32_t h,
                           float* out_xyz, float* out_weight, size_t n);
/* base64 of light_baker.cpp:330-373 (RFC 4648 alphabet, '=' padding).
 * Writes at most out_len bytes including the terminating NUL; *written gets
 * the encoded length.  VLB_ERR_INVALID if out_len is too small. */
vlb_status  vlb_base64_encode(const uint8_t* bytes, size_t n, char* out,
                              size_t out_len, size_t* written);
/* Baked-light glTF of LightBaker::serialize (light_baker.cpp:375-401): the
 * input glTF JSON text gets the coefficient buffer appended as a base64 data
 * URI, a bufferView, and light = {gridStep, lmax (= coefficient count),
 * bufferView}.  Writes at most out_len bytes (incl. NUL); *written gets the
 * length.  out may be NULL to query the size. */
vlb_status  vlb_gltf_bake_json(const char* gltf_json, const float* coeffs,
                               uint32_t num_probes, uint32_t num_coefs,
                               const float grid_step[3], char* out, size_t out_len,
                               size_t* written);
/* Reader of that file, Scene_t::loadBakedLight (scene_manager.cpp:613-648):
 * light.gridStep, light.lmax (the coefficient count) and the coefficient
 * floats of the buffer behind light.bufferView (its data URI decoded, the
 * buffer's byteLength bytes kept).  coeffs may be NULL to query *num_floats.
 * VLB_ERR_NOT_FOUND when the file has no "light". */
vlb_status  vlb_gltf_read_bake(const char* gltf_json, float grid_step[3], uint32_t* num_coefs,
                               float* coeffs, size_t coeffs_len, size_t* num_floats);


/* -------------------------------------------------------- synthetic scenes */
/* Seeded procedural scenes of the benchmark configurations.  kind:
 *  0 ROOM    : open-top room 10x4x10 m, 64 boxes U[0.2,1.5] m, 8 icospheres
 *              subdiv 3 (~11k tris)
 *  1 TERRAIN : n x n fBm heightfield (4 octaves) + n_spheres icospheres
 *              subdiv 5 (n=256, 50 spheres: ~1.15M tris)
 *  2 CITY    : 400 extruded, facade-tessellated buildings (~2M tris)
 *  3 CUBE    : the 12-triangle cube of default_blender_cube.gltf
 *  4 SPHERES : n_spheres icospheres subdiv `param` in a box (tests)
 * The scene lives in library memory until vlb_host_scene_free. */
typedef struct vlb_host_scene {
    uint32_t  nverts;
    uint32_t  ntris;
    float*    xyz;
    float*    normals;
    uint32_t* idx;
    float*    albedo;
    float     bounds_min[3];
    float     bounds_max[3];
} vlb_host_scene;
enum { VLB_SCENE_ROOM = 0, VLB_SCENE_TERRAIN = 1, VLB_SCENE_CITY = 2,
       VLB_SCENE_CUBE = 3, VLB_SCENE_SPHERES = 4 };
/* glTF scene ingest, Scene_t (src/scene_manager.cpp:31-66 file types,
 * 241-337 vertex/index fetch, 445-538 node matrices and bounds, 837-857 the
 * node tree of scenes[0]): .gltf (buffers as data URIs or files next to it)
 * or .glb.  Every mesh primitive of every node becomes world-space triangles
 * (node matrix T*R*S*M composed with its parents'); NORMAL (normalised at
 * fetch, as the reference) is carried through the inverse transpose and left
 * unnormalised (the shading normalises the interpolated normal, as
 * env_map.rchit:63-68 does after its WorldToObject product); albedo is the
 * material's baseColorFactor rgb, 1 when absent or all zero (textures are
 * not sampled).  bounds_min/max follow the reference's scene bounds
 * (scene_manager.cpp:497-507): they start at 0 and take each primitive's
 * local min and max corners through its node's own matrix only.  *json_text
 * (may be NULL) receives the glTF JSON (free with vlb_host_scene_free's
 * companion vlb_text_free). */
vlb_status  vlb_gltf_load(const char* path, vlb_host_scene** out, char** json_text);
void        vlb_text_free(char* text);

vlb_status  vlb_scene_generate(int32_t kind, uint64_t seed, uint32_t param0,
                               uint32_t param1, vlb_host_scene** out);
void        vlb_host_scene_free(vlb_host_scene* s);

#ifdef __cplusplus
}
#endif
#endif /* VLB_BAKE_H */
