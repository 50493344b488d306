/*
 * vlb_oracle.h -- CPU oracle of the probe-baking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * shaders (read as text) used to check the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only
 * as the checker or the timed CPU baseline -- never as the product path.
 *
 * What each function follows (reference file:line) and what pins it
 * (the check that would fail on a plausible mistake) is listed per function.
 * Functions whose pin is only partial say "parity unpinned" below.
 *
 * Semantics: "intended sum" -- every probe starts from zero (the reference
 * never re-zeroes its SSBO, light_baker.cpp:110,121) and the racy "+=" of
 * sh.comp:39 is an exact sum.
 */
#ifndef VLB_ORACLE_H
#define VLB_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/vlb_bake.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- SH basis */
/* Real SH of shaders/sh_common.h:26-223 (l <= 4, 25 functions, coef index
 * l(l+1)+m), evaluated in double at d (SH frame, z = "up").  exact = 0 uses
 * the literal 6-figure constants of sh_common.h, exact = 1 the closed-form
 * normalisations sqrt((2l+1)/(4pi) (l-|m|)!/(l+|m|)!) (x sqrt2 for m != 0).
 * Pinned by: scipy's complex Y_l^m converted to the real basis (independent
 * library routine, fixes signs/normalisation), orthonormality under a
 * Gauss-Legendre x uniform-phi product quadrature, literal/exact ratio
 * constant per function and within 1.4e-6 of 1 for l <= 3 (4.3e-6 at l = 4). */
void orc_sh_eval(int K, int exact, const double d[3], double* out);

/* --------------------------------------------------------------- directions */
/* Fibonacci sphere, R directions, y up: y_i = 1-(2i+1)/R, phi_i = i*pi*(3-sqrt5).
 * Computed in double, rounded to float.  Weight 4pi/R each.
 * Pinned by: unit length, sum of directions ~ 0, quadrature of Y_00^2 -> 1. */
void orc_fibonacci(uint32_t R, float* xyz);
/* shaders/env_map.rgen:21 + sh_common.h:8-22: toVector(x2phi(x,W),
 * y2theta(y,H)).xzy in float with PI = 3.1415926538f; weight of
 * sh.comp:32-33 (float).  Ray index y*W + x.
 * Pinned by: sum of weights -> 4pi (midpoint rule), closed-form projections. */
void orc_equirect_dirs(uint32_t W, uint32_t H, float* xyz, float* weight);

/* ------------------------------------------------------------------- grid */
/* src/baker/light_baker.cpp:80-101 (BFS order, repeated += step in float)
 * and a LINEAR order.  Pinned by: explicit index table for 7^3 and the
 * iterated-float positions recomputed independently in numpy float32. */
void orc_probe_grid(const float b0[3], const float b1[3], const uint32_t c[3],
                    int order, float* out, float* step);

/* ----------------------------------------------------------------- scenes */
typedef struct orc_scene orc_scene;
/* bvh = 0: brute force only; 1: also build a binned-SAH BVH (own layout). */
orc_scene* orc_scene_create(const float* xyz, uint32_t nverts, const uint32_t* idx,
                            uint32_t ntris, const float* normals,
                            const float* albedo, int bvh);
void       orc_scene_free(orc_scene* s);
uint32_t   orc_scene_bvh_depth(const orc_scene* s);

/* Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013) in
 * float with the double fallback for zero edge functions; inclusive interval
 * [tmin, tmax]; closest hit breaks ties on the lower prim id.  This replaces
 * the driver's traceRayEXT (env_map.rgen:26, env_map.rchit:87), whose
 * arithmetic is not under the reference tree: parity with the hardware is
 * unpinned; parity of the GPU library with THIS test is bit-exact.
 * Pinned by: analytic t for planes/boxes/spheres, watertightness on shared
 * edges/vertices, BVH result == brute force. */
int  orc_trace_closest(const orc_scene* s, const float o[3], const float d[3],
                       float tmin, float tmax, int brute, float* t, float* u,
                       float* v);
int  orc_trace_occluded(const orc_scene* s, const float o[3], const float d[3],
                        float tmin, float tmax, int brute);
void orc_trace_closest_batch(const orc_scene* s, const float* org, const float* dir,
                             uint32_t n, float tmin, float tmax, int brute,
                             int32_t* prim, float* t, float* u, float* v);
void orc_trace_occluded_batch(const orc_scene* s, const float* org, const float* dir,
                              const float* tmax, uint32_t n, float tmin, int brute,
                              uint8_t* occluded);

/* ------------------------------------------------------------------- bake */
/* LightBaker::bake (light_baker.cpp:287-328) per probe: env_map.rgen ->
 * env_map.rchit / shadow.rmiss / (procedural) main.rmiss -> sh.comp, with
 * per-probe reset.  mode64 = 0: shading in float (the timed CPU baseline);
 * mode64 = 1: the same geometric decisions (hits, N.L > 0, visibility, all
 * float) with radiance and projection in double (the error reference).
 * Projection always accumulates in double.  out: P*K*3 doubles.
 * radiance (P*R*3) and hit (P*R) are optional per-ray outputs in canonical
 * ray order.  nthreads <= 0: one thread per online core.
 * Returns 0 on success, -1 on bad parameters.
 * Pinned by (tests/test_oracle_bake.py, closed forms over one plane): sky-only
 * bakes = the quadrature sum_r (4 pi / R) Y_k L_r with the scipy-pinned basis;
 * Lambert with a directional light = albedo E cos / pi on every floor hit
 * (both windings: the two-sided normal); point light = albedo I cos /
 * (pi d^2) per analytic hit point, with analytic shadows of a square
 * occluder; REFERENCE = albedo (ambient + Cd sDotN + Cs max(R.d, 0)^gloss). */
int orc_bake(const orc_scene* s, const vlb_bake_params* p, int mode64, int nthreads,
             int brute, double* out, float* radiance, int32_t* hit,
             uint64_t* shadow_rays);
/* orc_bake with the multi-bounce indirect term of vlb_plan_set_indirect: at
 * each LAMBERT hit, radiance += albedo x scale x orc_probe_lookup(g, coeffs,
 * lp) at the hit point and the two-sided float shading normal; the lookup's
 * visibility rays count as shadow rays.  g describes the source buffer
 * (num_coefs = K); the baked probes may be any subset or other positions.
 * Pinned by: (1) zero albedo -> orc_bake exactly; (2) a constant source grid
 * (c00 only) in a closed room: result - orc_bake = albedo x scale x c00 x
 * Y00 x the projection of a constant over the same rays; (3) an l = 1 source
 * field L(d) = d.y over a floor whose triangles face away from the probes:
 * result = albedo x (2/3) x (N.y of the normal facing the ray) projected over
 * the downward rays (closed form of the clamped-cosine convolution and the
 * two-sided normal). */
int orc_bake_indirect(const orc_scene* s, const vlb_bake_params* p, const vlb_probe_grid_desc* g,
                      const float* coeffs, const vlb_lookup_params* lp, float scale, int mode64,
                      int nthreads, int brute, double* out, float* radiance, int32_t* hit,
                      uint64_t* shadow_rays);

/* Procedural sky of the benchmark configurations (replaces main.rmiss). */
void orc_sky(const vlb_sky* sky, const double d[3], double rgb[3]);

/* --------------------------------------------------------- SH projections */
/* Cubemap (6 x N x N RGBA32F, Vulkan face orientation, faces +X,-X,+Y,-Y,
 * +Z,-Z) to K coefs x RGB with exact texel solid angles, double.  SH frame
 * = world (x, z, y) as in the bake.  Pinned by: sum of solid angles = 4pi,
 * constant sky -> c00 = 2 sqrt(pi) c, zonal closed forms. */
void orc_skybox_sh(const float* cube, uint32_t N, int K, double* out);
/* shaders/sh.comp:23-42 (variant 0) and shaders/skybox_sh.comp:23-42
 * (variant 1: phi - pi/2, SH at dir.xzy) over an RGBA8 equirect image,
 * 16 coefs x RGB, double.  Pinned by closed forms on synthetic images. */
void orc_equirect_project(const uint8_t* rgba8, uint32_t W, uint32_t H,
                          int variant, double* out16x3);
/* shaders/sh_sum.comp:35-52 in float: 1250 * sum c_k Y_k -> RGBA8 UNORM.
 * Pinned by: projection of the reconstruction of a band-limited field. */
void orc_sh_reconstruct(const float* coeffs16x3, uint32_t W, uint32_t H,
                        uint8_t* out);

/* ------------------------------------------------------ probe-grid lookup */
/* shaders/main.rchit:124-168 + shaders/sh.rmiss:23-39 in the fixed form of
 * vlb_probe_lookup (grid as baked, bake's SH frame).  Visibility-ray geometry
 * in float (same operations as the HIP kernel), SH and sums in double.
 * s may be NULL (no visibility).  out: n x 3 doubles.  Returns 0, or -1 on
 * bad arguments.
 * Pinned by: partition of unity (equal probes -> that probe's value for any
 * point and weights), trilinear reproduction of a field linear in position,
 * irradiance of constant radiance L = L and of the l = 1 field L(d) = d.s ->
 * (2/3) n.s (closed forms of the clamped-cosine convolution), BFS index
 * table against orc_probe_grid's positions, and probes behind a wall
 * dropped by the visibility rays. */
int orc_probe_lookup(const orc_scene* s, const vlb_probe_grid_desc* g, const float* coeffs,
                     const vlb_lookup_params* lp, const float* pts, const float* nrm, uint32_t n,
                     double* out);

/* ---------------------------------------------------------------- format */
/* base64 of light_baker.cpp:330-373.  Pinned by Python's base64 (RFC 4648).
 * out must hold 4*ceil(n/3)+1 bytes; returns the encoded length. */
size_t orc_base64(const uint8_t* bytes, size_t n, char* out);

/* Colour helpers of env_map.rchit:27-34 and the UNORM8 store/load. */
float orc_srgb(float x);
float orc_unorm8(float x);

#ifdef __cplusplus
}
#endif
#endif
