// device_common.hpp -- device-side building blocks of the probe baker:
// BVH node layout, watertight ray/triangle test, conservative slab test,
// LDS-stack traversal (closest hit and shadow query), sky and shading.
//
// Every translation unit that includes this file is compiled with
// -ffp-contract=off: the intersection, hit-point, normal and shadow-ray
// arithmetic below is a fixed sequence of correctly rounded IEEE operations
// (HIP's default f32 division and sqrt are correctly rounded), so hit ids and
// visibility are bit-identical to a C restatement that does the same
// operations in the same order.  Only the slab test uses explicit fmaf: it
// decides which nodes are visited, never which triangle wins.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vlb_bake.h"
#include "probe_grid.hpp"

namespace vlb {

// ---------------------------------------------------------------- layout
// One BVH2 inner node = 64 bytes: the boxes of both children and their ids.
// child >= 0: inner node index; child < 0: leaf, ~child = leaf-order
// triangle index.  The root is node 0.
struct __align__(16) BNode {
    float4 a;  // lo0.x lo0.y lo0.z hi0.x
    float4 b;  // hi0.y hi0.z lo1.x lo1.y
    float4 c;  // lo1.z hi1.x hi1.y hi1.z
    int4   d;  // child0 child1 - -
};

// The traversal kernels walk a 4-wide BVH collapsed from it (half the
// dependent node fetches of the BVH2 per ray), stored twice with the same
// node ids: the root is node 0; child ids >= 0 are inner BVH4 nodes, < 0
// leaves (~ leaf-order triangle), kEmpty an unused slot.
//
// BNode4, 128 bytes (112 in use): the boxes of up to four children as float
// structure-of-arrays (lo.x of child c in lox[c], ...) and their ids.  Read
// by the fused kernels, which run at the 64-VGPR bound of 8 waves/SIMD where
// a decode step costs more than the bytes it saves.
//
// QNode4, 64 bytes: per axis a grid origin o (the node box's lo) and step s
// (its extent / 255, rounded up), the children's boxes as 8-bit grid
// coordinates q (plane = fma(q, s, o); lo codes rounded down, hi codes
// rounded up, each checked against that very decode when the node is
// built), and the ids.  A decoded box contains the child's float box, so the
// slab test sees conservative boxes, slightly larger than the BVH2's (2-6%
// more node visits).  Read by the wavefront kernels, which have registers to
// spare and whose node fetches are half the bytes through the texture path.
struct __align__(16) BNode4 {
    float4 lox, loy, loz;
    float4 hix, hiy, hiz;
    int4   child;
    int4   pad;
};
struct __align__(16) QNode4 {
    float4 a;      // o.x, o.y, o.z, s.x
    float4 b;      // s.y, s.z, lo.x codes, lo.y codes (as bits)
    uint4  q;      // lo.z, hi.x, hi.y, hi.z codes
    int4   child;
};
// (codes of one plane: child k in byte k)
constexpr int kEmpty = (int)0x80000000;

// Stack entries a BVH4 traversal can hold for a BVH2 of `depth` inner
// levels: at most 3 pushes per BVH4 level, ceil(depth / 2) levels.
__host__ __device__ inline uint32_t stack_bound(uint32_t depth)
{
    return 3u * ((depth + 1u) / 2u) + 1u;
}

// Leaf-order triangle = 3 x float4: (v0, prim id bits) (v1, 0) (v2, 0).
// Shading record per prim id = 3 x float4: (n0, albedo.r) (n1, .g) (n2, .b).

struct SceneView {
    const BNode4* nodes;      // float BVH4 (fused kernels)
    const QNode4* qnodes;     // quantized BVH4, same ids (wavefront kernels)
    const float4* tris;
    const float4* shade;
    int           ntris;
    float         maxabs;  // max |coordinate| of the scene's vertices
};

struct ShadeParams {
    int32_t      mode;         // VLB_SHADE_*
    int32_t      nlights;
    float        shadow_bias;  // LAMBERT
    float        t_max;        // shadow extent of directional lights
    vlb_light    lights[VLB_MAX_LIGHTS];
    vlb_sky      sky;
    vlb_ref_shading ref;
    // LAMBERT multi-bounce: a baked grid whose looked-up value at the hit
    // (times scale) adds albedo x that value; ind.coeffs null: direct only
    ProbeGridView ind;
};

// Stack entries kept in LDS per lane: 16 for the fused and single-query
// kernels, 12 for the wavefront kernels, whose leaf queues take another 8 KB
// per workgroup (12 + 8 entries x 256 lanes x 4 B = 20 KB lets them run 7
// waves/SIMD: the trace stage of a 1.15M-triangle terrain bake, 16,384
// probes x 16,384 rays, 2.5% faster; the fused kernels measured 1% slower at
// 12 on an 11k-triangle room bake).  Spill areas are sized for the smaller
// depth.
constexpr int kLdsDepth = 16;
constexpr int kLdsDepthWave = 12;
constexpr int kLdsDepthMin = kLdsDepthWave < kLdsDepth ? kLdsDepthWave : kLdsDepth;
// Work counters (persistent kernels' fetch positions) sit 128 bytes apart,
// one per cache line: device-scope atomics on one line serialise.
constexpr int kCtrStride = 16;

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ float dot3(const float3& a, const float3& b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

// -------------------------------------------------------- watertight test
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1),
// 2013: shear the triangle into ray space along the dominant axis kz, take
// the 2D edge functions U, V, W (recomputed in double when one is zero), and
// accept when they share a sign.
struct WtRay {
    float3 o;
    // axis permutation (kx, ky, kz) as select masks: kz = 0 -> r1, kz = 1 ->
    // r2, kz = 2 -> neither; swp exchanges kx and ky (d[kz] < 0)
    bool  r1, r2, swp;
    float Sx, Sy, Sz;
};

// (v[kx], v[ky], v[kz]) with selects only (no indexed access: a runtime
// component index would be lowered to divergent branches).
__device__ __forceinline__ float3 wt_perm(const WtRay& w, const float3& v)
{
    const float a = w.r1 ? v.y : (w.r2 ? v.z : v.x);
    const float b = w.r1 ? v.z : (w.r2 ? v.x : v.y);
    const float c = w.r1 ? v.x : (w.r2 ? v.y : v.z);
    return make_float3(w.swp ? b : a, w.swp ? a : b, c);
}

__device__ __forceinline__ WtRay wt_setup(const float3& o, const float3& d)
{
    WtRay w;
    w.o = o;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    const int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    w.r1 = kz == 0;
    w.r2 = kz == 1;
    w.swp = false;
    const float3 pd0 = wt_perm(w, d);
    w.swp = pd0.z < 0.0f;
    const float3 pd = wt_perm(w, d);
    w.Sx = pd.x / pd.z;
    w.Sy = pd.y / pd.z;
    w.Sz = 1.0f / pd.z;
    return w;
}

// Edge-function part of the test.  On acceptance returns t (correctly
// rounded T / det) with the numerators V, W and det kept for the
// barycentrics, which are divided out once for the final hit only.
struct WtHit {
    float t, V, W, det;
};

__device__ __forceinline__ bool wt_test(const WtRay& w, const float3& v0, const float3& v1,
                                        const float3& v2, float tmin, float tmax, WtHit& h)
{
    const float3 A = wt_perm(w, make_float3(v0.x - w.o.x, v0.y - w.o.y, v0.z - w.o.z));
    const float3 B = wt_perm(w, make_float3(v1.x - w.o.x, v1.y - w.o.y, v1.z - w.o.z));
    const float3 C = wt_perm(w, make_float3(v2.x - w.o.x, v2.y - w.o.y, v2.z - w.o.z));
    const float Ax = A.x - w.Sx * A.z;
    const float Ay = A.y - w.Sy * A.z;
    const float Bx = B.x - w.Sx * B.z;
    const float By = B.y - w.Sy * B.z;
    const float Cx = C.x - w.Sx * C.z;
    const float Cy = C.y - w.Sy * C.z;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By;
        double CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy;
        double AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay;
        double ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = w.Sz * A.z;
    const float Bz = w.Sz * B.z;
    const float Cz = w.Sz * C.z;
    const float T = U * Az + V * Bz + W * Cz;
    const float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    h.t = t;
    h.V = V;
    h.W = W;
    h.det = det;
    return true;
}

__device__ __forceinline__ bool wt_hit(const WtRay& w, const float3& v0, const float3& v1,
                                       const float3& v2, float tmin, float tmax, float& t_out,
                                       float& u_out, float& v_out)
{
    WtHit h;
    if (!wt_test(w, v0, v1, v2, tmin, tmax, h)) return false;
    t_out = h.t;
    u_out = h.V / h.det;
    v_out = h.W / h.det;
    return true;
}

// -------------------------------------------------------------- slab test
// Conservative: widened by 4e-6 relative plus an absolute margin of 64 ulps
// of the largest coordinate in play, so a box holding an accepted triangle
// is never culled.  Margins cost node visits, never change a result.
struct SlabRay {
    float3 inv;
    float3 oinv;  // o * inv
    float  apad;
};

__device__ __forceinline__ SlabRay slab_setup(const float3& o, const float3& d, float maxabs)
{
    SlabRay s;
    float dx = d.x == 0.0f ? copysignf(1e-20f, d.x) : d.x;
    float dy = d.y == 0.0f ? copysignf(1e-20f, d.y) : d.y;
    float dz = d.z == 0.0f ? copysignf(1e-20f, d.z) : d.z;
    s.inv = make_float3(1.0f / dx, 1.0f / dy, 1.0f / dz);
    s.oinv = make_float3(o.x * s.inv.x, o.y * s.inv.y, o.z * s.inv.z);
    float m = fmaxf(maxabs, fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z))));
    s.apad = m * (64.0f * 1.1920929e-7f);
    return s;
}

// returns the (widened) entry distance, for near-first ordering, or +inf on
// a miss
__device__ __forceinline__ float slab(const SlabRay& s, float lox, float loy, float loz,
                                      float hix, float hiy, float hiz, float tmin, float tbest)
{
    float t0x = fmaf(lox, s.inv.x, -s.oinv.x), t1x = fmaf(hix, s.inv.x, -s.oinv.x);
    float t0y = fmaf(loy, s.inv.y, -s.oinv.y), t1y = fmaf(hiy, s.inv.y, -s.oinv.y);
    float t0z = fmaf(loz, s.inv.z, -s.oinv.z), t1z = fmaf(hiz, s.inv.z, -s.oinv.z);
    float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
    float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    float hi = fminf(tf, tbest);
    hi = fmaf(fabsf(hi), 4e-6f, hi) + s.apad;
    const float lo = fmaf(-fabsf(tn), 4e-6f, tn) - s.apad;
    return (lo <= hi && hi >= tmin) ? lo : __builtin_huge_valf();
}

// ------------------------------------------------------------- traversal
// Stack: entries [0, depth) in LDS at lds[i * 256] (lane-interleaved,
// conflict-free), deeper entries in a global spill area at
// spill[(i - depth) * spill_stride].
struct Stack {
    uint32_t* lds;
    uint32_t* spill;
    uint32_t  spill_stride;
    int       depth = kLdsDepth;  // LDS entries per lane (a constant of the kernel)
};

__device__ __forceinline__ void push(const Stack& st, int& sp, uint32_t v)
{
    if (sp < st.depth) st.lds[sp * 256] = v;
    else st.spill[(size_t)(sp - st.depth) * st.spill_stride] = v;
    ++sp;
}
__device__ __forceinline__ uint32_t pop(const Stack& st, int& sp)
{
    --sp;
    if (sp < st.depth) return st.lds[sp * 256];
    return st.spill[(size_t)(sp - st.depth) * st.spill_stride];
}

struct TraceCount {
    uint32_t nodes = 0;
    uint32_t tris = 0;
};

__device__ __forceinline__ void load_tri(const float4* tris, int leaf, float3& v0, float3& v1,
                                         float3& v2, int& prim)
{
    const float4 a = tris[3 * leaf + 0];
    const float4 b = tris[3 * leaf + 1];
    const float4 c = tris[3 * leaf + 2];
    v0 = make_float3(a.x, a.y, a.z);
    v1 = make_float3(b.x, b.y, b.z);
    v2 = make_float3(c.x, c.y, c.z);
    prim = __float_as_int(a.w);
}

// Planes fma(q, s, o) of the four children for one code word (the decode
// the node builder checked its codes against).
__device__ __forceinline__ float4 decode4(uint32_t q, float s, float o)
{
    return make_float4(fmaf((float)(q & 255u), s, o), fmaf((float)((q >> 8) & 255u), s, o),
                       fmaf((float)((q >> 16) & 255u), s, o), fmaf((float)(q >> 24), s, o));
}

// Float node fetch: the 112 bytes in use (7 x 16 B).
__device__ __forceinline__ BNode4 load_node(const BNode4* nodes, uint32_t i)
{
    const BNode4* p = nodes + i;
    BNode4 n;
    n.lox = p->lox;
    n.loy = p->loy;
    n.loz = p->loz;
    n.hix = p->hix;
    n.hiy = p->hiy;
    n.hiz = p->hiz;
    n.child = p->child;
    return n;
}

// Quantized node fetch (4 x 16 B) and decode to the float form.
__device__ __forceinline__ BNode4 load_node(const QNode4* nodes, uint32_t i)
{
    const QNode4* p = nodes + i;
    const float4 a = p->a;
    const float4 b = p->b;
    const uint4 q = p->q;
    BNode4 n;
    n.lox = decode4(__float_as_uint(b.z), a.w, a.x);
    n.loy = decode4(__float_as_uint(b.w), b.x, a.y);
    n.loz = decode4(q.x, b.y, a.z);
    n.hix = decode4(q.y, a.w, a.x);
    n.hiy = decode4(q.z, b.x, a.y);
    n.hiz = decode4(q.w, b.y, a.z);
    n.child = p->child;
    return n;
}

// Node fetches of a wave whose active lanes are all at the same node (about
// half of the fused kernel's visits on coherent probe rays) come through the
// scalar cache into scalar registers, leaving the vector memory path, which
// the divergent fetches keep busy, to those.  Nodes are read-only while any
// trace kernel runs.  Used by the fused traversal (trace_closest, trace_any);
// the wavefront kernels refill lanes with unrelated rays, rarely share a
// node across a wave and measured slower with it (1.48 vs 1.41 ms on the
// 11k-triangle room bake, 2048 probes x 4096 rays).
typedef int sgpr16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ void scalar_load64(unsigned long long addr, sgpr16_t& a)
{
    __asm__ __volatile__("s_load_dwordx16 %0, %1, 0x0\n\t"
                         "s_waitcnt lgkmcnt(0)"
                         : "=s"(a)
                         : "s"(addr)
                         : "memory");
}

__device__ __forceinline__ BNode4 load_node_uniform(const BNode4* p)
{
    const unsigned long long addr = (unsigned long long)p;
    sgpr16_t a, b;
    scalar_load64(addr, a);
    scalar_load64(addr + 64, b);
    BNode4 n;
    n.lox = make_float4(__int_as_float(a[0]), __int_as_float(a[1]), __int_as_float(a[2]), __int_as_float(a[3]));
    n.loy = make_float4(__int_as_float(a[4]), __int_as_float(a[5]), __int_as_float(a[6]), __int_as_float(a[7]));
    n.loz = make_float4(__int_as_float(a[8]), __int_as_float(a[9]), __int_as_float(a[10]), __int_as_float(a[11]));
    n.hix = make_float4(__int_as_float(a[12]), __int_as_float(a[13]), __int_as_float(a[14]), __int_as_float(a[15]));
    n.hiy = make_float4(__int_as_float(b[0]), __int_as_float(b[1]), __int_as_float(b[2]), __int_as_float(b[3]));
    n.hiz = make_float4(__int_as_float(b[4]), __int_as_float(b[5]), __int_as_float(b[6]), __int_as_float(b[7]));
    n.child = make_int4(b[8], b[9], b[10], b[11]);
    return n;
}

__device__ __forceinline__ BNode4 load_node_uniform(const QNode4* p)
{
    sgpr16_t w;
    scalar_load64((unsigned long long)p, w);
    BNode4 n;
    const float ox = __int_as_float(w[0]), oy = __int_as_float(w[1]), oz = __int_as_float(w[2]);
    const float sx = __int_as_float(w[3]), sy = __int_as_float(w[4]), sz = __int_as_float(w[5]);
    n.lox = decode4((uint32_t)w[6], sx, ox);
    n.loy = decode4((uint32_t)w[7], sy, oy);
    n.loz = decode4((uint32_t)w[8], sz, oz);
    n.hix = decode4((uint32_t)w[9], sx, ox);
    n.hiy = decode4((uint32_t)w[10], sy, oy);
    n.hiz = decode4((uint32_t)w[11], sz, oz);
    n.child = make_int4(w[12], w[13], w[14], w[15]);
    return n;
}

// Watertight test of triangle `leaf` for the fused traversal.  SCALAR: when
// every active lane tests the same triangle, its record comes through the
// scalar cache (32 + 16 bytes, exactly the record) and the test runs on the
// scalar operands, sparing the vector memory path three 16-byte loads per
// lane; else, and without SCALAR, each lane loads its own.  Each path runs
// its own test, so the scalar words are never copied into vector registers.
// The multi-job kernel takes SCALAR in both queries (a batch of 64 jobs over
// four scenes: 196.4 vs 202.9 ms per trace stage; 200.0 with the shadow
// queries alone).  The single-bake kernel takes it in the shadow queries
// only (1.15M-triangle terrain, 16,384 probes x 16,384 rays: 38.25 vs 39.27
// ms; the 11k-triangle room bake unchanged): in its closest-hit query the
// extra scalar registers spill into vector-register lanes read back in the
// loop, and the room bake ran 1.03-1.04 vs 0.99-1.00 ms.
typedef int sgpr8_t __attribute__((ext_vector_type(8)));
typedef int sgpr4_t __attribute__((ext_vector_type(4)));

template <bool SCALAR>
__device__ __forceinline__ bool fetch_wt(const float4* tris, int leaf, const WtRay& w, float tmin,
                                         float tmax, WtHit& h, int& prim)
{
    if (!SCALAR) {
        float3 v0, v1, v2;
        load_tri(tris, leaf, v0, v1, v2, prim);
        return wt_test(w, v0, v1, v2, tmin, tmax, h);
    }
    const float4* p = tris + 3 * leaf;
    const unsigned long long a = (unsigned long long)p;
    const unsigned long long u =
        ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    if (__all(a == u)) {
        sgpr8_t x;
        sgpr4_t y;
        // early-clobber: the first load's destination never overlaps the
        // address the second one reads
        __asm__ __volatile__("s_load_dwordx8 %0, %2, 0x0\n\t"
                             "s_load_dwordx4 %1, %2, 0x20\n\t"
                             "s_waitcnt lgkmcnt(0)"
                             : "=&s"(x), "=&s"(y)
                             : "s"(u)
                             : "memory");
        prim = x[3];
        return wt_test(w, make_float3(__int_as_float(x[0]), __int_as_float(x[1]), __int_as_float(x[2])),
                       make_float3(__int_as_float(x[4]), __int_as_float(x[5]), __int_as_float(x[6])),
                       make_float3(__int_as_float(y[0]), __int_as_float(y[1]), __int_as_float(y[2])),
                       tmin, tmax, h);
    }
    float3 v0, v1, v2;
    load_tri(tris, leaf, v0, v1, v2, prim);
    return wt_test(w, v0, v1, v2, tmin, tmax, h);
}

// Child c[k] of a 4-way node with a runtime k, through selects (a register
// array indexed at run time would be placed in scratch memory).
__device__ __forceinline__ int pick4(const int c[4], int k)
{
    const int a = (k & 1) ? c[1] : c[0];
    const int b = (k & 1) ? c[3] : c[2];
    return (k & 2) ? b : a;
}

// Hit leaf children as a 4-bit mask; their entry distances are cleared so
// that only inner children are sorted and pushed.
__device__ __forceinline__ uint32_t take_leaves(float e[4], const int c[4])
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (e[k] != __builtin_huge_valf() && c[k] < 0) {
            m |= 1u << k;
            e[k] = __builtin_huge_valf();
        }
    return m;
}

// Entry distances of the four children (+inf: missed or empty slot).
// CHECK_EMPTY: the node's unused slots are marked kEmpty (quantized nodes);
// a float node's unused slots hold all-+inf boxes that the slab test misses.
template <bool CHECK_EMPTY>
__device__ __forceinline__ void slab4(const SlabRay& s, const BNode4& n, float tmin, float tbest,
                                      float e[4], int c[4])
{
    c[0] = n.child.x;
    c[1] = n.child.y;
    c[2] = n.child.z;
    c[3] = n.child.w;
    e[0] = slab(s, n.lox.x, n.loy.x, n.loz.x, n.hix.x, n.hiy.x, n.hiz.x, tmin, tbest);
    e[1] = slab(s, n.lox.y, n.loy.y, n.loz.y, n.hix.y, n.hiy.y, n.hiz.y, tmin, tbest);
    e[2] = slab(s, n.lox.z, n.loy.z, n.loz.z, n.hix.z, n.hiy.z, n.hiz.z, tmin, tbest);
    e[3] = slab(s, n.lox.w, n.loy.w, n.loz.w, n.hix.w, n.hiy.w, n.hiz.w, tmin, tbest);
    if (CHECK_EMPTY) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (c[k] == kEmpty) e[k] = __builtin_huge_valf();
    }
}

// Entry distances and ids of the children of node i, and its hit leaf
// children as a 4-bit mask (take_leaves).  The node comes through
// the scalar cache when its address is the same in every active lane (checked
// on the address, so lanes of a multi-job wave on different scenes take the
// vector path), else per lane.  Each path runs the slab tests on its own
// registers: merging the two fetches first would copy the 28 scalar node
// words into vector registers on every uniform visit.
template <typename NodeT>
__device__ __forceinline__ uint32_t fetch_slab4(const NodeT* nodes, uint32_t i, const SlabRay& s,
                                                float tmin, float tbest, float e[4], int c[4])
{
    const NodeT* p = nodes + i;
    const unsigned long long a = (unsigned long long)p;
    const unsigned long long u =
        ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
        (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    if (__all(a == u)) {
        const BNode4 n = load_node_uniform(reinterpret_cast<const NodeT*>(u));
        slab4<std::is_same<NodeT, QNode4>::value>(s, n, tmin, tbest, e, c);
        return take_leaves(e, c);
    }
    const BNode4 n = load_node(nodes, i);
    slab4<std::is_same<NodeT, QNode4>::value>(s, n, tmin, tbest, e, c);
    return take_leaves(e, c);
}

__device__ __forceinline__ void cswap(float& ea, int& ca, float& eb, int& cb)
{
    const bool sw = eb < ea;
    const float te = sw ? eb : ea;
    const int tc = sw ? cb : ca;
    eb = sw ? ea : eb;
    cb = sw ? ca : cb;
    ea = te;
    ca = tc;
}

// Orders next_node visits the inner children of a node in.
enum : int {
    kOrderNear = 0,  // nearest entry first (closest-hit queries)
    kOrderSlot = 1,  // slot order, no sort
    kOrderFar = 2,   // farthest entry first
};

// After the leaf tests (their e set to +inf): order the inner children, push
// all hit ones but the first (last first) and continue at the first, else
// pop.  Returns false when the stack is empty.  Closest-hit queries sort near
// first.  Shadow queries stop at any hit, so the order only moves the visit
// count: slot order drops the sort's 20 instructions per visit and visits
// about as many nodes as near first (0.1-0.3% fewer without the sort on
// rooms and terrains of 11k-1.15M triangles); far first (the sort on negated
// entries, misses staying +inf and last) reaches an occluder sooner on some
// trees and not on others, so it is a variant for the autotuners: on the
// 11k-triangle room it visited 5.1% fewer nodes and ran 4% faster, on a
// batch of 64 jobs over four scenes 1.4% fewer and 0.5% slower, both with
// trees built from per-axis Morton cells; on trees from cubic cells (the
// build's) it runs level with slot order or 1% behind.  The pushes below
// are right for any order (a missed child is written and then overwritten).
template <int ORDER = kOrderNear>
__device__ __forceinline__ bool next_node(float e[4], int c[4], const Stack& st, int& sp,
                                          uint32_t& node)
{
    if (ORDER == kOrderFar) {
        for (int k = 0; k < 4; ++k)
            if (e[k] != __builtin_huge_valf()) e[k] = -e[k];
    }
    if (ORDER != kOrderSlot) {
        cswap(e[0], c[0], e[1], c[1]);
        cswap(e[2], c[2], e[3], c[3]);
        cswap(e[0], c[0], e[2], c[2]);
        cswap(e[1], c[1], e[3], c[3]);
        cswap(e[1], c[1], e[2], c[2]);
    }
    const float inf = __builtin_huge_valf();
    if (__all(sp + 3 <= st.depth)) {
        // Every active lane has three free LDS entries (one wave-uniform
        // test): straight-line pushes with conditional increments.  A write
        // past the top lands at most where the all-hit case writes, and is
        // overwritten by the next push or left unused, so the pushes are the
        // same as below in any order.
        st.lds[sp * 256] = (uint32_t)c[3];
        sp += e[3] != inf ? 1 : 0;
        st.lds[sp * 256] = (uint32_t)c[2];
        sp += e[2] != inf ? 1 : 0;
        st.lds[sp * 256] = (uint32_t)c[1];
        sp += e[1] != inf ? 1 : 0;
        if (e[0] != inf) {
            node = (uint32_t)c[0];
            return true;
        }
        if (sp == 0) return false;
        --sp;
        node = st.lds[sp * 256];
        return true;
    }
    if (e[3] != inf) push(st, sp, (uint32_t)c[3]);
    if (e[2] != inf) push(st, sp, (uint32_t)c[2]);
    if (e[1] != inf) push(st, sp, (uint32_t)c[1]);
    if (e[0] != inf) {
        node = (uint32_t)c[0];
        return true;
    }
    if (sp == 0) return false;
    node = pop(st, sp);
    return true;
}

// Closest hit: lexicographic minimum of (t, prim id) over accepted
// triangles in [tmin, tmax].  Returns the prim id or -1.  Any BVH whose
// boxes are conservative gives the same answer: every triangle that could
// beat the current best is tested, and ties go to the lower prim id.
template <bool COUNT, bool STRI = false>
__device__ int trace_closest(const SceneView& sc, const float3& o, const float3& d, float tmin,
                             float tmax, float& t_out, float& u_out, float& v_out,
                             const Stack& st, TraceCount& cnt)
{
    if (sc.ntris <= 0) return -1;
    const WtRay w = wt_setup(o, d);
    const SlabRay s = slab_setup(o, d, sc.maxabs);
    float tbest = tmax, bV = 0.0f, bW = 0.0f, bdet = 1.0f;
    int best = -1;
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        if (COUNT) ++cnt.nodes;
        float e[4];
        int c[4];
        const uint32_t hit_leaves = fetch_slab4(sc.nodes, node, s, tmin, tbest, e, c);
        // leaves one at a time per lane: the wave runs as many tests as its
        // busiest lane has hit leaves, not one per child slot
        for (uint32_t leaves = hit_leaves; leaves != 0; leaves &= leaves - 1) {
            int prim;
            if (COUNT) ++cnt.tris;
            WtHit h;
            // t <= tbest holds on acceptance; an equal t keeps the lower prim id
            if (fetch_wt<STRI>(sc.tris, ~pick4(c, __builtin_ctz(leaves)), w, tmin, tbest, h, prim) &&
                (best < 0 || h.t < tbest || prim < best)) {
                tbest = h.t;
                best = prim;
                bV = h.V;
                bW = h.W;
                bdet = h.det;
            }
        }
        if (!next_node(e, c, st, sp, node)) break;
    }
    if (best >= 0) {
        t_out = tbest;
        u_out = bV / bdet;
        v_out = bW / bdet;
    }
    return best;
}

// Shadow query: true if any triangle is accepted in [tmin, tmax].  FAR: the
// inner children far first, else in slot order (same answer either way).
// STRI: uniform triangle records through the scalar cache (fetch_wt).
template <bool COUNT, bool FAR = false, bool STRI = false>
__device__ bool trace_any(const SceneView& sc, const float3& o, const float3& d, float tmin,
                          float tmax, const Stack& st, TraceCount& cnt)
{
    if (sc.ntris <= 0) return false;
    const WtRay w = wt_setup(o, d);
    const SlabRay s = slab_setup(o, d, sc.maxabs);
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        if (COUNT) ++cnt.nodes;
        float e[4];
        int c[4];
        const uint32_t hit_leaves = fetch_slab4(sc.nodes, node, s, tmin, tmax, e, c);
        for (uint32_t leaves = hit_leaves; leaves != 0; leaves &= leaves - 1) {
            int prim;
            if (COUNT) ++cnt.tris;
            WtHit h;
            if (fetch_wt<STRI>(sc.tris, ~pick4(c, __builtin_ctz(leaves)), w, tmin, tmax, h, prim)) return true;
        }
        if (!next_node<FAR ? kOrderFar : kOrderSlot>(e, c, st, sp, node)) break;
    }
    return false;
}

// ------------------------------------------------------------------- sky
__device__ __forceinline__ float3 sky_radiance(const vlb_sky& sky, const float3& d)
{
    float g = d.y > 0.0f ? d.y : 0.0f;
    float sd = d.x * sky.sun_dir[0] + d.y * sky.sun_dir[1] + d.z * sky.sun_dir[2];
    float sun = expf(sky.sun_sharpness * (sd - 1.0f));
    float3 r;
    r.x = sky.horizon_rgb[0] + (sky.zenith_rgb[0] - sky.horizon_rgb[0]) * g + sky.sun_rgb[0] * sun;
    r.y = sky.horizon_rgb[1] + (sky.zenith_rgb[1] - sky.horizon_rgb[1]) * g + sky.sun_rgb[1] * sun;
    r.z = sky.horizon_rgb[2] + (sky.zenith_rgb[2] - sky.horizon_rgb[2]) * g + sky.sun_rgb[2] * sun;
    return r;
}

// env_map.rchit:27-34
__device__ __forceinline__ float srgb(float x)
{
    if (x < 0.0031308f) return x * 12.92f;
    return 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}
// RGBA8 storage-image store followed by a load
__device__ __forceinline__ float unorm8(float x)
{
    float c = (x > 0.0f) ? (x < 1.0f ? x : 1.0f) : 0.0f;
    return rintf(c * 255.0f) / 255.0f;
}

// ---------------------------------------------------------------- shading
// Indirect term of a multi-bounce bake at hit point P with the two-sided
// shading normal N: the grid lookup of probe_grid.hpp (visibility rays to the
// corner probes traced here, counted as shadow rays), times g.scale.
template <bool COUNT, bool FAR = false, bool STRI = false>
__device__ float3 grid_indirect(const SceneView& sc, const ProbeGridView& g, const float3& P,
                                const float3& N, const Stack& st, TraceCount& cnt,
                                uint32_t& nshadow)
{
    const float p[3] = {P.x, P.y, P.z};
    const float n[3] = {N.x, N.y, N.z};
    float Y[16];
    grid_basis(g, n, Y);
    float sw = 0.0f, acc[3] = {0.0f, 0.0f, 0.0f};
    for (int v = 0; v < 8; ++v) {
        const GridCorner c = grid_corner(g, p, v);
        float w = c.w;
        if (corner_needs_ray(g, c, sc.ntris)) {
            ++nshadow;
            const float3 org = make_float3(P.x + g.bias * N.x, P.y + g.bias * N.y, P.z + g.bias * N.z);
            const float3 dir = make_float3(c.dir[0] / c.dist, c.dir[1] / c.dist, c.dir[2] / c.dist);
            if (trace_any<COUNT, FAR, STRI>(sc, org, dir, 0.0f, c.dist, st, cnt)) w = 0.0f;
        }
        if (w > 0.0f) {
            float val[3];
            corner_value(g, c.row, Y, val);
            acc[0] += w * val[0];
            acc[1] += w * val[1];
            acc[2] += w * val[2];
            sw += w;
        }
    }
    if (!(sw > 0.0f)) return make_float3(0.f, 0.f, 0.f);
    return make_float3(g.scale * (acc[0] / sw), g.scale * (acc[1] / sw), g.scale * (acc[2] / sw));
}

// Direction Ld and distance of light j seen from hit point P, and its
// irradiance E there (before the cosine).
__device__ __forceinline__ void light_geom(const ShadeParams& sp, int j, const float3& P,
                                           float3& Ld, float& dist, float3& E)
{
    const vlb_light& Lj = sp.lights[j];
    if (Lj.type == VLB_LIGHT_POINT) {
        const float3 Lv = make_float3(Lj.vec[0] - P.x, Lj.vec[1] - P.y, Lj.vec[2] - P.z);
        dist = sqrtf(dot3(Lv, Lv));
        Ld = dist > 0.0f ? make_float3(Lv.x / dist, Lv.y / dist, Lv.z / dist)
                         : make_float3(0.f, 0.f, 0.f);
        const float d2 = dist * dist;
        E = make_float3(Lj.rgb[0] / d2, Lj.rgb[1] / d2, Lj.rgb[2] / d2);
    } else {
        Ld = make_float3(-Lj.vec[0], -Lj.vec[1], -Lj.vec[2]);
        dist = sp.t_max;
        E = make_float3(Lj.rgb[0], Lj.rgb[1], Lj.rgb[2]);
    }
}

// Origin of the shadow ray from hit point P with the two-sided normal N.
__device__ __forceinline__ float3 shadow_origin(const ShadeParams& sp, const float3& P, const float3& N)
{
    return make_float3(P.x + sp.shadow_bias * N.x, P.y + sp.shadow_bias * N.y,
                       P.z + sp.shadow_bias * N.z);
}

// Lambert term of light j at hit point P with the two-sided normal N: the
// contribution E cos is formed before the shadow query, so only it (not E
// and the cosine) is live across the traversal; added to acc when unoccluded.
template <bool COUNT, bool FAR = false, bool STRI = false>
__device__ __forceinline__ void light_term(const SceneView& sc, const ShadeParams& sp, int j,
                                           const float3& P, const float3& N, float3& acc,
                                           const Stack& st, TraceCount& cnt, uint32_t& nshadow)
{
    float3 Ld, E;
    float dist;
    light_geom(sp, j, P, Ld, dist, E);
    const float cosNL = dot3(N, Ld);
    if (cosNL > 0.0f) {
        const float3 c = make_float3(E.x * cosNL, E.y * cosNL, E.z * cosNL);
        const float3 org = shadow_origin(sp, P, N);
        ++nshadow;
        if (!trace_any<COUNT, FAR, STRI>(sc, org, Ld, 0.0f, dist, st, cnt)) {
            acc.x = acc.x + c.x;
            acc.y = acc.y + c.y;
            acc.z = acc.z + c.z;
        }
    }
}

// Hit point and interpolated shading normal (env_map.rchit:63-68).
__device__ __forceinline__ void hit_point_normal(const SceneView& sc, int prim, const float3& o,
                                                 const float3& d, float t, float u, float v,
                                                 float3& P, float3& N)
{
    const float4 s0 = sc.shade[3 * prim + 0];
    const float4 s1 = sc.shade[3 * prim + 1];
    const float4 s2 = sc.shade[3 * prim + 2];
    const float b0 = 1.0f - u - v;
    P = make_float3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
    float3 m;
    m.x = s0.x * b0 + s1.x * u + s2.x * v;
    m.y = s0.y * b0 + s1.y * u + s2.y * v;
    m.z = s0.z * b0 + s1.z * u + s2.z * v;
    const float ml = sqrtf(dot3(m, m));
    N = ml > 0.0f ? make_float3(m.x / ml, m.y / ml, m.z / ml) : make_float3(0.f, 0.f, 0.f);
}

// Radiance carried by one primary ray.  `nshadow` counts traced shadow rays.
// IND compiles the multi-bounce term in (used only when sp.ind.coeffs is
// set); direct-only kernels are built without it and keep its registers free.
// FAR: shadow queries visit children far first (same radiance).  STRI:
// where uniform triangle records come through the scalar cache (fetch_wt):
// kStriClosest in the closest-hit query, kStriShadow in the shadow queries.
enum : int { kStriClosest = 1, kStriShadow = 2 };

template <bool COUNT, bool IND = false, bool FAR = false, int STRI = 0>
__device__ float3 shade_ray(const SceneView& sc, const ShadeParams& sp, const float3& o,
                            const float3& d, float tmin, float tmax, const Stack& st,
                            TraceCount& cnt, uint32_t& nshadow, int& hit_prim)
{
    float t = 0.0f, u = 0.0f, v = 0.0f;
    constexpr bool kShadowScalar = (STRI & kStriShadow) != 0;
    const int prim = trace_closest<COUNT, (STRI & kStriClosest) != 0>(sc, o, d, tmin, tmax, t, u, v, st, cnt);
    hit_prim = prim;
    const bool ref = sp.mode == VLB_SHADE_REFERENCE;
    if (prim < 0) {
        float3 s = sky_radiance(sp.sky, d);
        if (ref) {
            if (sp.ref.flags & VLB_REF_SRGB) {
                s.x = srgb(s.x);
                s.y = srgb(s.y);
                s.z = srgb(s.z);
            }
            if (sp.ref.flags & VLB_REF_UNORM8) {
                s.x = unorm8(s.x);
                s.y = unorm8(s.y);
                s.z = unorm8(s.z);
            }
        }
        return s;
    }
    float3 P, N;
    hit_point_normal(sc, prim, o, d, t, u, v, P, N);
    if (ref) {
        // env_map.rchit:73-101, main.rchit:95-116
        const vlb_ref_shading& r = sp.ref;
        const float3 L = make_float3(r.light_position[0] - P.x, r.light_position[1] - P.y,
                                     r.light_position[2] - P.z);
        const float Llen = sqrtf(dot3(L, L));
        const float3 Ln = Llen > 0.0f ? make_float3(L.x / Llen, L.y / Llen, L.z / Llen)
                                      : make_float3(0.f, 0.f, 0.f);
        const float sdn = dot3(Ln, N);
        const float sDotN = sdn > 0.0f ? sdn : 0.0f;
        const float3 org = make_float3(P.x + r.shadow_bias * N.x, P.y + r.shadow_bias * N.y,
                                       P.z + r.shadow_bias * N.z);
        bool in_shadow = true;
        if (sDotN != 0.0f) {
            ++nshadow;
            in_shadow = trace_any<COUNT, FAR, kShadowScalar>(sc, org, Ln, 0.0f, Llen, st, cnt);
        }
        float diffuse = 0.0f, spec = 0.0f;
        if (!in_shadow) {
            diffuse = r.c_diffuse * sDotN;
            const float dnl = dot3(N, Ln);
            const float3 R = make_float3(Ln.x - 2.0f * dnl * N.x, Ln.y - 2.0f * dnl * N.y,
                                         Ln.z - 2.0f * dnl * N.z);
            const float rd = dot3(R, d);
            spec = r.c_specular * powf(rd > 0.0f ? rd : 0.0f, r.glossiness);
        }
        // albedo read after the shadow query (see the Lambert path below)
        __asm__ __volatile__("" ::: "memory");
        const float3 alb = make_float3(sc.shade[3 * prim + 0].w, sc.shade[3 * prim + 1].w,
                                       sc.shade[3 * prim + 2].w);
        const bool zero = alb.x == 0.0f && alb.y == 0.0f && alb.z == 0.0f;
        float3 c;
        c.x = (zero ? 1.0f : alb.x) * (r.ambient + diffuse + spec);
        c.y = (zero ? 1.0f : alb.y) * (r.ambient + diffuse + spec);
        c.z = (zero ? 1.0f : alb.z) * (r.ambient + diffuse + spec);
        if (r.flags & VLB_REF_SRGB) {
            c.x = srgb(c.x);
            c.y = srgb(c.y);
            c.z = srgb(c.z);
        }
        if (r.flags & VLB_REF_UNORM8) {
            c.x = unorm8(c.x);
            c.y = unorm8(c.y);
            c.z = unorm8(c.z);
        }
        return c;
    }
    // LAMBERT, two-sided
    if (dot3(N, d) > 0.0f) N = make_float3(-N.x, -N.y, -N.z);
    float3 acc = make_float3(0.f, 0.f, 0.f);
    // the last light is peeled off the loop: after its shadow ray is set up,
    // neither the hit point nor the normal is live across its traversal
    // (unless the indirect term needs them)
    const int nl = sp.nlights;
    for (int j = 0; j + 1 < nl; ++j)
        light_term<COUNT, FAR, kShadowScalar>(sc, sp, j, P, N, acc, st, cnt, nshadow);
    if (nl > 0) light_term<COUNT, FAR, kShadowScalar>(sc, sp, nl - 1, P, N, acc, st, cnt, nshadow);
    // albedo read after the shadow queries rather than held across them (the
    // compiler barrier keeps it from reusing the loads of the normals, which
    // it would otherwise keep in scratch across the traversals)
    __asm__ __volatile__("" ::: "memory");
    const float3 alb = make_float3(sc.shade[3 * prim + 0].w, sc.shade[3 * prim + 1].w,
                                   sc.shade[3 * prim + 2].w);
    const float kInvPi = 0.318309886183790671538f;
    if (!IND || !sp.ind.coeffs)
        return make_float3(alb.x * acc.x * kInvPi, alb.y * acc.y * kInvPi, alb.z * acc.z * kInvPi);
    const float3 E = grid_indirect<COUNT, FAR, kShadowScalar>(sc, sp.ind, P, N, st, cnt, nshadow);
    return make_float3(alb.x * acc.x * kInvPi + alb.x * E.x, alb.y * acc.y * kInvPi + alb.y * E.y,
                       alb.z * acc.z * kInvPi + alb.z * E.z);
}

}  // namespace vlb
