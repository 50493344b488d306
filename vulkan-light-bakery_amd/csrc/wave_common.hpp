// wave_common.hpp -- device pieces of the wavefront pipeline shared by the
// single-bake kernels (wave.hip) and the multi-job kernels (multi.hip): hit
// geometry from a hit record, shadow-ray setup, one resumable traversal step,
// and the shading of a primary ray from its hit record and shadow
// visibility.  Every function performs the same operation sequence as
// shade_ray in device_common.hpp (and the CPU oracle), so the results do
// not depend on how rays are scheduled.
#pragma once

#include "vlb_internal.hpp"

namespace vlb {

// Geometry of a recorded primary hit (env_map.rchit:63-68): same operations
// as shade_ray in device_common.hpp.
struct HitGeom {
    float3 P, N, alb;
};

__device__ __forceinline__ HitGeom hit_geom(const float4* shade, const float3& o, const float3& d,
                                            int prim, float t, float u, float v)
{
    const float4 s0 = shade[3 * prim + 0];
    const float4 s1 = shade[3 * prim + 1];
    const float4 s2 = shade[3 * prim + 2];
    const float b0 = 1.0f - u - v;
    HitGeom g;
    g.P = make_float3(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
    float3 m;
    m.x = s0.x * b0 + s1.x * u + s2.x * v;
    m.y = s0.y * b0 + s1.y * u + s2.y * v;
    m.z = s0.z * b0 + s1.z * u + s2.z * v;
    const float ml = sqrtf(dot3(m, m));
    g.N = ml > 0.0f ? make_float3(m.x / ml, m.y / ml, m.z / ml) : make_float3(0.f, 0.f, 0.f);
    g.alb = make_float3(s0.w, s1.w, s2.w);
    return g;
}

// The shadow ray of light j for a hit (LAMBERT: the light list with a
// two-sided normal; REFERENCE: the single point light of env_map.rchit).
struct ShadowRay {
    float3 org, dir, E;
    float dist, cosv;
    bool needed;
};

// CORNERS: j >= nlights are the visibility rays of the multi-bounce term
// (compiled only into the kernels that trace them).
template <bool CORNERS = false>
__device__ __forceinline__ ShadowRay shadow_ray(const ShadeParams& sp, const HitGeom& g,
                                                const float3& d, int j)
{
    ShadowRay s;
    if (sp.mode == VLB_SHADE_REFERENCE) {
        const vlb_ref_shading& r = sp.ref;
        const float3 L = make_float3(r.light_position[0] - g.P.x, r.light_position[1] - g.P.y,
                                     r.light_position[2] - g.P.z);
        const float Llen = sqrtf(dot3(L, L));
        const float3 Ln = Llen > 0.0f ? make_float3(L.x / Llen, L.y / Llen, L.z / Llen)
                                      : make_float3(0.f, 0.f, 0.f);
        const float sdn = dot3(Ln, g.N);
        s.cosv = sdn > 0.0f ? sdn : 0.0f;
        s.org = make_float3(g.P.x + r.shadow_bias * g.N.x, g.P.y + r.shadow_bias * g.N.y,
                            g.P.z + r.shadow_bias * g.N.z);
        s.dir = Ln;
        s.dist = Llen;
        s.E = make_float3(0.f, 0.f, 0.f);
        s.needed = s.cosv != 0.0f;
        return s;
    }
    float3 N = g.N;
    if (dot3(N, d) > 0.0f) N = make_float3(-N.x, -N.y, -N.z);
    if (CORNERS && j >= sp.nlights) {
        // visibility ray to corner probe j - nlights of the indirect grid
        // (only when the grid asks for them; the hit implies a scene)
        const float p[3] = {g.P.x, g.P.y, g.P.z};
        const GridCorner c = grid_corner(sp.ind, p, j - sp.nlights);
        s.needed = corner_needs_ray(sp.ind, c, 1);
        s.org = make_float3(g.P.x + sp.ind.bias * N.x, g.P.y + sp.ind.bias * N.y,
                            g.P.z + sp.ind.bias * N.z);
        s.dist = c.dist;
        s.dir = s.needed ? make_float3(c.dir[0] / c.dist, c.dir[1] / c.dist, c.dir[2] / c.dist)
                         : make_float3(0.f, 0.f, 1.f);
        s.E = make_float3(0.f, 0.f, 0.f);
        s.cosv = 0.0f;
        return s;
    }
    const vlb_light& Lj = sp.lights[j];
    if (Lj.type == VLB_LIGHT_POINT) {
        const float3 Lv = make_float3(Lj.vec[0] - g.P.x, Lj.vec[1] - g.P.y, Lj.vec[2] - g.P.z);
        s.dist = sqrtf(dot3(Lv, Lv));
        s.dir = s.dist > 0.0f ? make_float3(Lv.x / s.dist, Lv.y / s.dist, Lv.z / s.dist)
                              : make_float3(0.f, 0.f, 0.f);
        const float d2 = s.dist * s.dist;
        s.E = make_float3(Lj.rgb[0] / d2, Lj.rgb[1] / d2, Lj.rgb[2] / d2);
    } else {
        s.dir = make_float3(-Lj.vec[0], -Lj.vec[1], -Lj.vec[2]);
        s.dist = sp.t_max;
        s.E = make_float3(Lj.rgb[0], Lj.rgb[1], Lj.rgb[2]);
    }
    s.cosv = dot3(N, s.dir);
    s.needed = s.cosv > 0.0f;
    s.org = make_float3(g.P.x + sp.shadow_bias * N.x, g.P.y + sp.shadow_bias * N.y,
                        g.P.z + sp.shadow_bias * N.z);
    return s;
}

// ------------------------------------------------------------ traversal
// Postponed leaf tests (Aila and Laine, "Understanding the Efficiency of Ray
// Traversal on GPUs", HPG 2009): a lane that hits leaf children queues their
// triangles and keeps visiting inner nodes; the wave tests the queues
// together once every active lane has queued triangles or no nodes left, or
// when a queue could overflow.  Testing leaves at every visit instead runs
// each triangle test with the few lanes that happened to hit a leaf.  A
// postponed test means some nodes are culled against an older tbest, which
// only adds visits: the closest hit is still the (t, prim id) minimum over a
// superset of the candidates, and any-hit still reports whether one exists.
constexpr int kLeafQueue = 8;  // entries per lane, in LDS

struct LeafQueue {
    uint32_t* lds;  // entry i of this lane at lds[i * 256]
};

struct TravState {
    WtRay w;
    SlabRay s;
    float tmin, tb;
    int best;
    float bV, bW, bdet;
    uint32_t node;
    int sp;
    int nq;     // queued triangles
    bool live;  // `node` is still to be visited
};

__device__ __forceinline__ void trav_init(TravState& T, float maxabs, const float3& o,
                                          const float3& d, float tmin, float tmax)
{
    T.w = wt_setup(o, d);
    T.s = slab_setup(o, d, maxabs);
    T.tmin = tmin;
    T.tb = tmax;
    T.best = -1;
    T.bV = 0.0f;
    T.bW = 0.0f;
    T.bdet = 1.0f;
    T.node = 0;
    T.sp = 0;
    T.nq = 0;
    T.live = true;
}

// One inner-node visit: hit leaf children are appended to the queue (four
// straight-line writes with conditional increments; needs nq <= kLeafQueue
// - 4, so a write past the end stays inside the queue), hit inner children
// go on the stack.
template <bool ANY, bool COUNT>
__device__ __forceinline__ void visit_node(TravState& T, const QNode4* nodes, const Stack& st,
                                           const LeafQueue& q, TraceCount& cnt)
{
    const BNode4 n = load_node(nodes, T.node);
    if (COUNT) ++cnt.nodes;
    float e[4];
    int c[4];
    slab4<true>(T.s, n, T.tmin, T.tb, e, c);
    const uint32_t leaves = take_leaves(e, c);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        q.lds[T.nq * 256] = (uint32_t)~c[k];
        T.nq += (int)((leaves >> k) & 1u);
    }
    T.live = next_node<ANY ? kOrderSlot : kOrderNear>(e, c, st, T.sp, T.node);
}

// Tests every queued triangle of the active lanes (the wave runs as many
// rounds as its longest queue) and empties the queues.
template <bool ANY, bool COUNT>
__device__ __forceinline__ void test_queue(TravState& T, const float4* tris, const LeafQueue& q,
                                           TraceCount& cnt)
{
    for (int i = 0; __any(i < T.nq); ++i) {
        if (i < T.nq && !(ANY && T.best == 0)) {
            float3 v0, v1, v2;
            int prim;
            load_tri(tris, (int)q.lds[i * 256], v0, v1, v2, prim);
            if (COUNT) ++cnt.tris;
            WtHit h;
            if (ANY) {
                if (wt_test(T.w, v0, v1, v2, T.tmin, T.tb, h)) T.best = 0;
            } else if (wt_test(T.w, v0, v1, v2, T.tmin, T.tb, h) &&
                       (T.best < 0 || h.t < T.tb || prim < T.best)) {
                // t <= tbest holds on acceptance; an equal t keeps the lower prim id
                T.tb = h.t;
                T.best = prim;
                T.bV = h.V;
                T.bW = h.W;
                T.bdet = h.det;
            }
        }
    }
    T.nq = 0;
    if (ANY && T.best == 0) T.live = false;
}

// One step of a lane's ray, called by the active lanes of a wave together:
// a node visit, then (wave-uniform) the queued tests when due.  Returns true
// when the ray is finished (ANY: T.best = 0 on an accepted triangle).
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool trav_step(TravState& T, const QNode4* nodes, const float4* tris,
                                          const Stack& st, const LeafQueue& q, TraceCount& cnt)
{
    if (T.live && T.nq <= kLeafQueue - 4) visit_node<ANY, COUNT>(T, nodes, st, q, cnt);
    const bool ready = T.nq > 0 || !T.live;
    if ((__all(ready) && __any(T.nq > 0)) || __any(T.nq > kLeafQueue - 4))
        test_queue<ANY, COUNT>(T, tris, q, cnt);
    return !T.live && T.nq == 0;
}

// Sky radiance of a missed primary ray, with the REFERENCE conversions.
__device__ __forceinline__ float3 sky_for(const ShadeParams& sp, const float3& d)
{
    float3 L = sky_radiance(sp.sky, d);
    if (sp.mode == VLB_SHADE_REFERENCE) {
        if (sp.ref.flags & VLB_REF_SRGB) {
            L.x = srgb(L.x);
            L.y = srgb(L.y);
            L.z = srgb(L.z);
        }
        if (sp.ref.flags & VLB_REF_UNORM8) {
            L.x = unorm8(L.x);
            L.y = unorm8(L.y);
            L.z = unorm8(L.z);
        }
    }
    return L;
}

// grid_indirect of device_common.hpp with the corner visibility rays already
// traced: vis[v] != 0 marks corner v occluded (read only where a ray was
// needed).
__device__ __forceinline__ float3 grid_indirect_rec(const ProbeGridView& g, const float3& P,
                                                    const float3& N, const uint8_t* vis)
{
    const float p[3] = {P.x, P.y, P.z};
    const float n[3] = {N.x, N.y, N.z};
    float Y[16];
    grid_basis(g, n, Y);
    float sw = 0.0f, acc[3] = {0.0f, 0.0f, 0.0f};
    for (int v = 0; v < 8; ++v) {
        const GridCorner c = grid_corner(g, p, v);
        float w = c.w;
        if (corner_needs_ray(g, c, 1) && vis[v] != 0) w = 0.0f;
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

// Radiance of a primary ray from direction d, origin *probe, with hit
// record h (prim >= 0, t, u, v as bits; -1 miss) and the visibility bytes
// vis[j] of its shadow rays (env_map.rchit:73-101 / Lambert).
template <bool IND = false>
__device__ __forceinline__ float3 shade_record(const ShadeParams& sp, const float4* shade,
                                               const float4* probe, const float3& d,
                                               const int4& h, const uint8_t* vis)
{
    if (h.x < 0) return sky_for(sp, d);
    const float4 pp = *probe;
    const float3 o = make_float3(pp.x, pp.y, pp.z);
    const HitGeom g = hit_geom(shade, o, d, h.x, __int_as_float(h.y), __int_as_float(h.z),
                               __int_as_float(h.w));
    float3 L;
    if (sp.mode == VLB_SHADE_REFERENCE) {
        const vlb_ref_shading& rs = sp.ref;
        const ShadowRay s = shadow_ray(sp, g, d, 0);
        const bool in_shadow = s.needed ? vis[0] != 0 : true;
        float diffuse = 0.0f, spec = 0.0f;
        if (!in_shadow) {
            diffuse = rs.c_diffuse * s.cosv;
            const float dnl = dot3(g.N, s.dir);
            const float3 R = make_float3(s.dir.x - 2.0f * dnl * g.N.x, s.dir.y - 2.0f * dnl * g.N.y,
                                         s.dir.z - 2.0f * dnl * g.N.z);
            const float rd = dot3(R, d);
            spec = rs.c_specular * powf(rd > 0.0f ? rd : 0.0f, rs.glossiness);
        }
        const bool zero = g.alb.x == 0.0f && g.alb.y == 0.0f && g.alb.z == 0.0f;
        L.x = (zero ? 1.0f : g.alb.x) * (rs.ambient + diffuse + spec);
        L.y = (zero ? 1.0f : g.alb.y) * (rs.ambient + diffuse + spec);
        L.z = (zero ? 1.0f : g.alb.z) * (rs.ambient + diffuse + spec);
        if (rs.flags & VLB_REF_SRGB) {
            L.x = srgb(L.x);
            L.y = srgb(L.y);
            L.z = srgb(L.z);
        }
        if (rs.flags & VLB_REF_UNORM8) {
            L.x = unorm8(L.x);
            L.y = unorm8(L.y);
            L.z = unorm8(L.z);
        }
        return L;
    }
    float3 acc = make_float3(0.f, 0.f, 0.f);
    for (int j = 0; j < sp.nlights; ++j) {
        const ShadowRay s = shadow_ray(sp, g, d, j);
        if (s.needed && vis[j] == 0) {
            acc.x = acc.x + s.E.x * s.cosv;
            acc.y = acc.y + s.E.y * s.cosv;
            acc.z = acc.z + s.E.z * s.cosv;
        }
    }
    const float kInvPi = 0.318309886183790671538f;
    if (!IND || !sp.ind.coeffs)
        return make_float3(g.alb.x * acc.x * kInvPi, g.alb.y * acc.y * kInvPi,
                           g.alb.z * acc.z * kInvPi);
    float3 N = g.N;
    if (dot3(N, d) > 0.0f) N = make_float3(-N.x, -N.y, -N.z);
    const float3 E = grid_indirect_rec(sp.ind, g.P, N, vis + sp.nlights);
    return make_float3(g.alb.x * acc.x * kInvPi + g.alb.x * E.x,
                       g.alb.y * acc.y * kInvPi + g.alb.y * E.y,
                       g.alb.z * acc.z * kInvPi + g.alb.z * E.z);
}

}  // namespace vlb
