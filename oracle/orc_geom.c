/*
 * orc_geom.c -- oracle ray queries: watertight ray/triangle test, a CPU
 * binned-SAH BVH and brute force.  TEST INFRASTRUCTURE ONLY (see
 * vlb_oracle.h).  Compiled with -ffp-contract=off: every float operation is
 * a separately rounded IEEE operation, in the order written.
 *
 * Reference: traceRayEXT of shaders/env_map.rgen:26 (closest hit, opaque,
 * tmin 0.001, tmax 1e4) and env_map.rchit:86-87 (TerminateOnFirstHit |
 * SkipClosestHit shadow query).  The reference's arithmetic lives in the
 * Vulkan driver and RT hardware; what this file fixes is the library's
 * contract: the closest hit is the lexicographic minimum of (t, prim id)
 * over all triangles that the watertight test accepts in [tmin, tmax], and a
 * shadow query is "some triangle accepted".  Both are independent of the
 * BVH, so the GPU's LBVH and this file's SAH BVH must agree bit for bit.
 */
#include "vlb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

/* ------------------------------------------------------ watertight test */
void orc_wt_setup(const float d[3], orc_wt* w)
{
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    w->kx = kx;
    w->ky = ky;
    w->kz = kz;
    w->Sx = d[kx] / d[kz];
    w->Sy = d[ky] / d[kz];
    w->Sz = 1.0f / d[kz];
}

/* Woop/Benthin/Wald 2013, section 3 + the double fallback of section 4. */
int orc_wt_hit(const float o[3], const orc_wt* w, const float* v0, const float* v1,
               const float* v2, float tmin, float tmax, float* t_out, float* u_out,
               float* v_out)
{
    const int kx = w->kx, ky = w->ky, kz = w->kz;
    float A[3] = {v0[0] - o[0], v0[1] - o[1], v0[2] - o[2]};
    float B[3] = {v1[0] - o[0], v1[1] - o[1], v1[2] - o[2]};
    float C[3] = {v2[0] - o[0], v2[1] - o[1], v2[2] - o[2]};
    float Ax = A[kx] - w->Sx * A[kz];
    float Ay = A[ky] - w->Sy * A[kz];
    float Bx = B[kx] - w->Sx * B[kz];
    float By = B[ky] - w->Sy * B[kz];
    float Cx = C[kx] - w->Sx * C[kz];
    float Cy = C[ky] - w->Sy * C[kz];
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
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return 0;
    float det = U + V + W;
    if (det == 0.0f) return 0;
    float Az = w->Sz * A[kz];
    float Bz = w->Sz * B[kz];
    float Cz = w->Sz * C[kz];
    float T = U * Az + V * Bz + W * Cz;
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return 0;
    *t_out = t;
    *u_out = V / det;
    *v_out = W / det;
    return 1;
}

/* ------------------------------------------------------------ slab test */
/* Conservative: the interval is widened by a relative and an absolute
 * margin, so a box that holds an accepted triangle is never culled.  The
 * margins only cost extra node visits; they never change a result. */
static inline int box_hit(const float* lo, const float* hi, const float o[3],
                          const float inv[3], float tmin, float tbest, float apad)
{
    float tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        float t0 = (lo[a] - o[a]) * inv[a];
        float t1 = (hi[a] - o[a]) * inv[a];
        float n = t0 < t1 ? t0 : t1;
        float f = t0 < t1 ? t1 : t0;
        tn = n > tn ? n : tn;
        tf = f < tf ? f : tf;
    }
    float hi_t = tf < tbest ? tf : tbest;
    hi_t = hi_t + fabsf(hi_t) * 4e-6f + apad;
    float lo_t = tn - fabsf(tn) * 4e-6f - apad;
    return lo_t <= hi_t && hi_t >= tmin;
}

static inline void inv_dir(const float d[3], float inv[3])
{
    for (int a = 0; a < 3; ++a) {
        float x = d[a];
        if (x == 0.0f) x = copysignf(1e-30f, x);
        inv[a] = 1.0f / x;
    }
}

/* ------------------------------------------------------------ BVH build */
typedef struct {
    float lo[3], hi[3];
    int32_t left;   /* inner: index of left child (right = left + 1); leaf: first */
    int32_t count;  /* 0 = inner, else number of triangles                       */
} onode;

typedef struct {
    float lo[3], hi[3];
    float c[3];
    uint32_t id;
} primref;

static void box_empty(float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
}
static void box_grow(float lo[3], float hi[3], const float* plo, const float* phi)
{
    for (int a = 0; a < 3; ++a) {
        if (plo[a] < lo[a]) lo[a] = plo[a];
        if (phi[a] > hi[a]) hi[a] = phi[a];
    }
}
static double box_area(const float lo[3], const float hi[3])
{
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

#define NBINS 16
#define LEAF_MAX 4

typedef struct {
    onode* nodes;
    uint32_t count, cap;
    primref* refs;
    uint32_t depth;
} builder;

static uint32_t new_node(builder* b)
{
    if (b->count == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 1024;
        b->nodes = (onode*)realloc(b->nodes, (size_t)b->cap * sizeof(onode));
    }
    return b->count++;
}

static void build_rec(builder* b, uint32_t ni, uint32_t first, uint32_t n, uint32_t level)
{
    if (level > b->depth) b->depth = level;
    onode* node = &b->nodes[ni];
    float lo[3], hi[3], clo[3], chi[3];
    box_empty(lo, hi);
    box_empty(clo, chi);
    for (uint32_t i = first; i < first + n; ++i) {
        box_grow(lo, hi, b->refs[i].lo, b->refs[i].hi);
        box_grow(clo, chi, b->refs[i].c, b->refs[i].c);
    }
    memcpy(node->lo, lo, sizeof lo);
    memcpy(node->hi, hi, sizeof hi);
    if (n <= LEAF_MAX) {
        node->left = (int32_t)first;
        node->count = (int32_t)n;
        return;
    }
    /* binned SAH over centroid bounds */
    double best_cost = INFINITY;
    int best_axis = -1, best_split = 0;
    for (int a = 0; a < 3; ++a) {
        float ext = chi[a] - clo[a];
        if (!(ext > 0.0f)) continue;
        uint32_t cnt[NBINS] = {0};
        float blo[NBINS][3], bhi[NBINS][3];
        for (int k = 0; k < NBINS; ++k) box_empty(blo[k], bhi[k]);
        for (uint32_t i = first; i < first + n; ++i) {
            int k = (int)((b->refs[i].c[a] - clo[a]) / ext * NBINS);
            if (k >= NBINS) k = NBINS - 1;
            if (k < 0) k = 0;
            cnt[k]++;
            box_grow(blo[k], bhi[k], b->refs[i].lo, b->refs[i].hi);
        }
        double left_area[NBINS];
        uint32_t left_cnt[NBINS];
        float l_lo[3], l_hi[3];
        box_empty(l_lo, l_hi);
        uint32_t lc = 0;
        for (int k = 0; k < NBINS - 1; ++k) {
            box_grow(l_lo, l_hi, blo[k], bhi[k]);
            lc += cnt[k];
            left_area[k] = box_area(l_lo, l_hi);
            left_cnt[k] = lc;
        }
        float r_lo[3], r_hi[3];
        box_empty(r_lo, r_hi);
        uint32_t rc = 0;
        for (int k = NBINS - 1; k > 0; --k) {
            box_grow(r_lo, r_hi, blo[k], bhi[k]);
            rc += cnt[k];
            uint32_t lcount = left_cnt[k - 1];
            if (lcount == 0 || rc == 0) continue;
            double cost = left_area[k - 1] * lcount + box_area(r_lo, r_hi) * rc;
            if (cost < best_cost) {
                best_cost = cost;
                best_axis = a;
                best_split = k;
            }
        }
    }
    uint32_t mid;
    if (best_axis < 0) {
        mid = first + n / 2; /* all centroids equal: split by index */
    } else {
        int a = best_axis;
        float ext = chi[a] - clo[a];
        uint32_t i = first, j = first + n;
        while (i < j) {
            int k = (int)((b->refs[i].c[a] - clo[a]) / ext * NBINS);
            if (k >= NBINS) k = NBINS - 1;
            if (k < 0) k = 0;
            if (k < best_split) {
                ++i;
            } else {
                --j;
                primref t = b->refs[i];
                b->refs[i] = b->refs[j];
                b->refs[j] = t;
            }
        }
        mid = i;
        if (mid == first || mid == first + n) mid = first + n / 2;
    }
    uint32_t l = new_node(b);
    uint32_t r = new_node(b);
    (void)r;
    node = &b->nodes[ni];
    node->left = (int32_t)l;
    node->count = 0;
    build_rec(b, l, first, mid - first, level + 1);
    build_rec(b, l + 1, mid, first + n - mid, level + 1);
}

/* ---------------------------------------------------------------- scene */
orc_scene* orc_scene_create(const float* xyz, uint32_t nverts, const uint32_t* idx,
                            uint32_t ntris, const float* normals, const float* albedo,
                            int bvh)
{
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->ntris = ntris;
    s->tri = (float*)malloc((size_t)ntris * 9 * sizeof(float));
    s->nrm = (float*)malloc((size_t)ntris * 9 * sizeof(float));
    s->albedo = (float*)malloc((size_t)ntris * 3 * sizeof(float));
    float maxabs = 0.0f;
    for (uint32_t i = 0; i < ntris; ++i) {
        for (int k = 0; k < 3; ++k) {
            uint32_t vi = idx[3 * i + k];
            if (vi >= nverts) vi = 0;
            for (int a = 0; a < 3; ++a) {
                float x = xyz[3 * vi + a];
                s->tri[9 * i + 3 * k + a] = x;
                if (fabsf(x) > maxabs) maxabs = fabsf(x);
            }
        }
        const float* v0 = &s->tri[9 * i];
        const float* v1 = v0 + 3;
        const float* v2 = v0 + 6;
        if (normals) {
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a)
                    s->nrm[9 * i + 3 * k + a] = normals[3 * idx[3 * i + k] + a];
        } else {
            float g[3];
            orc_geometric_normal(v0, v1, v2, g);
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) s->nrm[9 * i + 3 * k + a] = g[a];
        }
        for (int a = 0; a < 3; ++a) s->albedo[3 * i + a] = albedo ? albedo[3 * i + a] : 1.0f;
    }
    s->apad = maxabs * 64.0f * FLT_EPSILON;
    if (bvh && ntris > 0) {
        builder b = {0};
        b.refs = (primref*)malloc((size_t)ntris * sizeof(primref));
        for (uint32_t i = 0; i < ntris; ++i) {
            primref* r = &b.refs[i];
            box_empty(r->lo, r->hi);
            for (int k = 0; k < 3; ++k) box_grow(r->lo, r->hi, &s->tri[9 * i + 3 * k], &s->tri[9 * i + 3 * k]);
            for (int a = 0; a < 3; ++a) r->c[a] = 0.5f * (r->lo[a] + r->hi[a]);
            r->id = i;
        }
        uint32_t root = new_node(&b);
        build_rec(&b, root, 0, ntris, 1);
        s->nodes = b.nodes;
        s->nnodes = b.count;
        s->depth = b.depth;
        s->order = (uint32_t*)malloc((size_t)ntris * sizeof(uint32_t));
        for (uint32_t i = 0; i < ntris; ++i) s->order[i] = b.refs[i].id;
        free(b.refs);
    }
    return s;
}

void orc_scene_free(orc_scene* s)
{
    if (!s) return;
    free(s->tri);
    free(s->nrm);
    free(s->albedo);
    free(s->nodes);
    free(s->order);
    free(s);
}

uint32_t orc_scene_bvh_depth(const orc_scene* s) { return s ? s->depth : 0; }

void orc_geometric_normal(const float* v0, const float* v1, const float* v2, float g[3])
{
    float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                  e1[0] * e2[1] - e1[1] * e2[0]};
    float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0.0f) {
        g[0] = n[0] / l;
        g[1] = n[1] / l;
        g[2] = n[2] / l;
    } else {
        g[0] = 0.0f;
        g[1] = 1.0f;
        g[2] = 0.0f;
    }
}

/* ----------------------------------------------------------- traversal */
static inline void test_tri(const orc_scene* s, uint32_t prim, const float o[3],
                            const orc_wt* w, float tmin, float* tbest, int* best,
                            float* bu, float* bv)
{
    const float* v = &s->tri[9 * prim];
    float t, u, vv;
    if (orc_wt_hit(o, w, v, v + 3, v + 6, tmin, *tbest, &t, &u, &vv)) {
        /* t <= tbest holds here; an equal t keeps the lower prim id */
        if (*best < 0 || t < *tbest || (int)prim < *best) {
            *tbest = t;
            *best = (int)prim;
            *bu = u;
            *bv = vv;
        }
    }
}

int orc_trace_closest(const orc_scene* s, const float o[3], const float d[3], float tmin,
                      float tmax, int brute, float* t_out, float* u_out, float* v_out)
{
    orc_wt w;
    orc_wt_setup(d, &w);
    float tbest = tmax, bu = 0.0f, bv = 0.0f;
    int best = -1;
    if (brute || !s->nodes) {
        for (uint32_t i = 0; i < s->ntris; ++i) test_tri(s, i, o, &w, tmin, &tbest, &best, &bu, &bv);
    } else {
        float inv[3];
        inv_dir(d, inv);
        uint32_t stack[256];
        int sp = 0;
        stack[sp++] = 0;
        const onode* nodes = (const onode*)s->nodes;
        while (sp > 0) {
            const onode* n = &nodes[stack[--sp]];
            if (!box_hit(n->lo, n->hi, o, inv, tmin, tbest, s->apad)) continue;
            if (n->count > 0) {
                for (int32_t i = 0; i < n->count; ++i)
                    test_tri(s, s->order[n->left + i], o, &w, tmin, &tbest, &best, &bu, &bv);
            } else {
                stack[sp++] = (uint32_t)n->left + 1;
                stack[sp++] = (uint32_t)n->left;
            }
        }
    }
    if (best >= 0) {
        *t_out = tbest;
        *u_out = bu;
        *v_out = bv;
    }
    return best;
}

int orc_trace_occluded(const orc_scene* s, const float o[3], const float d[3], float tmin,
                       float tmax, int brute)
{
    orc_wt w;
    orc_wt_setup(d, &w);
    float t, u, v;
    if (brute || !s->nodes) {
        for (uint32_t i = 0; i < s->ntris; ++i) {
            const float* p = &s->tri[9 * i];
            if (orc_wt_hit(o, &w, p, p + 3, p + 6, tmin, tmax, &t, &u, &v)) return 1;
        }
        return 0;
    }
    float inv[3];
    inv_dir(d, inv);
    uint32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    const onode* nodes = (const onode*)s->nodes;
    while (sp > 0) {
        const onode* n = &nodes[stack[--sp]];
        if (!box_hit(n->lo, n->hi, o, inv, tmin, tmax, s->apad)) continue;
        if (n->count > 0) {
            for (int32_t i = 0; i < n->count; ++i) {
                const float* p = &s->tri[9 * s->order[n->left + i]];
                if (orc_wt_hit(o, &w, p, p + 3, p + 6, tmin, tmax, &t, &u, &v)) return 1;
            }
        } else {
            stack[sp++] = (uint32_t)n->left + 1;
            stack[sp++] = (uint32_t)n->left;
        }
    }
    return 0;
}

void orc_trace_closest_batch(const orc_scene* s, const float* org, const float* dir,
                             uint32_t n, float tmin, float tmax, int brute, int32_t* prim,
                             float* t, float* u, float* v)
{
    for (uint32_t i = 0; i < n; ++i) {
        float tt = 0.0f, uu = 0.0f, vv = 0.0f;
        int p = orc_trace_closest(s, &org[3 * i], &dir[3 * i], tmin, tmax, brute, &tt, &uu, &vv);
        prim[i] = p;
        if (t) t[i] = p >= 0 ? tt : 0.0f;
        if (u) u[i] = p >= 0 ? uu : 0.0f;
        if (v) v[i] = p >= 0 ? vv : 0.0f;
    }
}

void orc_trace_occluded_batch(const orc_scene* s, const float* org, const float* dir,
                              const float* tmax, uint32_t n, float tmin, int brute,
                              uint8_t* occluded)
{
    for (uint32_t i = 0; i < n; ++i)
        occluded[i] = (uint8_t)orc_trace_occluded(s, &org[3 * i], &dir[3 * i], tmin, tmax[i], brute);
}
