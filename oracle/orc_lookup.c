/*
 * orc_lookup.c -- oracle of the probe-grid lookup (TEST INFRASTRUCTURE ONLY;
 * see vlb_oracle.h).
 *
 * Restates shaders/main.rchit:124-168 (cell, 8 corner probes, shadow ray to
 * each, weight |gridStep| - distance, normalised sum) and
 * shaders/sh.rmiss:23-39 (sum_k c_k Y_k(normal)) in the fixed form of
 * vlb_probe_lookup (include/vlb_bake.h): the grid as baked (origin, counts,
 * probe order, K) and the bake's SH frame (x, z, y).  Geometry of the
 * visibility rays in float with the operation order of the HIP kernel, so
 * visibility decisions are bit-identical; SH evaluation and sums in double.
 */
#include <math.h>
#include <string.h>

#include "orc_internal.h"
#include "vlb_oracle.h"

/* index of grid point (x, y, z) in the buffer order */
static size_t grid_index(const vlb_probe_grid_desc* g, uint32_t x, uint32_t y, uint32_t z)
{
    const size_t cx = g->count[0], cy = g->count[1], cz = g->count[2];
    if (g->order == VLB_ORDER_REFERENCE_BFS) {
        /* light_baker.cpp:80-101: x-run first, then every point of it spawns
         * its y successors, then every point of the plane its z successors */
        const size_t r0 = y == 0 ? x : cx + (size_t)x * (cy - 1) + (y - 1);
        return z == 0 ? r0 : cx * cy + r0 * (cz - 1) + (z - 1);
    }
    return x + cx * (y + cy * (size_t)z);
}

static uint32_t clampu(long v, uint32_t hi)
{
    if (v < 0) return 0;
    if ((unsigned long)v > hi) return hi;
    return (uint32_t)v;
}

/* One point: the value of orc_probe_lookup at p with normal nn.  Returns
 * the number of visibility rays traced. */
int orc_lookup_point(const orc_scene* s, const vlb_probe_grid_desc* g, const float* coeffs,
                     const vlb_lookup_params* lp, const float p[3], const float nn[3],
                     double out[3])
{
    const int K = (int)g->num_coefs;
    const int exact = g->sh_constants == VLB_SH_EXACT;
    /* band factors A_l / pi of the clamped-cosine kernel */
    const double band[4] = {1.0, 2.0 / 3.0, 0.25, 0.0};
    const float wmax = sqrtf(g->step[0] * g->step[0] + g->step[1] * g->step[1] +
                             g->step[2] * g->step[2]);
    int rays = 0;
    uint32_t base[3];
    float fr[3];
    for (int a = 0; a < 3; ++a) {
        const float f = (p[a] - g->origin[a]) / g->step[a];
        const uint32_t hi = g->count[a] >= 2 ? g->count[a] - 2 : 0;
        base[a] = clampu((long)floorf(f), hi);
        float t = f - (float)base[a];
        fr[a] = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    }
    const float org[3] = {p[0] + lp->bias * nn[0], p[1] + lp->bias * nn[1],
                          p[2] + lp->bias * nn[2]};
    double Y[16];
    const double dsh[3] = {nn[0], nn[2], nn[1]};
    orc_sh_eval(K, exact, dsh, Y);
    double acc[3] = {0.0, 0.0, 0.0}, wsum = 0.0;
    for (int v = 0; v < 8; ++v) {
        /* corner order of main.rchit:128-137: bit 2 = x, bit 1 = y, bit 0 = z */
        const uint32_t off[3] = {(uint32_t)(v >> 2) & 1u, (uint32_t)(v >> 1) & 1u, (uint32_t)v & 1u};
        uint32_t gi[3];
        float q[3], dir[3];
        for (int a = 0; a < 3; ++a) {
            gi[a] = base[a] + off[a] < g->count[a] ? base[a] + off[a] : g->count[a] - 1;
            q[a] = g->origin[a] + g->step[a] * (float)gi[a];
            dir[a] = q[a] - p[a];
        }
        const float dist = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        float w;
        if (lp->mode == VLB_LOOKUP_REFERENCE) {
            w = wmax - dist;
            if (w < 0.0f) w = 0.0f;
        } else {
            w = (off[0] ? fr[0] : 1.0f - fr[0]) * (off[1] ? fr[1] : 1.0f - fr[1]) *
                (off[2] ? fr[2] : 1.0f - fr[2]);
        }
        if (!(w > 0.0f)) continue;
        if (lp->visibility && dist > 0.0f && s && s->ntris > 0) {
            const float d[3] = {dir[0] / dist, dir[1] / dist, dir[2] / dist};
            ++rays;
            if (orc_trace_occluded(s, org, d, 0.0f, dist, 0)) continue;
        }
        const float* c = &coeffs[grid_index(g, gi[0], gi[1], gi[2]) * (size_t)K * 3];
        for (int k = 0; k < K; ++k) {
            const int l = k < 1 ? 0 : (k < 4 ? 1 : (k < 9 ? 2 : 3));
            const double y = lp->mode == VLB_LOOKUP_IRRADIANCE ? band[l] * Y[k] : Y[k];
            for (int ch = 0; ch < 3; ++ch) acc[ch] += (double)w * (double)c[3 * k + ch] * y;
        }
        wsum += (double)w;
    }
    for (int ch = 0; ch < 3; ++ch) out[ch] = wsum > 0.0 ? acc[ch] / wsum : 0.0;
    return rays;
}

int orc_check_grid(const vlb_probe_grid_desc* g)
{
    if (!g || (g->num_coefs != 9 && g->num_coefs != 16)) return -1;
    if (g->count[0] == 0 || g->count[1] == 0 || g->count[2] == 0) return -1;
    return 0;
}

int orc_probe_lookup(const orc_scene* s, const vlb_probe_grid_desc* g, const float* coeffs,
                     const vlb_lookup_params* lp, const float* pts, const float* nrm, uint32_t n,
                     double* out)
{
    if (!g || !coeffs || !lp || !pts || !nrm || !out) return -1;
    if (orc_check_grid(g)) return -1;
    for (uint32_t i = 0; i < n; ++i)
        orc_lookup_point(s, g, coeffs, lp, &pts[3 * (size_t)i], &nrm[3 * (size_t)i], &out[3 * (size_t)i]);
    return 0;
}
