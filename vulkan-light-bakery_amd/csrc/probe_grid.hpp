// probe_grid.hpp -- a baked probe grid as the device reads it: the cell of a
// point, its 8 corner probes with their weights and coefficient rows, and the
// SH sum at a normal (shaders/main.rchit:124-168 with shaders/sh.rmiss:23-39,
// in the fixed form of vlb_probe_lookup: the grid as baked, with its origin,
// counts, probe order, K and the bake's SH frame (x, z, y)).  Shared by the
// lookup kernel and by the indirect term of a multi-bounce bake, so both
// decide weights and corner visibility rays with the same operations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vlb_bake.h"
#include "sh_eval.hpp"

namespace vlb {

struct ProbeGridView {
    const float* coeffs;   // [P][K][3] in the grid's order; null: no grid
    float origin[3], step[3];
    uint32_t count[3];
    int order, K, exact, mode, vis;
    float bias, wmax;      // visibility-ray offset, |step| (REFERENCE weights)
    float scale;           // factor on the looked-up value (bake's indirect term)
};

__device__ __forceinline__ size_t grid_index(const ProbeGridView& g, uint32_t x, uint32_t y,
                                             uint32_t z)
{
    const size_t cx = g.count[0], cy = g.count[1], cz = g.count[2];
    if (g.order == VLB_ORDER_REFERENCE_BFS) {
        // light_baker.cpp:80-101 breadth-first order
        const size_t r0 = y == 0 ? x : cx + (size_t)x * (cy - 1) + (y - 1);
        return z == 0 ? r0 : cx * cy + r0 * (cz - 1) + (z - 1);
    }
    return x + cx * (y + cy * (size_t)z);
}

// Corner v of the cell holding p: its weight (REFERENCE: |step| - distance,
// clamped at 0; IRRADIANCE: trilinear), the vector to its probe and that
// vector's length, and the probe's coefficient row.
struct GridCorner {
    float w, dist;
    float dir[3];
    size_t row;
};

__device__ __forceinline__ GridCorner grid_corner(const ProbeGridView& g, const float p[3], int v)
{
    // corner order of main.rchit:128-137: bit 2 = x, bit 1 = y, bit 0 = z
    const uint32_t off[3] = {(uint32_t)(v >> 2) & 1u, (uint32_t)(v >> 1) & 1u, (uint32_t)v & 1u};
    GridCorner c;
    uint32_t gi[3];
    float fr[3];
    for (int ax = 0; ax < 3; ++ax) {
        const float f = (p[ax] - g.origin[ax]) / g.step[ax];
        const int hi = g.count[ax] >= 2 ? (int)g.count[ax] - 2 : 0;
        const float fl = floorf(f);
        const int b = fl < 0.0f ? 0 : (fl > (float)hi ? hi : (int)fl);
        const float tt = f - (float)b;
        fr[ax] = tt < 0.0f ? 0.0f : (tt > 1.0f ? 1.0f : tt);
        gi[ax] = (uint32_t)b + off[ax] < g.count[ax] ? (uint32_t)b + off[ax] : g.count[ax] - 1;
        const float q = g.origin[ax] + g.step[ax] * (float)gi[ax];
        c.dir[ax] = q - p[ax];
    }
    c.dist = sqrtf(c.dir[0] * c.dir[0] + c.dir[1] * c.dir[1] + c.dir[2] * c.dir[2]);
    if (g.mode == VLB_LOOKUP_REFERENCE) {
        c.w = g.wmax - c.dist;
        if (c.w < 0.0f) c.w = 0.0f;
    } else {
        c.w = (off[0] ? fr[0] : 1.0f - fr[0]) * (off[1] ? fr[1] : 1.0f - fr[1]) *
              (off[2] ? fr[2] : 1.0f - fr[2]);
    }
    c.row = grid_index(g, gi[0], gi[1], gi[2]);
    return c;
}

// Basis at normal n in the bake's SH frame, times the band factors A_l / pi
// of the clamped-cosine kernel (1, 2/3, 1/4, 0) when the mode is IRRADIANCE.
__device__ __forceinline__ void grid_basis(const ProbeGridView& g, const float n[3], float Y[16])
{
    if (g.exact) {
        if (g.K > 9) sh_exact<16>(n[0], n[2], n[1], Y);
        else sh_exact<9>(n[0], n[2], n[1], Y);
    } else {
        sh_literal16(n[0], n[2], n[1], Y);
    }
    if (g.mode == VLB_LOOKUP_IRRADIANCE) {
        for (int k = 0; k < g.K; ++k) {
            const int l = k < 1 ? 0 : (k < 4 ? 1 : (k < 9 ? 2 : 3));
            Y[k] *= l == 0 ? 1.0f : (l == 1 ? 2.0f / 3.0f : (l == 2 ? 0.25f : 0.0f));
        }
    }
}

// sum_k c_k Y[k] of one probe row
__device__ __forceinline__ void corner_value(const ProbeGridView& g, size_t row, const float Y[16],
                                             float val[3])
{
    const float* c = g.coeffs + row * (size_t)g.K * 3;
    val[0] = val[1] = val[2] = 0.0f;
    for (int k = 0; k < g.K; ++k) {
        val[0] += c[3 * k + 0] * Y[k];
        val[1] += c[3 * k + 1] * Y[k];
        val[2] += c[3 * k + 2] * Y[k];
    }
}

// Whether corner c of a point needs a visibility ray.
__device__ __forceinline__ bool corner_needs_ray(const ProbeGridView& g, const GridCorner& c,
                                                 int ntris)
{
    return g.vis && c.w > 0.0f && c.dist > 0.0f && ntris > 0;
}

}  // namespace vlb
