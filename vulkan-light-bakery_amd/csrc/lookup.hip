// lookup.hip -- consumer side of a baked probe grid (vlb_probe_lookup):
// shaders/main.rchit:124-168 (cell, 8 corner probes, shadow ray to each,
// weight, normalised sum) with shaders/sh.rmiss:23-39 (SH at the normal), in
// the form that reads the grid as it was baked (origin, counts, probe order,
// K, SH frame (x, z, y)).
//
// One lane per (point, corner): the 8 lanes of a point sit in one wave and
// combine their weighted values with three xor shuffles in a fixed order.
// Cell, weights and basis come from probe_grid.hpp.
// The visibility rays use the same operation sequence as the CPU oracle, so
// which probes count is bit-identical; values are FP32.
#include "vlb_internal.hpp"

namespace vlb {

namespace {

__global__ __launch_bounds__(256) void k_lookup(LookupArgs a)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    Stack st;
    st.lds = &lds[threadIdx.x];
    st.spill = a.spill + (size_t)blockIdx.x * 256 + threadIdx.x;
    st.spill_stride = a.spill_stride;
    st.depth = kLdsDepth;
    TraceCount cnt;
    const ProbeGridView& g = a.g;
    const size_t total = (size_t)a.n * 8;
    for (size_t base = (size_t)blockIdx.x * 256; base < total; base += (size_t)gridDim.x * 256) {
        const size_t t = base + threadIdx.x;
        const bool live = t < total;
        const int i = live ? (int)(t >> 3) : 0;
        const int v = (int)(t & 7);
        float w = 0.0f, val[3] = {0.0f, 0.0f, 0.0f};
        if (live) {
            const float p[3] = {a.pts[3 * i], a.pts[3 * i + 1], a.pts[3 * i + 2]};
            const float nn[3] = {a.nrm[3 * i], a.nrm[3 * i + 1], a.nrm[3 * i + 2]};
            const GridCorner c = grid_corner(g, p, v);
            w = c.w;
            if (corner_needs_ray(g, c, a.sc.ntris)) {
                const float3 org = make_float3(p[0] + g.bias * nn[0], p[1] + g.bias * nn[1],
                                               p[2] + g.bias * nn[2]);
                const float3 d = make_float3(c.dir[0] / c.dist, c.dir[1] / c.dist, c.dir[2] / c.dist);
                if (trace_any<false>(a.sc, org, d, 0.0f, c.dist, st, cnt)) w = 0.0f;
            }
            if (w > 0.0f) {
                float Y[16];
                grid_basis(g, nn, Y);
                corner_value(g, c.row, Y, val);
                val[0] *= w;
                val[1] *= w;
                val[2] *= w;
            }
        }
        // the point's 8 lanes: fixed-order butterfly
        float sw = w;
        for (int m = 1; m < 8; m <<= 1) {
            sw += __shfl_xor(sw, m);
            val[0] += __shfl_xor(val[0], m);
            val[1] += __shfl_xor(val[1], m);
            val[2] += __shfl_xor(val[2], m);
        }
        if (live && v == 0) {
            a.out[3 * i + 0] = sw > 0.0f ? val[0] / sw : 0.0f;
            a.out[3 * i + 1] = sw > 0.0f ? val[1] / sw : 0.0f;
            a.out[3 * i + 2] = sw > 0.0f ? val[2] / sw : 0.0f;
        }
    }
}

}  // namespace

int lookup_grid_size(Context& ctx)
{
    return ctx.num_cus * 4;
}

void launch_lookup(const LookupArgs& a, int grid, hipStream_t st)
{
    k_lookup<<<grid, 256, 0, st>>>(a);
    VLB_HIP(hipGetLastError());
}

}  // namespace vlb
