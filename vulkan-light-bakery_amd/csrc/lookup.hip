// lookup.hip -- consumer side of a baked probe grid (vlb_probe_lookup):
// shaders/main.rchit:124-168 (cell, 8 corner probes, shadow ray to each,
// weight, normalised sum) with shaders/sh.rmiss:23-39 (SH at the normal), in
// the form that reads the grid as it was baked (origin, counts, probe order,
// K, SH frame (x, z, y)).
//
// One lane per (point, corner): the 8 lanes of a point sit in one wave and
// combine their weighted values with three xor shuffles in a fixed order.
// The visibility rays use the same operation sequence as the CPU oracle, so
// which probes count is bit-identical; values are FP32.
#include "sh_eval.hpp"

namespace vlb {

namespace {

__device__ __forceinline__ size_t grid_index(const LookupArgs& a, uint32_t x, uint32_t y, uint32_t z)
{
    const size_t cx = a.count[0], cy = a.count[1], cz = a.count[2];
    if (a.order == VLB_ORDER_REFERENCE_BFS) {
        // light_baker.cpp:80-101 breadth-first order
        const size_t r0 = y == 0 ? x : cx + (size_t)x * (cy - 1) + (y - 1);
        return z == 0 ? r0 : cx * cy + r0 * (cz - 1) + (z - 1);
    }
    return x + cx * (y + cy * (size_t)z);
}

__global__ __launch_bounds__(256) void k_lookup(LookupArgs a)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    Stack st;
    st.lds = &lds[threadIdx.x];
    st.spill = a.spill + (size_t)blockIdx.x * 256 + threadIdx.x;
    st.spill_stride = a.spill_stride;
    TraceCount cnt;
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
            // corner order of main.rchit:128-137: bit 2 = x, bit 1 = y, bit 0 = z
            const uint32_t off[3] = {(uint32_t)(v >> 2) & 1u, (uint32_t)(v >> 1) & 1u, (uint32_t)v & 1u};
            uint32_t gi[3];
            float dir[3], fr[3];
            for (int ax = 0; ax < 3; ++ax) {
                const float f = (p[ax] - a.origin[ax]) / a.step[ax];
                const int hi = a.count[ax] >= 2 ? (int)a.count[ax] - 2 : 0;
                const float fl = floorf(f);
                const int b = fl < 0.0f ? 0 : (fl > (float)hi ? hi : (int)fl);
                const float tt = f - (float)b;
                fr[ax] = tt < 0.0f ? 0.0f : (tt > 1.0f ? 1.0f : tt);
                gi[ax] = (uint32_t)b + off[ax] < a.count[ax] ? (uint32_t)b + off[ax] : a.count[ax] - 1;
                const float q = a.origin[ax] + a.step[ax] * (float)gi[ax];
                dir[ax] = q - p[ax];
            }
            const float dist = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
            if (a.mode == VLB_LOOKUP_REFERENCE) {
                w = a.wmax - dist;
                if (w < 0.0f) w = 0.0f;
            } else {
                w = (off[0] ? fr[0] : 1.0f - fr[0]) * (off[1] ? fr[1] : 1.0f - fr[1]) *
                    (off[2] ? fr[2] : 1.0f - fr[2]);
            }
            if (w > 0.0f && a.vis && dist > 0.0f && a.sc.ntris > 0) {
                const float3 org = make_float3(p[0] + a.bias * nn[0], p[1] + a.bias * nn[1],
                                               p[2] + a.bias * nn[2]);
                const float3 d = make_float3(dir[0] / dist, dir[1] / dist, dir[2] / dist);
                if (trace_any<false>(a.sc, org, d, 0.0f, dist, st, cnt)) w = 0.0f;
            }
            if (w > 0.0f) {
                float Y[16];
                if (a.exact) {
                    if (a.K > 9) sh_exact<16>(nn[0], nn[2], nn[1], Y);
                    else sh_exact<9>(nn[0], nn[2], nn[1], Y);
                } else {
                    sh_literal16(nn[0], nn[2], nn[1], Y);
                }
                const float* c = a.coeffs + grid_index(a, gi[0], gi[1], gi[2]) * (size_t)a.K * 3;
                const bool irr = a.mode == VLB_LOOKUP_IRRADIANCE;
                for (int k = 0; k < a.K; ++k) {
                    const int l = k < 1 ? 0 : (k < 4 ? 1 : (k < 9 ? 2 : 3));
                    const float band = l == 0 ? 1.0f : (l == 1 ? 2.0f / 3.0f : (l == 2 ? 0.25f : 0.0f));
                    const float y = irr ? band * Y[k] : Y[k];
                    val[0] += c[3 * k + 0] * y;
                    val[1] += c[3 * k + 1] * y;
                    val[2] += c[3 * k + 2] * y;
                }
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
