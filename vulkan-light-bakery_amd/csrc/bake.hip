// bake.hip -- the probe-baking hot path on gfx950.
//
// One fused, persistent kernel per bake replaces, per probe, the reference's
//   vkCmdTraceRaysKHR(3141x1000)  env_map.rgen -> env_map.rchit/shadow.rmiss
//                                 /main.rmiss -> RGBA8 image      (getMap,
//                                 src/baker/env_map_generator.cpp:325-359)
//   vkCmdDispatch(197x63) sh.comp (dispatchBakingKernel,
//                                 src/baker/light_baker.cpp:269-285)
//   mapMemory + memcpy 192 B      (light_baker.cpp:317-322)
// and its merge kernel replaces the host read-back.
//
// Work decomposition.  A tile is 16 probes x one slice of RS rays.  A
// 256-thread workgroup (4 waves) owns a tile; the waves split its 32-ray
// chunks round-robin.  In a wave, lane l traces for probe (l & 15) the rays
// 8*(l >> 4) + e, e = 0..7, of the chunk -- exactly the operand layout of
// the 16x16 MFMA shapes, so the SH projection
//     C[probe][coef] += sum_ray radiance[probe][ray] * Yw[ray][coef]
// is fed straight from registers: FP32 uses v_mfma_f32_16x16x4_f32 (8 per
// chunk and channel), BF16/FP16 use v_mfma_f32_16x16x32_{bf16,f16} (1 per
// chunk and channel) with FP32 accumulation.  Radiance never reaches HBM.
// Chunk results are Kahan-summed in FP32 across chunks and waves; each
// (slice, probe) writes one (sum, compensation) partial, and the merge
// kernel adds the slices in FP64 in a fixed order (deterministic; no atomics
// on the result path -- the reference's sh.comp:39 "+=" is a data race).
//
// Workgroups walk the tile list XCD by XCD: workgroup b runs on XCD b % 8
// (round-robin dispatch), and XCD x owns the x-th contiguous eighth of the
// tiles, so the slices of one probe group share that XCD's L2.
#include "vlb_internal.hpp"

namespace vlb {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

template <int PREC, bool RECORD, bool COUNT>
__global__ __launch_bounds__(256) void k_bake(BakeArgs a)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = lane >> 4;   // ray group (MFMA k-block)
    const int pl = lane & 15;  // probe row of A, coefficient column of B
    Stack st;
    st.lds = &lds[tid];
    st.spill = a.spill + (size_t)blockIdx.x * 256 + tid;
    st.spill_stride = a.spill_stride;
    TraceCount cnt;
    uint32_t nshadow = 0;

    const int G8 = gridDim.x >> 3;
    const int xcd = blockIdx.x & 7;
    const int local = blockIdx.x >> 3;
    const int per_xcd = (a.ntiles + 7) >> 3;
    const int nchunks = a.RS >> 5;

    for (int tl = local; tl < per_xcd; tl += G8) {
        const int tile = xcd * per_xcd + tl;
        if (tile >= a.ntiles) break;
        const int pg = tile / a.nslices;
        const int sl = tile - pg * a.nslices;
        const int probe = pg * 16 + pl;
        const bool pvalid = probe < a.P;
        const float4 pp = a.probes[probe];
        const float3 o = make_float3(pp.x, pp.y, pp.z);

        v4f S[3], Cp[3];
        for (int c = 0; c < 3; ++c) {
            S[c] = v4f{0.f, 0.f, 0.f, 0.f};
            Cp[c] = v4f{0.f, 0.f, 0.f, 0.f};
        }
        for (int ch = wave; ch < nchunks; ch += 4) {
            const int rbase = sl * a.RS + ch * 32 + 8 * g;  // first ray of this lane group
            v4f C[3];
            for (int c = 0; c < 3; ++c) C[c] = v4f{0.f, 0.f, 0.f, 0.f};
            // Radiance of the lane's 8 rays, kept in registers for the
            // 16x16x32 shapes (static indices only: no scratch).
            float lx[8], ly[8], lz[8];
#pragma unroll 1
            for (int e = 0; e < 8; ++e) {
                const int r = rbase + e;
                const float4 dd = a.dirs[r];
                float3 L = make_float3(0.f, 0.f, 0.f);
                int hp = -1;
                if (pvalid && dd.w != 0.0f) {
                    const float3 d = make_float3(dd.x, dd.y, dd.z);
                    L = shade_ray<COUNT>(a.sc, a.sp, o, d, a.tmin, a.tmax, st, cnt, nshadow, hp);
                    if (RECORD) {
                        const int op = a.probe_perm[probe];
                        const int cr = a.dir_perm[r];
                        const size_t ri = (size_t)op * a.R + cr;
                        if (a.rad_out) {
                            a.rad_out[3 * ri + 0] = L.x;
                            a.rad_out[3 * ri + 1] = L.y;
                            a.rad_out[3 * ri + 2] = L.z;
                        }
                        if (a.hit_out) a.hit_out[ri] = hp;
                    }
                }
                if (PREC == VLB_PREC_FP32) {
                    const float* yt = static_cast<const float*>(a.ytab);
                    const float y = yt[(((size_t)(r >> 3) * 16) + pl) * 8 + (r & 7)];
                    C[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(L.x, y, C[0], 0, 0, 0);
                    C[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(L.y, y, C[1], 0, 0, 0);
                    C[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(L.z, y, C[2], 0, 0, 0);
                } else {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (q == e) {
                            lx[q] = L.x;
                            ly[q] = L.y;
                            lz[q] = L.z;
                        }
                }
            }
            if (PREC == VLB_PREC_BF16) {
                const v8bf* yt = static_cast<const v8bf*>(a.ytab);
                const v8bf y = yt[(size_t)(rbase >> 3) * 16 + pl];
                v8bf A0, A1, A2;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    A0[e] = (__bf16)lx[e];
                    A1[e] = (__bf16)ly[e];
                    A2[e] = (__bf16)lz[e];
                }
                C[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, y, C[0], 0, 0, 0);
                C[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, y, C[1], 0, 0, 0);
                C[2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2, y, C[2], 0, 0, 0);
            } else if (PREC == VLB_PREC_FP16) {
                const v8h* yt = static_cast<const v8h*>(a.ytab);
                const v8h y = yt[(size_t)(rbase >> 3) * 16 + pl];
                v8h A0, A1, A2;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    A0[e] = (_Float16)lx[e];
                    A1[e] = (_Float16)ly[e];
                    A2[e] = (_Float16)lz[e];
                }
                C[0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A0, y, C[0], 0, 0, 0);
                C[1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1, y, C[1], 0, 0, 0);
                C[2] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A2, y, C[2], 0, 0, 0);
            }
            // Kahan: value = S - Cp
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float y = C[c][i] - Cp[c][i];
                    const float t = S[c][i] + y;
                    Cp[c][i] = (t - S[c][i]) - y;
                    S[c][i] = t;
                }
        }
        // Cross-wave combine through LDS (the stack area is free now).
        __syncthreads();
        float* red = reinterpret_cast<float*>(lds);
        if (wave > 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    red[((wave - 1) * 24 + c * 4 + i) * 64 + lane] = S[c][i];
                    red[((wave - 1) * 24 + 12 + c * 4 + i) * 64 + lane] = Cp[c][i];
                }
        }
        __syncthreads();
        if (wave == 0) {
            for (int w = 0; w < 3; ++w)
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float s2 = red[(w * 24 + c * 4 + i) * 64 + lane];
                        const float c2 = red[(w * 24 + 12 + c * 4 + i) * 64 + lane];
                        float y = s2 - Cp[c][i];
                        float t = S[c][i] + y;
                        Cp[c][i] = (t - S[c][i]) - y;
                        S[c][i] = t;
                        y = -c2 - Cp[c][i];
                        t = S[c][i] + y;
                        Cp[c][i] = (t - S[c][i]) - y;
                        S[c][i] = t;
                    }
            // D layout: lane holds rows 4g..4g+3 (probes) of column pl (coef)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const size_t row = (size_t)sl * a.Ppad + pg * 16 + 4 * g + i;
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    a.partials[(row * 3 + c) * 16 + pl] = make_float2(S[c][i], Cp[c][i]);
            }
        }
        __syncthreads();
    }
    if (COUNT) {
        atomicAdd(&a.counters[0], (unsigned long long)cnt.nodes);
        atomicAdd(&a.counters[1], (unsigned long long)cnt.tris);
        atomicAdd(&a.counters[2], (unsigned long long)nshadow);
    }
}

// FP64 cross-slice merge: out[p][k][c] = scale * sum_s (sum_s - comp_s).
__global__ __launch_bounds__(256) void k_merge(MergeArgs m)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= m.Ppad * 48) return;
    const int ps = idx / 48;
    const int rem = idx - ps * 48;
    const int c = rem >> 4;
    const int k = rem & 15;
    if (k >= m.K) return;
    const int op = m.probe_perm[ps];
    if (op < 0) return;
    double s = 0.0;
    for (int sl = 0; sl < m.nslices; ++sl) {
        const float2 v = m.partials[((size_t)sl * m.Ppad + ps) * 48 + rem];
        s += (double)v.x - (double)v.y;
    }
    s *= m.scale;
    const size_t o = ((size_t)op * m.K + k) * 3 + c;
    if (m.out) m.out[o] = (float)s;
    if (m.out64) m.out64[o] = s;
}

__global__ __launch_bounds__(256) void k_trace_closest(SceneView sc, const float* __restrict__ org,
                                                       const float* __restrict__ dir, int n,
                                                       float tmin, float tmax, int* prim,
                                                       float* tout, float* uout, float* vout,
                                                       uint32_t* spill, uint32_t stride)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    Stack st;
    st.lds = &lds[threadIdx.x];
    st.spill = spill + (size_t)blockIdx.x * 256 + threadIdx.x;
    st.spill_stride = stride;
    TraceCount cnt;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float3 o = make_float3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        const float3 d = make_float3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        float t = 0.f, u = 0.f, v = 0.f;
        const int p = trace_closest<false>(sc, o, d, tmin, tmax, t, u, v, st, cnt);
        prim[i] = p;
        tout[i] = p >= 0 ? t : 0.0f;
        uout[i] = p >= 0 ? u : 0.0f;
        vout[i] = p >= 0 ? v : 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_trace_occluded(SceneView sc, const float* __restrict__ org,
                                                        const float* __restrict__ dir,
                                                        const float* __restrict__ tmax, int n,
                                                        float tmin, uint8_t* occ,
                                                        uint32_t* spill, uint32_t stride)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    Stack st;
    st.lds = &lds[threadIdx.x];
    st.spill = spill + (size_t)blockIdx.x * 256 + threadIdx.x;
    st.spill_stride = stride;
    TraceCount cnt;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float3 o = make_float3(org[3 * i], org[3 * i + 1], org[3 * i + 2]);
        const float3 d = make_float3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
        occ[i] = trace_any<false>(sc, o, d, tmin, tmax[i], st, cnt) ? 1 : 0;
    }
}

template <int PREC, bool RECORD, bool COUNT>
int occupancy_of(int cus)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_bake<PREC, RECORD, COUNT>, 256, 0) !=
            hipSuccess ||
        nb < 1)
        nb = 1;
    return nb * cus;
}

}  // namespace

int bake_grid_size(Context& ctx, int precision)
{
    int g;
    switch (precision) {
    case VLB_PREC_BF16: g = occupancy_of<VLB_PREC_BF16, false, false>(ctx.num_cus); break;
    case VLB_PREC_FP16: g = occupancy_of<VLB_PREC_FP16, false, false>(ctx.num_cus); break;
    default: g = occupancy_of<VLB_PREC_FP32, false, false>(ctx.num_cus); break;
    }
    return (g + 7) & ~7;
}

void launch_bake(const BakeArgs& a, int precision, bool record, bool count, int grid,
                 hipStream_t st)
{
    VLB_REQUIRE(grid % 8 == 0, VLB_ERR_INVALID, "bake grid must be a multiple of 8");
    if (record) {
        k_bake<VLB_PREC_FP32, true, false><<<grid, 256, 0, st>>>(a);
    } else if (count) {
        k_bake<VLB_PREC_FP32, false, true><<<grid, 256, 0, st>>>(a);
    } else if (precision == VLB_PREC_BF16) {
        k_bake<VLB_PREC_BF16, false, false><<<grid, 256, 0, st>>>(a);
    } else if (precision == VLB_PREC_FP16) {
        k_bake<VLB_PREC_FP16, false, false><<<grid, 256, 0, st>>>(a);
    } else {
        k_bake<VLB_PREC_FP32, false, false><<<grid, 256, 0, st>>>(a);
    }
    VLB_HIP(hipGetLastError());
}

void launch_merge(const MergeArgs& m, hipStream_t st)
{
    const int n = m.Ppad * 48;
    k_merge<<<(n + 255) / 256, 256, 0, st>>>(m);
    VLB_HIP(hipGetLastError());
}

void launch_trace_closest(const SceneView& sc, const float* org, const float* dir, int n,
                          float tmin, float tmax, int* prim, float* t, float* u, float* v,
                          uint32_t* spill, uint32_t stride, int grid, hipStream_t st)
{
    k_trace_closest<<<grid, 256, 0, st>>>(sc, org, dir, n, tmin, tmax, prim, t, u, v, spill,
                                          stride);
    VLB_HIP(hipGetLastError());
}

void launch_trace_occluded(const SceneView& sc, const float* org, const float* dir,
                           const float* tmax, int n, float tmin, uint8_t* occ, uint32_t* spill,
                           uint32_t stride, int grid, hipStream_t st)
{
    k_trace_occluded<<<grid, 256, 0, st>>>(sc, org, dir, tmax, n, tmin, occ, spill, stride);
    VLB_HIP(hipGetLastError());
}

}  // namespace vlb
