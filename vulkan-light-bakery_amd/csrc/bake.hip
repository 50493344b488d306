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
// Two kernels per bake plus the merge:
//  k_trace   one ray per lane: closest hit, shading, shadow rays, sky; writes
//            FP32 radiance once, in the MFMA operand order of k_project.
//            Few live registers (no accumulators across the traversal), so
//            occupancy is bounded by the LDS stack, not by VGPRs.
//  k_project the SH projection as a dense contraction on MFMA with FP32
//            Kahan accumulation across chunks (HBM-bound, 12 B per ray).
//  k_merge   FP64 cross-slice merge in a fixed order (deterministic; no
//            atomics on the result path -- the reference's sh.comp:39 "+="
//            is a data race).
//
// k_trace walks its rays XCD by XCD: workgroup b runs on XCD b % 8
// (round-robin dispatch), and XCD x owns the x-th contiguous eighth of the
// rays, so neighbouring probes and directions share that XCD's L2.
#include "vlb_internal.hpp"

namespace vlb {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mma(const v8bf& a, const v8bf& b, const v4f& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mma(const v8h& a, const v8h& b, const v4f& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// 16-bit operand types of the projection forms.
template <int PREC> struct Half;
template <> struct Half<VLB_PREC_BF16> {
    using V = v8bf;
    using T = __bf16;
    static constexpr bool split = true;
    static constexpr float scale = 1.0f;
};
template <> struct Half<VLB_PREC_BF16_FAST> {
    using V = v8bf;
    using T = __bf16;
    static constexpr bool split = false;
    static constexpr float scale = 1.0f;
};
template <> struct Half<VLB_PREC_FP16> {
    using V = v8h;
    using T = _Float16;
    static constexpr bool split = true;
    static constexpr float scale = 1.0f / 16.0f;
};
template <> struct Half<VLB_PREC_FP16_FAST> {
    using V = v8h;
    using T = _Float16;
    static constexpr bool split = false;
    static constexpr float scale = 1.0f / 16.0f;
};

// Project 8 rays x 3 channels held by this lane onto the 16x16 tile C.
// Split forms: x = hi + lo with hi = T(x), lo = T(x - hi); products
// lo*Yhi + hi*Ylo + hi*Yhi, smallest first.
template <int PREC>
__device__ __forceinline__ void project16(const float* lx, const float* ly, const float* lz,
                                          const void* ytab, size_t lo_off, size_t vidx, v4f* C)
{
    using H = Half<PREC>;
    using V = typename H::V;
    using T = typename H::T;
    const V* yt = static_cast<const V*>(ytab);
    const V yhi = yt[vidx];
    const float* src[3] = {lx, ly, lz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        V ahi, alo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float x = src[c][e] * H::scale;
            if (PREC == VLB_PREC_FP16 || PREC == VLB_PREC_FP16_FAST)
                x = fminf(fmaxf(x, -65504.0f), 65504.0f);
            const T h = (T)x;
            ahi[e] = h;
            if constexpr (H::split) alo[e] = (T)(x - (float)h);
        }
        if constexpr (H::split) {
            const V ylo = yt[lo_off + vidx];
            C[c] = mma(alo, yhi, C[c]);
            C[c] = mma(ahi, ylo, C[c]);
        }
        C[c] = mma(ahi, yhi, C[c]);
    }
}

// ----------------------------------------------------------- trace + shade
// One ray per lane, persistent grid-stride over "trace tiles" (16 probes x
// 32 rays = 512 rays).  Thread t of a tile traces for probe 16*pg + (l & 15)
// the ray 32*ch + 8*(l >> 4) + e, with l = t >> 3 and e = t & 7, and stores
// its radiance at rad[((tile*3 + c)*64 + l)*8 + e]: the 16x16 MFMA A-operand
// order of the projection kernel, written fully coalesced.  A wave covers 8
// neighbouring probes x 8 neighbouring directions (both Morton sorted).
// Lane -> ray maps inside a 512-ray trace tile (t = thread in tile):
//   MAP 0: a wave = 8 probes x 8 directions
//   MAP 1: a wave = 16 probes x 4 directions
//   MAP 2: a wave = 2 probes x 32 directions
// The radiance slot (l, e) is the same for every map.
__device__ __forceinline__ void tile_map(int map, int t, int& l, int& e)
{
    if (map == 1) {
        const int p = t & 15, ray = t >> 4;
        l = p + 16 * (ray >> 3);
        e = ray & 7;
    } else if (map == 2) {
        const int ray = t & 31, p = t >> 5;
        l = p + 16 * (ray >> 3);
        e = ray & 7;
    } else {
        l = t >> 3;
        e = t & 7;
    }
}

// OCC > 0 asks the compiler for OCC waves per SIMD (VGPR budget 512/OCC).
template <bool RECORD, bool COUNT, int OCC>
__global__ __launch_bounds__(256, OCC > 0 ? OCC : 1) void k_trace(BakeArgs a)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    const int tid = threadIdx.x;
    Stack st;
    st.lds = &lds[tid];
    st.spill = a.spill + (size_t)blockIdx.x * 256 + tid;
    st.spill_stride = a.spill_stride;
    TraceCount cnt;
    uint32_t nshadow = 0;
    // XCD x (= blockIdx % 8) owns the x-th contiguous eighth of the rays
    const size_t nitems = (size_t)a.ntrace * 512;
    const size_t per_xcd = (nitems + 7) >> 3;
    const size_t xbeg = (size_t)(blockIdx.x & 7) * per_xcd;
    const size_t xend = min(nitems, xbeg + per_xcd);
    const size_t stride = (size_t)(gridDim.x >> 3) * 256;
    for (size_t it = xbeg + (size_t)(blockIdx.x >> 3) * 256 + tid; it < xend; it += stride) {
        const int tile = (int)(it >> 9);
        int l, e;
        tile_map(a.map, (int)(it & 511), l, e);
        const int t = l * 8 + e;
        const int pg = tile / a.nchunk;
        const int ch = tile - pg * a.nchunk;
        const int probe = pg * 16 + (l & 15);
        const int r = ch * 32 + 8 * (l >> 4) + e;
        const float4 dd = a.dirs[r];
        float3 L = make_float3(0.f, 0.f, 0.f);
        if (probe < a.P && dd.w != 0.0f) {
            const float4 pp = a.probes[probe];
            int hp = -1;
            L = shade_ray<COUNT>(a.sc, a.sp, make_float3(pp.x, pp.y, pp.z),
                                 make_float3(dd.x, dd.y, dd.z), a.tmin, a.tmax, st, cnt, nshadow, hp);
            if (RECORD) {
                const size_t ri = (size_t)a.probe_perm[probe] * a.R + a.dir_perm[r];
                if (a.rad_out) {
                    a.rad_out[3 * ri + 0] = L.x;
                    a.rad_out[3 * ri + 1] = L.y;
                    a.rad_out[3 * ri + 2] = L.z;
                }
                if (a.hit_out) a.hit_out[ri] = hp;
            }
        }
        float* out = a.rad + (size_t)tile * 1536 + t;
        out[0] = L.x;
        out[512] = L.y;
        out[1024] = L.z;
    }
    if (COUNT) {
        atomicAdd(&a.counters[0], (unsigned long long)cnt.nodes);
        atomicAdd(&a.counters[1], (unsigned long long)cnt.tris);
        atomicAdd(&a.counters[2], (unsigned long long)nshadow);
    }
}

// ------------------------------------------------------------- projection
// C[probe][coef] += sum_ray rad[probe][ray] * Yw[ray][coef] on MFMA.  A
// workgroup (4 waves) owns 16 probes x one slice of RS rays; the waves take
// its 32-ray chunks round-robin.  Lane l = 16*g + pl holds, per channel, the
// radiance of probe pl for rays 8g..8g+7 of the chunk (A operand) and the
// basis values Y_pl of the same rays (B operand).  FP32 uses 8 x
// v_mfma_f32_16x16x4_f32 per chunk and channel; the 16-bit forms 1 or 3 x
// v_mfma_f32_16x16x32_{bf16,f16}.  Chunk results are Kahan-summed in FP32
// across chunks and waves into one (sum, compensation) partial per
// (slice, probe, channel, coef).  HBM-bound: 12 B of radiance per ray.
template <int PREC>
__global__ __launch_bounds__(256) void k_project(BakeArgs a)
{
    __shared__ float red[3 * 24 * 64];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = lane >> 4;
    const int pl = lane & 15;
    const int nch = a.RS >> 5;
    for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const int pg = tile / a.nslices;
        const int sl = tile - pg * a.nslices;
        v4f S[3], Cp[3];
        for (int c = 0; c < 3; ++c) {
            S[c] = v4f{0.f, 0.f, 0.f, 0.f};
            Cp[c] = v4f{0.f, 0.f, 0.f, 0.f};
        }
        for (int k = wave; k < nch; k += 4) {
            const int ch = sl * nch + k;                 // chunk within the probe group
            const int ttile = pg * a.nchunk + ch;        // trace tile
            // streamed once: non-temporal 16-byte loads
            const v4f* src = reinterpret_cast<const v4f*>(a.rad + (size_t)ttile * 1536) + lane * 2;
            float lx[8], ly[8], lz[8];
            {
                const v4f x0 = __builtin_nontemporal_load(src), x1 = __builtin_nontemporal_load(src + 1);
                const v4f y0 = __builtin_nontemporal_load(src + 128), y1 = __builtin_nontemporal_load(src + 129);
                const v4f z0 = __builtin_nontemporal_load(src + 256), z1 = __builtin_nontemporal_load(src + 257);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    lx[q] = x0[q];
                    lx[4 + q] = x1[q];
                    ly[q] = y0[q];
                    ly[4 + q] = y1[q];
                    lz[q] = z0[q];
                    lz[4 + q] = z1[q];
                }
            }
            const size_t vidx = (size_t)(ch * 4 + g) * 16 + pl;  // ray octet 4*ch + g, coef pl
            v4f C[3];
            for (int c = 0; c < 3; ++c) C[c] = v4f{0.f, 0.f, 0.f, 0.f};
            if constexpr (PREC == VLB_PREC_FP32) {
                const float4* yt = static_cast<const float4*>(a.ytab) + vidx * 2;
                const float4 y0 = yt[0], y1 = yt[1];
                const float y[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    C[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(lx[e], y[e], C[0], 0, 0, 0);
                    C[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ly[e], y[e], C[1], 0, 0, 0);
                    C[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(lz[e], y[e], C[2], 0, 0, 0);
                }
            } else {
                project16<PREC>(lx, ly, lz, a.ytab, (size_t)a.Rpad * 2, vidx, C);
            }
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
        // cross-wave Kahan combine through LDS, fixed order
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

template <class K>
int occupancy_of(K kernel, int cus)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return nb * cus;
}

}  // namespace

int trace_grid_size(Context& ctx, int occ)
{
    int g;
    if (occ == 8) g = occupancy_of(k_trace<false, false, 8>, ctx.num_cus);
    else if (occ == 6) g = occupancy_of(k_trace<false, false, 6>, ctx.num_cus);
    else g = occupancy_of(k_trace<false, false, 0>, ctx.num_cus);
    return (g + 7) & ~7;
}

void launch_trace(const BakeArgs& a, bool record, bool count, int occ, int grid, hipStream_t st)
{
    VLB_REQUIRE(grid % 8 == 0, VLB_ERR_INVALID, "trace grid must be a multiple of 8");
    if (record) k_trace<true, false, 0><<<grid, 256, 0, st>>>(a);
    else if (count) k_trace<false, true, 0><<<grid, 256, 0, st>>>(a);
    else if (occ == 8) k_trace<false, false, 8><<<grid, 256, 0, st>>>(a);
    else if (occ == 6) k_trace<false, false, 6><<<grid, 256, 0, st>>>(a);
    else k_trace<false, false, 0><<<grid, 256, 0, st>>>(a);
    VLB_HIP(hipGetLastError());
}

void launch_project(const BakeArgs& a, int precision, int grid, hipStream_t st)
{
    grid = std::max(1, std::min(grid, a.ntiles));
    switch (precision) {
    case VLB_PREC_BF16: k_project<VLB_PREC_BF16><<<grid, 256, 0, st>>>(a); break;
    case VLB_PREC_FP16: k_project<VLB_PREC_FP16><<<grid, 256, 0, st>>>(a); break;
    case VLB_PREC_BF16_FAST: k_project<VLB_PREC_BF16_FAST><<<grid, 256, 0, st>>>(a); break;
    case VLB_PREC_FP16_FAST: k_project<VLB_PREC_FP16_FAST><<<grid, 256, 0, st>>>(a); break;
    default: k_project<VLB_PREC_FP32><<<grid, 256, 0, st>>>(a); break;
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
