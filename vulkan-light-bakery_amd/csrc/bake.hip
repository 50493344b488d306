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
// rays, so neighbouring probes and directions share that XCD's L2.  Work is
// taken 256 rays at a time from the XCD's counter, then from the other
// XCDs' once it is drained, so the launch ends when the work does rather
// than when its unluckiest workgroup does.
#include "project_common.hpp"

namespace vlb {
namespace {

// Clears the 32 work counters of the fused kernel (a kernel, not a memset, so that captured graphs replay it).
__global__ void k_zero_ctr8(unsigned long long* c)
{
    if (threadIdx.x < 32) c[threadIdx.x * kCtrStride] = 0ull;
}

// ----------------------------------------------------------- trace + shade
// One ray per lane, persistent, over "trace tiles" (16 probes x 32 rays =
// 512 rays) taken half a tile per workgroup and fetch.  Thread t of a tile traces for probe 16*pg + (l & 15)
// the ray 32*ch + 8*(l >> 4) + e, with l = t >> 3 and e = t & 7, and stores
// its radiance at rad[((tile*3 + c)*64 + l)*8 + e]: the 16x16 MFMA A-operand
// order of the projection kernel, written fully coalesced.  The lane map
// (tile_map, a.map) sets how many neighbouring probes x directions (probes
// Morton sorted, directions on a Hilbert curve) one wave covers.
// OCC > 0 asks the compiler for OCC waves per SIMD (VGPR budget 512/OCC).
// IND: plans with the multi-bounce term.  FAR: shadow queries visit children
// far first (same radiance; fewer node visits on some scenes).
template <bool RECORD, bool COUNT, int OCC, bool IND, bool FAR = false>
__global__ __launch_bounds__(256, OCC > 0 ? OCC : 1) void k_trace(BakeArgs a)
{
    __shared__ uint32_t lds[kLdsDepth * 256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    Stack st;
    st.lds = &lds[tid];
    st.spill = a.spill + (size_t)blockIdx.x * 256 + tid;
    st.spill_stride = a.spill_stride;
    st.depth = kLdsDepth;
    TraceCount cnt;
    uint32_t nshadow = 0;
    // 32 64-aligned ranges, range 4x + w owned by wave w of the workgroups on
    // XCD x (= blockIdx % 8); a wave takes 64 rays per fetch from its own
    // range, then from its XCD's other ranges, then from the other XCDs'
    // (nitems is a multiple of 512, so a fetched chunk is whole)
    const size_t nitems = (size_t)a.ntrace * 512;
    const size_t per_rng = ((nitems + 31) / 32 + 63) / 64 * 64;
    const int own = (blockIdx.x & 7) * 4 + (tid >> 6);
    int victim = 0;  // ranges tried (wave-uniform)
    for (;;) {
        unsigned long long b = ~0ull;
        if (lane == 0) {
            while (victim < 32) {
                const int v = (own & ~3) + ((own + victim) & 3) + 4 * (victim >> 2);
                const int vv = v & 31;
                const size_t vbeg = (size_t)vv * per_rng;
                const size_t vend = min(nitems, vbeg + per_rng);
                if (vbeg < vend) {
                    const unsigned long long base = atomicAdd(&a.wctr[vv * kCtrStride], 64ull);
                    if (vbeg + base < vend) {
                        b = vbeg + base;
                        break;
                    }
                }
                ++victim;
            }
        }
        // every lane is active here, so the first active lane is lane 0: the
        // fetch position and range count become scalars.  Broadcast with
        // __shfl they stayed per-lane values: 34 instead of 26 spilled VGPRs
        // at 8 waves/SIMD and a trace stage 0.5-2% longer on the
        // 11k-triangle room bake, 2048 probes x 4096 rays.  (The
        // multi-job kernel measured the same either way and keeps __shfl.)
        b = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(b >> 32)) << 32) |
            (unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)b);
        victim = __builtin_amdgcn_readfirstlane(victim);
        if (b == ~0ull) break;
        const size_t it = (size_t)b + lane;
        int tile, l, e;
        tile_map(a.map, it, tile, l, e);
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
            L = shade_ray<COUNT, IND, FAR, kStriShadow>(a.sc, a.sp, make_float3(pp.x, pp.y, pp.z),
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
// Tile = (probe group pg, slice sl): chunks sl*nch .. sl*nch + nch - 1 of the
// group's trace tiles (consecutive in `rad`), partials of slice sl.
template <int PREC>
__global__ __launch_bounds__(256) void k_project(BakeArgs a)
{
    __shared__ float red[3 * 24 * 64];
    const int nch = a.RS >> 5;
    for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const int pg = tile / a.nslices;
        const int sl = tile - pg * a.nslices;
        const int ch0 = sl * nch;
        const float* rad0 = a.rad + ((size_t)pg * a.nchunk + ch0) * 1536;
        float2* prow = a.partials + ((size_t)sl * a.Ppad + (size_t)pg * 16) * 48;
        project_tile<PREC>(rad0, nch, ch0, a.ytab, (size_t)a.Rpad * 2, red, prow);
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
    st.depth = kLdsDepth;
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
    st.depth = kLdsDepth;
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

int trace_grid_size(Context& ctx, int occ, bool far)
{
    int g;
    if (far) {
        if (occ == 8) g = occupancy_of(k_trace<false, false, 8, false, true>, ctx.num_cus);
        else if (occ == 6) g = occupancy_of(k_trace<false, false, 6, false, true>, ctx.num_cus);
        else g = occupancy_of(k_trace<false, false, 0, false, true>, ctx.num_cus);
    } else if (occ == 8) g = occupancy_of(k_trace<false, false, 8, false>, ctx.num_cus);
    else if (occ == 6) g = occupancy_of(k_trace<false, false, 6, false>, ctx.num_cus);
    else g = occupancy_of(k_trace<false, false, 0, false>, ctx.num_cus);
    return (g + 7) & ~7;
}

void launch_trace(const BakeArgs& a, bool record, bool count, int occ, bool far, int grid,
                  hipStream_t st)
{
    VLB_REQUIRE(grid % 8 == 0, VLB_ERR_INVALID, "trace grid must be a multiple of 8");
    k_zero_ctr8<<<1, 64, 0, st>>>(a.wctr);
    // recording (hit ids are the same in any shadow order) and multi-bounce
    // plans keep slot-order shadows
    if (a.sp.ind.coeffs) {
        if (record) k_trace<true, false, 0, true><<<grid, 256, 0, st>>>(a);
        else if (count) k_trace<false, true, 0, true><<<grid, 256, 0, st>>>(a);
        else k_trace<false, false, 0, true><<<grid, 256, 0, st>>>(a);
    } else if (record) k_trace<true, false, 0, false><<<grid, 256, 0, st>>>(a);
    else if (count && far) k_trace<false, true, 0, false, true><<<grid, 256, 0, st>>>(a);
    else if (count) k_trace<false, true, 0, false><<<grid, 256, 0, st>>>(a);
    else if (far) {
        if (occ == 8) k_trace<false, false, 8, false, true><<<grid, 256, 0, st>>>(a);
        else if (occ == 6) k_trace<false, false, 6, false, true><<<grid, 256, 0, st>>>(a);
        else k_trace<false, false, 0, false, true><<<grid, 256, 0, st>>>(a);
    } else if (occ == 8) k_trace<false, false, 8, false><<<grid, 256, 0, st>>>(a);
    else if (occ == 6) k_trace<false, false, 6, false><<<grid, 256, 0, st>>>(a);
    else k_trace<false, false, 0, false><<<grid, 256, 0, st>>>(a);
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

namespace {
__global__ void k_copy_f32(float* __restrict__ dst, const float* __restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}
}  // namespace

void launch_copy_f32(float* dst, const float* src, size_t n, hipStream_t st)
{
    if (n == 0) return;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_copy_f32<<<grid, 256, 0, st>>>(dst, src, n);
    VLB_HIP(hipGetLastError());
}

}  // namespace vlb
