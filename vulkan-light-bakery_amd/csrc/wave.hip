// wave.hip -- wavefront form of the trace + shade stage.
//
// The fused k_trace (bake.hip) runs primary traversal, shading and the shadow
// query back to back in one lane: a wave waits for its slowest lane in each
// phase, and lanes whose primary ray missed sit idle during shadow rays.
// Here the same work is three kernels:
//   k_rays<false>  closest hit for every primary ray -> 16 B hit records
//   k_rays<true>   shadow query for every (light, ray) pair that needs one
//                  -> one visibility byte per pair
//   k_shade        per ray: sky or shading, visibility combined in light
//                  order -> radiance in the MFMA operand order of k_project
// Both traversal kernels are persistent with dynamic ray fetch: when at
// least `refill` lanes of a wave are idle, the wave hands them the next rays
// of its XCD's work range (one atomic per chunk of rays), then of the other
// XCDs' ranges once its own is drained, so lanes stay busy until all the
// work is taken.  Items that need no traversal (padding, shadow
// pairs with N.L <= 0 or a missed primary) are retired at fetch time.
//
// Determinism: every decision (hit, N.L > 0, occlusion) is a function of the
// ray alone, computed with the same operation sequence as the fused kernel
// and the CPU oracle, so the results do not depend on which lane or wave
// processed which ray.
#include "wave_common.hpp"

namespace vlb {
namespace {

// item (in the trace-tile space of bake.hip) -> radiance slot, probe, ray
__device__ __forceinline__ void item_ray(const BakeArgs& a, size_t it, int& slot, int& probe,
                                         int& r)
{
    int tile, l, e;
    // same maps as k_trace
    if (a.map == 3) {
        const int i = (int)(it & 1023);
        const int p = i >> 6, rr = i & 63, rc = rr & 31;
        tile = (int)(it >> 10) * 2 + (rr >> 5);
        l = p + 16 * (rc >> 3);
        e = rc & 7;
    } else {
        tile = (int)(it >> 9);
        const int t = (int)(it & 511);
        if (a.map == 1) {
            const int p = t & 15, ray = t >> 4;
            l = p + 16 * (ray >> 3);
            e = ray & 7;
        } else if (a.map == 2) {
            const int ray = t & 31, p = t >> 5;
            l = p + 16 * (ray >> 3);
            e = ray & 7;
        } else {
            l = t >> 3;
            e = t & 7;
        }
    }
    slot = tile * 512 + l * 8 + e;
    const int pg = tile / a.nchunk;
    const int ch = tile - pg * a.nchunk;
    probe = pg * 16 + (l & 15);
    r = ch * 32 + 8 * (l >> 4) + e;
}

__device__ __forceinline__ void slot_ray(const BakeArgs& a, int slot, int& probe, int& r)
{
    const int tile = slot >> 9, t = slot & 511;
    const int l = t >> 3, e = t & 7;
    const int pg = tile / a.nchunk;
    const int ch = tile - pg * a.nchunk;
    probe = pg * 16 + (l & 15);
    r = ch * 32 + 8 * (l >> 4) + e;
}

// Number of rays a wave takes from its XCD's counter at once.
constexpr int kChunk = 256;

// Sets up item `it` for this lane.  Returns false when the item needs no
// traversal (its result is written here).
template <bool ANY, bool CORNERS>
__device__ __forceinline__ bool item_setup(const BakeArgs& a, size_t it, TravState& T, int& slot_out,
                                           int& light_out)
{
    if (!ANY) {
        int slot, probe, r;
        item_ray(a, it, slot, probe, r);
        slot_out = slot;
        const float4 dd = a.dirs[r];
        if (probe >= a.P || dd.w == 0.0f || a.sc.ntris <= 0) {
            a.hits[slot] = make_int4(probe >= a.P || dd.w == 0.0f ? -2 : -1, 0, 0, 0);
            return false;
        }
        const float4 pp = a.probes[probe];
        trav_init(T, a.sc.maxabs, make_float3(pp.x, pp.y, pp.z), make_float3(dd.x, dd.y, dd.z), a.tmin,
                  a.tmax);
        return true;
    }
    // shadow items are light-major: it = j * nslots + slot
    const size_t nslots = (size_t)a.ntrace * 512;
    const int j = (int)(it / nslots);
    const int slot = (int)(it - (size_t)j * nslots);
    slot_out = slot;
    light_out = j;
    const int4 h = a.hits[slot];
    uint8_t* vis = a.vis + (size_t)slot * a.nvis + j;
    if (h.x < 0) {
        *vis = 0;
        return false;
    }
    int probe, r;
    slot_ray(a, slot, probe, r);
    const float4 pp = a.probes[probe];
    const float4 dd = a.dirs[r];
    const float3 o = make_float3(pp.x, pp.y, pp.z), d = make_float3(dd.x, dd.y, dd.z);
    const HitGeom g = hit_geom(a.sc.shade, o, d, h.x, __int_as_float(h.y), __int_as_float(h.z),
                               __int_as_float(h.w));
    const ShadowRay s = shadow_ray<CORNERS>(a.sp, g, d, j);
    if (!s.needed) {
        *vis = 0;
        return false;
    }
    trav_init(T, a.sc.maxabs, s.org, s.dir, 0.0f, s.dist);
    return true;
}

// CORNERS (shadow pass only): the item space also holds the corner
// visibility rays of the multi-bounce term.
template <bool ANY, bool COUNT, bool CORNERS>
__global__ __launch_bounds__(256) void k_rays(BakeArgs a, size_t nitems, unsigned long long* ctr)
{
    __shared__ uint32_t lds[kLdsDepthWave * 256];
    __shared__ uint32_t qlds[kLeafQueue * 256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    Stack st;
    st.lds = &lds[tid];
    LeafQueue q;
    q.lds = &qlds[tid];
    st.spill = a.spill + (size_t)blockIdx.x * 256 + tid;
    st.spill_stride = a.spill_stride;
    st.depth = kLdsDepthWave;
    TraceCount cnt;
    uint32_t ntraced = 0;
    // eight kChunk-aligned ranges of the item space, one per XCD; a wave
    // takes chunks from its XCD's range, then from the others' once that is
    // drained, so the launch ends when the work does
    const int xcd = blockIdx.x & 7;
    const size_t per_xcd = ((nitems + 7) / 8 + kChunk - 1) / kChunk * kChunk;
    int victim = 0;              // ranges tried: own first, then xcd+1, ...
    size_t pnext = 0, pend = 0;  // wave's current chunk [pnext, pend)
    bool drained = nitems == 0;

    TravState T;
    int slot = 0, light = 0;
    bool act = false;
    for (;;) {
        // --- refill idle lanes (wave-uniform branch)
        uint64_t idle = __ballot(!act);
        if (!drained && (int)__popcll(idle) >= a.refill) {
            while (idle != 0 && !drained) {
                if (pnext >= pend) {
                    for (;;) {
                        const int v = (xcd + victim) & 7;
                        const size_t vbeg = (size_t)v * per_xcd;
                        const size_t vend = min(nitems, vbeg + per_xcd);
                        unsigned long long base = 0;
                        if (vbeg < vend && lane == 0)
                            base = atomicAdd(&ctr[v * kCtrStride], (unsigned long long)kChunk);
                        // every lane is active in this wave-uniform loop, so
                        // the first active lane is lane 0: the chunk base, and
                        // with it the wave's chunk bounds, become scalars
                        // (broadcast with __shfl they took 10-11 more VGPRs;
                        // the 1.15M-triangle terrain bake 4% and a batch of
                        // 64 jobs over four scenes 11% slower)
                        base = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32)) << 32) |
                               (unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)base);
                        if (vbeg < vend && vbeg + base < vend) {
                            pnext = vbeg + base;
                            pend = min(vend, pnext + kChunk);
                            break;
                        }
                        if (++victim == 8) {
                            drained = true;
                            break;
                        }
                    }
                    if (drained) break;
                }
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const size_t avail = pend - pnext;
                const bool mine = !act && rank < avail;
                const size_t take = min((size_t)__popcll(idle), avail);
                if (mine) {
                    act = item_setup<ANY, CORNERS>(a, pnext + rank, T, slot, light);
                    if (act && ANY) ++ntraced;
                }
                pnext += take;
                idle = __ballot(!act);
                if ((int)__popcll(idle) < a.refill) break;
            }
        }
        if (__ballot(act) == 0) {
            if (drained) break;
            continue;
        }
        // --- one traversal step for active lanes
        if (act) {
            if (trav_step<ANY, COUNT>(T, a.sc.qnodes, a.sc.tris, st, q, cnt)) {
                if (ANY) {
                    a.vis[(size_t)slot * a.nvis + light] = (uint8_t)(T.best == 0 ? 1 : 0);
                } else if (T.best >= 0) {
                    a.hits[slot] = make_int4(T.best, __float_as_int(T.tb), __float_as_int(T.bV / T.bdet),
                                             __float_as_int(T.bW / T.bdet));
                } else {
                    a.hits[slot] = make_int4(-1, 0, 0, 0);
                }
                act = false;
            }
        }
    }
    if (COUNT) {
        atomicAdd(&a.counters[0], (unsigned long long)cnt.nodes);
        atomicAdd(&a.counters[1], (unsigned long long)cnt.tris);
        if (ANY) atomicAdd(&a.counters[2], (unsigned long long)ntraced);
    }
}

// Per ray: sky on a miss, else the shading of env_map.rchit (REFERENCE) or
// Lambert, with the recorded visibility; radiance in MFMA operand order.
template <bool RECORD, bool IND>
__global__ __launch_bounds__(256) void k_shade(BakeArgs a)
{
    const size_t nslots = (size_t)a.ntrace * 512;
    for (size_t slot = (size_t)blockIdx.x * 256 + threadIdx.x; slot < nslots;
         slot += (size_t)gridDim.x * 256) {
        const int4 h = a.hits[slot];
        int probe, r;
        slot_ray(a, (int)slot, probe, r);
        float3 L = make_float3(0.f, 0.f, 0.f);
        if (h.x != -2) {
            const float4 dd = a.dirs[r];
            const float3 d = make_float3(dd.x, dd.y, dd.z);
            L = shade_record<IND>(a.sp, a.sc.shade, a.probes + probe, d, h, a.vis + slot * a.nvis);
            if (RECORD) {
                const size_t ri = (size_t)a.probe_perm[probe] * a.R + a.dir_perm[r];
                if (a.rad_out) {
                    a.rad_out[3 * ri + 0] = L.x;
                    a.rad_out[3 * ri + 1] = L.y;
                    a.rad_out[3 * ri + 2] = L.z;
                }
                if (a.hit_out) a.hit_out[ri] = h.x;
            }
        }
        const size_t tile = slot >> 9, t = slot & 511;
        float* out = a.rad + tile * 1536 + t;
        out[0] = L.x;
        out[512] = L.y;
        out[1024] = L.z;
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[(size_t)i * kCtrStride] = 0ull;
}

template <class K>
int occupancy_of(K kernel, int cus)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return nb * cus;
}

}  // namespace

int wave_grid_size(Context& ctx)
{
    const int g = std::max(occupancy_of(k_rays<false, false, false>, ctx.num_cus),
                           occupancy_of(k_rays<true, false, false>, ctx.num_cus));
    return (g + 7) & ~7;
}

void launch_wave(const BakeArgs& a, bool record, bool count, int grid, hipStream_t st)
{
    VLB_REQUIRE(grid % 8 == 0, VLB_ERR_INVALID, "wave grid must be a multiple of 8");
    const size_t nslots = (size_t)a.ntrace * 512;
    // work counters 0..7 primary, 8..15 shadow, kCtrStride apart (cleared by
    // a kernel so the sequence can be captured in a graph)
    k_zero_u64<<<1, 64, 0, st>>>(a.wctr, 16);  // primary and shadow counters
    if (count) k_rays<false, true, false><<<grid, 256, 0, st>>>(a, nslots, a.wctr);
    else k_rays<false, false, false><<<grid, 256, 0, st>>>(a, nslots, a.wctr);
    VLB_HIP(hipGetLastError());
    const size_t nshadow = nslots * (size_t)a.nvis;
    if (a.nvis > 0) {
        const bool corners = a.sp.ind.coeffs && a.sp.ind.vis;
        if (corners) {
            if (count) k_rays<true, true, true><<<grid, 256, 0, st>>>(a, nshadow, a.wctr + 8 * kCtrStride);
            else k_rays<true, false, true><<<grid, 256, 0, st>>>(a, nshadow, a.wctr + 8 * kCtrStride);
        } else if (count) {
            k_rays<true, true, false><<<grid, 256, 0, st>>>(a, nshadow, a.wctr + 8 * kCtrStride);
        } else {
            k_rays<true, false, false><<<grid, 256, 0, st>>>(a, nshadow, a.wctr + 8 * kCtrStride);
        }
        VLB_HIP(hipGetLastError());
    }
    const int sgrid = (int)std::min<size_t>((nslots + 255) / 256, 8192);
    if (a.sp.ind.coeffs) {
        if (record) k_shade<true, true><<<sgrid, 256, 0, st>>>(a);
        else k_shade<false, true><<<sgrid, 256, 0, st>>>(a);
    } else if (record) {
        k_shade<true, false><<<sgrid, 256, 0, st>>>(a);
    } else {
        k_shade<false, false><<<sgrid, 256, 0, st>>>(a);
    }
    VLB_HIP(hipGetLastError());
}

}  // namespace vlb
