// multi.hip -- the probe-baking hot path for many jobs at once.
//
// A round is a list of runs ("segments") of projection tiles taken from any
// number of jobs, each job with its own scene BVH, probes, direction set,
// lights and output.  One round is four persistent launches, whatever the
// number of jobs:
//   k_rays_jobs<false>  closest hit of every primary ray of the round
//   k_rays_jobs<true>   shadow query of every (light, ray) pair
//   k_shade_jobs        radiance per ray (MFMA operand order)
//   k_project_jobs      SH projection of every projection tile (MFMA)
// (k_fused_jobs replaces the first three for trace_variant 1..12, 25..36) and
// k_merge_jobs merges the FP32 (sum, compensation) partials of finished
// jobs in FP64.  The counts of a round live in device memory (WorkHeader),
// so one captured hipGraph replays rounds of any size.
//
// Work distribution: the item space of a pass is cut into eight contiguous
// ranges, one per XCD (workgroup b runs on XCD b % 8); a wave takes `chunk`
// items at a time from its XCD's counter and, once that range is drained,
// from the other XCDs' counters in turn, so a round whose jobs differ in
// cost per ray (scenes of 10k to 2M triangles) still ends evenly.  A chunk
// never straddles two segments (chunks divide a pair of trace tiles, and
// segments hold whole pairs), so the job of a chunk is found once per fetch
// by a wave-uniform binary search; each lane keeps the BVH pointers of the
// ray it holds, and lanes of one wave may trace different scenes.  The fused
// kernel cuts the item space into 32 ranges (four per XCD, one per wave
// slot of a workgroup) and fetches 64 items per wave.  Every counter sits on
// its own cache line: device-scope atomics on one line serialise.
//
// Per ray, every decision and every value is computed by the same device
// functions as the single-bake kernels (wave_common.hpp, project_common.hpp),
// so a job's coefficients do not depend on what else runs in its round.
#include "project_common.hpp"
#include "wave_common.hpp"

namespace vlb {
namespace {

// Last segment whose first trace tile is <= tile (tile wave-uniform).
__device__ __forceinline__ int seg_of_tile(const Segment* __restrict__ seg, int n, int tile)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].tb <= tile) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// Last segment whose first projection tile is <= pt.
__device__ __forceinline__ int seg_of_proj(const Segment* __restrict__ seg, int n, int pt)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid].pb <= pt) lo = mid;
        else hi = mid - 1;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

// Round trace tile + slot in tile -> (probe, ray) of the job, given the
// segment (first trace tile tb, first projection tile pt0).
__device__ __forceinline__ void decode_slot(const JobDesc& J, int tb, int pt0, int tile, int l,
                                            int e, int& probe, int& r)
{
    const int nch = J.RS >> 5;
    const int lt = tile - tb;
    const int q = lt / nch;
    const int kk = lt - q * nch;
    const int pt = pt0 + q;
    const int pg = pt / J.nslices;
    const int sl = pt - pg * J.nslices;
    probe = pg * 16 + (l & 15);
    r = (sl * nch + kk) * 32 + 8 * (l >> 4) + e;
}

// Per-lane ray of a traversal pass: its state and the BVH it walks.
struct LaneRay {
    TravState T;
    const QNode4* nodes;
    const float4* tris;
    int slot, light;
};

template <bool ANY>
__device__ __forceinline__ bool setup_item(const MultiArgs& a, const JobDesc& J, int tb, int pt0,
                                           size_t it, size_t nslots, LaneRay& R)
{
    if (!ANY) {
        int tile, l, e;
        tile_map(a.map, it, tile, l, e);
        int probe, r;
        decode_slot(J, tb, pt0, tile, l, e, probe, r);
        R.slot = tile * 512 + l * 8 + e;
        const float4 dd = J.dirs[r];
        if (probe >= J.P || dd.w == 0.0f || J.sc.ntris <= 0) {
            a.hits[R.slot] = make_int4(probe >= J.P || dd.w == 0.0f ? -2 : -1, 0, 0, 0);
            return false;
        }
        const float4 pp = J.probes[probe];
        trav_init(R.T, J.sc.maxabs, make_float3(pp.x, pp.y, pp.z), make_float3(dd.x, dd.y, dd.z),
                  J.tmin, J.tmax);
        R.nodes = J.sc.qnodes;
        R.tris = J.sc.tris;
        return true;
    }
    // shadow items are light-major: it = j * nslots + slot
    const int j = (int)(it / nslots);
    const int slot = (int)(it - (size_t)j * nslots);
    R.slot = slot;
    R.light = j;
    uint8_t* vis = a.vis + (size_t)slot * a.nvis + j;
    if (j >= J.nvis) {
        *vis = 0;
        return false;
    }
    const int4 h = a.hits[slot];
    if (h.x < 0) {
        *vis = 0;
        return false;
    }
    const int tile = slot >> 9, t = slot & 511;
    int probe, r;
    decode_slot(J, tb, pt0, tile, t >> 3, t & 7, probe, r);
    const float4 pp = J.probes[probe];
    const float4 dd = J.dirs[r];
    const float3 o = make_float3(pp.x, pp.y, pp.z), d = make_float3(dd.x, dd.y, dd.z);
    const HitGeom g = hit_geom(J.sc.shade, o, d, h.x, __int_as_float(h.y), __int_as_float(h.z),
                               __int_as_float(h.w));
    const ShadowRay s = shadow_ray(J.sp, g, d, j);
    if (!s.needed) {
        *vis = 0;
        return false;
    }
    trav_init(R.T, J.sc.maxabs, s.org, s.dir, 0.0f, s.dist);
    R.nodes = J.sc.qnodes;
    R.tris = J.sc.tris;
    return true;
}

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(256) void k_rays_jobs(MultiArgs a, const JobDesc* __restrict__ jobs,
                                                   const Segment* __restrict__ seg,
                                                   const WorkHeader* __restrict__ hdr)
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

    const int nseg = hdr->nseg;
    const size_t nslots = (size_t)hdr->ntrace * 512;
    const size_t nitems = ANY ? nslots * (size_t)min(hdr->nvis, a.nvis) : nslots;
    unsigned long long* ctr = a.wctr + (ANY ? 8 * kCtrStride : 0);
    const size_t chunk = (size_t)a.chunk;
    // eight chunk-aligned ranges, one per XCD
    const size_t per_xcd = ((nitems + 7) / 8 + chunk - 1) / chunk * chunk;
    const int xcd = blockIdx.x & 7;
    int victim = 0;                  // ranges tried: own first, then xcd+1, ...
    size_t pnext = 0, pend = 0;      // wave's current chunk
    bool drained = nitems == 0;
    int cj = 0, ctb = 0, cpt0 = 0;   // job and segment of the current chunk

    LaneRay R;
    R.slot = 0;
    R.light = 0;
    R.nodes = nullptr;
    R.tris = nullptr;
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
                            base = atomicAdd(&ctr[v * kCtrStride], (unsigned long long)chunk);
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
                            pend = min(vend, pnext + chunk);
                            break;
                        }
                        if (++victim == 8) {
                            drained = true;
                            break;
                        }
                    }
                    if (drained) break;
                    // the chunk's segment: pairs of trace tiles never straddle two
                    const size_t slot0 = ANY ? pnext % nslots : pnext;
                    const int tile0 = __builtin_amdgcn_readfirstlane((int)(slot0 >> 10) * 2);
                    const int s = seg_of_tile(seg, nseg, tile0);
                    cj = __builtin_amdgcn_readfirstlane(seg[s].job);
                    ctb = __builtin_amdgcn_readfirstlane(seg[s].tb);
                    cpt0 = __builtin_amdgcn_readfirstlane(seg[s].pt0);
                }
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const size_t avail = pend - pnext;
                const bool mine = !act && rank < avail;
                const size_t take = min((size_t)__popcll(idle), avail);
                if (mine) {
                    act = setup_item<ANY>(a, jobs[cj], ctb, cpt0, pnext + rank, nslots, R);
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
            if (trav_step<ANY, COUNT>(R.T, R.nodes, R.tris, st, q, cnt)) {
                if (ANY) {
                    a.vis[(size_t)R.slot * a.nvis + R.light] = (uint8_t)(R.T.best == 0 ? 1 : 0);
                } else if (R.T.best >= 0) {
                    a.hits[R.slot] = make_int4(R.T.best, __float_as_int(R.T.tb),
                                               __float_as_int(R.T.bV / R.T.bdet),
                                               __float_as_int(R.T.bW / R.T.bdet));
                } else {
                    a.hits[R.slot] = make_int4(-1, 0, 0, 0);
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

// Fused form of the three passes above (trace_variant 1..12; 25..36 with FAR,
// far-first shadow queries): each lane takes one primary ray and runs closest
// hit, shading and its shadow rays with the single-bake kernel's shade_ray,
// writing radiance in MFMA operand order.  A
// wave fetches 64 items at a time (a chunk lies within a pair of trace tiles,
// hence one segment and one job).  Same values as the wavefront passes, bit
// for bit.
template <bool COUNT, int OCC, bool FAR = false>
__global__ __launch_bounds__(256, OCC > 0 ? OCC : 1) void k_fused_jobs(MultiArgs a,
                                                                       const JobDesc* __restrict__ jobs,
                                                                       const Segment* __restrict__ seg,
                                                                       const WorkHeader* __restrict__ hdr)
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
    const int nseg = hdr->nseg;
    const size_t nitems = (size_t)hdr->ntrace * 512;
    // 32 64-aligned ranges, range 4x + w owned by wave w of the workgroups on
    // XCD x; a wave takes 64 items per fetch from its own range, then from
    // its XCD's other ranges, then from the other XCDs' (nitems is a multiple
    // of 512, so a fetched chunk is whole and lies in one pair of tiles)
    const size_t per_rng = ((nitems + 31) / 32 + 63) / 64 * 64;
    const int own = (blockIdx.x & 7) * 4 + (tid >> 6);
    int victim = 0;  // ranges tried (wave-uniform)
    for (;;) {
        unsigned long long b = ~0ull;
        if (lane == 0) {
            while (victim < 32) {
                const int v = ((own & ~3) + ((own + victim) & 3) + 4 * (victim >> 2)) & 31;
                const size_t vbeg = (size_t)v * per_rng;
                const size_t vend = min(nitems, vbeg + per_rng);
                if (vbeg < vend) {
                    const unsigned long long base = atomicAdd(&a.wctr[v * kCtrStride], 64ull);
                    if (vbeg + base < vend) {
                        b = vbeg + base;
                        break;
                    }
                }
                ++victim;
            }
        }
        b = ((unsigned long long)__shfl((unsigned)(b >> 32), 0) << 32) | (unsigned long long)__shfl((unsigned)b, 0);
        victim = __shfl(victim, 0);
        if (b == ~0ull) break;
        const size_t it = (size_t)b + lane;
        int tile, l, e;
        tile_map(a.map, it, tile, l, e);
        const int tile0 = __builtin_amdgcn_readfirstlane((int)(b >> 10) * 2);
        const int s = seg_of_tile(seg, nseg, tile0);
        const JobDesc& J = jobs[__builtin_amdgcn_readfirstlane(seg[s].job)];
        int probe, r;
        decode_slot(J, seg[s].tb, seg[s].pt0, tile, l, e, probe, r);
        const float4 dd = J.dirs[r];
        float3 L = make_float3(0.f, 0.f, 0.f);
        if (probe < J.P && dd.w != 0.0f) {
            const float4 pp = J.probes[probe];
            int hp = -1;
            L = shade_ray<COUNT, false, FAR, kStriClosest | kStriShadow>(J.sc, J.sp, make_float3(pp.x, pp.y, pp.z), make_float3(dd.x, dd.y, dd.z),
                                 J.tmin, J.tmax, st, cnt, nshadow, hp);
        }
        float* out = a.rad + (size_t)tile * 1536 + l * 8 + e;
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

// Radiance per ray of the round, in MFMA operand order.  A workgroup's 256
// slots lie in one trace tile, hence one segment and one job.
__global__ __launch_bounds__(256) void k_shade_jobs(MultiArgs a, const JobDesc* __restrict__ jobs,
                                                    const Segment* __restrict__ seg,
                                                    const WorkHeader* __restrict__ hdr)
{
    const int nseg = hdr->nseg;
    const size_t nslots = (size_t)hdr->ntrace * 512;
    for (size_t base = (size_t)blockIdx.x * 256; base < nslots; base += (size_t)gridDim.x * 256) {
        const int tile = __builtin_amdgcn_readfirstlane((int)(base >> 9));
        const int s = seg_of_tile(seg, nseg, tile);
        const JobDesc& J = jobs[__builtin_amdgcn_readfirstlane(seg[s].job)];
        const size_t slot = base + threadIdx.x;
        const int t = (int)(slot & 511);
        int probe, r;
        decode_slot(J, seg[s].tb, seg[s].pt0, tile, t >> 3, t & 7, probe, r);
        const int4 h = a.hits[slot];
        float3 L = make_float3(0.f, 0.f, 0.f);
        if (h.x != -2) {
            const float4 dd = J.dirs[r];
            L = shade_record(J.sp, J.sc.shade, J.probes + probe, make_float3(dd.x, dd.y, dd.z), h,
                             a.vis + slot * a.nvis);
        }
        float* out = a.rad + (size_t)tile * 1536 + t;
        out[0] = L.x;
        out[512] = L.y;
        out[1024] = L.z;
    }
}

// SH projection of every projection tile of the round.
template <int PREC>
__global__ __launch_bounds__(256) void k_project_jobs(MultiArgs a, const JobDesc* __restrict__ jobs,
                                                      const Segment* __restrict__ seg,
                                                      const WorkHeader* __restrict__ hdr)
{
    __shared__ float red[3 * 24 * 64];
    const int nseg = hdr->nseg;
    const int nproj = hdr->nproj;
    for (int ptr = blockIdx.x; ptr < nproj; ptr += gridDim.x) {
        const int s = seg_of_proj(seg, nseg, ptr);
        const JobDesc& J = jobs[__builtin_amdgcn_readfirstlane(seg[s].job)];
        const int q = ptr - seg[s].pb;
        const int pt = seg[s].pt0 + q;
        const int pg = pt / J.nslices;
        const int sl = pt - pg * J.nslices;
        const int nch = J.RS >> 5;
        const float* rad0 = a.rad + ((size_t)seg[s].tb + (size_t)q * nch) * 1536;
        float2* prow = J.partials + ((size_t)sl * J.Ppad + (size_t)pg * 16) * 48;
        project_tile<PREC>(rad0, nch, sl * nch, J.ytab, (size_t)J.Rpad * 2, red, prow);
    }
}

// FP64 cross-slice merge of the jobs listed in the round's merge table:
// out[p][k][c] = scale * sum_s (sum_s - comp_s), slices in order.
__global__ __launch_bounds__(256) void k_merge_jobs(MultiArgs a, const JobDesc* __restrict__ jobs,
                                                    const MergeItem* __restrict__ merge,
                                                    const WorkHeader* __restrict__ hdr)
{
    const int nmerge = hdr->nmerge;
    const int total = hdr->merge_threads;
    for (int base = blockIdx.x * 256; base < total; base += gridDim.x * 256) {
        int lo = 0, hi = nmerge - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (merge[mid].base <= base) lo = mid;
            else hi = mid - 1;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        const JobDesc& J = jobs[__builtin_amdgcn_readfirstlane(merge[lo].job)];
        const int local = base - merge[lo].base + (int)threadIdx.x;
        if (local >= J.Ppad * 48) continue;
        const int ps = local / 48;
        const int rem = local - ps * 48;
        const int c = rem >> 4;
        const int k = rem & 15;
        if (k >= J.K) continue;
        const int op = J.probe_perm[ps];
        if (op < 0) continue;
        double s = 0.0;
        for (int sl = 0; sl < J.nslices; ++sl) {
            const float2 v = J.partials[((size_t)sl * J.Ppad + ps) * 48 + rem];
            s += (double)v.x - (double)v.y;
        }
        s *= J.scale;
        const size_t o = ((size_t)op * J.K + k) * 3 + c;
        J.out[o] = (float)s;
        if (J.out64) J.out64[o] = s;
    }
}

// Clears the first n work counters (128 bytes apart).
__global__ void k_zero_ctr(unsigned long long* p, int n)
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

int multi_fused_grid_size(Context& ctx, int occ, bool far)
{
    int g;
    if (far) {
        if (occ == 8) g = occupancy_of(k_fused_jobs<false, 8, true>, ctx.num_cus);
        else if (occ == 6) g = occupancy_of(k_fused_jobs<false, 6, true>, ctx.num_cus);
        else g = occupancy_of(k_fused_jobs<false, 0, true>, ctx.num_cus);
    } else if (occ == 8) g = occupancy_of(k_fused_jobs<false, 8>, ctx.num_cus);
    else if (occ == 6) g = occupancy_of(k_fused_jobs<false, 6>, ctx.num_cus);
    else g = occupancy_of(k_fused_jobs<false, 0>, ctx.num_cus);
    return (g + 7) & ~7;
}

int multi_grid_size(Context& ctx)
{
    const int g = std::max(occupancy_of(k_rays_jobs<false, false>, ctx.num_cus),
                           occupancy_of(k_rays_jobs<true, false>, ctx.num_cus));
    return (g + 7) & ~7;
}

void launch_multi_trace(const MultiArgs& a, bool count, int grid, int shade_grid, hipStream_t st)
{
    VLB_REQUIRE(grid % 8 == 0 && grid > 0, VLB_ERR_INVALID, "multi-job grid must be a multiple of 8");
    VLB_REQUIRE(a.chunk >= 64 && a.chunk <= 1024 && (1024 % a.chunk) == 0, VLB_ERR_INVALID,
                "work chunk must divide 1024");
    k_zero_ctr<<<1, 64, 0, st>>>(a.wctr, 32);
    if (a.focc >= 0) {
        VLB_REQUIRE(a.fgrid % 8 == 0 && a.fgrid > 0, VLB_ERR_INVALID, "fused grid must be a multiple of 8");
        if (count && a.far) k_fused_jobs<true, 0, true><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        else if (count) k_fused_jobs<true, 0><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        else if (a.far) {
            if (a.focc == 8) k_fused_jobs<false, 8, true><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
            else if (a.focc == 6) k_fused_jobs<false, 6, true><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
            else k_fused_jobs<false, 0, true><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        } else if (a.focc == 8) k_fused_jobs<false, 8><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        else if (a.focc == 6) k_fused_jobs<false, 6><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        else k_fused_jobs<false, 0><<<a.fgrid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        VLB_HIP(hipGetLastError());
        return;
    }
    if (count) k_rays_jobs<false, true><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
    else k_rays_jobs<false, false><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
    VLB_HIP(hipGetLastError());
    if (a.nvis > 0) {
        if (count) k_rays_jobs<true, true><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        else k_rays_jobs<true, false><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        VLB_HIP(hipGetLastError());
    }
    k_shade_jobs<<<shade_grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
    VLB_HIP(hipGetLastError());
}

void launch_multi_project(const MultiArgs& a, int precision, int grid, hipStream_t st)
{
    switch (precision) {
    case VLB_PREC_BF16: k_project_jobs<VLB_PREC_BF16><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr); break;
    case VLB_PREC_FP16: k_project_jobs<VLB_PREC_FP16><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr); break;
    case VLB_PREC_BF16_FAST:
        k_project_jobs<VLB_PREC_BF16_FAST><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        break;
    case VLB_PREC_FP16_FAST:
        k_project_jobs<VLB_PREC_FP16_FAST><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr);
        break;
    default: k_project_jobs<VLB_PREC_FP32><<<grid, 256, 0, st>>>(a, a.jobs, a.seg, a.hdr); break;
    }
    VLB_HIP(hipGetLastError());
}

void launch_multi_merge(const MultiArgs& a, int grid, hipStream_t st)
{
    k_merge_jobs<<<grid, 256, 0, st>>>(a, a.jobs, a.merge, a.hdr);
    VLB_HIP(hipGetLastError());
}

}  // namespace vlb
