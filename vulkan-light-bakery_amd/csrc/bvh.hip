// bvh.hip -- GPU BVH build for the probe baker (replaces the Vulkan BLAS/TLAS
// build of src/scene_manager.cpp:339-443; the reference hands triangles to
// the driver with ePreferFastTrace and culling disabled).
//
// 30-bit Morton codes of triangle centroids, a stable radix sort, then a
// PLOC hierarchy (agglomerative clustering over the Morton order, below)
// with one triangle per leaf, collapsed to the 4-wide BVH the traversal
// kernels walk.  Node boxes are unions of their children's boxes, i.e. the
// exact min/max of their triangles' vertices (conservative for the
// watertight test), and every step is deterministic.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "vlb_internal.hpp"

namespace vlb {
namespace {

// Gather triangles, per-prim boxes and centroids, shading records.
__global__ void k_gather(const float* __restrict__ xyz, const uint32_t* __restrict__ idx,
                         const float* __restrict__ normals, const float* __restrict__ albedo,
                         int ntris, int nverts, float4* __restrict__ tri_raw,
                         float4* __restrict__ lo, float4* __restrict__ hi,
                         float4* __restrict__ cen, float4* __restrict__ shade)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntris) return;
    uint32_t i0 = idx[3 * i], i1 = idx[3 * i + 1], i2 = idx[3 * i + 2];
    if (i0 >= (uint32_t)nverts) i0 = 0;
    if (i1 >= (uint32_t)nverts) i1 = 0;
    if (i2 >= (uint32_t)nverts) i2 = 0;
    float3 v0 = make_float3(xyz[3 * i0], xyz[3 * i0 + 1], xyz[3 * i0 + 2]);
    float3 v1 = make_float3(xyz[3 * i1], xyz[3 * i1 + 1], xyz[3 * i1 + 2]);
    float3 v2 = make_float3(xyz[3 * i2], xyz[3 * i2 + 1], xyz[3 * i2 + 2]);
    tri_raw[3 * i + 0] = make_float4(v0.x, v0.y, v0.z, __int_as_float(i));
    tri_raw[3 * i + 1] = make_float4(v1.x, v1.y, v1.z, 0.0f);
    tri_raw[3 * i + 2] = make_float4(v2.x, v2.y, v2.z, 0.0f);
    float4 l = make_float4(fminf(v0.x, fminf(v1.x, v2.x)), fminf(v0.y, fminf(v1.y, v2.y)),
                           fminf(v0.z, fminf(v1.z, v2.z)), 0.0f);
    float4 h = make_float4(fmaxf(v0.x, fmaxf(v1.x, v2.x)), fmaxf(v0.y, fmaxf(v1.y, v2.y)),
                           fmaxf(v0.z, fmaxf(v1.z, v2.z)), 0.0f);
    lo[i] = l;
    hi[i] = h;
    cen[i] = make_float4(0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z), 0.0f);
    float3 n0, n1, n2;
    if (normals) {
        n0 = make_float3(normals[3 * i0], normals[3 * i0 + 1], normals[3 * i0 + 2]);
        n1 = make_float3(normals[3 * i1], normals[3 * i1 + 1], normals[3 * i1 + 2]);
        n2 = make_float3(normals[3 * i2], normals[3 * i2 + 1], normals[3 * i2 + 2]);
    } else {
        // geometric normal, same operation sequence as the host helper
        float3 e1 = make_float3(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z);
        float3 e2 = make_float3(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
        float3 n = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z,
                               e1.x * e2.y - e1.y * e2.x);
        float l2 = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
        float3 g = l2 > 0.0f ? make_float3(n.x / l2, n.y / l2, n.z / l2)
                             : make_float3(0.0f, 1.0f, 0.0f);
        n0 = n1 = n2 = g;
    }
    float ar = albedo ? albedo[3 * i] : 1.0f;
    float ag = albedo ? albedo[3 * i + 1] : 1.0f;
    float ab = albedo ? albedo[3 * i + 2] : 1.0f;
    shade[3 * i + 0] = make_float4(n0.x, n0.y, n0.z, ar);
    shade[3 * i + 1] = make_float4(n1.x, n1.y, n1.z, ag);
    shade[3 * i + 2] = make_float4(n2.x, n2.y, n2.z, ab);
}

// Bounds reduction: centroid box (for Morton), scene box, max |coordinate|.
// Stage 1: one partial of 13 floats per block; stage 2: one block.
__device__ __forceinline__ void red13(float* v, float* sh)
{
    // v[0..2] min, v[3..5] max, v[6..8] min, v[9..11] max, v[12] max
    const int t = threadIdx.x;
    for (int k = 0; k < 13; ++k) sh[k * 256 + t] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 13; ++k) {
                float a = sh[k * 256 + t], b = sh[k * 256 + t + s];
                bool is_min = (k < 3) || (k >= 6 && k < 9);
                sh[k * 256 + t] = is_min ? fminf(a, b) : fmaxf(a, b);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_bounds1(const float4* __restrict__ lo,
                                                 const float4* __restrict__ hi,
                                                 const float4* __restrict__ cen, int n,
                                                 float* __restrict__ part)
{
    __shared__ float sh[13 * 256];
    float v[13];
    for (int k = 0; k < 3; ++k) {
        v[k] = v[6 + k] = __builtin_huge_valf();
        v[3 + k] = v[9 + k] = -__builtin_huge_valf();
    }
    v[12] = 0.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float4 c = cen[i], l = lo[i], h = hi[i];
        v[0] = fminf(v[0], c.x);
        v[1] = fminf(v[1], c.y);
        v[2] = fminf(v[2], c.z);
        v[3] = fmaxf(v[3], c.x);
        v[4] = fmaxf(v[4], c.y);
        v[5] = fmaxf(v[5], c.z);
        v[6] = fminf(v[6], l.x);
        v[7] = fminf(v[7], l.y);
        v[8] = fminf(v[8], l.z);
        v[9] = fmaxf(v[9], h.x);
        v[10] = fmaxf(v[10], h.y);
        v[11] = fmaxf(v[11], h.z);
        float m = fmaxf(fmaxf(fabsf(l.x), fabsf(l.y)), fmaxf(fabsf(l.z), fabsf(h.x)));
        m = fmaxf(m, fmaxf(fabsf(h.y), fabsf(h.z)));
        v[12] = fmaxf(v[12], m);
    }
    red13(v, sh);
    if (threadIdx.x < 13) part[blockIdx.x * 13 + threadIdx.x] = sh[threadIdx.x * 256];
}

__global__ __launch_bounds__(256) void k_bounds2(const float* __restrict__ part, int nb,
                                                 float* __restrict__ out)
{
    __shared__ float sh[13 * 256];
    float v[13];
    for (int k = 0; k < 3; ++k) {
        v[k] = v[6 + k] = __builtin_huge_valf();
        v[3 + k] = v[9 + k] = -__builtin_huge_valf();
    }
    v[12] = 0.0f;
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int k = 0; k < 13; ++k) {
            float x = part[b * 13 + k];
            bool is_min = (k < 3) || (k >= 6 && k < 9);
            v[k] = is_min ? fminf(v[k], x) : fmaxf(v[k], x);
        }
    red13(v, sh);
    if (threadIdx.x < 13) out[threadIdx.x] = sh[threadIdx.x * 256];
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// 30-bit Morton code of each triangle's box centre on a grid of cubic cells:
// 1024 cells along the longest extent of the centre bounds, the same cell
// size on the other axes.  With cells stretched to 1024 per axis (2.5-6x
// finer vertically on the flat terrains and rooms) the PLOC pass built
// trees that traced 2-6% slower on three of the four scenes of the batched
// jobs and the same on the fourth (208.3 against 202.4 ms per batch); the
// 11k-triangle room bake of 2048 probes traces 2% faster with them (0.99
// against 1.01 ms).
__global__ void k_morton(const float4* __restrict__ cen, int n, const float* __restrict__ bounds,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = cen[i];
    float q[3];
    const float cc[3] = {c.x, c.y, c.z};
    const float ext = fmaxf(fmaxf(bounds[3] - bounds[0], bounds[4] - bounds[1]), bounds[5] - bounds[2]);
    for (int a = 0; a < 3; ++a) {
        float x = ext > 0.0f ? (cc[a] - bounds[a]) / ext * 1024.0f : 0.0f;
        q[a] = fminf(fmaxf(x, 0.0f), 1023.0f);
    }
    keys[i] = (expand_bits((uint32_t)q[0]) << 2) | (expand_bits((uint32_t)q[1]) << 1) |
              expand_bits((uint32_t)q[2]);
    vals[i] = (uint32_t)i;
}

// ------------------------------------------------------------- PLOC
// Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding
// Volume Hierarchy Construction", IEEE TVCG 24(3), 2018.  Clusters start as
// the Morton-ordered leaves; each pass every cluster finds the neighbour
// within kPlocRadius positions whose union box has the least surface area,
// mutual pairs merge into a new inner node at the lower position, and the
// cluster list is compacted.  Candidate pairs are ordered by (area, distance
// in positions, parity of the lower position, lower position): a strict total
// order, so the least pair of a pass is always mutual and every pass merges
// at least once; and where areas tie (coincident or degenerate triangles),
// neighbours pair up as (0,1), (2,3), ... instead of chaining, so such runs
// halve each pass.  The build is deterministic (no atomics; positions and
// node ids come from one scan).  Node ids are handed out downward, so the
// final merge -- the root -- is node 0.  Search radius 32: on the batched
// jobs (rooms and terrains of 11k-1.15M triangles) 1.5% fewer node visits
// than radius 16 and a 1.3% shorter trace stage; 24 and 48 were no better,
// and the 11k-triangle room and the 1.15M-triangle terrain bakes trace in
// the same time at 16-64.
constexpr int kPlocRadius = 32;

// Leaf-order triangles and the initial clusters (leaf j = ~j).
__global__ void k_leaves(const uint32_t* __restrict__ order, int n,
                         const float4* __restrict__ tri_raw, const float4* __restrict__ lo,
                         const float4* __restrict__ hi, float4* __restrict__ tris,
                         float4* __restrict__ clo, float4* __restrict__ chi, int* __restrict__ cid)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t p = order[j];
    tris[3 * j + 0] = tri_raw[3 * p + 0];
    tris[3 * j + 1] = tri_raw[3 * p + 1];
    tris[3 * j + 2] = tri_raw[3 * p + 2];
    clo[j] = lo[p];
    chi[j] = hi[p];
    cid[j] = ~j;
}

// Candidate j = i + o against the best so far.  A NaN area (non-finite
// input) counts as +inf so the pair order stays total.  Tie key for this i:
// distance, then the lower position's parity, then the lower position (the
// left candidate is the lower one).
__device__ __forceinline__ void ploc_consider(const float4& a, const float4& b, const float4& l,
                                              const float4& h, int i, int o, float& best,
                                              int& bj, int& bk)
{
    const float dx = fmaxf(b.x, h.x) - fminf(a.x, l.x);
    const float dy = fmaxf(b.y, h.y) - fminf(a.y, l.y);
    const float dz = fmaxf(b.z, h.z) - fminf(a.z, l.z);
    float ar = dx * dy + dy * dz + dz * dx;
    ar = ar == ar ? ar : __builtin_huge_valf();
    const int k = 4 * abs(o) + 2 * (min(i, i + o) & 1) + (o > 0 ? 1 : 0);
    if (bj < 0 || ar < best || (ar == best && k < bk)) {
        best = ar;
        bj = i + o;
        bk = k;
    }
}

// Nearest neighbour of every cluster within kPlocRadius positions; the
// block's window of boxes is staged in LDS.
__global__ __launch_bounds__(256) void k_ploc_nn(const float4* __restrict__ clo,
                                                 const float4* __restrict__ chi, int m,
                                                 int* __restrict__ nn)
{
    constexpr int W = 256 + 2 * kPlocRadius;
    __shared__ float4 slo[W], shi[W];
    const int base = blockIdx.x * 256;
    for (int t = threadIdx.x; t < W; t += 256) {
        const int g = base - kPlocRadius + t;
        if (g >= 0 && g < m) {
            slo[t] = clo[g];
            shi[t] = chi[g];
        }
    }
    __syncthreads();
    const int i = base + threadIdx.x;
    if (i >= m) return;
    const int li = threadIdx.x + kPlocRadius;
    const float4 a = slo[li], b = shi[li];
    float best = __builtin_huge_valf();
    int bj = -1, bk = 0;
    const int o0 = max(-kPlocRadius, -i), o1 = min(kPlocRadius, m - 1 - i);
    for (int o = o0; o <= o1; ++o)
        if (o != 0) ploc_consider(a, b, slo[li + o], shi[li + o], i, o, best, bj, bk);
    nn[i] = bj;
}

// The last passes (m <= kPlocFull clusters, the top of the tree): exact
// nearest neighbour over all clusters, one workgroup.
constexpr int kPlocFull = 1024;
__global__ __launch_bounds__(kPlocFull) void k_ploc_nn_full(const float4* __restrict__ clo,
                                                            const float4* __restrict__ chi,
                                                            int m, int* __restrict__ nn)
{
    __shared__ float4 slo[kPlocFull], shi[kPlocFull];
    const int i = threadIdx.x;
    if (i < m) {
        slo[i] = clo[i];
        shi[i] = chi[i];
    }
    __syncthreads();
    if (i >= m) return;
    const float4 a = slo[i], b = shi[i];
    float best = __builtin_huge_valf();
    int bj = -1, bk = 0;
    for (int j = 0; j < m; ++j)
        if (j != i) ploc_consider(a, b, slo[j], shi[j], i, j - i, best, bj, bk);
    nn[i] = bj;
}

// Per cluster: low word 1 if it survives the pass (unpaired, or the lower
// of a mutual pair), high word 1 if it is the lower of a mutual pair.
__global__ void k_ploc_flags(const int* __restrict__ nn, int m,
                             unsigned long long* __restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    unsigned long long f = 0;
    if (!mutual) f = 1ull;
    else if (i < j) f = (1ull << 32) | 1ull;
    flag[i] = f;
}

// Compaction and merge.  scan = inclusive sum of the flags; with M merges
// in a pass over m clusters, the pass's nodes are [m - 1 - M, m - 1).
__global__ void k_ploc_merge(const int* __restrict__ nn, int m,
                             const unsigned long long* __restrict__ scan,
                             const float4* __restrict__ clo, const float4* __restrict__ chi,
                             const int* __restrict__ cid, float4* __restrict__ olo,
                             float4* __restrict__ ohi, int* __restrict__ ocid,
                             BNode* __restrict__ nodes, int* __restrict__ iparent,
                             int* __restrict__ lparent)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = nn[j] == i;
    if (mutual && j < i) return;  // absorbed by its partner
    const unsigned long long f = mutual ? (1ull << 32) | 1ull : 1ull;
    const unsigned long long ex = scan[i] - f;
    const int pos = (int)(uint32_t)ex;
    if (!mutual) {
        olo[pos] = clo[i];
        ohi[pos] = chi[i];
        ocid[pos] = cid[i];
        return;
    }
    const int merges = (int)(scan[m - 1] >> 32);
    const int id = m - 1 - merges + (int)(ex >> 32);
    const float4 l0 = clo[i], h0 = chi[i], l1 = clo[j], h1 = chi[j];
    const int c0 = cid[i], c1 = cid[j];
    BNode n;
    n.a = make_float4(l0.x, l0.y, l0.z, h0.x);
    n.b = make_float4(h0.y, h0.z, l1.x, l1.y);
    n.c = make_float4(l1.z, h1.x, h1.y, h1.z);
    n.d = make_int4(c0, c1, 0, 0);
    nodes[id] = n;
    if (c0 >= 0) iparent[c0] = 2 * id + 0;
    else lparent[~c0] = 2 * id + 0;
    if (c1 >= 0) iparent[c1] = 2 * id + 1;
    else lparent[~c1] = 2 * id + 1;
    olo[pos] = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.0f);
    ohi[pos] = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.0f);
    ocid[pos] = id;
}

// One triangle: a root whose two children are the same leaf; testing it
// twice cannot change a result.
__global__ void k_single(const float4* __restrict__ clo, const float4* __restrict__ chi,
                         BNode* __restrict__ nodes, int* __restrict__ iparent,
                         int* __restrict__ lparent)
{
    const float4 l = clo[0], h = chi[0];
    BNode n;
    n.a = make_float4(l.x, l.y, l.z, h.x);
    n.b = make_float4(h.y, h.z, l.x, l.y);
    n.c = make_float4(l.z, h.x, h.y, h.z);
    n.d = make_int4(~0, ~0, 0, 0);
    nodes[0] = n;
    iparent[0] = -1;
    lparent[0] = 0;
}

// Depth = inner nodes on the longest root-to-leaf path.
__global__ void k_depth(const int* __restrict__ iparent, const int* __restrict__ lparent, int n,
                        unsigned int* __restrict__ depth)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int p = lparent[j] >> 1;
    unsigned int d = 1;
    while (p != 0) {
        p = iparent[p] >> 1;
        ++d;
        if (d > 4096) break;  // malformed tree guard
    }
    atomicMax(depth, d);
}

// Depth of every inner node (root 0), for the BVH4 collapse.
__global__ void k_inner_depth(const int* __restrict__ iparent, int nin, int* __restrict__ depth,
                              int* __restrict__ even)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nin) return;
    int d = 0, p = i;
    while (p != 0 && d < 8192) {
        p = iparent[p] >> 1;
        ++d;
    }
    depth[i] = d;
    even[i] = (d & 1) == 0 ? 1 : 0;
}

__device__ __forceinline__ void child_box(const BNode& a, int k, float b[6])
{
    if (k == 0) {
        b[0] = a.a.x; b[1] = a.a.y; b[2] = a.a.z; b[3] = a.a.w; b[4] = a.b.x; b[5] = a.b.y;
    } else {
        b[0] = a.b.z; b[1] = a.b.w; b[2] = a.c.x; b[3] = a.c.y; b[4] = a.c.z; b[5] = a.c.w;
    }
}

// Grid of one axis of a BVH4 node's box [lo, hi] (finite, |x| <= 2^100):
// origin o = lo and the smallest float step s with fma(255, s, o) >= hi.  A
// plane q decodes as fma(q, s, o) here and in the traversal (the same
// correctly rounded operation, non-decreasing in q), so codes chosen by
// comparing that decoded value with the float plane are conservative
// whatever the rounding.
__host__ __device__ void axis_grid(float lo, float hi, float& o, float& s)
{
    o = lo;
    const double want = ((double)hi - (double)lo) / 255.0;
    s = (float)want;
    if ((double)s < want) s = nextafterf(s, __builtin_huge_valf());
    // s >= the exact extent / 255 makes fma(255, s, o) >= hi already; the
    // loop only covers a rounded `want`, first in ulps, then by doubling
    for (int i = 0; fmaf(255.0f, s, o) < hi; ++i)
        s = i < 8 ? nextafterf(s, __builtin_huge_valf()) : fmaxf(2.0f * s, 0x1p-100f);
}

// 8-bit codes of one plane of the four children: the largest q whose plane
// is <= the child's lo (UP false), or the smallest q whose plane is >= its
// hi (UP true).
template <bool UP>
__host__ __device__ uint32_t plane_codes(const float v[4], int m, float o, float s)
{
    uint32_t w = 0;
    for (int k = 0; k < m; ++k) {
        const float x = v[k];
        int q = 0;
        if (s > 0.0f) {
            const double r = ((double)x - (double)o) / (double)s;
            q = (int)fmin(fmax(UP ? ceil(r) : floor(r), 0.0), 255.0);
            if (UP) {
                while (q < 255 && fmaf((float)q, s, o) < x) ++q;
                while (q > 0 && fmaf((float)(q - 1), s, o) >= x) --q;
            } else {
                while (q > 0 && fmaf((float)q, s, o) > x) --q;
                while (q < 255 && fmaf((float)(q + 1), s, o) <= x) ++q;
            }
        }
        w |= (uint32_t)q << (8 * k);
    }
    return w;
}

// BVH2 -> BVH4: every even-depth BVH2 node becomes a BVH4 node whose
// children are its grandchildren through inner children, or the child
// itself when that is a leaf (2 to 4 entries).  The float node copies the
// boxes (exactly as conservative as the BVH2); the quantized node rounds
// them outward on the node's grid.
__global__ void k_collapse(const BNode* __restrict__ n2, const int* __restrict__ depth,
                           const int* __restrict__ idx4, int nin, BNode4* __restrict__ n4,
                           QNode4* __restrict__ q4)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nin || (depth[i] & 1)) return;
    float bx[4][6];
    int ch[4];
    int m = 0;
    const BNode a = n2[i];
    for (int k = 0; k < 2; ++k) {
        const int c = k == 0 ? a.d.x : a.d.y;
        if (c < 0) {
            child_box(a, k, bx[m]);
            ch[m++] = c;
            continue;
        }
        const BNode g = n2[c];
        for (int q = 0; q < 2; ++q) {
            const int gc = q == 0 ? g.d.x : g.d.y;
            child_box(g, q, bx[m]);
            ch[m++] = gc < 0 ? gc : idx4[gc];
        }
    }
    const int used = m;
    // Unused slots: in the float node an all-+inf box, which the slab test
    // misses for every ray (each axis gives t = +inf or -inf, or NaN, which
    // min/max drop, and a unit direction keeps one axis finite), with leaf 0
    // as its id, so that even a box it did not miss would only re-test a
    // real triangle; in the quantized node kEmpty (its decoded planes are
    // finite, slab4<true> checks the id).
    int chf[4];
    for (int k = 0; k < used; ++k) chf[k] = ch[k];
    for (; m < 4; ++m) {
        for (int t = 0; t < 6; ++t) bx[m][t] = __builtin_huge_valf();
        ch[m] = kEmpty;
        chf[m] = ~0;
    }
    BNode4 o;
    o.lox = make_float4(bx[0][0], bx[1][0], bx[2][0], bx[3][0]);
    o.loy = make_float4(bx[0][1], bx[1][1], bx[2][1], bx[3][1]);
    o.loz = make_float4(bx[0][2], bx[1][2], bx[2][2], bx[3][2]);
    o.hix = make_float4(bx[0][3], bx[1][3], bx[2][3], bx[3][3]);
    o.hiy = make_float4(bx[0][4], bx[1][4], bx[2][4], bx[3][4]);
    o.hiz = make_float4(bx[0][5], bx[1][5], bx[2][5], bx[3][5]);
    o.child = make_int4(chf[0], chf[1], chf[2], chf[3]);
    o.pad = make_int4(0, 0, 0, 0);
    n4[idx4[i]] = o;
    float org[3], stp[3];
    uint32_t code[6];
    for (int a = 0; a < 3; ++a) {
        float lo[4], hi[4];
        float blo = bx[0][a], bhi = bx[0][3 + a];
        for (int k = 0; k < used; ++k) {
            lo[k] = bx[k][a];
            hi[k] = bx[k][3 + a];
            blo = fminf(blo, lo[k]);
            bhi = fmaxf(bhi, hi[k]);
        }
        axis_grid(blo, bhi, org[a], stp[a]);
        code[a] = plane_codes<false>(lo, used, org[a], stp[a]);
        code[3 + a] = plane_codes<true>(hi, used, org[a], stp[a]);
    }
    QNode4 q;
    q.a = make_float4(org[0], org[1], org[2], stp[0]);
    q.b = make_float4(stp[1], stp[2], __uint_as_float(code[0]), __uint_as_float(code[1]));
    q.q = make_uint4(code[2], code[3], code[4], code[5]);
    q.child = make_int4(ch[0], ch[1], ch[2], ch[3]);
    q4[idx4[i]] = q;
}

inline int nblk(int n, int b = 256) { return (n + b - 1) / b; }

}  // namespace

void build_scene_bvh(Context& ctx, Scene& s, const float* xyz, uint32_t nverts,
                     const uint32_t* idx, uint32_t ntris, const float* normals,
                     const float* albedo)
{
    hipStream_t st = ctx.stream;
    const int n = (int)ntris;
    s.ntris = n;
    if (n == 0) {
        s.nnodes = 0;
        s.depth = 0;
        return;
    }
    hipEvent_t e0, e1;
    VLB_HIP(hipEventCreate(&e0));
    VLB_HIP(hipEventCreate(&e1));
    // uploads
    DevBuf d_xyz = dev_alloc(ctx, (size_t)nverts * 3 * sizeof(float));
    DevBuf d_idx = dev_alloc(ctx, (size_t)n * 3 * sizeof(uint32_t));
    DevBuf d_nrm, d_alb;
    VLB_HIP(hipMemcpyAsync(d_xyz.p, xyz, d_xyz.bytes, hipMemcpyHostToDevice, st));
    VLB_HIP(hipMemcpyAsync(d_idx.p, idx, d_idx.bytes, hipMemcpyHostToDevice, st));
    if (normals) {
        d_nrm = dev_alloc(ctx, (size_t)nverts * 3 * sizeof(float));
        VLB_HIP(hipMemcpyAsync(d_nrm.p, normals, d_nrm.bytes, hipMemcpyHostToDevice, st));
    }
    if (albedo) {
        d_alb = dev_alloc(ctx, (size_t)n * 3 * sizeof(float));
        VLB_HIP(hipMemcpyAsync(d_alb.p, albedo, d_alb.bytes, hipMemcpyHostToDevice, st));
    }
    VLB_HIP(hipEventRecord(e0, st));
    DevBuf tri_raw = dev_alloc(ctx, (size_t)n * 3 * sizeof(float4));
    DevBuf lo = dev_alloc(ctx, (size_t)n * sizeof(float4));
    DevBuf hi = dev_alloc(ctx, (size_t)n * sizeof(float4));
    DevBuf cen = dev_alloc(ctx, (size_t)n * sizeof(float4));
    s.shade = dev_alloc(ctx, (size_t)n * 3 * sizeof(float4));
    k_gather<<<nblk(n), 256, 0, st>>>(d_xyz.as<float>(), d_idx.as<uint32_t>(), d_nrm.as<float>(),
                                      d_alb.as<float>(), n, (int)nverts, tri_raw.as<float4>(),
                                      lo.as<float4>(), hi.as<float4>(), cen.as<float4>(),
                                      s.shade.as<float4>());
    VLB_HIP(hipGetLastError());
    const int nb = std::min(1024, nblk(n));
    DevBuf part = dev_alloc(ctx, (size_t)nb * 13 * sizeof(float));
    DevBuf bnd = dev_alloc(ctx, 16 * sizeof(float));
    k_bounds1<<<nb, 256, 0, st>>>(lo.as<float4>(), hi.as<float4>(), cen.as<float4>(), n,
                                  part.as<float>());
    k_bounds2<<<1, 256, 0, st>>>(part.as<float>(), nb, bnd.as<float>());
    VLB_HIP(hipGetLastError());
    // Morton codes + sort
    DevBuf keys = dev_alloc(ctx, (size_t)n * sizeof(uint32_t));
    DevBuf vals = dev_alloc(ctx, (size_t)n * sizeof(uint32_t));
    DevBuf keys2 = dev_alloc(ctx, (size_t)n * sizeof(uint32_t));
    DevBuf vals2 = dev_alloc(ctx, (size_t)n * sizeof(uint32_t));
    k_morton<<<nblk(n), 256, 0, st>>>(cen.as<float4>(), n, bnd.as<float>(), keys.as<uint32_t>(),
                                      vals.as<uint32_t>());
    VLB_HIP(hipGetLastError());
    size_t tmp_bytes = 0;
    VLB_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint32_t>(), keys2.as<uint32_t>(),
                                      vals.as<uint32_t>(), vals2.as<uint32_t>(), (size_t)n, 0, 30, st));
    DevBuf tmp = dev_alloc(ctx, tmp_bytes > 0 ? tmp_bytes : 16);
    VLB_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint32_t>(), keys2.as<uint32_t>(),
                                      vals.as<uint32_t>(), vals2.as<uint32_t>(), (size_t)n, 0, 30, st));
    // hierarchy (PLOC), one host round trip per pass for the cluster count
    const int nin = n > 1 ? n - 1 : 1;
    s.nodes = dev_alloc(ctx, (size_t)nin * sizeof(BNode));
    s.tris = dev_alloc(ctx, (size_t)n * 3 * sizeof(float4));
    DevBuf iparent = dev_alloc(ctx, (size_t)nin * sizeof(int));
    DevBuf lparent = dev_alloc(ctx, (size_t)n * sizeof(int));
    DevBuf clo[2], chi[2], cid[2];
    for (int k = 0; k < 2; ++k) {
        clo[k] = dev_alloc(ctx, (size_t)n * sizeof(float4));
        chi[k] = dev_alloc(ctx, (size_t)n * sizeof(float4));
        cid[k] = dev_alloc(ctx, (size_t)n * sizeof(int));
    }
    k_leaves<<<nblk(n), 256, 0, st>>>(vals2.as<uint32_t>(), n, tri_raw.as<float4>(),
                                      lo.as<float4>(), hi.as<float4>(), s.tris.as<float4>(),
                                      clo[0].as<float4>(), chi[0].as<float4>(), cid[0].as<int>());
    VLB_HIP(hipGetLastError());
    DevBuf dep = dev_alloc(ctx, sizeof(unsigned int));
    VLB_HIP(hipMemsetAsync(dep.p, 0, sizeof(unsigned int), st));
    if (n == 1) {
        k_single<<<1, 1, 0, st>>>(clo[0].as<float4>(), chi[0].as<float4>(), s.nodes.as<BNode>(),
                                  iparent.as<int>(), lparent.as<int>());
        unsigned int one = 1;
        VLB_HIP(hipMemcpyAsync(dep.p, &one, sizeof one, hipMemcpyHostToDevice, st));
    } else {
        DevBuf nn = dev_alloc(ctx, (size_t)n * sizeof(int));
        DevBuf flag = dev_alloc(ctx, (size_t)n * sizeof(unsigned long long));
        DevBuf scan = dev_alloc(ctx, (size_t)n * sizeof(unsigned long long));
        size_t ps_bytes = 0;
        VLB_HIP(rocprim::inclusive_scan(nullptr, ps_bytes, flag.as<unsigned long long>(),
                                        scan.as<unsigned long long>(), (size_t)n,
                                        rocprim::plus<unsigned long long>(), st));
        DevBuf ps_tmp = dev_alloc(ctx, ps_bytes > 0 ? ps_bytes : 16);
        int m = n, cur = 0;
        // every pass merges at least one pair, so at most n - 1 passes
        for (int pass = 0; m > 1; ++pass) {
            VLB_REQUIRE(pass < n, VLB_ERR_DEVICE, "PLOC build made no progress");
            if (m <= kPlocFull)
                k_ploc_nn_full<<<1, kPlocFull, 0, st>>>(clo[cur].as<float4>(), chi[cur].as<float4>(),
                                                        m, nn.as<int>());
            else
                k_ploc_nn<<<nblk(m), 256, 0, st>>>(clo[cur].as<float4>(), chi[cur].as<float4>(), m,
                                                   nn.as<int>());
            k_ploc_flags<<<nblk(m), 256, 0, st>>>(nn.as<int>(), m, flag.as<unsigned long long>());
            VLB_HIP(rocprim::inclusive_scan(ps_tmp.p, ps_bytes, flag.as<unsigned long long>(),
                                            scan.as<unsigned long long>(), (size_t)m,
                                            rocprim::plus<unsigned long long>(), st));
            k_ploc_merge<<<nblk(m), 256, 0, st>>>(
                nn.as<int>(), m, scan.as<unsigned long long>(), clo[cur].as<float4>(),
                chi[cur].as<float4>(), cid[cur].as<int>(), clo[cur ^ 1].as<float4>(),
                chi[cur ^ 1].as<float4>(), cid[cur ^ 1].as<int>(), s.nodes.as<BNode>(),
                iparent.as<int>(), lparent.as<int>());
            VLB_HIP(hipGetLastError());
            unsigned long long tot = 0;
            VLB_HIP(hipMemcpyAsync(&tot, scan.as<unsigned long long>() + (m - 1), sizeof tot,
                                   hipMemcpyDeviceToHost, st));
            VLB_HIP(hipStreamSynchronize(st));
            const int m2 = (int)(uint32_t)tot;
            VLB_REQUIRE(m2 >= 1 && m2 < m, VLB_ERR_DEVICE, "PLOC build made no progress");
            m = m2;
            cur ^= 1;
        }
        const int root_parent = -1;
        VLB_HIP(hipMemcpyAsync(iparent.p, &root_parent, sizeof root_parent, hipMemcpyHostToDevice, st));
        k_depth<<<nblk(n), 256, 0, st>>>(iparent.as<int>(), lparent.as<int>(), n,
                                         dep.as<unsigned int>());
    }
    VLB_HIP(hipGetLastError());
    // BVH4: even-depth BVH2 nodes, numbered by an exclusive scan
    DevBuf idepth = dev_alloc(ctx, (size_t)nin * sizeof(int));
    DevBuf even = dev_alloc(ctx, (size_t)nin * sizeof(int));
    DevBuf idx4 = dev_alloc(ctx, (size_t)nin * sizeof(int));
    if (n == 1) {
        const int zero = 0, one = 1;
        VLB_HIP(hipMemcpyAsync(idepth.p, &zero, sizeof zero, hipMemcpyHostToDevice, st));
        VLB_HIP(hipMemcpyAsync(even.p, &one, sizeof one, hipMemcpyHostToDevice, st));
    } else {
        k_inner_depth<<<nblk(nin), 256, 0, st>>>(iparent.as<int>(), nin, idepth.as<int>(),
                                                 even.as<int>());
        VLB_HIP(hipGetLastError());
    }
    size_t scan_bytes = 0;
    VLB_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, even.as<int>(), idx4.as<int>(), 0, (size_t)nin,
                                    rocprim::plus<int>(), st));
    DevBuf scan_tmp = dev_alloc(ctx, scan_bytes > 0 ? scan_bytes : 16);
    VLB_HIP(rocprim::exclusive_scan(scan_tmp.p, scan_bytes, even.as<int>(), idx4.as<int>(), 0, (size_t)nin,
                                    rocprim::plus<int>(), st));
    int last[2] = {0, 0};
    VLB_HIP(hipMemcpyAsync(&last[0], idx4.as<int>() + (nin - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    VLB_HIP(hipMemcpyAsync(&last[1], even.as<int>() + (nin - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    VLB_HIP(hipStreamSynchronize(st));
    const int n4 = last[0] + last[1];
    s.nodes4 = dev_alloc(ctx, (size_t)n4 * sizeof(BNode4));
    s.qnodes4 = dev_alloc(ctx, (size_t)n4 * sizeof(QNode4));
    k_collapse<<<nblk(nin), 256, 0, st>>>(s.nodes.as<BNode>(), idepth.as<int>(), idx4.as<int>(), nin,
                                          s.nodes4.as<BNode4>(), s.qnodes4.as<QNode4>());
    VLB_HIP(hipGetLastError());
    VLB_HIP(hipEventRecord(e1, st));
    float hb[16];
    unsigned int depth = 0;
    VLB_HIP(hipMemcpyAsync(hb, bnd.p, 13 * sizeof(float), hipMemcpyDeviceToHost, st));
    VLB_HIP(hipMemcpyAsync(&depth, dep.p, sizeof depth, hipMemcpyDeviceToHost, st));
    VLB_HIP(hipStreamSynchronize(st));
    float ms = 0.0f;
    VLB_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s.build_ms = ms;
    s.nnodes = nin;
    s.nnodes4 = n4;
    s.depth = depth;
    for (int a = 0; a < 3; ++a) {
        s.bmin[a] = hb[6 + a];
        s.bmax[a] = hb[9 + a];
    }
    s.maxabs = hb[12];
}

}  // namespace vlb
