// bvh.hip -- GPU BVH build for the probe baker (replaces the Vulkan BLAS/TLAS
// build of src/scene_manager.cpp:339-443; the reference hands triangles to
// the driver with ePreferFastTrace and culling disabled).
//
// LBVH (Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", HPG 2012): 30-bit Morton codes of triangle centroids,
// radix sort, one thread per inner node finds its key range and split, then
// every inner node's child boxes come from a segment-tree pyramid over the
// sorted leaf boxes -- no atomics or inter-workgroup fences, so the build is
// deterministic.  Node boxes are the exact min/max of their triangles'
// vertices (conservative for the watertight test).
#include <hipcub/hipcub.hpp>

#include "vlb_internal.hpp"

namespace vlb {
namespace {

struct Box6 {
    float lo[3], hi[3];
};

__device__ __forceinline__ void box_union(Box6& a, const float4& lo, const float4& hi)
{
    a.lo[0] = fminf(a.lo[0], lo.x);
    a.lo[1] = fminf(a.lo[1], lo.y);
    a.lo[2] = fminf(a.lo[2], lo.z);
    a.hi[0] = fmaxf(a.hi[0], hi.x);
    a.hi[1] = fmaxf(a.hi[1], hi.y);
    a.hi[2] = fmaxf(a.hi[2], hi.z);
}

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

__global__ void k_morton(const float4* __restrict__ cen, int n, const float* __restrict__ bounds,
                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = cen[i];
    float q[3];
    const float cc[3] = {c.x, c.y, c.z};
    for (int a = 0; a < 3; ++a) {
        float ext = bounds[3 + a] - bounds[a];
        float x = ext > 0.0f ? (cc[a] - bounds[a]) / ext * 1024.0f : 0.0f;
        q[a] = fminf(fmaxf(x, 0.0f), 1023.0f);
    }
    keys[i] = (expand_bits((uint32_t)q[0]) << 2) | (expand_bits((uint32_t)q[1]) << 1) |
              expand_bits((uint32_t)q[2]);
    vals[i] = (uint32_t)i;
}

__device__ __forceinline__ int delta(const uint32_t* __restrict__ keys, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// One thread per inner node i in [0, n-2] (Karras 2012, Figure 4).
// Writes children into nodes[i].d, the covered leaf range, and the parent
// link (2*parent + slot) of both children.
__global__ void k_karras(const uint32_t* __restrict__ keys, int n, BNode* __restrict__ nodes,
                         int2* __restrict__ range, int* __restrict__ iparent,
                         int* __restrict__ lparent)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int first = min(i, j), last = max(i, j);
    int c0, c1;
    if (first == gamma) {
        c0 = ~gamma;
        lparent[gamma] = 2 * i + 0;
    } else {
        c0 = gamma;
        iparent[gamma] = 2 * i + 0;
    }
    if (last == gamma + 1) {
        c1 = ~(gamma + 1);
        lparent[gamma + 1] = 2 * i + 1;
    } else {
        c1 = gamma + 1;
        iparent[gamma + 1] = 2 * i + 1;
    }
    nodes[i].d = make_int4(c0, c1, 0, 0);
    range[i] = make_int2(first, last);
    if (i == 0) iparent[0] = -1;
}

// Leaf-order triangles and level 0 of the box pyramid.
__global__ void k_leaves(const uint32_t* __restrict__ order, int n,
                         const float4* __restrict__ tri_raw, const float4* __restrict__ lo,
                         const float4* __restrict__ hi, float4* __restrict__ tris,
                         float4* __restrict__ plo, float4* __restrict__ phi)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t p = order[j];
    tris[3 * j + 0] = tri_raw[3 * p + 0];
    tris[3 * j + 1] = tri_raw[3 * p + 1];
    tris[3 * j + 2] = tri_raw[3 * p + 2];
    plo[j] = lo[p];
    phi[j] = hi[p];
}

// Pyramid level k from level k-1 (m_prev entries).
__global__ void k_pyramid(const float4* __restrict__ lo_prev, const float4* __restrict__ hi_prev,
                          int m_prev, float4* __restrict__ lo, float4* __restrict__ hi)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int m = (m_prev + 1) >> 1;
    if (i >= m) return;
    float4 l = lo_prev[2 * i], h = hi_prev[2 * i];
    if (2 * i + 1 < m_prev) {
        float4 l2 = lo_prev[2 * i + 1], h2 = hi_prev[2 * i + 1];
        l = make_float4(fminf(l.x, l2.x), fminf(l.y, l2.y), fminf(l.z, l2.z), 0.0f);
        h = make_float4(fmaxf(h.x, h2.x), fmaxf(h.y, h2.y), fmaxf(h.z, h2.z), 0.0f);
    }
    lo[i] = l;
    hi[i] = h;
}

struct Pyramid {
    const float4* lo[32];
    const float4* hi[32];
    int levels;
};

__device__ Box6 range_box(const Pyramid& py, int a, int b)
{
    Box6 r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = __builtin_huge_valf();
        r.hi[k] = -__builtin_huge_valf();
    }
    // Segment-tree walk: level k entry i covers leaves [i*2^k, (i+1)*2^k).
    // An odd left end or an even right end is taken alone; what remains is
    // [even, odd], whose parents are [a/2, b/2] one level up.
    int lvl = 0;
    while (a <= b && lvl < py.levels) {
        if (a & 1) {
            box_union(r, py.lo[lvl][a], py.hi[lvl][a]);
            ++a;
        }
        if (a <= b && !(b & 1)) {
            box_union(r, py.lo[lvl][b], py.hi[lvl][b]);
            --b;
        }
        if (a > b) break;
        a >>= 1;
        b >>= 1;
        ++lvl;
    }
    return r;
}

__global__ void k_node_boxes(BNode* __restrict__ nodes, int nin, const int2* __restrict__ range,
                             Pyramid py)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nin) return;
    int4 c = nodes[i].d;
    Box6 b[2];
    const int ch[2] = {c.x, c.y};
    for (int k = 0; k < 2; ++k) {
        if (ch[k] < 0) {
            int leaf = ~ch[k];
            float4 l = py.lo[0][leaf], h = py.hi[0][leaf];
            b[k].lo[0] = l.x;
            b[k].lo[1] = l.y;
            b[k].lo[2] = l.z;
            b[k].hi[0] = h.x;
            b[k].hi[1] = h.y;
            b[k].hi[2] = h.z;
        } else {
            int2 r = range[ch[k]];
            b[k] = range_box(py, r.x, r.y);
        }
    }
    BNode n;
    n.a = make_float4(b[0].lo[0], b[0].lo[1], b[0].lo[2], b[0].hi[0]);
    n.b = make_float4(b[0].hi[1], b[0].hi[2], b[1].lo[0], b[1].lo[1]);
    n.c = make_float4(b[1].lo[2], b[1].hi[0], b[1].hi[1], b[1].hi[2]);
    n.d = c;
    nodes[i] = n;
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

// BVH2 -> BVH4: every even-depth BVH2 node becomes a BVH4 node whose
// children are its grandchildren through inner children, or the child
// itself when that is a leaf (2 to 4 entries).  Boxes are copied, so the
// BVH4 is exactly as conservative as the BVH2.
__global__ void k_collapse(const BNode* __restrict__ n2, const int* __restrict__ depth,
                           const int* __restrict__ idx4, int nin, BNode4* __restrict__ n4)
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
    for (; m < 4; ++m) {
        for (int t = 0; t < 6; ++t) bx[m][t] = 0.0f;
        ch[m] = kEmpty;
    }
    BNode4 o;
    o.lox = make_float4(bx[0][0], bx[1][0], bx[2][0], bx[3][0]);
    o.loy = make_float4(bx[0][1], bx[1][1], bx[2][1], bx[3][1]);
    o.loz = make_float4(bx[0][2], bx[1][2], bx[2][2], bx[3][2]);
    o.hix = make_float4(bx[0][3], bx[1][3], bx[2][3], bx[3][3]);
    o.hiy = make_float4(bx[0][4], bx[1][4], bx[2][4], bx[3][4]);
    o.hiz = make_float4(bx[0][5], bx[1][5], bx[2][5], bx[3][5]);
    o.child = make_int4(ch[0], ch[1], ch[2], ch[3]);
    o.pad = make_int4(0, 0, 0, 0);
    n4[idx4[i]] = o;
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
    VLB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.as<uint32_t>(),
                                               keys2.as<uint32_t>(), vals.as<uint32_t>(),
                                               vals2.as<uint32_t>(), n, 0, 30, st));
    DevBuf tmp = dev_alloc(ctx, tmp_bytes > 0 ? tmp_bytes : 16);
    VLB_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys.as<uint32_t>(),
                                               keys2.as<uint32_t>(), vals.as<uint32_t>(),
                                               vals2.as<uint32_t>(), n, 0, 30, st));
    // hierarchy
    const int nin = n > 1 ? n - 1 : 1;
    s.nodes = dev_alloc(ctx, (size_t)nin * sizeof(BNode));
    s.tris = dev_alloc(ctx, (size_t)n * 3 * sizeof(float4));
    DevBuf range = dev_alloc(ctx, (size_t)nin * sizeof(int2));
    DevBuf iparent = dev_alloc(ctx, (size_t)nin * sizeof(int));
    DevBuf lparent = dev_alloc(ctx, (size_t)n * sizeof(int));
    // pyramid storage
    std::vector<DevBuf> plo, phi;
    std::vector<int> sizes;
    for (int m = n;; m = (m + 1) >> 1) {
        sizes.push_back(m);
        plo.push_back(dev_alloc(ctx, (size_t)m * sizeof(float4)));
        phi.push_back(dev_alloc(ctx, (size_t)m * sizeof(float4)));
        if (m == 1) break;
    }
    VLB_REQUIRE(sizes.size() <= 32, VLB_ERR_INVALID, "scene too large for the box pyramid");
    k_leaves<<<nblk(n), 256, 0, st>>>(vals2.as<uint32_t>(), n, tri_raw.as<float4>(),
                                      lo.as<float4>(), hi.as<float4>(), s.tris.as<float4>(),
                                      plo[0].as<float4>(), phi[0].as<float4>());
    for (size_t k = 1; k < sizes.size(); ++k)
        k_pyramid<<<nblk(sizes[k]), 256, 0, st>>>(plo[k - 1].as<float4>(), phi[k - 1].as<float4>(),
                                                  sizes[k - 1], plo[k].as<float4>(),
                                                  phi[k].as<float4>());
    VLB_HIP(hipGetLastError());
    Pyramid py{};
    py.levels = (int)sizes.size();
    for (size_t k = 0; k < sizes.size(); ++k) {
        py.lo[k] = plo[k].as<float4>();
        py.hi[k] = phi[k].as<float4>();
    }
    DevBuf dep = dev_alloc(ctx, sizeof(unsigned int));
    VLB_HIP(hipMemsetAsync(dep.p, 0, sizeof(unsigned int), st));
    if (n == 1) {
        // single triangle: one inner node whose two children are the same
        // leaf; testing it twice cannot change a result.
        BNode h;
        h.d = make_int4(~0, ~0, 0, 0);
        VLB_HIP(hipMemcpyAsync(s.nodes.p, &h, sizeof h, hipMemcpyHostToDevice, st));
        int2 r0 = make_int2(0, 0);
        VLB_HIP(hipMemcpyAsync(range.p, &r0, sizeof r0, hipMemcpyHostToDevice, st));
        k_node_boxes<<<1, 256, 0, st>>>(s.nodes.as<BNode>(), 1, range.as<int2>(), py);
        unsigned int one = 1;
        VLB_HIP(hipMemcpyAsync(dep.p, &one, sizeof one, hipMemcpyHostToDevice, st));
    } else {
        k_karras<<<nblk(nin), 256, 0, st>>>(keys2.as<uint32_t>(), n, s.nodes.as<BNode>(),
                                            range.as<int2>(), iparent.as<int>(),
                                            lparent.as<int>());
        k_node_boxes<<<nblk(nin), 256, 0, st>>>(s.nodes.as<BNode>(), nin, range.as<int2>(), py);
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
    VLB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, even.as<int>(), idx4.as<int>(), nin, st));
    DevBuf scan_tmp = dev_alloc(ctx, scan_bytes > 0 ? scan_bytes : 16);
    VLB_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_bytes, even.as<int>(), idx4.as<int>(),
                                             nin, st));
    int last[2] = {0, 0};
    VLB_HIP(hipMemcpyAsync(&last[0], idx4.as<int>() + (nin - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    VLB_HIP(hipMemcpyAsync(&last[1], even.as<int>() + (nin - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    VLB_HIP(hipStreamSynchronize(st));
    const int n4 = last[0] + last[1];
    s.nodes4 = dev_alloc(ctx, (size_t)n4 * sizeof(BNode4));
    k_collapse<<<nblk(nin), 256, 0, st>>>(s.nodes.as<BNode>(), idepth.as<int>(), idx4.as<int>(), nin,
                                          s.nodes4.as<BNode4>());
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
