// sh_image.hip -- SH projection of skies and images.
//
//  k_skybox    : this library's form of skybox_sh.comp (shaders/skybox_sh.comp:
//                23-42, host src/skybox_manager.cpp:107-130): batched
//                6 x N x N RGBA32F cubemaps -> K coefs x RGB, exact texel
//                solid angles.  HBM-bound: 16 B per texel streamed once.
//  k_equirect  : shaders/sh.comp:23-42 and skybox_sh.comp:23-42 over an RGBA8
//                equirect image (intended sum: no race, fresh zero).
//  k_sh_recon  : shaders/sh_sum.comp:35-52 (1250 * sum_k c_k Y_k -> RGBA8).
//
// Reductions: per-thread FP32 accumulators, a wave butterfly (xor shuffles),
// an LDS combine across the 4 waves, one partial per workgroup, then an FP64
// merge kernel that adds the partials in a fixed order.  No float atomics,
// so results are deterministic (sh.comp's shared "+=" is a data race).
#include "vlb_internal.hpp"
#include "sh_eval.hpp"

#include <mutex>

namespace vlb {
namespace {

// Sum NV values over the workgroup; thread 0's acc holds the result.
template <int NV>
__device__ __forceinline__ void block_reduce(float* acc, float* lds)
{
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        float v = acc[k];
        v += __shfl_xor(v, 32);
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 8);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 1);
        acc[k] = v;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NV; ++k) lds[wave * NV + k] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) acc[k] = ((lds[k] + lds[NV + k]) + lds[2 * NV + k]) + lds[3 * NV + k];
}

// Vulkan cube-map face orientation (faces +X,-X,+Y,-Y,+Z,-Z; s = u, t = v).
__device__ __forceinline__ float3 cube_dir(int f, float u, float v)
{
    float3 d;
    switch (f) {
    case 0: d = make_float3(1.0f, -v, -u); break;
    case 1: d = make_float3(-1.0f, -v, u); break;
    case 2: d = make_float3(u, 1.0f, v); break;
    case 3: d = make_float3(u, -1.0f, -v); break;
    case 4: d = make_float3(u, -v, 1.0f); break;
    default: d = make_float3(-u, -v, -1.0f); break;
    }
    // hardware reciprocal square root (~1 ulp): the projection is checked to
    // 1e-5 relative against the FP64 oracle, and this is ~20 instructions
    // cheaper per texel than a correctly rounded divide of a sqrt
    const float inv = rsqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    return make_float3(d.x * inv, d.y * inv, d.z * inv);
}

// Direction (world frame, unit) and solid angle of texel t of a cube map.
__device__ __forceinline__ float4 skybox_dir_weight(int t, int N, int logN, int NN, float s2, float s1,
                                                    const float* __restrict__ wtab)
{
    int f, rem, j, i;
    if (logN >= 0) {
        f = t >> (2 * logN);
        rem = t & (NN - 1);
        j = rem >> logN;
        i = rem & (N - 1);
    } else {
        f = t / NN;
        rem = t - f * NN;
        j = rem / N;
        i = rem - j * N;
    }
    const float u = fmaf((float)i, s2, s1);
    const float v = fmaf((float)j, s2, s1);
    const float3 d = cube_dir(f, u, v);
    return make_float4(d.x, d.y, d.z, wtab[rem]);
}

// Table of skybox_dir_weight over the 6 N^2 texels of one cube map.
__global__ __launch_bounds__(256) void k_skybox_table(int N, int logN, const float* __restrict__ wtab,
                                                      float4* __restrict__ tab)
{
    const int NN = N * N;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 6 * NN) return;
    tab[t] = skybox_dir_weight(t, N, logN, NN, 2.0f / (float)N, 1.0f / (float)N - 1.0f, wtab);
}

// One texel: SH basis at its direction times its solid angle, into acc.
// TAB: direction and weight read from the table (the same values), else
// computed from the texel index.
template <int K, bool TAB>
__device__ __forceinline__ void skybox_texel(float* acc, const float4& px, int t, int N, int logN,
                                             int NN, float s2, float s1,
                                             const float* __restrict__ wtab,
                                             const float4* __restrict__ tab)
{
    const float4 dw = TAB ? tab[t] : skybox_dir_weight(t, N, logN, NN, s2, s1, wtab);
    const float3 d = make_float3(dw.x, dw.y, dw.z);
    const float w = dw.w;
    const float r = px.x * w, g = px.y * w, b = px.z * w;
    float Y[16];
    sh_exact<K>(d.x, d.z, d.y, Y);  // SH frame: world (x, z, y)
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc[3 * k + 0] = fmaf(Y[k], r, acc[3 * k + 0]);
        acc[3 * k + 1] = fmaf(Y[k], g, acc[3 * k + 1]);
        acc[3 * k + 2] = fmaf(Y[k], b, acc[3 * k + 2]);
    }
}

// Four texels per thread per round, their loads issued before any math so
// that each lane keeps 64 B in flight.
template <int K, bool TAB>
__global__ __launch_bounds__(256) void k_skybox(const float4* __restrict__ cubes, int N, int logN,
                                                int per_block, int blocks_per_cube,
                                                const float* __restrict__ wtab,
                                                const float4* __restrict__ tab,
                                                float* __restrict__ partial)
{
    __shared__ float lds[4 * K * 3];
    const int cube = blockIdx.x / blocks_per_cube;
    const int blk = blockIdx.x - cube * blocks_per_cube;
    const int NN = N * N;
    const int T = 6 * NN;
    const int start = blk * per_block;
    const int end = min(T, start + per_block);
    const float4* src = cubes + (size_t)cube * T;
    const float s2 = 2.0f / (float)N, s1 = 1.0f / (float)N - 1.0f;
    float acc[K * 3];
#pragma unroll
    for (int k = 0; k < K * 3; ++k) acc[k] = 0.0f;
    typedef float nf4 __attribute__((ext_vector_type(4)));
    int t = start + (int)threadIdx.x;
    for (; t + 768 < end; t += 1024) {
        nf4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            q[u] = __builtin_nontemporal_load(reinterpret_cast<const nf4*>(src + t + 256 * u));
#pragma unroll
        for (int u = 0; u < 4; ++u)
            skybox_texel<K, TAB>(acc, make_float4(q[u][0], q[u][1], q[u][2], q[u][3]), t + 256 * u, N,
                                 logN, NN, s2, s1, wtab, tab);
    }
    for (; t < end; t += 256) {
        const nf4 q = __builtin_nontemporal_load(reinterpret_cast<const nf4*>(src + t));
        skybox_texel<K, TAB>(acc, make_float4(q[0], q[1], q[2], q[3]), t, N, logN, NN, s2, s1, wtab, tab);
    }
    block_reduce<K * 3>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < K * 3; ++k) partial[(size_t)blockIdx.x * K * 3 + k] = acc[k];
}

// FP64 merge of per-workgroup partials: out[g][v] = sum_b partial[g][b][v],
// one thread per output, partials added in order.
__global__ __launch_bounds__(256) void k_merge_partials(const float* __restrict__ partial,
                                                        int groups, int per_group, int nv,
                                                        float* __restrict__ out,
                                                        double* __restrict__ out64)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= groups * nv) return;
    const int gi = idx / nv, v = idx - gi * nv;
    double s = 0.0;
    for (int b = 0; b < per_group; ++b) s += (double)partial[((size_t)gi * per_group + b) * nv + v];
    if (out) out[idx] = (float)s;
    if (out64) out64[idx] = s;
}

// The same merge with one wave per output, for many partials per output
// (one or a few images split over hundreds of workgroups): lane l adds
// partials l, l + 64, ... in order, then a fixed butterfly adds the lanes, so
// the sum is the same on every run.
__global__ __launch_bounds__(256) void k_merge_partials_wide(const float* __restrict__ partial,
                                                             int groups, int per_group, int nv,
                                                             float* __restrict__ out,
                                                             double* __restrict__ out64)
{
    const int idx = (int)((blockIdx.x * 256u + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (idx >= groups * nv) return;  // wave-uniform
    const int gi = idx / nv, v = idx - gi * nv;
    double s = 0.0;
    for (int b = lane; b < per_group; b += 64) s += (double)partial[((size_t)gi * per_group + b) * nv + v];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) {
        if (out) out[idx] = (float)s;
        if (out64) out64[idx] = s;
    }
}

void launch_merge_partials(const float* partial, int groups, int per_group, int nv, float* out,
                           double* out64, hipStream_t st)
{
    const int n = groups * nv;
    if (per_group > 64)
        k_merge_partials_wide<<<(n + 3) / 4, 256, 0, st>>>(partial, groups, per_group, nv, out, out64);
    else
        k_merge_partials<<<(n + 255) / 256, 256, 0, st>>>(partial, groups, per_group, nv, out, out64);
    VLB_HIP(hipGetLastError());
}

// sh.comp (variant 0) / skybox_sh.comp (variant 1) over an RGBA8 image.
__global__ __launch_bounds__(256) void k_equirect(const uint8_t* __restrict__ img, int W, int H,
                                                  int variant, int per_block,
                                                  float* __restrict__ partial)
{
    __shared__ float lds[4 * 48];
    const float PI = 3.1415926538f;  // sh_common.h:1
    const int n = W * H;
    const int start = blockIdx.x * per_block;
    const int end = min(n, start + per_block);
    const float area = (2.0f * PI / (float)W) * (PI / (float)H);
    float acc[48];
#pragma unroll
    for (int k = 0; k < 48; ++k) acc[k] = 0.0f;
    for (int p = start + (int)threadIdx.x; p < end; p += 256) {
        const int y = p / W, x = p - y * W;
        float phi = 2.f * PI * ((float)x + 0.5f) / (float)W;
        if (variant == VLB_EQUIRECT_SKYBOX_COMP) phi -= PI / 2.0f;
        const float theta = PI * ((float)y + 0.5f) / (float)H;
        float st, ct, sp, cp;
        sincosf(theta, &st, &ct);
        sincosf(phi, &sp, &cp);
        float dx = st * cp, dy = st * sp, dz = ct;
        const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
        dx *= inv;
        dy *= inv;
        dz *= inv;
        const uchar4 px = reinterpret_cast<const uchar4*>(img)[p];
        const float w = area * st;
        const float r = (float)px.x / 255.0f * w, g = (float)px.y / 255.0f * w,
                    b = (float)px.z / 255.0f * w;
        float Y[16];
        if (variant == VLB_EQUIRECT_SKYBOX_COMP) sh_literal16(dx, dz, dy, Y);  // SH(dir.xzy)
        else sh_literal16(dx, dy, dz, Y);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            acc[3 * k + 0] = fmaf(Y[k], r, acc[3 * k + 0]);
            acc[3 * k + 1] = fmaf(Y[k], g, acc[3 * k + 1]);
            acc[3 * k + 2] = fmaf(Y[k], b, acc[3 * k + 2]);
        }
    }
    block_reduce<48>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 48; ++k) partial[(size_t)blockIdx.x * 48 + k] = acc[k];
}

// sh_sum.comp:35-52 in float, operation order of the GLSL.
__global__ __launch_bounds__(256) void k_sh_recon(const float* __restrict__ coeffs, int W, int H,
                                                  uint8_t* __restrict__ out)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= W * H) return;
    const float PI = 3.1415926538f;
    const int y = p / W, x = p - y * W;
    const float phi = 2.f * PI * ((float)x + 0.5f) / (float)W;
    const float theta = PI * ((float)y + 0.5f) / (float)H;
    const float r = sinf(theta);
    const float a = r * cosf(phi), b = r * sinf(phi), c = cosf(theta);
    const float l = sqrtf(a * a + b * b + c * c);
    float v[16];
    sh_literal16(a / l, b / l, c / l, v);
    float col[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 16; ++k)
        for (int ch = 0; ch < 3; ++ch) col[ch] = col[ch] + coeffs[3 * k + ch] * v[k];
    uchar4 o;
    float s[3];
    for (int ch = 0; ch < 3; ++ch) {
        float t = 1250.0f * col[ch];
        float cl = (t > 0.0f) ? (t < 1.0f ? t : 1.0f) : 0.0f;
        s[ch] = rintf(cl * 255.0f);
    }
    o.x = (uint8_t)s[0];
    o.y = (uint8_t)s[1];
    o.z = (uint8_t)s[2];
    o.w = 255;
    reinterpret_cast<uchar4*>(out)[p] = o;
}

// Exact solid angle of each texel of a cube-map face, the host table of
// solid_angle_table (api.cpp) computed on the device in double: the
// allocation-free entry point builds it in its scratch instead of uploading.
__device__ __forceinline__ double corner_area(double x, double y)
{
    return atan2(x * y, sqrt(x * x + y * y + 1.0));
}

__global__ __launch_bounds__(256) void k_solid_angle(int N, float* __restrict__ w)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= N * N) return;
    const int j = t / N, i = t - j * N;
    const double x0 = 2.0 * i / N - 1.0, x1 = 2.0 * (i + 1) / N - 1.0;
    const double y0 = 2.0 * j / N - 1.0, y1 = 2.0 * (j + 1) / N - 1.0;
    w[t] = (float)(corner_area(x0, y0) - corner_area(x0, y1) - corner_area(x1, y0) + corner_area(x1, y1));
}

// Plain streaming copy, 16 B per lane per iteration (bandwidth reference).
__global__ __launch_bounds__(256) void k_copy(float4* __restrict__ dst,
                                              const float4* __restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

// Read-only streaming pass (bandwidth reference for read-bound kernels such
// as k_skybox): each workgroup streams its own contiguous range with four
// 16-byte nontemporal loads per lane per round, issued before they are
// used, as k_skybox does.  The sum is handed to an empty asm statement,
// which keeps the loads from being removed; nothing is stored.
__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ src, size_t n, size_t per_block)
{
    typedef float nf4 __attribute__((ext_vector_type(4)));
    const size_t start = (size_t)blockIdx.x * per_block;
    const size_t end = min(n, start + per_block);
    nf4 a = {0.0f, 0.0f, 0.0f, 0.0f};
    size_t t = start + threadIdx.x;
    for (; t + 768 < end; t += 1024) {
        nf4 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = __builtin_nontemporal_load(reinterpret_cast<const nf4*>(src + t + 256 * u));
#pragma unroll
        for (int u = 0; u < 4; ++u) a += q[u];
    }
    for (; t < end; t += 256) a += __builtin_nontemporal_load(reinterpret_cast<const nf4*>(src + t));
    const float sum = (a.x + a.y) + (a.z + a.w);
    __asm__ __volatile__("" ::"v"(sum));
}

// Scratch of the allocation-free skybox projection (vlb_skybox_sh_noalloc):
// [solid-angle table | direction table | partial sums] of the current call.
// The name is a host-side stand-in in host code, never a device address:
// skybox_scratch_init resolves this device's address once, and launches take
// that pointer.  Every call on a GPU shares the array (calls that overlap on
// one GPU race) and overwrites all of it that it reads.
static __device__ __align__(256) unsigned char scratch[kSkyScratchBytes];

}  // namespace

// Device address of `scratch` on `device` (the current device), resolved
// with hipGetSymbolAddress the first time and kept per device.
void* skybox_scratch_init(int device)
{
    static std::mutex mu;
    static void* addr[64] = {};
    VLB_REQUIRE(device >= 0 && device < 64, VLB_ERR_INVALID, "device index beyond 63");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        void* p = nullptr;
        VLB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        VLB_REQUIRE(p != nullptr, VLB_ERR_DEVICE, "hipGetSymbolAddress returned no address");
        addr[device] = p;
    }
    return addr[device];
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0) return;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 256 * 64);
    k_copy<<<grid, 256, 0, st>>>(static_cast<float4*>(dst), static_cast<const float4*>(src), n);
    VLB_HIP(hipGetLastError());
}

void launch_read(const void* src, size_t bytes, int grid, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (n == 0) return;
    const size_t per_block = ((n + grid - 1) / grid + 255) / 256 * 256;
    const int g = (int)((n + per_block - 1) / per_block);
    k_read<<<g, 256, 0, st>>>(static_cast<const float4*>(src), n, per_block);
    VLB_HIP(hipGetLastError());
}

static int log2_or_minus1(int N)
{
    if ((N & (N - 1)) != 0) return -1;
    int logN = 0;
    while ((1 << logN) < N) ++logN;
    return logN;
}

void launch_skybox_table(int N, const float* wtab, float4* tab, hipStream_t st)
{
    const int T = 6 * N * N;
    k_skybox_table<<<(T + 255) / 256, 256, 0, st>>>(N, log2_or_minus1(N), wtab, tab);
    VLB_HIP(hipGetLastError());
}

void launch_skybox_sh(const float4* cubes, int N, int batch, int K, const float* wtab,
                      const float4* tab, float* partial, int blocks_per_cube, float* out,
                      hipStream_t st, hipEvent_t e0, hipEvent_t e1, hipEvent_t e2)
{
    const int T = 6 * N * N;
    const int per_block = (T + blocks_per_cube - 1) / blocks_per_cube;
    const int logN = log2_or_minus1(N);
    if (e0) VLB_HIP(hipEventRecord(e0, st));
    const int grid = batch * blocks_per_cube;
    if (K == 16 && tab)
        k_skybox<16, true><<<grid, 256, 0, st>>>(cubes, N, logN, per_block, blocks_per_cube, wtab, tab, partial);
    else if (K == 16)
        k_skybox<16, false><<<grid, 256, 0, st>>>(cubes, N, logN, per_block, blocks_per_cube, wtab, tab, partial);
    else if (tab)
        k_skybox<9, true><<<grid, 256, 0, st>>>(cubes, N, logN, per_block, blocks_per_cube, wtab, tab, partial);
    else
        k_skybox<9, false><<<grid, 256, 0, st>>>(cubes, N, logN, per_block, blocks_per_cube, wtab, tab, partial);
    VLB_HIP(hipGetLastError());
    if (e1) VLB_HIP(hipEventRecord(e1, st));
    const int nv = K * 3;
    launch_merge_partials(partial, batch, blocks_per_cube, nv, out, nullptr, st);
    if (e2) VLB_HIP(hipEventRecord(e2, st));
}

void launch_skybox_sh_noalloc(const float4* cubes, int N, int batch, int K, void* scr,
                              size_t tab_off, size_t part_off, bool use_tab, int blocks_per_cube,
                              float* out, hipStream_t st)
{
    uint8_t* base = static_cast<uint8_t*>(scr);
    float* wtab = reinterpret_cast<float*>(base);
    k_solid_angle<<<(N * N + 255) / 256, 256, 0, st>>>(N, wtab);
    VLB_HIP(hipGetLastError());
    float4* tab = use_tab ? reinterpret_cast<float4*>(base + tab_off) : nullptr;
    if (tab) launch_skybox_table(N, wtab, tab, st);
    launch_skybox_sh(cubes, N, batch, K, wtab, tab, reinterpret_cast<float*>(base + part_off),
                     blocks_per_cube, out, st, nullptr, nullptr, nullptr);
}

void launch_equirect_project(const uint8_t* img, int W, int H, int variant, double* out,
                             float* partial, int nblocks, hipStream_t st)
{
    const int n = W * H;
    const int per_block = (n + nblocks - 1) / nblocks;
    k_equirect<<<nblocks, 256, 0, st>>>(img, W, H, variant, per_block, partial);
    VLB_HIP(hipGetLastError());
    launch_merge_partials(partial, 1, nblocks, 48, nullptr, out, st);
}

void launch_sh_reconstruct(const float* coeffs, int W, int H, uint8_t* out, hipStream_t st)
{
    const int n = W * H;
    k_sh_recon<<<(n + 255) / 256, 256, 0, st>>>(coeffs, W, H, out);
    VLB_HIP(hipGetLastError());
}

}  // namespace vlb
