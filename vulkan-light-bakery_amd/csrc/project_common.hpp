// project_common.hpp -- the SH projection of one probe tile on MFMA, shared
// by the single-bake kernel (bake.hip) and the multi-job kernels (multi.hip).
//
// C[probe][coef] += sum_ray rad[probe][ray] * Yw[ray][coef].  A workgroup (4
// waves) owns 16 probes x one slice of nch 32-ray chunks; the waves take the
// chunks round-robin.  Lane l = 16*g + pl holds, per channel, the radiance of
// probe pl for rays 8g..8g+7 of the chunk (A operand) and the basis values
// Y_pl of the same rays (B operand).  FP32 uses 8 x v_mfma_f32_16x16x4_f32
// per chunk and channel; the 16-bit forms 1 or 3 x
// v_mfma_f32_16x16x32_{bf16,f16}.  Chunk results are Kahan-summed in FP32
// across chunks and waves into one (sum, compensation) pair per (probe,
// channel, coef), combined across waves in a fixed order, so the result does
// not depend on scheduling.  HBM-bound: 12 B of radiance per ray.
#pragma once

#include "vlb_internal.hpp"

namespace vlb {

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

// Kahan step on element i of (s, comp)
__device__ __forceinline__ void kahan_add(v4f& s, v4f& comp, int i, float x)
{
    const float y = x - comp[i];
    const float t = s[i] + y;
    comp[i] = (t - s[i]) - y;
    s[i] = t;
}

// One probe tile: radiance chunks k = 0..nch-1 at rad0 + 1536*k (MFMA A
// order, written by the trace kernels), basis octets of chunk ch0 + k of the
// ray space.  Writes prow[((4g + i)*3 + c)*16 + pl] = (sum, compensation)
// for probe 4g + i of the tile.  `red` is 3*24*64 floats of LDS.  Called by
// all 256 threads of the workgroup; ends with a barrier.
template <int PREC>
__device__ __forceinline__ void project_tile(const float* rad0, int nch, int ch0,
                                             const void* ytab, size_t lo_off, float* red,
                                             float2* prow)
{
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = lane >> 4;
    const int pl = lane & 15;
    v4f S[3], Cp[3];
    for (int c = 0; c < 3; ++c) {
        S[c] = v4f{0.f, 0.f, 0.f, 0.f};
        Cp[c] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    for (int k = wave; k < nch; k += 4) {
        // streamed once: non-temporal 16-byte loads
        const v4f* src = reinterpret_cast<const v4f*>(rad0 + (size_t)k * 1536) + lane * 2;
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
        const size_t vidx = ((size_t)(ch0 + k) * 4 + g) * 16 + pl;  // ray octet, coef pl
        v4f C[3];
        for (int c = 0; c < 3; ++c) C[c] = v4f{0.f, 0.f, 0.f, 0.f};
        if constexpr (PREC == VLB_PREC_FP32) {
            const float4* yt = static_cast<const float4*>(ytab) + vidx * 2;
            const float4 y0 = yt[0], y1 = yt[1];
            const float y[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                C[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(lx[e], y[e], C[0], 0, 0, 0);
                C[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ly[e], y[e], C[1], 0, 0, 0);
                C[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(lz[e], y[e], C[2], 0, 0, 0);
            }
        } else {
            project16<PREC>(lx, ly, lz, ytab, lo_off, vidx, C);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) kahan_add(S[c], Cp[c], i, C[c][i]);
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
                    kahan_add(S[c], Cp[c], i, red[(w * 24 + c * 4 + i) * 64 + lane]);
                    kahan_add(S[c], Cp[c], i, -red[(w * 24 + 12 + c * 4 + i) * 64 + lane]);
                }
        // D layout: lane holds rows 4g..4g+3 (probes) of column pl (coef)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                prow[((4 * g + i) * 3 + c) * 16 + pl] = make_float2(S[c][i], Cp[c][i]);
    }
    __syncthreads();
}

// Lane -> ray maps of the trace kernels.  Trace tiles are 16 probes x 32
// rays (512 rays); item `it` of the ray space decodes to (tile, l, e), the
// radiance slot of the MFMA operand order above, which is the same for every
// map: thread t = 8l + e of a tile traces, for probe (l & 15) of the tile,
// ray 8*(l >> 4) + e of the tile's chunk.
//   MAP 0: a wave = 8 probes x 8 directions
//   MAP 1: a wave = 16 probes x 4 directions
//   MAP 2: a wave = 2 probes x 32 directions
//   MAP 3: a wave = 1 probe x 64 directions (pairs of tiles; needs an even
//          number of chunks per probe group)
__device__ __forceinline__ void tile_map(int map, size_t it, int& tile, int& l, int& e)
{
    if (map == 3) {
        const int i = (int)(it & 1023);
        const int p = i >> 6, rr = i & 63, rc = rr & 31;
        tile = (int)(it >> 10) * 2 + (rr >> 5);
        l = p + 16 * (rc >> 3);
        e = rc & 7;
        return;
    }
    tile = (int)(it >> 9);
    const int t = (int)(it & 511);
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

}  // namespace vlb
