// tables.cpp -- host-side preparation shared by single bakes, batches and the
// request server: parameter validation, the direction set with its SH basis
// table (directions on a Hilbert curve, 16-bit split tables for the MFMA
// forms), probe sets in Morton order, and shading parameters.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "vlb_internal.hpp"

namespace vlb {

namespace {

uint32_t expand10(uint32_t v)
{
    v &= 1023u;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

uint64_t hilbert_index(uint32_t n, uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

uint16_t to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
uint16_t to_fp16(float f)
{
    _Float16 h = (_Float16)f;
    uint16_t u;
    memcpy(&u, &h, 2);
    return u;
}
double from_bf16(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
double from_fp16(uint16_t u)
{
    _Float16 h;
    memcpy(&h, &u, 2);
    return (double)(float)h;
}

}  // namespace

// Morton order of 3D points (stable on equal codes).
std::vector<uint32_t> morton_order(const float* xyz, size_t n)
{
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], xyz[3 * i + a]);
            hi[a] = std::max(hi[a], xyz[3 * i + a]);
        }
    std::vector<uint32_t> code(n), order(n);
    for (size_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) {
            float ext = hi[a] - lo[a];
            float x = ext > 0.0f ? (xyz[3 * i + a] - lo[a]) / ext * 1023.0f : 0.0f;
            q[a] = (uint32_t)std::min(1023.0f, std::max(0.0f, x));
        }
        code[i] = (expand10(q[0]) << 2) | (expand10(q[1]) << 1) | expand10(q[2]);
    }
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    return order;
}

// Directions ordered along a Hilbert curve over their octahedral map (pole
// = world +y): consecutive rays look at neighbouring patches of the sphere.
std::vector<uint32_t> direction_order(const float* xyz, size_t n)
{
    std::vector<uint64_t> code(n);
    std::vector<uint32_t> order(n);
    const uint32_t N = 1u << 16;
    for (size_t i = 0; i < n; ++i) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const float s = fabsf(x) + fabsf(y) + fabsf(z);
        float u = s > 0.0f ? x / s : 0.0f, v = s > 0.0f ? z / s : 0.0f;
        if (y < 0.0f) {
            const float uu = (1.0f - fabsf(v)) * (u >= 0.0f ? 1.0f : -1.0f);
            const float vv = (1.0f - fabsf(u)) * (v >= 0.0f ? 1.0f : -1.0f);
            u = uu;
            v = vv;
        }
        const uint32_t qx = (uint32_t)std::min<float>(N - 1, std::max(0.0f, (u * 0.5f + 0.5f) * N));
        const uint32_t qy = (uint32_t)std::min<float>(N - 1, std::max(0.0f, (v * 0.5f + 0.5f) * N));
        code[i] = hilbert_index(N, qx, qy);
    }
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    return order;
}

void validate(const vlb_bake_params& p)
{
    VLB_REQUIRE(p.num_probes >= 1 && p.probe_xyz, VLB_ERR_INVALID, "num_probes must be >= 1 with probe_xyz");
    VLB_REQUIRE(p.num_probes <= (1u << 26), VLB_ERR_INVALID, "num_probes too large");
    VLB_REQUIRE(p.sh_order == 2 || p.sh_order == 3, VLB_ERR_INVALID, "sh_order must be 2 or 3");
    VLB_REQUIRE(p.sh_constants == VLB_SH_LITERAL || p.sh_constants == VLB_SH_EXACT, VLB_ERR_INVALID,
                "bad sh_constants");
    VLB_REQUIRE(p.precision >= VLB_PREC_FP32 && p.precision <= VLB_PREC_FP16_FAST, VLB_ERR_INVALID,
                "bad precision");
    VLB_REQUIRE(p.shade_mode == VLB_SHADE_LAMBERT || p.shade_mode == VLB_SHADE_REFERENCE,
                VLB_ERR_INVALID, "bad shade_mode");
    VLB_REQUIRE(p.num_lights <= VLB_MAX_LIGHTS, VLB_ERR_INVALID, "too many lights");
    VLB_REQUIRE(p.num_lights == 0 || p.lights, VLB_ERR_INVALID, "lights pointer is NULL");
    for (uint32_t i = 0; i < p.num_lights; ++i)
        VLB_REQUIRE(p.lights[i].type == VLB_LIGHT_DIRECTIONAL || p.lights[i].type == VLB_LIGHT_POINT,
                    VLB_ERR_INVALID, "bad light type");
    if (p.dir_mode == VLB_DIRS_FIBONACCI) {
        VLB_REQUIRE(p.rays_per_probe >= 1 && p.rays_per_probe <= (1u << 26), VLB_ERR_INVALID,
                    "rays_per_probe out of range");
    } else if (p.dir_mode == VLB_DIRS_EQUIRECT) {
        VLB_REQUIRE(p.equirect_w >= 1 && p.equirect_h >= 1 &&
                        (uint64_t)p.equirect_w * p.equirect_h <= (1u << 26),
                    VLB_ERR_INVALID, "equirect size out of range");
    } else {
        throw Error{VLB_ERR_INVALID, "bad dir_mode"};
    }
    VLB_REQUIRE(p.t_min >= 0.0f && p.t_max > p.t_min, VLB_ERR_INVALID, "need 0 <= t_min < t_max");
    VLB_REQUIRE((uint64_t)p.num_probes * 48ull * 8ull < (1ull << 40), VLB_ERR_INVALID, "bake too large");
}


size_t scene_spill_bytes(const Scene& s, uint32_t stride)
{
    const uint32_t sb = stack_bound(s.depth);
    const uint32_t e = sb > (uint32_t)kLdsDepthMin ? sb - kLdsDepthMin : 0;
    return std::max<size_t>(16, (size_t)e * stride * 4);
}


Scene& find_scene(Context& ctx, uint32_t id)
{
    auto it = ctx.scenes.find(id);
    if (it == ctx.scenes.end()) throw Error{VLB_ERR_NOT_FOUND, "unknown scene id " + std::to_string(id)};
    return *it->second;
}

uint32_t rays_of(const vlb_bake_params& p)
{
    return p.dir_mode == VLB_DIRS_FIBONACCI ? p.rays_per_probe : p.equirect_w * p.equirect_h;
}

// Rays per projection slice: the caller's value rounded up to 64, else
// enough slices that a bake of npg probe groups gives the projection grid
// about two tiles per workgroup (an even chunk count per slice).
uint32_t slice_rays(const vlb_bake_params& p, uint32_t R, int npg, int pgrid)
{
    if (p.rays_per_slice) return (p.rays_per_slice + 63) & ~63u;
    uint64_t nsl = ((uint64_t)2 * pgrid + npg - 1) / npg;
    const uint64_t max_sl = (R + 31) / 32;
    nsl = std::max<uint64_t>(1, std::min(nsl, max_sl));
    return (uint32_t)(((R + nsl - 1) / nsl + 63) & ~63ull);
}

void build_ray_table(Context& ctx, const vlb_bake_params& p, int K, uint32_t RS, RayTable& t)
{
    uint32_t R;
    std::vector<float> dxyz, dw;
    if (p.dir_mode == VLB_DIRS_FIBONACCI) {
        R = p.rays_per_probe;
        dxyz.resize((size_t)R * 3);
        dw.assign(R, (float)(4.0 * M_PI / R));
        fibonacci_dirs(R, dxyz.data());
    } else {
        R = p.equirect_w * p.equirect_h;
        dxyz.resize((size_t)R * 3);
        dw.resize(R);
        equirect_dirs(p.equirect_w, p.equirect_h, dxyz.data(), dw.data());
    }
    VLB_REQUIRE(RS >= 64 && RS % 64 == 0, VLB_ERR_INVALID, "rays per slice must be a multiple of 64");
    const double wref = 4.0 * M_PI / R;
    // spatial order of directions: neighbouring lanes trace neighbouring rays
    const std::vector<uint32_t> dorder = direction_order(dxyz.data(), R);
    const int nslices = (int)((R + RS - 1) / RS);
    const int Rpad = nslices * (int)RS;

    std::vector<float4> dirs4(Rpad, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<int> dperm(Rpad, 0);
    // 16-bit forms: [hi table][lo table] (lo only for the split forms)
    const bool is16 = p.precision != VLB_PREC_FP32;
    const bool split = p.precision == VLB_PREC_BF16 || p.precision == VLB_PREC_FP16;
    const bool bf = p.precision == VLB_PREC_BF16 || p.precision == VLB_PREC_BF16_FAST;
    std::vector<float> yt32;
    std::vector<uint16_t> yt16;
    if (!is16) yt32.assign((size_t)Rpad * 16, 0.0f);
    else yt16.assign((size_t)Rpad * 16 * (split ? 2 : 1), 0);
    const bool exact = p.sh_constants == VLB_SH_EXACT;
    for (int r = 0; r < (int)R; ++r) {
        const uint32_t c = dorder[r];
        const float* d = &dxyz[3 * (size_t)c];
        dirs4[r] = make_float4(d[0], d[1], d[2], 1.0f);
        dperm[r] = (int)c;
        double sh[3] = {d[0], d[2], d[1]};  // SH frame: world (x, z, y)
        double Y[16];
        sh_eval_double(K, exact, sh, Y);
        const double wr = (double)dw[c] / wref;
        for (int k = 0; k < K; ++k) {
            const size_t ti = ((size_t)(r >> 3) * 16 + k) * 8 + (r & 7);
            const double v = Y[k] * wr;
            if (!is16) {
                yt32[ti] = (float)v;
            } else {
                const uint16_t hi = bf ? to_bf16((float)v) : to_fp16((float)v);
                yt16[ti] = hi;
                if (split) {
                    const double rest = v - (bf ? from_bf16(hi) : from_fp16(hi));
                    yt16[(size_t)Rpad * 16 + ti] = bf ? to_bf16((float)rest) : to_fp16((float)rest);
                }
            }
        }
    }
    t.R = R;
    t.RS = RS;
    t.Rpad = (uint32_t)Rpad;
    t.nslices = nslices;
    t.wref = wref;
    t.K = K;
    t.precision = p.precision;
    t.dirs = upload_vec(ctx, dirs4);
    t.dir_perm = upload_vec(ctx, dperm);
    if (!is16) t.ytab = upload_vec(ctx, yt32);
    else t.ytab = upload_vec(ctx, yt16);
}

void build_probe_set(Context& ctx, const float* xyz, uint32_t P, ProbeSet& s)
{
    const int npg = (int)((P + 15) / 16);
    const int Ppad = npg * 16;
    const std::vector<uint32_t> porder = morton_order(xyz, P);
    std::vector<float4> probes4(Ppad, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<int> pperm(Ppad, -1);
    for (uint32_t i = 0; i < P; ++i) {
        const float* q = &xyz[3 * (size_t)porder[i]];
        probes4[i] = make_float4(q[0], q[1], q[2], 1.0f);
        pperm[i] = (int)porder[i];
    }
    for (int i = (int)P; i < Ppad; ++i) probes4[i] = probes4[P - 1];
    s.P = (int)P;
    s.Ppad = Ppad;
    s.probes = upload_vec(ctx, probes4);
    s.probe_perm = upload_vec(ctx, pperm);
}

ShadeParams shade_params(const vlb_bake_params& p)
{
    ShadeParams sp;
    memset(&sp, 0, sizeof sp);
    sp.mode = p.shade_mode;
    sp.nlights = (int)p.num_lights;
    sp.shadow_bias = p.shadow_bias;
    sp.t_max = p.t_max;
    for (uint32_t i = 0; i < p.num_lights; ++i) sp.lights[i] = p.lights[i];
    sp.sky = p.sky;
    sp.ref = p.ref;
    return sp;
}

}  // namespace vlb
