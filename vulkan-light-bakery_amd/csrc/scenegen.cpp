// scenegen.cpp -- seeded procedural scenes of the benchmark configurations
// (the reference loads glTF assets through tinygltf, src/scene_manager.cpp:
// 32-67, 241-337; its asset pack is downloaded at configure time and is not
// available, so the workloads are synthetic and fully determined by a seed).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "../../include/vlb_bake.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next()
    {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uni() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    float uni(float a, float b) { return a + (b - a) * uni(); }
};

struct Builder {
    std::vector<float> xyz, nrm, alb;
    std::vector<uint32_t> idx;
    uint32_t vert(float x, float y, float z, float nx, float ny, float nz)
    {
        xyz.insert(xyz.end(), {x, y, z});
        nrm.insert(nrm.end(), {nx, ny, nz});
        return (uint32_t)(xyz.size() / 3 - 1);
    }
    void tri(uint32_t a, uint32_t b, uint32_t c, const float rgb[3])
    {
        idx.insert(idx.end(), {a, b, c});
        alb.insert(alb.end(), {rgb[0], rgb[1], rgb[2]});
    }
    // quad a-b-c-d counter-clockwise seen from the normal side
    void quad(const float* a, const float* b, const float* c, const float* d, const float n[3],
              const float rgb[3])
    {
        uint32_t i0 = vert(a[0], a[1], a[2], n[0], n[1], n[2]);
        uint32_t i1 = vert(b[0], b[1], b[2], n[0], n[1], n[2]);
        uint32_t i2 = vert(c[0], c[1], c[2], n[0], n[1], n[2]);
        uint32_t i3 = vert(d[0], d[1], d[2], n[0], n[1], n[2]);
        tri(i0, i1, i2, rgb);
        tri(i0, i2, i3, rgb);
    }
    // axis-aligned box with outward flat normals, 12 triangles
    void box(const float lo[3], const float hi[3], const float rgb[3])
    {
        const float x0 = lo[0], y0 = lo[1], z0 = lo[2], x1 = hi[0], y1 = hi[1], z1 = hi[2];
        const float p[8][3] = {{x0, y0, z0}, {x1, y0, z0}, {x1, y1, z0}, {x0, y1, z0},
                               {x0, y0, z1}, {x1, y0, z1}, {x1, y1, z1}, {x0, y1, z1}};
        const int f[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {3, 7, 6, 2},
                             {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
        const float n[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int k = 0; k < 6; ++k) quad(p[f[k][0]], p[f[k][1]], p[f[k][2]], p[f[k][3]], n[k], rgb);
    }
    // icosphere, 20 * 4^sub triangles, smooth normals
    void icosphere(const float c[3], float r, int sub, const float rgb[3])
    {
        const float t = (1.0f + sqrtf(5.0f)) / 2.0f;
        std::vector<float> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                                0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
        std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11,
                                   1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
                                   3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9,
                                   4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
        for (size_t i = 0; i < v.size(); i += 3) {
            float l = sqrtf(v[i] * v[i] + v[i + 1] * v[i + 1] + v[i + 2] * v[i + 2]);
            v[i] /= l;
            v[i + 1] /= l;
            v[i + 2] /= l;
        }
        for (int s = 0; s < sub; ++s) {
            std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
            auto midpoint = [&](uint32_t a, uint32_t b) {
                auto key = std::make_pair(std::min(a, b), std::max(a, b));
                auto it = mid.find(key);
                if (it != mid.end()) return it->second;
                float m[3] = {v[3 * a] + v[3 * b], v[3 * a + 1] + v[3 * b + 1], v[3 * a + 2] + v[3 * b + 2]};
                float l = sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
                v.insert(v.end(), {m[0] / l, m[1] / l, m[2] / l});
                uint32_t id = (uint32_t)(v.size() / 3 - 1);
                mid[key] = id;
                return id;
            };
            std::vector<uint32_t> nf;
            nf.reserve(f.size() * 4);
            for (size_t i = 0; i < f.size(); i += 3) {
                uint32_t a = f[i], b = f[i + 1], cc = f[i + 2];
                uint32_t ab = midpoint(a, b), bc = midpoint(b, cc), ca = midpoint(cc, a);
                nf.insert(nf.end(), {a, ab, ca, b, bc, ab, cc, ca, bc, ab, bc, ca});
            }
            f.swap(nf);
        }
        const uint32_t base = (uint32_t)(xyz.size() / 3);
        for (size_t i = 0; i < v.size(); i += 3)
            vert(c[0] + r * v[i], c[1] + r * v[i + 1], c[2] + r * v[i + 2], v[i], v[i + 1], v[i + 2]);
        for (size_t i = 0; i < f.size(); i += 3) tri(base + f[i], base + f[i + 1], base + f[i + 2], rgb);
    }
};

// value noise with smoothstep interpolation, lattice hash of (ix, iz, octave)
float lattice(int ix, int iz, int oct, uint64_t seed)
{
    uint64_t h = seed ^ ((uint64_t)(uint32_t)ix * 0x9E3779B1ull) ^ ((uint64_t)(uint32_t)iz * 0x85EBCA77ull << 1) ^
                 ((uint64_t)oct << 56);
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    h *= 0xC4CEB9FE1A85EC53ull;
    h ^= h >> 33;
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}
float vnoise(float x, float z, int oct, uint64_t seed)
{
    int ix = (int)floorf(x), iz = (int)floorf(z);
    float fx = x - ix, fz = z - iz;
    fx = fx * fx * (3.0f - 2.0f * fx);
    fz = fz * fz * (3.0f - 2.0f * fz);
    float a = lattice(ix, iz, oct, seed), b = lattice(ix + 1, iz, oct, seed);
    float c = lattice(ix, iz + 1, oct, seed), d = lattice(ix + 1, iz + 1, oct, seed);
    return (a + (b - a) * fx) + ((c + (d - c) * fx) - (a + (b - a) * fx)) * fz;
}
float fbm(float x, float z, uint64_t seed)
{
    float s = 0.0f, amp = 1.0f, freq = 1.0f / 48.0f;
    for (int o = 0; o < 4; ++o) {
        s += amp * (vnoise(x * freq, z * freq, o, seed) - 0.5f);
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return 24.0f * s;
}

void terrain(Builder& B, Rng& rng, uint64_t seed, int n, float extent)
{
    const float h = extent / (float)(n - 1);
    const uint32_t base = (uint32_t)(B.xyz.size() / 3);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            float x = i * h, z = j * h;
            float y = fbm(x, z, seed);
            float e = 0.25f * h;
            float dx = (fbm(x + e, z, seed) - fbm(x - e, z, seed)) / (2 * e);
            float dz = (fbm(x, z + e, seed) - fbm(x, z - e, seed)) / (2 * e);
            float l = sqrtf(dx * dx + 1.0f + dz * dz);
            B.vert(x, y, z, -dx / l, 1.0f / l, -dz / l);
        }
    const float rgb0[3] = {rng.uni(0.3f, 0.5f), rng.uni(0.4f, 0.6f), rng.uni(0.2f, 0.35f)};
    for (int j = 0; j + 1 < n; ++j)
        for (int i = 0; i + 1 < n; ++i) {
            uint32_t a = base + j * n + i, b = a + 1, c = a + n, d = c + 1;
            B.tri(a, c, b, rgb0);
            B.tri(b, c, d, rgb0);
        }
}

vlb_host_scene* finish(Builder& B)
{
    vlb_host_scene* s = (vlb_host_scene*)calloc(1, sizeof(vlb_host_scene));
    s->nverts = (uint32_t)(B.xyz.size() / 3);
    s->ntris = (uint32_t)(B.idx.size() / 3);
    s->xyz = (float*)malloc(B.xyz.size() * sizeof(float));
    s->normals = (float*)malloc(B.nrm.size() * sizeof(float));
    s->idx = (uint32_t*)malloc(B.idx.size() * sizeof(uint32_t));
    s->albedo = (float*)malloc(B.alb.size() * sizeof(float));
    memcpy(s->xyz, B.xyz.data(), B.xyz.size() * sizeof(float));
    memcpy(s->normals, B.nrm.data(), B.nrm.size() * sizeof(float));
    memcpy(s->idx, B.idx.data(), B.idx.size() * sizeof(uint32_t));
    memcpy(s->albedo, B.alb.data(), B.alb.size() * sizeof(float));
    for (int a = 0; a < 3; ++a) {
        s->bounds_min[a] = INFINITY;
        s->bounds_max[a] = -INFINITY;
    }
    for (uint32_t i = 0; i < s->nverts; ++i)
        for (int a = 0; a < 3; ++a) {
            s->bounds_min[a] = std::min(s->bounds_min[a], s->xyz[3 * i + a]);
            s->bounds_max[a] = std::max(s->bounds_max[a], s->xyz[3 * i + a]);
        }
    return s;
}

}  // namespace

extern "C" {

vlb_status vlb_scene_generate(int32_t kind, uint64_t seed, uint32_t param0, uint32_t param1,
                              vlb_host_scene** out)
{
    if (!out) return VLB_ERR_INVALID;
    *out = nullptr;
    Builder B;
    Rng rng(seed);
    switch (kind) {
    case VLB_SCENE_ROOM: {
        // open-top room 10 x 4 x 10 m (floor + 4 inward walls)
        const float wall[3] = {rng.uni(0.5f, 0.8f), rng.uni(0.5f, 0.8f), rng.uni(0.5f, 0.8f)};
        const float floor_[3] = {rng.uni(0.3f, 0.6f), rng.uni(0.3f, 0.6f), rng.uni(0.3f, 0.6f)};
        const float p[8][3] = {{0, 0, 0}, {10, 0, 0}, {10, 4, 0}, {0, 4, 0},
                               {0, 0, 10}, {10, 0, 10}, {10, 4, 10}, {0, 4, 10}};
        const float nu[3] = {0, 1, 0}, nx0[3] = {1, 0, 0}, nx1[3] = {-1, 0, 0},
                    nz0[3] = {0, 0, 1}, nz1[3] = {0, 0, -1};
        B.quad(p[0], p[4], p[5], p[1], nu, floor_);
        B.quad(p[0], p[3], p[7], p[4], nx0, wall);
        B.quad(p[1], p[5], p[6], p[2], nx1, wall);
        B.quad(p[0], p[1], p[2], p[3], nz0, wall);
        B.quad(p[4], p[7], p[6], p[5], nz1, wall);
        const uint32_t nboxes = param0 ? param0 : 64, nsph = param1 ? param1 : 8;
        for (uint32_t b = 0; b < nboxes; ++b) {
            float sz[3] = {rng.uni(0.2f, 1.5f), rng.uni(0.2f, 1.5f), rng.uni(0.2f, 1.5f)};
            float lo[3] = {rng.uni(0.0f, 10.0f - sz[0]), 0.0f, rng.uni(0.0f, 10.0f - sz[2])};
            if (b & 1) lo[1] = rng.uni(0.0f, 4.0f - sz[1]);
            float hi[3] = {lo[0] + sz[0], lo[1] + sz[1], lo[2] + sz[2]};
            float rgb[3] = {rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)};
            B.box(lo, hi, rgb);
        }
        for (uint32_t k = 0; k < nsph; ++k) {
            float r = rng.uni(0.3f, 0.8f);
            float c[3] = {rng.uni(r, 10.0f - r), rng.uni(r, 4.0f - r), rng.uni(r, 10.0f - r)};
            float rgb[3] = {rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)};
            B.icosphere(c, r, 3, rgb);
        }
        break;
    }
    case VLB_SCENE_TERRAIN: {
        const int n = param0 ? (int)param0 : 256;
        const uint32_t nsph = param1 == 0xFFFFFFFFu ? 0 : (param1 ? param1 : 50);
        if (n < 2) return VLB_ERR_INVALID;
        const float extent = 256.0f;
        terrain(B, rng, seed, n, extent);
        for (uint32_t k = 0; k < nsph; ++k) {
            float r = rng.uni(2.0f, 6.0f);
            float cx = rng.uni(10.0f, extent - 10.0f), cz = rng.uni(10.0f, extent - 10.0f);
            float c[3] = {cx, fbm(cx, cz, seed) + r + rng.uni(0.0f, 10.0f), cz};
            float rgb[3] = {rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)};
            B.icosphere(c, r, 5, rgb);
        }
        break;
    }
    case VLB_SCENE_CITY: {
        const uint32_t nb = param0 ? param0 : 400;
        const float cell = param1 ? (float)param1 / 100.0f : 1.1f;  // facade tile size (m)
        const int side = (int)ceilf(sqrtf((float)nb));
        const float lot = 40.0f;
        const float g0[3] = {0, 0, 0}, g1[3] = {side * lot, 0, 0}, g2[3] = {side * lot, 0, side * lot},
                    g3[3] = {0, 0, side * lot}, up[3] = {0, 1, 0};
        const float grey[3] = {0.35f, 0.35f, 0.35f};
        B.quad(g0, g3, g2, g1, up, grey);
        for (uint32_t b = 0; b < nb; ++b) {
            const int gx = (int)(b % side), gz = (int)(b / side);
            const float w = rng.uni(5.0f, 30.0f), d = rng.uni(5.0f, 30.0f), h = rng.uni(5.0f, 80.0f);
            const float x0 = gx * lot + rng.uni(0.0f, lot - w), z0 = gz * lot + rng.uni(0.0f, lot - d);
            const float rgb[3] = {rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)};
            const float cs[4][2] = {{x0, z0}, {x0 + w, z0}, {x0 + w, z0 + d}, {x0, z0 + d}};
            for (int wi = 0; wi < 4; ++wi) {
                const float* A = cs[wi];
                const float* C = cs[(wi + 1) & 3];
                const float len = sqrtf((C[0] - A[0]) * (C[0] - A[0]) + (C[1] - A[1]) * (C[1] - A[1]));
                const int nu = std::max(1, (int)ceilf(len / cell)), nv = std::max(1, (int)ceilf(h / cell));
                const float ex = (C[0] - A[0]) / len, ez = (C[1] - A[1]) / len;
                const float n[3] = {ez, 0.0f, -ex};  // outward for counter-clockwise footprint
                for (int iv = 0; iv < nv; ++iv)
                    for (int iu = 0; iu < nu; ++iu) {
                        const float u0 = len * iu / nu, u1 = len * (iu + 1) / nu;
                        const float y0 = h * iv / nv, y1 = h * (iv + 1) / nv;
                        const float a[3] = {A[0] + ex * u0, y0, A[1] + ez * u0};
                        const float bq[3] = {A[0] + ex * u1, y0, A[1] + ez * u1};
                        const float c[3] = {A[0] + ex * u1, y1, A[1] + ez * u1};
                        const float dq[3] = {A[0] + ex * u0, y1, A[1] + ez * u0};
                        B.quad(a, dq, c, bq, n, rgb);
                    }
            }
            const float r0[3] = {x0, h, z0}, r1[3] = {x0 + w, h, z0}, r2[3] = {x0 + w, h, z0 + d},
                        r3[3] = {x0, h, z0 + d};
            B.quad(r0, r3, r2, r1, up, rgb);
        }
        break;
    }
    case VLB_SCENE_CUBE: {
        // the mesh of src/vendor/default_blender_cube.gltf: [-1,1]^3, 12 tris
        const float lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1}, rgb[3] = {0.8f, 0.8f, 0.8f};
        B.box(lo, hi, rgb);
        break;
    }
    case VLB_SCENE_SPHERES: {
        const uint32_t ns = param0 ? param0 : 4, sub = param1;
        for (uint32_t k = 0; k < ns; ++k) {
            float r = rng.uni(0.5f, 1.5f);
            float c[3] = {rng.uni(r, 10.0f - r), rng.uni(r, 10.0f - r), rng.uni(r, 10.0f - r)};
            float rgb[3] = {rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f), rng.uni(0.2f, 0.8f)};
            B.icosphere(c, r, (int)sub, rgb);
        }
        break;
    }
    default:
        return VLB_ERR_INVALID;
    }
    *out = finish(B);
    return VLB_OK;
}

void vlb_host_scene_free(vlb_host_scene* s)
{
    if (!s) return;
    free(s->xyz);
    free(s->normals);
    free(s->idx);
    free(s->albedo);
    free(s);
}

}  // extern "C"
