// gltf.cpp -- glTF scene ingest for the baker: the part of Scene_t
// (src/scene_manager.cpp) that turns a .gltf / .glb file into triangles.
//
//   file types        scene_manager.cpp:31-66   .gltf (JSON) or .glb
//   node tree         scene_manager.cpp:837-857  the nodes of scenes[0]
//   node matrix       scene_manager.cpp:462-477  T * R * S * M (glTF TRS +
//                                                matrix), composed with the
//                                                parents' (445-460)
//   vertex fetch      scene_manager.cpp:241-291  POSITION, NORMAL
//                                                (normalised), min/max
//   index fetch       scene_manager.cpp:293-337  u8 / u16 / u32
//   scene bounds      scene_manager.cpp:497-507  start at 0; each
//                                                primitive's local min and
//                                                max corner through its
//                                                node's own matrix
//   base colour       env_map.rchit:36-49        factor unless all four
//                                                components are 0, else 1
//
// The reference hands local vertices plus a per-instance transform to the
// driver; here every primitive is flattened into world space (one triangle
// soup per scene, the form the GPU BVH builder takes).  Normals go through
// the inverse transpose and stay unnormalised: the shading normalises the
// barycentric mix, as env_map.rchit:63-68 does after its WorldToObject
// product.  Textures are not sampled (a material's base colour is its
// factor), and a material whose factor is black in rgb but not in alpha
// shades as white here (the albedo carries no alpha).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/vlb_bake.h"
#include "json.hpp"

using namespace vlb::json;

namespace {

struct Fail {
    vlb_status code;
};

const JVal* member(const JVal& o, const char* k)
{
    if (o.kind != JVal::Obj) return nullptr;
    auto it = o.o.find(k);
    return it == o.o.end() ? nullptr : &it->second;
}

double num(const JVal* v, double dflt)
{
    if (!v) return dflt;
    if (v->kind == JVal::Int) return (double)v->i;
    if (v->kind == JVal::UInt) return (double)v->u;
    if (v->kind == JVal::Float) return v->f;
    throw Fail{VLB_ERR_INVALID};
}

size_t index_of(const JVal* v)
{
    const double x = num(v, -1.0);
    if (x < 0) throw Fail{VLB_ERR_INVALID};
    return (size_t)x;
}

const JVal& arr_at(const JVal& root, const char* name, size_t i)
{
    const JVal* a = member(root, name);
    if (!a || a->kind != JVal::Arr || i >= a->a.size()) throw Fail{VLB_ERR_INVALID};
    return a->a[i];
}

bool read_file(const std::string& path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

std::vector<uint8_t> b64(const std::string& s, size_t from)
{
    auto val = [](unsigned char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        const int v = val((unsigned char)s[i]);
        if (v < 0) break;
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)((acc >> bits) & 0xFFu));
        }
    }
    return out;
}

// column-major 4x4, as glTF and glm store it
struct Mat4 {
    double m[16];
    static Mat4 id()
    {
        Mat4 r{};
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
        return r;
    }
    double& at(int row, int col) { return m[col * 4 + row]; }
    double at(int row, int col) const { return m[col * 4 + row]; }
};

Mat4 mul(const Mat4& a, const Mat4& b)
{
    Mat4 r{};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a.at(i, k) * b.at(k, j);
            r.at(i, j) = s;
        }
    return r;
}

// scene_manager.cpp:462-477: translate(T) * mat4(quat R) * scale(S) * matrix
Mat4 node_matrix(const JVal& node)
{
    Mat4 T = Mat4::id(), R = Mat4::id(), S = Mat4::id(), M = Mat4::id();
    const JVal* t = member(node, "translation");
    if (t && t->kind == JVal::Arr && t->a.size() == 3)
        for (int k = 0; k < 3; ++k) T.at(k, 3) = num(&t->a[k], 0.0);
    const JVal* q = member(node, "rotation");
    if (q && q->kind == JVal::Arr && q->a.size() == 4) {
        const double x = num(&q->a[0], 0), y = num(&q->a[1], 0), z = num(&q->a[2], 0), w = num(&q->a[3], 1);
        R.at(0, 0) = 1 - 2 * (y * y + z * z);
        R.at(0, 1) = 2 * (x * y - w * z);
        R.at(0, 2) = 2 * (x * z + w * y);
        R.at(1, 0) = 2 * (x * y + w * z);
        R.at(1, 1) = 1 - 2 * (x * x + z * z);
        R.at(1, 2) = 2 * (y * z - w * x);
        R.at(2, 0) = 2 * (x * z - w * y);
        R.at(2, 1) = 2 * (y * z + w * x);
        R.at(2, 2) = 1 - 2 * (x * x + y * y);
    }
    const JVal* s = member(node, "scale");
    if (s && s->kind == JVal::Arr && s->a.size() == 3)
        for (int k = 0; k < 3; ++k) S.at(k, k) = num(&s->a[k], 1.0);
    const JVal* mm = member(node, "matrix");
    if (mm && mm->kind == JVal::Arr && mm->a.size() == 16)
        for (int k = 0; k < 16; ++k) M.m[k] = num(&mm->a[k], 0.0);
    return mul(mul(mul(T, R), S), M);
}

// inverse transpose of the upper 3x3 (row-major result)
void normal_matrix(const Mat4& a, double n[9])
{
    const double m00 = a.at(0, 0), m01 = a.at(0, 1), m02 = a.at(0, 2);
    const double m10 = a.at(1, 0), m11 = a.at(1, 1), m12 = a.at(1, 2);
    const double m20 = a.at(2, 0), m21 = a.at(2, 1), m22 = a.at(2, 2);
    const double c00 = m11 * m22 - m12 * m21, c01 = m12 * m20 - m10 * m22, c02 = m10 * m21 - m11 * m20;
    const double c10 = m02 * m21 - m01 * m22, c11 = m00 * m22 - m02 * m20, c12 = m01 * m20 - m00 * m21;
    const double c20 = m01 * m12 - m02 * m11, c21 = m02 * m10 - m00 * m12, c22 = m00 * m11 - m01 * m10;
    const double det = m00 * c00 + m01 * c01 + m02 * c02;
    const double inv = det != 0.0 ? 1.0 / det : 0.0;
    // (M^-1)^T = cofactor matrix / det
    const double c[9] = {c00, c01, c02, c10, c11, c12, c20, c21, c22};
    for (int k = 0; k < 9; ++k) n[k] = c[k] * inv;
}

struct Loader {
    JVal root;
    std::string dir;
    std::vector<std::vector<uint8_t>> buffers;
    std::vector<float> xyz, nrm, alb;
    std::vector<uint32_t> idx;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};

    void load_buffers(const std::string* glb_bin)
    {
        const JVal* bufs = member(root, "buffers");
        if (!bufs) return;
        if (bufs->kind != JVal::Arr) throw Fail{VLB_ERR_INVALID};
        for (size_t i = 0; i < bufs->a.size(); ++i) {
            const JVal& b = bufs->a[i];
            const JVal* uri = member(b, "uri");
            std::vector<uint8_t> data;
            if (!uri) {
                if (!glb_bin || i != 0) throw Fail{VLB_ERR_INVALID};
                data.assign(glb_bin->begin(), glb_bin->end());
            } else {
                if (uri->kind != JVal::Str) throw Fail{VLB_ERR_INVALID};
                const std::string& u = uri->s;
                if (u.compare(0, 5, "data:") == 0) {
                    const size_t comma = u.find(',');
                    if (comma == std::string::npos || u.find(";base64") == std::string::npos ||
                        u.find(";base64") > comma)
                        throw Fail{VLB_ERR_INVALID};
                    data = b64(u, comma + 1);
                } else {
                    std::string file;
                    if (!read_file(dir + u, file)) throw Fail{VLB_ERR_NOT_FOUND};
                    data.assign(file.begin(), file.end());
                }
            }
            const size_t len = (size_t)num(member(b, "byteLength"), (double)data.size());
            if (data.size() < len) throw Fail{VLB_ERR_INVALID};
            buffers.push_back(std::move(data));
        }
    }

    // accessor -> pointer to element i, component size and count per element
    struct Acc {
        const uint8_t* base = nullptr;
        size_t stride = 0, count = 0;
        int ctype = 0, ncomp = 0;
    };
    Acc accessor(size_t ai)
    {
        const JVal& a = arr_at(root, "accessors", ai);
        if (member(a, "sparse")) throw Fail{VLB_ERR_UNSUPPORTED};
        Acc r;
        r.count = (size_t)num(member(a, "count"), 0.0);
        r.ctype = (int)num(member(a, "componentType"), 0.0);
        const JVal* ty = member(a, "type");
        if (!ty || ty->kind != JVal::Str) throw Fail{VLB_ERR_INVALID};
        const std::string& t = ty->s;
        r.ncomp = t == "SCALAR" ? 1 : t == "VEC2" ? 2 : t == "VEC3" ? 3 : t == "VEC4" ? 4 : t == "MAT4" ? 16 : 0;
        if (!r.ncomp) throw Fail{VLB_ERR_UNSUPPORTED};
        const int csize = (r.ctype == 5120 || r.ctype == 5121) ? 1 : (r.ctype == 5122 || r.ctype == 5123) ? 2
                          : (r.ctype == 5125 || r.ctype == 5126) ? 4 : 0;
        if (!csize) throw Fail{VLB_ERR_UNSUPPORTED};
        const JVal* bvi = member(a, "bufferView");
        if (!bvi) throw Fail{VLB_ERR_UNSUPPORTED};
        const JVal& bv = arr_at(root, "bufferViews", index_of(bvi));
        const size_t bi = index_of(member(bv, "buffer"));
        if (bi >= buffers.size()) throw Fail{VLB_ERR_INVALID};
        const size_t off = (size_t)num(member(bv, "byteOffset"), 0.0) + (size_t)num(member(a, "byteOffset"), 0.0);
        const size_t elem = (size_t)csize * r.ncomp;
        r.stride = (size_t)num(member(bv, "byteStride"), 0.0);
        if (r.stride == 0) r.stride = elem;
        const size_t need = r.count ? off + (r.count - 1) * r.stride + elem : off;
        if (need > buffers[bi].size()) throw Fail{VLB_ERR_INVALID};
        r.base = buffers[bi].data() + off;
        return r;
    }

    static float f32(const uint8_t* p)
    {
        float f;
        memcpy(&f, p, 4);
        return f;
    }

    void primitive(const JVal& prim, const Mat4& local, const Mat4& world)
    {
        const JVal* attrs = member(prim, "attributes");
        const JVal* pos = attrs ? member(*attrs, "POSITION") : nullptr;
        if (!pos) return;
        const Acc P = accessor(index_of(pos));
        if (P.ctype != 5126 || P.ncomp != 3) throw Fail{VLB_ERR_UNSUPPORTED};
        Acc N;
        const JVal* nrmv = member(*attrs, "NORMAL");
        if (nrmv) {
            N = accessor(index_of(nrmv));
            if (N.ctype != 5126 || N.ncomp != 3 || N.count != P.count) throw Fail{VLB_ERR_INVALID};
        }
        double nm[9];
        normal_matrix(world, nm);
        const uint32_t v0 = (uint32_t)(xyz.size() / 3);
        float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
        float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
        for (size_t v = 0; v < P.count; ++v) {
            const uint8_t* pp = P.base + v * P.stride;
            const float p[3] = {f32(pp), f32(pp + 4), f32(pp + 8)};
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], p[k]);
                hi[k] = std::max(hi[k], p[k]);
            }
            for (int r = 0; r < 3; ++r)
                xyz.push_back((float)(world.at(r, 0) * p[0] + world.at(r, 1) * p[1] + world.at(r, 2) * p[2] +
                                      world.at(r, 3)));
            float n[3] = {0.0f, 0.0f, 0.0f};
            if (nrmv) {
                const uint8_t* np = N.base + v * N.stride;
                n[0] = f32(np);
                n[1] = f32(np + 4);
                n[2] = f32(np + 8);
                const float l = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (l > 0.0f)
                    for (float& c : n) c = c / l;
            }
            for (int r = 0; r < 3; ++r)
                nrm.push_back((float)(nm[3 * r] * n[0] + nm[3 * r + 1] * n[1] + nm[3 * r + 2] * n[2]));
        }
        // scene bounds: the local corners through the node's own matrix
        for (int r = 0; r < 3; ++r) {
            const float g0 = (float)(local.at(r, 0) * lo[0] + local.at(r, 1) * lo[1] + local.at(r, 2) * lo[2] +
                                     local.at(r, 3));
            const float g1 = (float)(local.at(r, 0) * hi[0] + local.at(r, 1) * hi[1] + local.at(r, 2) * hi[2] +
                                     local.at(r, 3));
            bmin[r] = std::min(bmin[r], g0);
            bmax[r] = std::max(bmax[r], g1);
        }
        // indices
        std::vector<uint32_t> ind;
        const JVal* iv = member(prim, "indices");
        if (iv) {
            const Acc I = accessor(index_of(iv));
            if (I.ncomp != 1) throw Fail{VLB_ERR_INVALID};
            for (size_t i = 0; i < I.count; ++i) {
                const uint8_t* q = I.base + i * I.stride;
                uint32_t x;
                if (I.ctype == 5125) memcpy(&x, q, 4);
                else if (I.ctype == 5123) {
                    uint16_t h;
                    memcpy(&h, q, 2);
                    x = h;
                } else if (I.ctype == 5121) x = *q;
                else throw Fail{VLB_ERR_UNSUPPORTED};  // "Index component type not supported!"
                if (x >= P.count) throw Fail{VLB_ERR_INVALID};
                ind.push_back(x);
            }
        } else {
            for (size_t i = 0; i < P.count; ++i) ind.push_back((uint32_t)i);
        }
        // base colour of the primitive's material (env_map.rchit:36-49)
        float rgb[3] = {1.0f, 1.0f, 1.0f};
        const JVal* mi = member(prim, "material");
        if (mi) {
            const JVal& mat = arr_at(root, "materials", index_of(mi));
            const JVal* pbr = member(mat, "pbrMetallicRoughness");
            const JVal* f = pbr ? member(*pbr, "baseColorFactor") : nullptr;
            if (f && f->kind == JVal::Arr && f->a.size() == 4) {
                double c[4];
                for (int k = 0; k < 4; ++k) c[k] = num(&f->a[k], 0.0);
                if (c[0] != 0.0 || c[1] != 0.0 || c[2] != 0.0 || c[3] != 0.0)
                    for (int k = 0; k < 3; ++k) rgb[k] = (float)c[k];
            }
        }
        for (size_t t = 0; t + 2 < ind.size(); t += 3) {
            for (int k = 0; k < 3; ++k) idx.push_back(v0 + ind[t + k]);
            for (int k = 0; k < 3; ++k) alb.push_back(rgb[k]);
        }
    }

    void node(size_t ni, const Mat4& parent, int depth)
    {
        if (depth > 256) throw Fail{VLB_ERR_INVALID};
        const JVal& n = arr_at(root, "nodes", ni);
        const Mat4 local = node_matrix(n);
        const Mat4 world = mul(parent, local);
        const JVal* mesh = member(n, "mesh");
        if (mesh) {
            const JVal& m = arr_at(root, "meshes", index_of(mesh));
            const JVal* prims = member(m, "primitives");
            if (prims && prims->kind == JVal::Arr)
                for (const JVal& p : prims->a) primitive(p, local, world);
        }
        const JVal* ch = member(n, "children");
        if (ch && ch->kind == JVal::Arr)
            for (const JVal& c : ch->a) node(index_of(&c), world, depth + 1);
    }
};

uint32_t rd32(const std::string& s, size_t off)
{
    uint32_t x;
    memcpy(&x, s.data() + off, 4);
    return x;
}

}  // namespace

extern "C" {

vlb_status vlb_gltf_load(const char* path, vlb_host_scene** out, char** json_text)
{
    if (!path || !out) return VLB_ERR_INVALID;
    *out = nullptr;
    if (json_text) *json_text = nullptr;
    try {
        const std::string p(path);
        std::string file;
        if (!read_file(p, file)) return VLB_ERR_NOT_FOUND;
        Loader L;
        const size_t slash = p.find_last_of('/');
        L.dir = slash == std::string::npos ? std::string() : p.substr(0, slash + 1);
        std::string json, bin;
        const bool glb = p.size() >= 4 && p.compare(p.size() - 4, 4, ".glb") == 0;
        const bool gltf = p.size() >= 5 && p.compare(p.size() - 5, 5, ".gltf") == 0;
        if (!glb && !gltf) return VLB_ERR_UNSUPPORTED;  // "Not supported file format"
        if (glb) {
            if (file.size() < 20 || rd32(file, 0) != 0x46546C67u || rd32(file, 4) != 2u) return VLB_ERR_INVALID;
            size_t off = 12;
            while (off + 8 <= file.size()) {
                const uint32_t len = rd32(file, off), type = rd32(file, off + 4);
                if (off + 8 + (size_t)len > file.size()) return VLB_ERR_INVALID;
                if (type == 0x4E4F534Au && json.empty()) json = file.substr(off + 8, len);
                else if (type == 0x004E4942u && bin.empty()) bin = file.substr(off + 8, len);
                off += 8 + (size_t)len;
            }
            if (json.empty()) return VLB_ERR_INVALID;
        } else {
            json = file;
        }
        if (!parse(json.data(), json.size(), L.root) || L.root.kind != JVal::Obj) return VLB_ERR_INVALID;
        L.load_buffers(glb ? &bin : nullptr);
        const JVal* scenes = member(L.root, "scenes");
        if (scenes && scenes->kind == JVal::Arr && !scenes->a.empty()) {
            const JVal* roots = member(scenes->a[0], "nodes");
            if (roots && roots->kind == JVal::Arr)
                for (const JVal& r : roots->a) L.node(index_of(&r), Mat4::id(), 0);
        }
        vlb_host_scene* s = (vlb_host_scene*)calloc(1, sizeof(vlb_host_scene));
        if (!s) return VLB_ERR_OOM;
        s->nverts = (uint32_t)(L.xyz.size() / 3);
        s->ntris = (uint32_t)(L.idx.size() / 3);
        s->xyz = (float*)malloc(std::max<size_t>(1, L.xyz.size()) * sizeof(float));
        s->normals = (float*)malloc(std::max<size_t>(1, L.nrm.size()) * sizeof(float));
        s->idx = (uint32_t*)malloc(std::max<size_t>(1, L.idx.size()) * sizeof(uint32_t));
        s->albedo = (float*)malloc(std::max<size_t>(1, L.alb.size()) * sizeof(float));
        if (!s->xyz || !s->normals || !s->idx || !s->albedo) {
            vlb_host_scene_free(s);
            return VLB_ERR_OOM;
        }
        if (!L.xyz.empty()) memcpy(s->xyz, L.xyz.data(), L.xyz.size() * sizeof(float));
        if (!L.nrm.empty()) memcpy(s->normals, L.nrm.data(), L.nrm.size() * sizeof(float));
        if (!L.idx.empty()) memcpy(s->idx, L.idx.data(), L.idx.size() * sizeof(uint32_t));
        if (!L.alb.empty()) memcpy(s->albedo, L.alb.data(), L.alb.size() * sizeof(float));
        for (int a = 0; a < 3; ++a) {
            s->bounds_min[a] = L.bmin[a];
            s->bounds_max[a] = L.bmax[a];
        }
        if (json_text) {
            char* t = (char*)malloc(json.size() + 1);
            if (!t) {
                vlb_host_scene_free(s);
                return VLB_ERR_OOM;
            }
            memcpy(t, json.data(), json.size());
            t[json.size()] = '\0';
            *json_text = t;
        }
        *out = s;
        return VLB_OK;
    } catch (const Fail& f) {
        return f.code;
    } catch (...) {
        return VLB_ERR_INVALID;
    }
}

void vlb_text_free(char* text) { free(text); }

}  // extern "C"
