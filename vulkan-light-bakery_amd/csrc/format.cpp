// format.cpp -- the baked-light glTF: output of LightBaker::serialize
// (src/baker/light_baker.cpp:375-401) and its reader, Scene_t::loadBakedLight
// (src/scene_manager.cpp:613-648).
//
// The reference reads the scene's glTF JSON with nlohmann::json, appends the
// coefficient bytes as a base64 data-URI buffer plus a bufferView, sets
// json["light"] = {gridStep: {x,y,z}, lmax: <coefficient count>, bufferView},
// and writes it with std::setw(4).  This file does the same with the JSON DOM
// of json.hpp, whose printer follows nlohmann's dump(4) layout.
#include <string.h>

#include <string>
#include <vector>

#include "../../include/vlb_bake.h"
#include "json.hpp"

using namespace vlb::json;

namespace {

// RFC 4648 base64 decode (the alphabet of light_baker.cpp:330-373); stops at
// '=' or the first byte outside the alphabet, as scene_manager.cpp:561-609.
std::vector<uint8_t> b64_decode(const std::string& s)
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
    out.reserve(s.size() / 4 * 3);
    uint32_t acc = 0;
    int bits = 0;
    for (unsigned char c : s) {
        const int v = val(c);
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

const JVal* member(const JVal& o, const char* k)
{
    if (o.kind != JVal::Obj) return nullptr;
    auto it = o.o.find(k);
    return it == o.o.end() ? nullptr : &it->second;
}

bool as_number(const JVal* v, double& x)
{
    if (!v) return false;
    if (v->kind == JVal::Int) x = (double)v->i;
    else if (v->kind == JVal::UInt) x = (double)v->u;
    else if (v->kind == JVal::Float) x = v->f;
    else return false;
    return true;
}

}  // namespace


extern "C" vlb_status vlb_gltf_bake_json(const char* gltf_json, const float* coeffs,
                                         uint32_t num_probes, uint32_t num_coefs,
                                         const float grid_step[3], char* out, size_t out_len,
                                         size_t* written)
{
    if (!gltf_json || !grid_step || (num_probes && !coeffs)) return VLB_ERR_INVALID;
    try {
        JVal root;
        if (!parse(gltf_json, strlen(gltf_json), root) || root.kind != JVal::Obj) return VLB_ERR_INVALID;
        const size_t nbytes = (size_t)num_probes * num_coefs * 3 * sizeof(float);
        size_t enc_len = 0;
        vlb_base64_encode(reinterpret_cast<const uint8_t*>(coeffs), nbytes, nullptr, 0, &enc_len);
        std::string enc(enc_len + 1, '\0');
        if (vlb_base64_encode(reinterpret_cast<const uint8_t*>(coeffs), nbytes, &enc[0], enc.size(),
                              &enc_len) != VLB_OK)
            return VLB_ERR_INVALID;
        enc.resize(enc_len);
        // json["buffers"].push_back({byteLength, uri})
        JVal& buffers = root.o["buffers"];
        if (buffers.kind == JVal::Null) buffers.kind = JVal::Arr;
        if (buffers.kind != JVal::Arr) return VLB_ERR_INVALID;
        JVal buf;
        buf.kind = JVal::Obj;
        buf.o["byteLength"] = juint(nbytes);
        buf.o["uri"] = jstr("data:application/octet-stream;base64," + enc);
        buffers.a.push_back(std::move(buf));
        // json["bufferViews"].push_back({buffer, byteLength, byteOffset})
        JVal& views = root.o["bufferViews"];
        if (views.kind == JVal::Null) views.kind = JVal::Arr;
        if (views.kind != JVal::Arr) return VLB_ERR_INVALID;
        JVal bv;
        bv.kind = JVal::Obj;
        bv.o["buffer"] = juint(buffers.a.size() - 1);
        bv.o["byteLength"] = juint(nbytes);
        bv.o["byteOffset"] = juint(0);
        views.a.push_back(std::move(bv));
        // json["light"] = {gridStep, lmax, bufferView}
        JVal& light = root.o["light"];
        if (light.kind == JVal::Null) light.kind = JVal::Obj;
        if (light.kind != JVal::Obj) return VLB_ERR_INVALID;
        JVal gs;
        gs.kind = JVal::Obj;
        gs.o["x"] = jfloat((double)grid_step[0]);
        gs.o["y"] = jfloat((double)grid_step[1]);
        gs.o["z"] = jfloat((double)grid_step[2]);
        light.o["gridStep"] = std::move(gs);
        light.o["lmax"] = juint(num_coefs);
        light.o["bufferView"] = juint(views.a.size() - 1);
        std::string text;
        dump(text, root, 0);
        text += "\n";
        if (written) *written = text.size();
        if (!out) return VLB_OK;
        if (out_len < text.size() + 1) return VLB_ERR_INVALID;
        memcpy(out, text.c_str(), text.size() + 1);
        return VLB_OK;
    } catch (...) {
        return VLB_ERR_INVALID;
    }
}

extern "C" vlb_status vlb_gltf_read_bake(const char* gltf_json, float grid_step[3], uint32_t* num_coefs,
                                         float* coeffs, size_t coeffs_len, size_t* num_floats)
{
    if (!gltf_json) return VLB_ERR_INVALID;
    try {
        JVal root;
        if (!parse(gltf_json, strlen(gltf_json), root) || root.kind != JVal::Obj) return VLB_ERR_INVALID;
        const JVal* light = member(root, "light");
        if (!light || light->kind != JVal::Obj || light->o.empty()) return VLB_ERR_NOT_FOUND;
        const JVal* gs = member(*light, "gridStep");
        double g[3], lmax, bvi;
        if (!gs || !as_number(member(*gs, "x"), g[0]) || !as_number(member(*gs, "y"), g[1]) ||
            !as_number(member(*gs, "z"), g[2]) || !as_number(member(*light, "lmax"), lmax) ||
            !as_number(member(*light, "bufferView"), bvi))
            return VLB_ERR_INVALID;
        const JVal* views = member(root, "bufferViews");
        if (!views || views->kind != JVal::Arr || bvi < 0 || (size_t)bvi >= views->a.size())
            return VLB_ERR_INVALID;
        double bi;
        if (!as_number(member(views->a[(size_t)bvi], "buffer"), bi)) return VLB_ERR_INVALID;
        const JVal* bufs = member(root, "buffers");
        if (!bufs || bufs->kind != JVal::Arr || bi < 0 || (size_t)bi >= bufs->a.size()) return VLB_ERR_INVALID;
        const JVal& buf = bufs->a[(size_t)bi];
        const JVal* uri = member(buf, "uri");
        double blen;
        if (!uri || uri->kind != JVal::Str || !as_number(member(buf, "byteLength"), blen) || blen < 0)
            return VLB_ERR_INVALID;
        // the data-URI header is dropped and the buffer's byteLength bytes kept
        const std::string header = "data:application/octet-stream;base64,";
        std::string enc = uri->s;
        if (enc.compare(0, header.size(), header) == 0) enc.erase(0, header.size());
        std::vector<uint8_t> bytes = b64_decode(enc);
        const size_t size = (size_t)blen;
        if (bytes.size() < size) return VLB_ERR_INVALID;
        for (int a = 0; a < 3; ++a)
            if (grid_step) grid_step[a] = (float)g[a];
        if (num_coefs) *num_coefs = (uint32_t)lmax;
        const size_t nf = size / sizeof(float);
        if (num_floats) *num_floats = nf;
        if (!coeffs) return VLB_OK;
        if (coeffs_len < nf) return VLB_ERR_INVALID;
        memcpy(coeffs, bytes.data(), nf * sizeof(float));
        return VLB_OK;
    } catch (...) {
        return VLB_ERR_INVALID;
    }
}
