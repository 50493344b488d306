// host_util.cpp -- host helpers of the C ABI: direction sets, probe grid,
// double-precision SH basis, base64.
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/vlb_bake.h"

namespace vlb {

static const double kPi = 3.14159265358979323846;

// Fibonacci sphere, y up: y_i = 1 - (2i+1)/R, phi_i = i * pi * (3 - sqrt 5).
void fibonacci_dirs(uint32_t R, float* xyz)
{
    const double ga = kPi * (3.0 - sqrt(5.0));
    for (uint32_t i = 0; i < R; ++i) {
        double y = 1.0 - (2.0 * i + 1.0) / (double)R;
        double r = sqrt(fmax(0.0, 1.0 - y * y));
        double phi = fmod((double)i * ga, 2.0 * kPi);
        xyz[3 * i + 0] = (float)(r * cos(phi));
        xyz[3 * i + 1] = (float)y;
        xyz[3 * i + 2] = (float)(r * sin(phi));
    }
}

// shaders/env_map.rgen:21 with sh_common.h:8-22 (float, PI = 3.1415926538f),
// weight of shaders/sh.comp:32-33.
void equirect_dirs(uint32_t W, uint32_t H, float* xyz, float* weight)
{
    const float PI = 3.1415926538f;
    const float area = (2.0f * PI / (float)W) * (PI / (float)H);
    for (uint32_t y = 0; y < H; ++y) {
        const float theta = PI * ((float)y + 0.5f) / (float)H;
        const float wgt = area * sinf(theta);
        for (uint32_t x = 0; x < W; ++x) {
            const float phi = 2.f * PI * ((float)x + 0.5f) / (float)W;
            const float r = sinf(theta);
            const float a = r * cosf(phi), b = r * sinf(phi), c = cosf(theta);
            const float l = sqrtf(a * a + b * b + c * c);
            const size_t i = (size_t)y * W + x;
            xyz[3 * i + 0] = a / l;
            xyz[3 * i + 1] = c / l;
            xyz[3 * i + 2] = b / l;
            if (weight) weight[i] = wgt;
        }
    }
}

// Real SH basis of shaders/sh_common.h:26-152 (l <= 3 here), double.
void sh_eval_double(int K, bool exact, const double d[3], double* out)
{
    const double x = d[0], y = d[1], z = d[2];
    double c0, c1, c2a, c20, c22, c33, c32, c31, c30, c3p2;
    if (exact) {
        c0 = 0.5 * sqrt(1.0 / kPi);
        c1 = sqrt(3.0 / (4.0 * kPi));
        c2a = 0.5 * sqrt(15.0 / kPi);
        c20 = 0.25 * sqrt(5.0 / kPi);
        c22 = 0.25 * sqrt(15.0 / kPi);
        c33 = 0.25 * sqrt(35.0 / (2.0 * kPi));
        c32 = 0.5 * sqrt(105.0 / kPi);
        c31 = 0.25 * sqrt(21.0 / (2.0 * kPi));
        c30 = 0.25 * sqrt(7.0 / kPi);
        c3p2 = 0.25 * sqrt(105.0 / kPi);
    } else {
        c0 = 0.282095;
        c1 = 0.488603;
        c2a = 1.092548;
        c20 = 0.315392;
        c22 = 0.546274;
        c33 = 0.590044;
        c32 = 2.890611;
        c31 = 0.457046;
        c30 = 0.373176;
        c3p2 = 1.445306;
    }
    double v[16];
    v[0] = c0;
    v[1] = -c1 * y;
    v[2] = c1 * z;
    v[3] = -c1 * x;
    v[4] = c2a * x * y;
    v[5] = -c2a * y * z;
    v[6] = c20 * (-x * x - y * y + 2.0 * z * z);
    v[7] = -c2a * x * z;
    v[8] = c22 * (x * x - y * y);
    v[9] = -c33 * y * (3.0 * x * x - y * y);
    v[10] = c32 * x * y * z;
    v[11] = -c31 * y * (4.0 * z * z - x * x - y * y);
    v[12] = c30 * z * (2.0 * z * z - 3.0 * x * x - 3.0 * y * y);
    v[13] = -c31 * x * (4.0 * z * z - x * x - y * y);
    v[14] = c3p2 * z * (x * x - y * y);
    v[15] = -c33 * x * (x * x - 3.0 * y * y);
    for (int k = 0; k < K && k < 16; ++k) out[k] = v[k];
}

}  // namespace vlb

extern "C" {

vlb_status vlb_probe_grid(const float bmin[3], const float bmax[3], const uint32_t counts[3],
                          int32_t order, float* out_xyz, float* out_step)
{
    if (!bmin || !bmax || !counts || !out_xyz) return VLB_ERR_INVALID;
    if (counts[0] == 0 || counts[1] == 0 || counts[2] == 0) return VLB_ERR_INVALID;
    if (order != VLB_ORDER_LINEAR && order != VLB_ORDER_REFERENCE_BFS) return VLB_ERR_INVALID;
    // gridStep = (bounds[1] - bounds[0]) / (probesCount3D - 1.f)
    float step[3];
    for (int a = 0; a < 3; ++a) step[a] = (bmax[a] - bmin[a]) / ((float)counts[a] - 1.0f);
    if (out_step) memcpy(out_step, step, sizeof step);
    if (order == VLB_ORDER_REFERENCE_BFS) {
        // light_baker.cpp:80-101: start from bounds[0]; for each dimension,
        // every position found so far spawns counts[dim]-1 successors by
        // repeated "position[dim] += gridStep[dim]".
        size_t n = 1;
        memcpy(out_xyz, bmin, 3 * sizeof(float));
        for (int dim = 0; dim < 3; ++dim) {
            const size_t ncopy = n;
            for (size_t q = 0; q < ncopy; ++q) {
                float p[3] = {out_xyz[3 * q], out_xyz[3 * q + 1], out_xyz[3 * q + 2]};
                for (int i = 0; (float)i < (float)counts[dim] - 1.0f; ++i) {
                    p[dim] += step[dim];
                    memcpy(&out_xyz[3 * n], p, sizeof p);
                    ++n;
                }
            }
        }
    } else {
        size_t i = 0;
        for (uint32_t z = 0; z < counts[2]; ++z)
            for (uint32_t y = 0; y < counts[1]; ++y)
                for (uint32_t x = 0; x < counts[0]; ++x, ++i) {
                    out_xyz[3 * i + 0] = bmin[0] + (float)x * step[0];
                    out_xyz[3 * i + 1] = bmin[1] + (float)y * step[1];
                    out_xyz[3 * i + 2] = bmin[2] + (float)z * step[2];
                }
    }
    return VLB_OK;
}

vlb_status vlb_directions(int32_t dir_mode, uint32_t rays, uint32_t w, uint32_t h, float* out_xyz,
                          float* out_weight, size_t n)
{
    if (!out_xyz) return VLB_ERR_INVALID;
    if (dir_mode == VLB_DIRS_FIBONACCI) {
        if (rays == 0 || n < rays) return VLB_ERR_INVALID;
        vlb::fibonacci_dirs(rays, out_xyz);
        if (out_weight)
            for (uint32_t i = 0; i < rays; ++i) out_weight[i] = (float)(4.0 * vlb::kPi / rays);
        return VLB_OK;
    }
    if (dir_mode == VLB_DIRS_EQUIRECT) {
        if (w == 0 || h == 0 || n < (size_t)w * h) return VLB_ERR_INVALID;
        vlb::equirect_dirs(w, h, out_xyz, out_weight);
        return VLB_OK;
    }
    return VLB_ERR_INVALID;
}

vlb_status vlb_base64_encode(const uint8_t* b, size_t n, char* out, size_t out_len,
                             size_t* written)
{
    // base64_encode of src/baker/light_baker.cpp:330-373: RFC 4648 alphabet,
    // 3 bytes -> 4 symbols, '=' padding of the last group.
    static const char tbl[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    const size_t need = 4 * ((n + 2) / 3);
    if (written) *written = need;
    if (!out) return VLB_OK;
    if (out_len < need + 1 || (n && !b)) return VLB_ERR_INVALID;
    size_t o = 0, i = 0;
    for (; i + 3 <= n; i += 3) {
        const uint32_t v = ((uint32_t)b[i] << 16) | ((uint32_t)b[i + 1] << 8) | b[i + 2];
        out[o++] = tbl[(v >> 18) & 63];
        out[o++] = tbl[(v >> 12) & 63];
        out[o++] = tbl[(v >> 6) & 63];
        out[o++] = tbl[v & 63];
    }
    const size_t rem = n - i;
    if (rem) {
        uint32_t v = (uint32_t)b[i] << 16;
        if (rem == 2) v |= (uint32_t)b[i + 1] << 8;
        out[o++] = tbl[(v >> 18) & 63];
        out[o++] = tbl[(v >> 12) & 63];
        out[o++] = rem == 2 ? tbl[(v >> 6) & 63] : '=';
        out[o++] = '=';
    }
    out[o] = 0;
    return VLB_OK;
}

}  // extern "C"
