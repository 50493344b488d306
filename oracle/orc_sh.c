/*
 * orc_sh.c -- oracle SH basis, direction sets, probe grid, image/sky
 * projections, colour helpers and base64.  TEST INFRASTRUCTURE ONLY (see
 * vlb_oracle.h).  Compiled with -ffp-contract=off.
 */
#include "vlb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

/* ------------------------------------------------------------- SH basis */
/* shaders/sh_common.h:26-152, same polynomials, coef index l(l+1)+m. */
void orc_sh_eval(int K, int exact, const double d[3], double* out)
{
    const double x = d[0], y = d[1], z = d[2];
    const double pi = ORC_PI_D;
    double c0, c1, c2a, c20, c22, c33, c32, c31, c30, c3p2, c44, c43, c42, c41, c40, c4p2, c4p4;
    if (exact) {
        c0 = 0.5 * sqrt(1.0 / pi);
        c1 = sqrt(3.0 / (4.0 * pi));
        c2a = 0.5 * sqrt(15.0 / pi);
        c20 = 0.25 * sqrt(5.0 / pi);
        c22 = 0.25 * sqrt(15.0 / pi);
        c33 = 0.25 * sqrt(35.0 / (2.0 * pi));
        c32 = 0.5 * sqrt(105.0 / pi);
        c31 = 0.25 * sqrt(21.0 / (2.0 * pi));
        c30 = 0.25 * sqrt(7.0 / pi);
        c3p2 = 0.25 * sqrt(105.0 / pi);
        c44 = 0.75 * sqrt(35.0 / pi);
        c43 = 0.75 * sqrt(35.0 / (2.0 * pi));
        c42 = 0.75 * sqrt(5.0 / pi);
        c41 = 0.75 * sqrt(5.0 / (2.0 * pi));
        c40 = 3.0 / 16.0 * sqrt(1.0 / pi);
        c4p2 = 3.0 / 8.0 * sqrt(5.0 / pi);
        c4p4 = 3.0 / 16.0 * sqrt(35.0 / pi);
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
        c44 = 2.503343;
        c43 = 1.770131;
        c42 = 0.946175;
        c41 = 0.669047;
        c40 = 0.105786;
        c4p2 = 0.473087;
        c4p4 = 0.625836;
    }
    double v[25];
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
    v[16] = c44 * x * y * (x * x - y * y);
    v[17] = -c43 * y * z * (3.0 * x * x - y * y);
    v[18] = c42 * x * y * (7.0 * z * z - 1.0);
    v[19] = -c41 * y * z * (7.0 * z * z - 3.0);
    double z2 = z * z;
    v[20] = c40 * (35.0 * z2 * z2 - 30.0 * z2 + 3.0);
    v[21] = -c41 * x * z * (7.0 * z * z - 3.0);
    v[22] = c4p2 * (x * x - y * y) * (7.0 * z * z - 1.0);
    v[23] = -c43 * x * z * (x * x - 3.0 * y * y);
    double x2 = x * x, y2 = y * y;
    v[24] = c4p4 * (x2 * (x2 - 3.0 * y2) - y2 * (3.0 * x2 - y2));
    for (int k = 0; k < K && k < 25; ++k) out[k] = v[k];
}

/* Float evaluation with the literal constants, in the operation order of the
 * GLSL source (used by the sh_sum.comp restatement). */
static void sh16_f(const float d[3], float v[16])
{
    const float x = d[0], y = d[1], z = d[2];
    v[0] = 0.282095f;
    v[1] = -0.488603f * y;
    v[2] = 0.488603f * z;
    v[3] = -0.488603f * x;
    v[4] = 1.092548f * x * y;
    v[5] = -1.092548f * y * z;
    v[6] = 0.315392f * (-x * x - y * y + 2.0f * z * z);
    v[7] = -1.092548f * x * z;
    v[8] = 0.546274f * (x * x - y * y);
    v[9] = -0.590044f * y * (3.0f * x * x - y * y);
    v[10] = 2.890611f * x * y * z;
    v[11] = -0.457046f * y * (4.0f * z * z - x * x - y * y);
    v[12] = 0.373176f * z * (2.0f * z * z - 3.0f * x * x - 3.0f * y * y);
    v[13] = -0.457046f * x * (4.0f * z * z - x * x - y * y);
    v[14] = 1.445306f * z * (x * x - y * y);
    v[15] = -0.590044f * x * (x * x - 3.0f * y * y);
}

/* ------------------------------------------------------------ directions */
void orc_fibonacci(uint32_t R, float* xyz)
{
    const double ga = ORC_PI_D * (3.0 - sqrt(5.0));
    for (uint32_t i = 0; i < R; ++i) {
        double y = 1.0 - (2.0 * i + 1.0) / (double)R;
        double r = sqrt(fmax(0.0, 1.0 - y * y));
        double phi = fmod((double)i * ga, 2.0 * ORC_PI_D);
        xyz[3 * i + 0] = (float)(r * cos(phi));
        xyz[3 * i + 1] = (float)y;
        xyz[3 * i + 2] = (float)(r * sin(phi));
    }
}

/* sh_common.h:8-22 in float */
static void to_vector_f(float phi, float theta, float v[3])
{
    float r = sinf(theta);
    float a = r * cosf(phi), b = r * sinf(phi), c = cosf(theta);
    float l = sqrtf(a * a + b * b + c * c);
    v[0] = a / l;
    v[1] = b / l;
    v[2] = c / l;
}
static float x2phi_f(int x, int w) { return 2.f * ORC_PI_F * ((float)x + 0.5f) / (float)w; }
static float y2theta_f(int y, int h) { return ORC_PI_F * ((float)y + 0.5f) / (float)h; }

void orc_equirect_dirs(uint32_t W, uint32_t H, float* xyz, float* weight)
{
    const float area = (2.0f * ORC_PI_F / (float)W) * (ORC_PI_F / (float)H);
    for (uint32_t y = 0; y < H; ++y) {
        float theta = y2theta_f((int)y, (int)H);
        float wgt = area * sinf(theta);
        for (uint32_t x = 0; x < W; ++x) {
            float v[3];
            to_vector_f(x2phi_f((int)x, (int)W), theta, v);
            size_t r = (size_t)y * W + x;
            xyz[3 * r + 0] = v[0]; /* .xzy swizzle of env_map.rgen:21 */
            xyz[3 * r + 1] = v[2];
            xyz[3 * r + 2] = v[1];
            if (weight) weight[r] = wgt;
        }
    }
}

/* ------------------------------------------------------------------ grid */
void orc_probe_grid(const float b0[3], const float b1[3], const uint32_t c[3], int order,
                    float* out, float* step)
{
    float s[3];
    for (int a = 0; a < 3; ++a) s[a] = (b1[a] - b0[a]) / ((float)c[a] - 1.0f);
    if (step) memcpy(step, s, sizeof s);
    size_t P = (size_t)c[0] * c[1] * c[2];
    if (order == VLB_ORDER_REFERENCE_BFS) {
        size_t n = 0;
        memcpy(&out[0], b0, 3 * sizeof(float));
        n = 1;
        for (int dim = 0; dim < 3; ++dim) {
            size_t ncopy = n;
            for (size_t q = 0; q < ncopy; ++q) {
                float p[3] = {out[3 * q], out[3 * q + 1], out[3 * q + 2]};
                for (int i = 0; (float)i < (float)c[dim] - 1.0f; ++i) {
                    p[dim] += s[dim];
                    memcpy(&out[3 * n], p, sizeof p);
                    ++n;
                }
            }
        }
        (void)P;
    } else {
        size_t i = 0;
        for (uint32_t z = 0; z < c[2]; ++z)
            for (uint32_t y = 0; y < c[1]; ++y)
                for (uint32_t x = 0; x < c[0]; ++x, ++i) {
                    out[3 * i + 0] = b0[0] + (float)x * s[0];
                    out[3 * i + 1] = b0[1] + (float)y * s[1];
                    out[3 * i + 2] = b0[2] + (float)z * s[2];
                }
    }
}

/* ------------------------------------------------------------- colour */
float orc_srgb(float x)
{
    /* env_map.rchit:27-34: mix(higher, lower, x < 0.0031308) */
    if (x < 0.0031308f) return x * 12.92f;
    return 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

float orc_unorm8(float x)
{
    /* Vulkan float -> UNORM8 store (round to nearest) and UNORM8 -> float load */
    float c = (x > 0.0f) ? (x < 1.0f ? x : 1.0f) : 0.0f;
    float q = rintf(c * 255.0f);
    return q / 255.0f;
}

/* -------------------------------------------------------------- skybox */
static double area_elem(double x, double y) { return atan2(x * y, sqrt(x * x + y * y + 1.0)); }

static void cube_dir(int face, double u, double v, double d[3])
{
    switch (face) {
    case 0: d[0] = 1.0; d[1] = -v; d[2] = -u; break;
    case 1: d[0] = -1.0; d[1] = -v; d[2] = u; break;
    case 2: d[0] = u; d[1] = 1.0; d[2] = v; break;
    case 3: d[0] = u; d[1] = -1.0; d[2] = -v; break;
    case 4: d[0] = u; d[1] = -v; d[2] = 1.0; break;
    default: d[0] = -u; d[1] = -v; d[2] = -1.0; break;
    }
    double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    d[0] /= l;
    d[1] /= l;
    d[2] /= l;
}

void orc_skybox_sh(const float* cube, uint32_t N, int K, double* out)
{
    memset(out, 0, sizeof(double) * 3 * (size_t)K);
    double Y[25];
    for (int f = 0; f < 6; ++f)
        for (uint32_t j = 0; j < N; ++j)
            for (uint32_t i = 0; i < N; ++i) {
                double x0 = 2.0 * i / N - 1.0, x1 = 2.0 * (i + 1) / N - 1.0;
                double y0 = 2.0 * j / N - 1.0, y1 = 2.0 * (j + 1) / N - 1.0;
                double w = area_elem(x0, y0) - area_elem(x0, y1) - area_elem(x1, y0) + area_elem(x1, y1);
                double u = 2.0 * (i + 0.5) / N - 1.0, v = 2.0 * (j + 0.5) / N - 1.0;
                double d[3];
                cube_dir(f, u, v, d);
                double sh[3] = {d[0], d[2], d[1]};
                orc_sh_eval(K, 1, sh, Y);
                const float* t = &cube[(((size_t)f * N + j) * N + i) * 4];
                for (int k = 0; k < K; ++k)
                    for (int c = 0; c < 3; ++c) out[3 * k + c] += Y[k] * w * (double)t[c];
            }
}

void orc_equirect_project(const uint8_t* img, uint32_t W, uint32_t H, int variant,
                          double* out)
{
    memset(out, 0, sizeof(double) * 48);
    const double pi = ORC_PI_D;
    const double area = (2.0 * pi / W) * (pi / H);
    double Y[16];
    for (uint32_t y = 0; y < H; ++y) {
        double theta = pi * (y + 0.5) / H;
        double w = area * sin(theta);
        for (uint32_t x = 0; x < W; ++x) {
            double phi = 2.0 * pi * (x + 0.5) / W;
            if (variant == VLB_EQUIRECT_SKYBOX_COMP) phi -= pi / 2.0;
            double d[3] = {sin(theta) * cos(phi), sin(theta) * sin(phi), cos(theta)};
            double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            d[0] /= l;
            d[1] /= l;
            d[2] /= l;
            if (variant == VLB_EQUIRECT_SKYBOX_COMP) {
                double s[3] = {d[0], d[2], d[1]}; /* SH(l, m, dir.xzy) */
                orc_sh_eval(16, 0, s, Y);
            } else {
                orc_sh_eval(16, 0, d, Y);
            }
            const uint8_t* px = &img[((size_t)y * W + x) * 4];
            for (int k = 0; k < 16; ++k)
                for (int c = 0; c < 3; ++c) out[3 * k + c] += Y[k] * (px[c] / 255.0) * w;
        }
    }
}

void orc_sh_reconstruct(const float* coeffs, uint32_t W, uint32_t H, uint8_t* out)
{
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            float d[3], v[16];
            to_vector_f(x2phi_f((int)x, (int)W), y2theta_f((int)y, (int)H), d);
            sh16_f(d, v);
            float col[3] = {0.f, 0.f, 0.f};
            for (int k = 0; k < 16; ++k)
                for (int c = 0; c < 3; ++c) col[c] = col[c] + coeffs[3 * k + c] * v[k];
            uint8_t* px = &out[((size_t)y * W + x) * 4];
            for (int c = 0; c < 3; ++c) {
                float s = 1250.0f * col[c];
                float cl = (s > 0.0f) ? (s < 1.0f ? s : 1.0f) : 0.0f;
                px[c] = (uint8_t)rintf(cl * 255.0f);
            }
            px[3] = 255;
        }
}

/* -------------------------------------------------------------- base64 */
size_t orc_base64(const uint8_t* b, size_t n, char* out)
{
    static const char tbl[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    size_t o = 0, i = 0;
    for (; i + 3 <= n; i += 3) {
        uint32_t v = ((uint32_t)b[i] << 16) | ((uint32_t)b[i + 1] << 8) | b[i + 2];
        out[o++] = tbl[(v >> 18) & 63];
        out[o++] = tbl[(v >> 12) & 63];
        out[o++] = tbl[(v >> 6) & 63];
        out[o++] = tbl[v & 63];
    }
    size_t rem = n - i;
    if (rem) {
        uint32_t v = (uint32_t)b[i] << 16;
        if (rem == 2) v |= (uint32_t)b[i + 1] << 8;
        out[o++] = tbl[(v >> 18) & 63];
        out[o++] = tbl[(v >> 12) & 63];
        out[o++] = rem == 2 ? tbl[(v >> 6) & 63] : '=';
        out[o++] = '=';
    }
    out[o] = 0;
    return o;
}
