// sh_eval.hpp -- real SH basis (l <= 3) in float on the device: the exact
// normalisations and the literal 6-figure constants of shaders/sh_common.h.
#pragma once

#include <hip/hip_runtime.h>

namespace vlb {

// Real SH with exact normalisation, l <= 3, SH frame argument (x, y, z).
template <int K>
__device__ __forceinline__ void sh_exact(float x, float y, float z, float* Y)
{
    Y[0] = 0.28209479177387814f;
    Y[1] = -0.48860251190291992f * y;
    Y[2] = 0.48860251190291992f * z;
    Y[3] = -0.48860251190291992f * x;
    Y[4] = 1.0925484305920792f * x * y;
    Y[5] = -1.0925484305920792f * y * z;
    Y[6] = 0.31539156525252005f * (2.0f * z * z - x * x - y * y);
    Y[7] = -1.0925484305920792f * x * z;
    Y[8] = 0.54627421529603959f * (x * x - y * y);
    if (K > 9) {
        Y[9] = -0.59004358992664352f * y * (3.0f * x * x - y * y);
        Y[10] = 2.8906114426405538f * x * y * z;
        Y[11] = -0.45704579946446572f * y * (4.0f * z * z - x * x - y * y);
        Y[12] = 0.37317633259011540f * z * (2.0f * z * z - 3.0f * x * x - 3.0f * y * y);
        Y[13] = -0.45704579946446572f * x * (4.0f * z * z - x * x - y * y);
        Y[14] = 1.4453057213202769f * z * (x * x - y * y);
        Y[15] = -0.59004358992664352f * x * (x * x - 3.0f * y * y);
    }
}

// The literal 6-figure constants of shaders/sh_common.h:26-116, l <= 3, in the
// GLSL operation order.
__device__ __forceinline__ void sh_literal16(float x, float y, float z, float* v)
{
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

}  // namespace vlb
