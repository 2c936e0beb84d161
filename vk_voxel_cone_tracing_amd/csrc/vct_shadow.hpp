// vct_shadow.hpp - shadow-map visibility (shadow.glsl:8-36) shared by the
// radiance injector and the SVO fragment voxelizer. FP32 without contraction
// in the reference's operation order, so its result is reproducible bit for
// bit.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/vct.h"

namespace vct {

struct LightParams {
    float dir[3];       // normalize(-direction)
    float color[3];
    float intensity;
    float view[16];
    float proj[16];
    float z_near, z_far;
    const float* depth;
    int width, height;
};

__device__ __forceinline__ float fetch_depth(const LightParams& L, int i, int j)
{
    if (i < 0 || j < 0 || i >= L.width || j >= L.height) return 0.0f;  // CLAMP_TO_BORDER, opaque black
    return L.depth[(size_t)j * (size_t)L.width + (size_t)i];
}

__device__ __forceinline__ float shadow_tap(const LightParams& L, float s, float t, bool compare, float ref)
{
#pragma clang fp contract(off)
    float uu = s * (float)L.width - 0.5f;
    float vv = t * (float)L.height - 0.5f;
    float fu = floorf(uu), fv = floorf(vv);
    float a = uu - fu, b = vv - fv;
    fu = fminf(fmaxf(fu, -4.0f), (float)L.width + 4.0f);
    fv = fminf(fmaxf(fv, -4.0f), (float)L.height + 4.0f);
    int i0 = (int)fu, j0 = (int)fv;
    float t00 = fetch_depth(L, i0, j0), t10 = fetch_depth(L, i0 + 1, j0);
    float t01 = fetch_depth(L, i0, j0 + 1), t11 = fetch_depth(L, i0 + 1, j0 + 1);
    if (compare) {
        bool in00 = i0 >= 0 && j0 >= 0 && i0 < L.width && j0 < L.height;
        bool in10 = i0 + 1 >= 0 && j0 >= 0 && i0 + 1 < L.width && j0 < L.height;
        bool in01 = i0 >= 0 && j0 + 1 >= 0 && i0 < L.width && j0 + 1 < L.height;
        bool in11 = i0 + 1 >= 0 && j0 + 1 >= 0 && i0 + 1 < L.width && j0 + 1 < L.height;
        t00 = (!in00 || ref <= t00) ? 1.0f : 0.0f;
        t10 = (!in10 || ref <= t10) ? 1.0f : 0.0f;
        t01 = (!in01 || ref <= t01) ? 1.0f : 0.0f;
        t11 = (!in11 || ref <= t11) ? 1.0f : 0.0f;
    }
    float r = ((1.0f - a) * (1.0f - b)) * t00;
    r = r + (a * (1.0f - b)) * t10;
    r = r + ((1.0f - a) * b) * t01;
    r = r + (a * b) * t11;
    return r;
}

// shadow.glsl:28-36 calcVisibility + PCF16 (:14-26)
__device__ __forceinline__ float visibility(const LightParams& L, const float pos[3], uint32_t quirks)
{
#pragma clang fp contract(off)
    const float* V = L.view;
    const float* Pm = L.proj;
    float ls[3];
    for (int k = 0; k < 3; ++k) ls[k] = V[0 + k] * pos[0] + V[4 + k] * pos[1] + V[8 + k] * pos[2] + V[12 + k];
    const bool compare = !(quirks & VCT_QUIRK_SHADOW_RAW_DEPTH);
    float ref = 0.0f;
    if (compare) {
        ref = Pm[2] * ls[0] + Pm[6] * ls[1] + Pm[10] * ls[2] + Pm[14];
        ref = ref - 0.001f;
    }
    ls[2] = ls[2] / (L.z_far - L.z_near);
    float xy[2];
    for (int k = 0; k < 2; ++k) xy[k] = Pm[0 + k] * ls[0] + Pm[4 + k] * ls[1] + Pm[12 + k];
    xy[0] = xy[0] * 0.5f + 0.5f;
    xy[1] = xy[1] * 0.5f + 0.5f;
    const float sx = 1.0f / (float)L.width;
    const float sy = 1.0f / (float)L.height;
    // offsets -1.5, -0.5, 0.5, 1.5 (exact in float), y outer as in the
    // reference; integer counters let the 16 taps unroll so their fetches
    // are in flight together
    float sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float y = -1.5f + (float)j;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float x = -1.5f + (float)i;
            sum = sum + shadow_tap(L, xy[0] + x * sx, xy[1] + y * sy, compare, ref);
        }
    }
    return sum * 0.0625f;
}


}  // namespace vct
