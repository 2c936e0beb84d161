// vct_textures.hpp - scene texture sampling of the voxel passes.
//
// The reference's texture samplers (GLTFScene.cpp:339 -> Sampler.cpp:29-50)
// use REPEAT addressing, LINEAR filtering and maxLod 0, so both texture()
// (msaaVoxelizer.frag:64) and textureLod() (msaaInjectRadiance.frag:80,134,
// voxelizer.frag:66,76) reduce to a bilinear sample of level 0:
//   x = u * W - 0.5, i0 = floor(x) mod W, i1 = (i0 + 1) mod W, a = x - floor(x)
// (likewise y), with texels decoded as UNORM8 (b / 255). The weights are
// applied in FP32 without contraction in a fixed order, so the result is
// reproducible bit for bit; hardware filtering would quantize the weights.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "vct_texel.hpp"

namespace vct {

struct DevTexture {
    const uint32_t* texels;
    uint32_t width, height;
};

// Texture coordinates, materials and textures of the soup for the texture
// lookups of the fragment passes; tex is null (count 0) when the mesh has no
// texture table.
struct TexInputs {
    const float* uv;
    const uint32_t* mat;
    const vct_gltf_material* mats;
    const DevTexture* tex;
    uint32_t count;
};

// Whether a material's texture index names a texture of the table (the host
// validates indices; this keeps a stale or unchecked one from being read).
__device__ __forceinline__ bool has_texture(const TexInputs& T, int32_t idx)
{
    return idx > -1 && (uint32_t)idx < T.count;
}

// Texture coordinate of a fragment from the triangle's three vertex texcoords
// and the fragment's barycentrics (interpolated in double, rounded once).
__device__ __forceinline__ void interp_uv(const float* __restrict__ uv6, const double lam[3], float uv[2])
{
#pragma clang fp contract(off)
    for (int c = 0; c < 2; ++c)
        uv[c] = (float)(lam[0] * (double)uv6[c] + lam[1] * (double)uv6[2 + c] + lam[2] * (double)uv6[4 + c]);
}

__device__ __forceinline__ int wrap_texel(float f, uint32_t n)
{
    // f is an integral float; keep it in int range, then a non-negative modulo
    const float c = fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
    const int m = (int)c % (int)n;
    return m < 0 ? m + (int)n : m;
}

__device__ __forceinline__ f4 sample_texture(const DevTexture& T, const float uv[2])
{
#pragma clang fp contract(off)
    const float x = uv[0] * (float)T.width - 0.5f;
    const float y = uv[1] * (float)T.height - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int i0 = wrap_texel(fx, T.width), i1 = i0 + 1 == (int)T.width ? 0 : i0 + 1;
    const int j0 = wrap_texel(fy, T.height), j1 = j0 + 1 == (int)T.height ? 0 : j0 + 1;
    const f4 t00 = unpack_unorm(T.texels[(size_t)j0 * T.width + (size_t)i0]);
    const f4 t10 = unpack_unorm(T.texels[(size_t)j0 * T.width + (size_t)i1]);
    const f4 t01 = unpack_unorm(T.texels[(size_t)j1 * T.width + (size_t)i0]);
    const f4 t11 = unpack_unorm(T.texels[(size_t)j1 * T.width + (size_t)i1]);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    f4 r;
    r.x = ((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x;
    r.y = ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y;
    r.z = ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z;
    r.w = ((w00 * t00.w + w10 * t10.w) + w01 * t01.w) + w11 * t11.w;
    return r;
}

// msaaVoxelizer.frag:64 (and msaaInjectRadiance.frag:73, voxelizer.frag:52):
// a fragment whose occlusion texture reads below 0.1 in red is discarded.
__device__ __forceinline__ bool occluded(const TexInputs& T, const vct_gltf_material& m, const float uv[2])
{
    return has_texture(T, m.occlusion_texture) && sample_texture(T.tex[m.occlusion_texture], uv).x < 0.1f;
}

}  // namespace vct
