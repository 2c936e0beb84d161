// vct_texel.hpp - texel encode/decode shared by the gfx950 kernels.
//
// UNORM8 follows the Vulkan conversion rules: decode b / 255 (IEEE division),
// encode clamp to [0,1] (NaN -> 0), scale by 255, round to nearest even.
// Half floats use the hardware round-to-nearest-even conversion.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vct {

__device__ __forceinline__ float u8_to_f(uint32_t b)
{
    return (float)b / 255.0f;
}

__device__ __forceinline__ uint32_t f_to_u8(float f)
{
    float c = f > 0.0f ? (f < 1.0f ? f : 1.0f) : 0.0f;
    return (uint32_t)__builtin_rintf(c * 255.0f);
}

struct f4 {
    float x, y, z, w;
};

__device__ __forceinline__ f4 unpack_unorm(uint32_t t)
{
    return {u8_to_f(t & 0xffu), u8_to_f((t >> 8) & 0xffu), u8_to_f((t >> 16) & 0xffu), u8_to_f(t >> 24)};
}

__device__ __forceinline__ uint32_t pack_unorm(f4 v)
{
    return f_to_u8(v.x) | (f_to_u8(v.y) << 8) | (f_to_u8(v.z) << 16) | (f_to_u8(v.w) << 24);
}

__device__ __forceinline__ f4 unpack_half(uint2 t)
{
    _Float16 h[4];
    __builtin_memcpy(h, &t, 8);
    return {(float)h[0], (float)h[1], (float)h[2], (float)h[3]};
}

__device__ __forceinline__ uint2 pack_half(f4 v)
{
    _Float16 h[4] = {(_Float16)v.x, (_Float16)v.y, (_Float16)v.z, (_Float16)v.w};
    uint2 t;
    __builtin_memcpy(&t, h, 8);
    return t;
}

// atomic.glsl:4-22 convRGBA8ToVec4 / convVec4ToRGBA8 (raw byte values)
__device__ __forceinline__ f4 unpack_raw(uint32_t v)
{
    return {(float)(v & 0xffu), (float)((v >> 8) & 0xffu), (float)((v >> 16) & 0xffu), (float)(v >> 24)};
}

__device__ __forceinline__ uint32_t pack_raw(f4 v)
{
    return (((uint32_t)v.w & 0xffu) << 24) | (((uint32_t)v.z & 0xffu) << 16) | (((uint32_t)v.y & 0xffu) << 8) |
           ((uint32_t)v.x & 0xffu);
}

// One uncontended pass of the RGBA8 compare-and-swap running average
// (msaaInjectRadiance.frag:185-201, octreeNodeLeafWrite.comp:73-89).
__device__ __forceinline__ uint32_t rgba8_avg(uint32_t stored, f4 value)
{
#pragma clang fp contract(off)
    if (stored == 0u) return pack_raw(value);
    f4 r = unpack_raw(stored);
    r.x = r.x * r.w;
    r.y = r.y * r.w;
    r.z = r.z * r.w;
    f4 c = {r.x + value.x, r.y + value.y, r.z + value.z, r.w + value.w};
    c.x = c.x / c.w;
    c.y = c.y / c.w;
    c.z = c.z / c.w;
    return pack_raw(c);
}

}  // namespace vct
