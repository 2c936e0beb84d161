// vct_dense_mip.hpp - one destination texel of the dense grid's anisotropic
// mip (radianceDownSample.comp:95-139 without the clipmap transition band),
// shared by vct_dense_mipmap and the batched launch so both run the same
// arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vct_internal.hpp"
#include "vct_texel.hpp"

namespace vct {

template <bool kF32>
__device__ __forceinline__ f4 dense_load(const void* base, size_t i)
{
    if (kF32) {
        const float4 t = ((const float4*)base)[i];
        return {t.x, t.y, t.z, t.w};
    }
    return unpack_unorm(((const uint32_t*)base)[i]);
}

template <bool kF32>
__device__ __forceinline__ void dense_store(void* base, size_t i, f4 v)
{
    if (kF32) ((float4*)base)[i] = make_float4(v.x, v.y, v.z, v.w);
    else ((uint32_t*)base)[i] = pack_unorm(v);
}

// The face-f value of a parent from its 8 children: the mean of 4 front-to-back
// pairs a + (1 - a.alpha) * b, accumulated in pair order.
__device__ __forceinline__ f4 dense_mip_face(const f4 v[8], int f)  // f must unroll to a constant
{
#pragma clang fp contract(off)
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f4 a = v[face_pair_child(f, p, 0)];
        const f4 b = v[face_pair_child(f, p, 1)];
        const float t = 1.0f - a.w;
        acc[0] = acc[0] + a.x;
        acc[0] = acc[0] + t * b.x;
        acc[1] = acc[1] + a.y;
        acc[1] = acc[1] + t * b.y;
        acc[2] = acc[2] + a.z;
        acc[2] = acc[2] + t * b.z;
        acc[3] = acc[3] + a.w;
        acc[3] = acc[3] + t * b.w;
    }
    return f4{acc[0] * 0.25f, acc[1] * 0.25f, acc[2] * 0.25f, acc[3] * 0.25f};
}

// Destination texel (x, y, z) of a level with n^3 interior texels from the
// level below (2n)^3, which has one face (level 0) or six. Both levels are
// bordered blocks of (n+2)^3 texels per face, x fastest.
template <bool kF32>
__device__ __forceinline__ void dense_mip_texel(const void* src, void* dst, int n, int x, int y, int z,
                                                bool src_single_face)
{
    const size_t Ds = (size_t)(2 * n) + 2, Dd = (size_t)n + 2;
    const size_t sblock = Ds * Ds * Ds, dblock = Dd * Dd * Dd;
#pragma unroll
    for (int f = 0; f < kFaces; ++f) {
        const size_t sface = src_single_face ? 0 : (size_t)f * sblock;
        f4 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const size_t si = sface + ((size_t)(2 * z + (i >> 2) + 1) * Ds + (size_t)(2 * y + ((i >> 1) & 1) + 1)) * Ds +
                              (size_t)(2 * x + (i & 1) + 1);
            v[i] = dense_load<kF32>(src, si);
        }
        const size_t di = (size_t)f * dblock + ((size_t)(z + 1) * Dd + (size_t)(y + 1)) * Dd + (size_t)(x + 1);
        dense_store<kF32>(dst, di, dense_mip_face(v, f));
    }
}

// dense_mip_texel's arithmetic with one face at a time (the face loop is not
// unrolled and each pair's children are loaded by computed index), for
// kernels that must keep few registers live: the persistent batched launch,
// whose pixel tiles would otherwise inherit the unrolled version's registers.
template <bool kF32>
__device__ __forceinline__ void dense_mip_texel_lean(const void* src, void* dst, int n, int x, int y, int z,
                                                     bool src_single_face)
{
#pragma clang fp contract(off)
    const size_t Ds = (size_t)(2 * n) + 2, Dd = (size_t)n + 2;
    const size_t sblock = Ds * Ds * Ds, dblock = Dd * Dd * Dd;
    auto child = [&](size_t sface, int i) {
        return sface + ((size_t)(2 * z + (i >> 2) + 1) * Ds + (size_t)(2 * y + ((i >> 1) & 1) + 1)) * Ds +
               (size_t)(2 * x + (i & 1) + 1);
    };
#pragma unroll 1
    for (int f = 0; f < kFaces; ++f) {
        const size_t sface = src_single_face ? 0 : (size_t)f * sblock;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const f4 a = dense_load<kF32>(src, child(sface, face_pair_child(f, p, 0)));
            const f4 b = dense_load<kF32>(src, child(sface, face_pair_child(f, p, 1)));
            const float t = 1.0f - a.w;
            acc[0] = acc[0] + a.x;
            acc[0] = acc[0] + t * b.x;
            acc[1] = acc[1] + a.y;
            acc[1] = acc[1] + t * b.y;
            acc[2] = acc[2] + a.z;
            acc[2] = acc[2] + t * b.z;
            acc[3] = acc[3] + a.w;
            acc[3] = acc[3] + t * b.w;
        }
        const size_t di = (size_t)f * dblock + ((size_t)(z + 1) * Dd + (size_t)(y + 1)) * Dd + (size_t)(x + 1);
        dense_store<kF32>(dst, di, f4{acc[0] * 0.25f, acc[1] * 0.25f, acc[2] * 0.25f, acc[3] * 0.25f});
    }
}

}  // namespace vct
