// vct_trace.hip - per-pixel voxel cone tracing (K8) for gfx950.
//
// One lane per pixel; a 64-lane wavefront covers an 8x8 pixel tile so that
// neighbouring lanes march nearby cones and share cache lines. Texture
// filtering is done in registers: the sampler of the reference
// (voxelConeTracing.frag:313-327, REPEAT + LINEAR over the atlas) reduces to
// a trilinear fetch inside one (level, face) block at the block coordinate
// u = fract(p / extent) * res + 0.5, whose 2x2x2 footprint never leaves the
// block's [0, res+1] range. The three faces selected by the cone direction
// share the footprint offsets, so one address computation serves 24 loads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "vct_internal.hpp"
#include "vct_texel.hpp"

namespace vct {

__constant__ float c_dirs16[16][3] = {
    {0.57735f, 0.57735f, 0.57735f},       {0.57735f, -0.57735f, -0.57735f},
    {-0.57735f, 0.57735f, -0.57735f},     {-0.57735f, -0.57735f, 0.57735f},
    {-0.903007f, -0.182696f, -0.388844f}, {-0.903007f, 0.182696f, 0.388844f},
    {0.903007f, -0.182696f, 0.388844f},   {0.903007f, 0.182696f, -0.388844f},
    {-0.388844f, -0.903007f, -0.182696f}, {0.388844f, -0.903007f, 0.182696f},
    {0.388844f, 0.903007f, -0.182696f},   {-0.388844f, 0.903007f, 0.182696f},
    {-0.182696f, -0.388844f, -0.903007f}, {0.182696f, 0.388844f, -0.903007f},
    {-0.182696f, 0.388844f, 0.903007f},   {0.182696f, -0.388844f, 0.903007f}};

__constant__ float c_dirs32[32][3] = {
    {0.898904f, 0.435512f, 0.0479745f},   {0.898904f, -0.435512f, -0.0479745f},
    {0.898904f, 0.0479745f, -0.435512f},  {0.898904f, -0.0479745f, 0.435512f},
    {-0.898904f, 0.435512f, -0.0479745f}, {-0.898904f, -0.435512f, 0.0479745f},
    {-0.898904f, 0.0479745f, 0.435512f},  {-0.898904f, -0.0479745f, -0.435512f},
    {0.0479745f, 0.898904f, 0.435512f},   {-0.0479745f, 0.898904f, -0.435512f},
    {-0.435512f, 0.898904f, 0.0479745f},  {0.435512f, 0.898904f, -0.0479745f},
    {-0.0479745f, -0.898904f, 0.435512f}, {0.0479745f, -0.898904f, -0.435512f},
    {0.435512f, -0.898904f, 0.0479745f},  {-0.435512f, -0.898904f, -0.0479745f},
    {0.435512f, 0.0479745f, 0.898904f},   {-0.435512f, -0.0479745f, 0.898904f},
    {0.0479745f, -0.435512f, 0.898904f},  {-0.0479745f, 0.435512f, 0.898904f},
    {0.435512f, -0.0479745f, -0.898904f}, {-0.435512f, 0.0479745f, -0.898904f},
    {0.0479745f, 0.435512f, -0.898904f},  {-0.0479745f, -0.435512f, -0.898904f},
    {0.57735f, 0.57735f, 0.57735f},       {0.57735f, 0.57735f, -0.57735f},
    {0.57735f, -0.57735f, 0.57735f},      {0.57735f, -0.57735f, -0.57735f},
    {-0.57735f, 0.57735f, 0.57735f},      {-0.57735f, 0.57735f, -0.57735f},
    {-0.57735f, -0.57735f, 0.57735f},     {-0.57735f, -0.57735f, -0.57735f}};

struct TraceParams {
    float center[3];
    uint32_t mode;
    float voxel_size;
    float dim;
    float start_offset;
    float diffuse_intensity;
    float ao_factor;
    float step_factor;  // max(0.2, minTraceStepFactor)
    float specular_intensity;
    float decay;
    int cones32;
    int levels;
    int res;
    int D;
    float vp_inv[16];
    float eye[3];
    int width, height;
    uint32_t quirks;
    // light
    float light_dir[3];  // normalize(-direction)
    float light_view[16];
    float light_proj[16];
    float z_near, z_far;
    const float* shadow;
    int sw, sh;
    // reciprocals of the march's constant divisors
    float inv_extent0;     // 1 / (voxel_size * dim): level-0 clip extent
    float inv_voxel_size;  // 1 / voxel_size
    float inv_min_radius;  // 1 / (voxel_size * dim / 2)
    uint32_t vol_bytes;    // size of the radiance volume (buffer-load range)
    // occupancy of the radiance volume: bit (z*D + y)*D + x of level m's
    // occ_words words is set when the 2x2x2 footprint with that corner holds
    // a nonzero texel in any face
    const uint32_t* occ;
    uint32_t occ_words;
};

template <bool kHalf>
struct Vol;

template <>
struct Vol<false> {
    using T = uint32_t;
    __device__ static __forceinline__ f4 load(const T* __restrict__ v, size_t i)
    {
        const uint32_t t = v[i];
        const float s = 1.0f / 255.0f;
        return {(float)(t & 0xffu) * s, (float)((t >> 8) & 0xffu) * s, (float)((t >> 16) & 0xffu) * s,
                (float)(t >> 24) * s};
    }
};

template <>
struct Vol<true> {
    using T = uint2;
    __device__ static __forceinline__ f4 load(const T* __restrict__ v, size_t i) { return unpack_half(v[i]); }
};

typedef unsigned int vct_u4 __attribute__((ext_vector_type(4)));
typedef unsigned int vct_u2 __attribute__((ext_vector_type(2)));

// The trilinear footprint of `pos` at `level`: the block coordinate
// u = fract(p / extent) * res + 0.5 per axis, whose floor is the corner texel
// (clamped to [0, res], so the 2x2x2 footprint stays inside the block) and
// whose fraction is the weight. Returns the corner's texel index within a
// (level, face) block.
__device__ __forceinline__ uint32_t footprint(const TraceParams& P, const float pos[3], int level, float fr[3])
{
    const float inv_extent = ldexpf(P.inv_extent0, -level);  // 1 / (voxel_size * 2^level * dim), exact
    int i0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float q = pos[k] * inv_extent;
        const float u = (q - floorf(q)) * (float)P.res + 0.5f;
        const float fu = floorf(u);
        fr[k] = u - fu;
        // no clamp needed: q - floor(q) lies in [0, 1] up to one rounding of
        // the product (the fused form can leave it a hair below 0 or above
        // 1), so u is within [0.5, res + 0.5] but for res x that hair, and
        // floor(u) within [0, res]; the oracle's clamp to [0, res] never acts
        i0[k] = (int)fu;
    }
    const uint32_t D = (uint32_t)P.D;
    return ((uint32_t)i0[2] * D + (uint32_t)i0[1]) * D + (uint32_t)i0[0];
}

// Whether the footprint with corner o000 of `level` holds a nonzero texel.
__device__ __forceinline__ bool occupied(const TraceParams& P, int level, uint32_t o000)
{
    // a 32-bit word index: the mask is levels x occ_words words, far below 2^32
    return (P.occ[(uint32_t)level * P.occ_words + (o000 >> 5)] >> (o000 & 31u)) & 1u;
}

// Trilinear sample of the three selected faces of one level, weighted.
// Texels are fetched with buffer loads: a 32-bit byte offset per lane (the
// volume is far below 4 GiB) and the y / z / yz neighbour strides as scalar
// offsets, so the 2x2x2 footprint of a face costs one offset computation and
// four loads of a texel pair (x, x + 1) each. A fetch past the end of the
// volume reads zero instead of faulting.
template <bool kHalf>
__device__ __forceinline__ f4 sample_level(const typename Vol<kHalf>::T* __restrict__ vol, const TraceParams& P,
                                           int level, uint32_t o000, const float fr[3], const int face[3],
                                           const float w[3])
{
    constexpr uint32_t TB = kHalf ? 8u : 4u;  // bytes per texel
    const __amdgpu_buffer_rsrc_t R =
        __builtin_amdgcn_make_buffer_rsrc((void*)vol, (short)0, (int)P.vol_bytes, 0x00020000);
    const uint32_t D = (uint32_t)P.D;
    const int sy = (int)(D * TB), sz = (int)(D * D * TB), syz = sy + sz;
    const float ax = fr[0], ay = fr[1], az = fr[2];
    const float w000 = (1.0f - ax) * (1.0f - ay) * (1.0f - az), w100 = ax * (1.0f - ay) * (1.0f - az);
    const float w010 = (1.0f - ax) * ay * (1.0f - az), w110 = ax * ay * (1.0f - az);
    const float w001 = (1.0f - ax) * (1.0f - ay) * az, w101 = ax * (1.0f - ay) * az;
    const float w011 = (1.0f - ax) * ay * az, w111 = ax * ay * az;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int voff = (int)(((((uint32_t)level * (uint32_t)kFaces + (uint32_t)face[a]) * D * D * D) + o000) * TB);
        f4 t000, t100, t010, t110, t001, t101, t011, t111;
        if (kHalf) {
            const vct_u4 q0 = __builtin_amdgcn_raw_buffer_load_b128(R, voff, 0, 0);
            const vct_u4 q1 = __builtin_amdgcn_raw_buffer_load_b128(R, voff, sy, 0);
            const vct_u4 q2 = __builtin_amdgcn_raw_buffer_load_b128(R, voff, sz, 0);
            const vct_u4 q3 = __builtin_amdgcn_raw_buffer_load_b128(R, voff, syz, 0);
            t000 = unpack_half(make_uint2(q0.x, q0.y));
            t100 = unpack_half(make_uint2(q0.z, q0.w));
            t010 = unpack_half(make_uint2(q1.x, q1.y));
            t110 = unpack_half(make_uint2(q1.z, q1.w));
            t001 = unpack_half(make_uint2(q2.x, q2.y));
            t101 = unpack_half(make_uint2(q2.z, q2.w));
            t011 = unpack_half(make_uint2(q3.x, q3.y));
            t111 = unpack_half(make_uint2(q3.z, q3.w));
        } else {
            // raw bytes; the 1/255 is applied once to the level's sum below
            auto u8 = [](uint32_t t) -> f4 {
                return {(float)(t & 0xffu), (float)((t >> 8) & 0xffu), (float)((t >> 16) & 0xffu), (float)(t >> 24)};
            };
            const vct_u2 q0 = __builtin_amdgcn_raw_buffer_load_b64(R, voff, 0, 0);
            const vct_u2 q1 = __builtin_amdgcn_raw_buffer_load_b64(R, voff, sy, 0);
            const vct_u2 q2 = __builtin_amdgcn_raw_buffer_load_b64(R, voff, sz, 0);
            const vct_u2 q3 = __builtin_amdgcn_raw_buffer_load_b64(R, voff, syz, 0);
            t000 = u8(q0.x);
            t100 = u8(q0.y);
            t010 = u8(q1.x);
            t110 = u8(q1.y);
            t001 = u8(q2.x);
            t101 = u8(q2.y);
            t011 = u8(q3.x);
            t111 = u8(q3.y);
        }
        const float wa = w[a];
#define VCT_TRI(c)                                                                                        \
    acc.c += wa * (w000 * t000.c + w100 * t100.c + w010 * t010.c + w110 * t110.c + w001 * t001.c +       \
                   w101 * t101.c + w011 * t011.c + w111 * t111.c)
        VCT_TRI(x);
        VCT_TRI(y);
        VCT_TRI(z);
        VCT_TRI(w);
#undef VCT_TRI
    }
    if (!kHalf) {
        const float sc = 1.0f / 255.0f;
        acc.x *= sc;
        acc.y *= sc;
        acc.z *= sc;
        acc.w *= sc;
    }
    return acc;
}

// The per-step level of traceCone (voxelConeTracing.frag:357-363):
// clamp(max(startLevel, log2(d / vs), ceil(log2(|p - c| / r0))), ., L - 1).
__device__ __forceinline__ float cone_level(const TraceParams& P, const float pos[3], float diameter,
                                            float start_level, float max_level)
{
    const float dx = P.center[0] - pos[0], dy = P.center[1] - pos[1], dz = P.center[2] - pos[2];
    const float dist = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
    // the hardware log2 (v_log_f32): log2f is the same instruction for every
    // argument but denormals, which it first scales by 2^32; neither argument
    // is ever one (0 gives -inf either way)
    const float min_level = ceilf(__builtin_amdgcn_logf(dist * P.inv_min_radius));
    const float cur = __builtin_amdgcn_logf(diameter * P.inv_voxel_size);
    return fminf(fmaxf(fmaxf(start_level, cur), min_level), max_level);
}

// sampleClipmapLinear (voxelConeTracing.frag:329-339). When no lane of the
// wave has a nonzero texel in its footprints, every lane's sample is exactly
// zero (a sum of zero texels), which the march accumulates as no change at
// all (opacity 1 - (1 - 0)^c = 0, colour + 0), so the fetches and the
// filtering are skipped and zero is returned.
template <bool kHalf>
__device__ __forceinline__ f4 cone_sample(const typename Vol<kHalf>::T* __restrict__ vol, const TraceParams& P,
                                          const float pos[3], float cur, const int face[3], const float w[3])
{
    const int lower = (int)floorf(cur);
    const int upper = (int)ceilf(cur);
    const float fl = cur - floorf(cur);
    float fr_lo[3], fr_hi[3];
    const uint32_t o_lo = footprint(P, pos, lower, fr_lo);
    bool occ = occupied(P, lower, o_lo);
    uint32_t o_hi = 0;
    if (upper != lower) {
        o_hi = footprint(P, pos, upper, fr_hi);
        occ = occ || occupied(P, upper, o_hi);
    }
    f4 s = {0.0f, 0.0f, 0.0f, 0.0f};
    if (!__any(occ)) return s;
    s = sample_level<kHalf>(vol, P, lower, o_lo, fr_lo, face, w);
    if (upper != lower) {
        const f4 hi = sample_level<kHalf>(vol, P, upper, o_hi, fr_hi, face, w);
        s.x = s.x * (1.0f - fl) + hi.x * fl;
        s.y = s.y * (1.0f - fl) + hi.y * fl;
        s.z = s.z * (1.0f - fl) + hi.z * fl;
        s.w = s.w * (1.0f - fl) + hi.w * fl;
    }
    return s;
}

// Front-to-back accumulation of one sample (voxelConeTracing.frag:370-386).
// pow(1 - a, c) is evaluated as exp2(c * log2(1 - a)) on the hardware
// transcendental units (a = 1 gives log2(0) = -inf and exp2(-inf) = 0).
__device__ __forceinline__ void cone_accumulate(const TraceParams& P, f4& result, float& occlusion, const f4& s,
                                                float cur, float step, float seg)
{
    // the hardware exp2 (v_exp_f32): exp2f is the same instruction for every
    // argument whose result is a normal number, and cur lies in [0, levels - 1]
    const float voxel_size = P.voxel_size * __builtin_amdgcn_exp2f(cur);
    const float correction = seg * (P.inv_voxel_size * __builtin_amdgcn_exp2f(-cur));
    // 1 - a clamped at 0: a filtered alpha of texels <= 1 can exceed 1 by an
    // ulp in FP32 (the reference's fixed-point filter weights sum to 1
    // exactly), and pow of a negative base is undefined
    float opacity = 1.0f - __builtin_amdgcn_exp2f(correction * __builtin_amdgcn_logf(fmaxf(1.0f - s.w, 0.0f)));
    opacity = fminf(fmaxf(opacity, 0.0f), 1.0f);
    const float blend = fminf(fmaxf(1.0f - result.w, 0.0f), 1.0f);
    result.x += blend * (s.x * correction);
    result.y += blend * (s.y * correction);
    result.z += blend * (s.z * correction);
    result.w += blend * opacity;
    occlusion += (1.0f - occlusion) * opacity * __builtin_amdgcn_rcpf(1.0f + (step + voxel_size) * P.decay);
}

// voxelConeTracing.frag:341-392 traceCone. The step sequence does not depend
// on the data (only the exit on full occlusion does), so the samples of two
// consecutive steps are fetched together and the second is accumulated only
// if the loop would have reached it: the arithmetic and its order are those
// of the one-step loop, with twice the loads in flight per lane.
template <bool kHalf>
__device__ f4 trace_cone(const typename Vol<kHalf>::T* __restrict__ vol, const TraceParams& P, const float start_in[3],
                         const float dir[3], float aperture, float max_dist, float start_level, float step_factor,
                         uint32_t& steps)
{
    f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
    const float cone_coef = 2.0f * tanf(aperture * 0.5f);
    const float voxel_size0 = P.voxel_size * exp2f(start_level);
    float start[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * voxel_size0 * P.start_offset * 0.5f;
    float step = 0.0f;
    float diameter = fmaxf(step * cone_coef, P.voxel_size);
    float occlusion = 0.0f;
    const int face[3] = {dir[0] > 0.0f ? 0 : 1, dir[1] > 0.0f ? 2 : 3, dir[2] > 0.0f ? 4 : 5};
    const float w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
    float seg = voxel_size0;
    const float max_level = (float)(P.levels - 1);
    while (step < max_dist && occlusion < 1.0f) {
        float pos_a[3], pos_b[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pos_a[k] = start[k] + dir[k] * step;
        const float cur_a = cone_level(P, pos_a, diameter, start_level, max_level);
        const float step_b = step + fmaxf(diameter, P.voxel_size) * step_factor;
        const float seg_b = step_b - step;
        const float diam_b = step_b * cone_coef;
#pragma unroll
        for (int k = 0; k < 3; ++k) pos_b[k] = start[k] + dir[k] * step_b;
        const float cur_b = cone_level(P, pos_b, diam_b, start_level, max_level);
        const f4 sa = cone_sample<kHalf>(vol, P, pos_a, cur_a, face, w);
        const f4 sb = cone_sample<kHalf>(vol, P, pos_b, cur_b, face, w);
        cone_accumulate(P, result, occlusion, sa, cur_a, step, seg);
        ++steps;
        step = step_b;
        seg = seg_b;
        diameter = diam_b;
        if (!(step < max_dist && occlusion < 1.0f)) break;
        cone_accumulate(P, result, occlusion, sb, cur_b, step, seg);
        ++steps;
        const float prev = step;
        step += fmaxf(diameter, P.voxel_size) * step_factor;
        seg = step - prev;
        diameter = step * cone_coef;
    }
    return {result.x, result.y, result.z, 1.0f - occlusion};
}

__device__ __forceinline__ float shadow_fetch(const TraceParams& P, int i, int j)
{
    if (i < 0 || j < 0 || i >= P.sw || j >= P.sh) return 0.0f;
    return P.shadow[(size_t)j * (size_t)P.sw + (size_t)i];
}

// shadow.glsl:28-36 (PCF16 of bilinear taps)
__device__ float trace_visibility(const TraceParams& P, const float pos[3])
{
    const float* V = P.light_view;
    const float* Pm = P.light_proj;
    float ls[3];
    for (int k = 0; k < 3; ++k) ls[k] = V[0 + k] * pos[0] + V[4 + k] * pos[1] + V[8 + k] * pos[2] + V[12 + k];
    const bool compare = !(P.quirks & VCT_QUIRK_SHADOW_RAW_DEPTH);
    float ref = 0.0f;
    if (compare) ref = Pm[2] * ls[0] + Pm[6] * ls[1] + Pm[10] * ls[2] + Pm[14] - 0.001f;
    ls[2] = ls[2] / (P.z_far - P.z_near);
    float xy[2];
    for (int k = 0; k < 2; ++k) xy[k] = Pm[0 + k] * ls[0] + Pm[4 + k] * ls[1] + Pm[12 + k];
    xy[0] = xy[0] * 0.5f + 0.5f;
    xy[1] = xy[1] * 0.5f + 0.5f;
    const float sx = 1.0f / (float)P.sw, sy = 1.0f / (float)P.sh;
    float sum = 0.0f;
    for (int yi = 0; yi < 4; ++yi)
        for (int xi = 0; xi < 4; ++xi) {
            const float s = xy[0] + ((float)xi - 1.5f) * sx, t = xy[1] + ((float)yi - 1.5f) * sy;
            float uu = s * (float)P.sw - 0.5f, vv = t * (float)P.sh - 0.5f;
            float fu = floorf(uu), fv = floorf(vv);
            const float a = uu - fu, b = vv - fv;
            fu = fminf(fmaxf(fu, -4.0f), (float)P.sw + 4.0f);
            fv = fminf(fmaxf(fv, -4.0f), (float)P.sh + 4.0f);
            const int i0 = (int)fu, j0 = (int)fv;
            float tq[4];
            const int ii[4] = {i0, i0 + 1, i0, i0 + 1}, jj[4] = {j0, j0, j0 + 1, j0 + 1};
            for (int q = 0; q < 4; ++q) {
                float d = shadow_fetch(P, ii[q], jj[q]);
                if (compare) {
                    const bool in = ii[q] >= 0 && jj[q] >= 0 && ii[q] < P.sw && jj[q] < P.sh;
                    d = (!in || ref <= d) ? 1.0f : 0.0f;
                }
                tq[q] = d;
            }
            sum += (1.0f - a) * (1.0f - b) * tq[0] + a * (1.0f - b) * tq[1] + (1.0f - a) * b * tq[2] + a * b * tq[3];
        }
    return sum * 0.0625f;
}

// brdf.glsl:28-78 microfacetBRDF
__device__ void brdf(const float n[3], const float v[3], const float l[3], const float dc[3], const float sc[3],
                     float rough, float out[3])
{
    const float PI = 3.141592f;
    const float alpha = rough * rough;
    const float reflectance = fmaxf(fmaxf(sc[0], sc[1]), sc[2]);
    const float r90 = fminf(fmaxf(reflectance * 50.0f, 0.0f), 1.0f);
    float h[3] = {l[0] + v[0], l[1] + v[1], l[2] + v[2]};
    const float hl = sqrtf(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    h[0] /= hl;
    h[1] /= hl;
    h[2] /= hl;
    const float ndl = fminf(fmaxf(n[0] * l[0] + n[1] * l[1] + n[2] * l[2], 0.001f), 1.0f);
    const float ndv = fminf(fmaxf(fabsf(n[0] * v[0] + n[1] * v[1] + n[2] * v[2]), 0.001f), 1.0f);
    const float ndh = fminf(fmaxf(n[0] * h[0] + n[1] * h[1] + n[2] * h[2], 0.0f), 1.0f);
    const float vdh = fminf(fmaxf(v[0] * h[0] + v[1] * h[1] + v[2] * h[2], 0.0f), 1.0f);
    const float r = alpha;
    const float att_l = 2.0f * ndl / (ndl + sqrtf(r * r + (1.0f - r * r) * (ndl * ndl)));
    const float att_v = 2.0f * ndv / (ndv + sqrtf(r * r + (1.0f - r * r) * (ndv * ndv)));
    const float G = att_l * att_v;
    const float rsq = alpha * alpha;
    const float f = (ndh * rsq - ndh) * ndh + 1.0f;
    const float Dd = rsq / (PI * f * f);
    const float omv = fminf(fmaxf(1.0f - vdh, 0.0f), 1.0f);
    const float fres = omv * omv * omv * omv * omv;
    for (int k = 0; k < 3; ++k) {
        const float F = sc[k] + (r90 - sc[k]) * fres;
        const float diffuse_contrib = (1.0f - F) * (dc[k] / PI);
        const float spec_contrib = F * G * Dd / (4.0f * ndl * ndv);
        out[k] = ndl * (diffuse_contrib + spec_contrib);
    }
}

__device__ __forceinline__ float half_to_float(uint16_t h)
{
    _Float16 f;
    __builtin_memcpy(&f, &h, 2);
    return (float)f;
}

// Per-pixel inputs of voxelConeTracing.frag main (:143-200): world position
// from depth, G-buffer decode, calcMinLevel and the cone start point.
struct PixelIn {
    float world[3], view[3], normal[3], start[3];
    float dc[3], sc[3], ef[3];
    float rough, metallic, min_level;
    bool has_emission;
};

__device__ __forceinline__ void pixel_in(const TraceParams& P, size_t idx, int px, int py, float depth,
                                         const uint32_t* __restrict__ g_diffuse, const uint16_t* __restrict__ g_normal,
                                         const uint32_t* __restrict__ g_specular,
                                         const uint16_t* __restrict__ g_emission, PixelIn& I)
{
    const float tcx = ((float)px + 0.5f) / (float)P.width, tcy = ((float)py + 0.5f) / (float)P.height;
    const float q[4] = {tcx * 2.0f - 1.0f, tcy * 2.0f - 1.0f, depth, 1.0f};
    float wp[4];
    for (int k = 0; k < 4; ++k)
        wp[k] = P.vp_inv[k] * q[0] + P.vp_inv[4 + k] * q[1] + P.vp_inv[8 + k] * q[2] + P.vp_inv[12 + k] * q[3];
    for (int k = 0; k < 3; ++k) I.world[k] = wp[k] / wp[3];
    for (int k = 0; k < 3; ++k) I.view[k] = P.eye[k] - I.world[k];
    {
        const float l = sqrtf(I.view[0] * I.view[0] + I.view[1] * I.view[1] + I.view[2] * I.view[2]);
        for (int k = 0; k < 3; ++k) I.view[k] /= l;
    }
    const uint32_t dt = g_diffuse[idx], st = g_specular[idx];
    I.dc[0] = u8_to_f(dt & 0xffu);
    I.dc[1] = u8_to_f((dt >> 8) & 0xffu);
    I.dc[2] = u8_to_f((dt >> 16) & 0xffu);
    I.sc[0] = u8_to_f(st & 0xffu);
    I.sc[1] = u8_to_f((st >> 8) & 0xffu);
    I.sc[2] = u8_to_f((st >> 16) & 0xffu);
    I.rough = u8_to_f(dt >> 24);
    I.metallic = u8_to_f(st >> 24);
    const f4 nf = unpack_half(*(const uint2*)(g_normal + 4 * idx));
    const f4 ef = unpack_half(*(const uint2*)(g_emission + 4 * idx));
    I.normal[0] = nf.x * 2.0f - 1.0f;
    I.normal[1] = nf.y * 2.0f - 1.0f;
    I.normal[2] = nf.z * 2.0f - 1.0f;
    {
        const float l = sqrtf(I.normal[0] * I.normal[0] + I.normal[1] * I.normal[1] + I.normal[2] * I.normal[2]);
        for (int k = 0; k < 3; ++k) I.normal[k] /= l;
    }
    I.ef[0] = ef.x;
    I.ef[1] = ef.y;
    I.ef[2] = ef.z;
    I.has_emission = ef.x > 0.0f || ef.y > 0.0f || ef.z > 0.0f;
    // calcMinLevel (voxelConeTracing.frag:394-414)
    {
        const float dx = P.center[0] - I.world[0], dy = P.center[1] - I.world[1], dz = P.center[2] - I.world[2];
        const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        const float min_radius = P.voxel_size * P.dim * 0.5f;
        const float ml = fmaxf(log2f(dist / min_radius), 0.0f);
        const float radius = min_radius * exp2f(ceilf(ml));
        const float f = dist / radius;
        I.min_level = f > 0.5f ? ceilf(ml) + (f - 0.5f) * 2.0f : ceilf(ml);
    }
    const float vs = P.voxel_size * exp2f(I.min_level);
    for (int k = 0; k < 3; ++k) I.start[k] = I.world[k] + I.normal[k] * vs * P.start_offset;
}

__device__ __forceinline__ bool wants_specular(const TraceParams& P, const PixelIn& I)
{
    return (P.mode == 6u || P.mode == 8u) && (I.sc[0] > 1e-6f || I.sc[1] > 1e-6f || I.sc[2] > 1e-6f) &&
           I.metallic > 1e-6f;
}

// The specular cone (voxelConeTracing.frag:205-216): reflect, aperture
// max(roughness, 0.05), step factor = voxel size.
template <bool kHalf>
__device__ __forceinline__ float4 specular_cone(const typename Vol<kHalf>::T* __restrict__ vol, const TraceParams& P,
                                                const PixelIn& I, uint32_t& steps)
{
    const float in[3] = {-I.view[0], -I.view[1], -I.view[2]};
    const float dni = I.normal[0] * in[0] + I.normal[1] * in[1] + I.normal[2] * in[2];
    const float dir[3] = {in[0] - 2.0f * dni * I.normal[0], in[1] - 2.0f * dni * I.normal[1],
                          in[2] - 2.0f * dni * I.normal[2]};
    const f4 c = trace_cone<kHalf>(vol, P, I.start, dir, fmaxf(I.rough, 0.05f), 30.0f, I.min_level, P.voxel_size, steps);
    return make_float4(c.x * I.sc[0] * P.specular_intensity, c.y * I.sc[1] * P.specular_intensity,
                       c.z * I.sc[2] * P.specular_intensity, 1.0f);
}

__device__ __forceinline__ void add_samples(unsigned long long* samples, uint32_t steps)
{
    if (!samples) return;
    unsigned long long v = steps;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(samples, v);
}

// Diffuse cones, direct term and mode composition, one lane per pixel in 8x8
// wave tiles. With kSplitSpec the specular cone is left to k_trace_specular
// (which writes out_s); otherwise it is traced here.
template <bool kHalf, bool kSplitSpec>
__global__ __launch_bounds__(256) void k_trace(const typename Vol<kHalf>::T* __restrict__ vol, TraceParams P,
                                               const uint32_t* __restrict__ g_diffuse,
                                               const uint16_t* __restrict__ g_normal,
                                               const uint32_t* __restrict__ g_specular,
                                               const uint16_t* __restrict__ g_emission,
                                               const float* __restrict__ g_depth, float4* __restrict__ out_d,
                                               float4* __restrict__ out_s, unsigned long long* __restrict__ samples)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned tiles_x = ((unsigned)P.width + 15u) / 16u;
    const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    uint32_t steps = 0;
    if (px < P.width && py < P.height) {
        const size_t idx = (size_t)py * (size_t)P.width + (size_t)px;
        float4 dO = make_float4(0.0f, 0.0f, 0.0f, 1.0f), sO = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        const float depth = g_depth[idx];
        if (depth != 1.0f) {
            PixelIn I;
            pixel_in(P, idx, px, py, depth, g_diffuse, g_normal, g_specular, g_emission, I);
            f4 ind = {0.0f, 0.0f, 0.0f, 1.0f};
            const int ncones = P.cones32 ? 32 : 16;
            const float aperture = P.cones32 ? 0.628319f : 0.872665f;
            const bool want_diffuse = (P.mode == 4u || P.mode == 5u || P.mode == 7u || P.mode == 8u);
            for (int i = 0; want_diffuse && i < ncones; ++i) {
                const float* d = P.cones32 ? c_dirs32[i] : c_dirs16[i];
                const float ct = I.normal[0] * d[0] + I.normal[1] * d[1] + I.normal[2] * d[2];
                if (ct < 0.0f) continue;
                const f4 c = trace_cone<kHalf>(vol, P, I.start, d, aperture, 30.0f, I.min_level, P.step_factor, steps);
                ind.x += c.x * ct;
                ind.y += c.y * ct;
                ind.z += c.z * ct;
                ind.w += c.w * ct;
            }
            const float inv_n = 1.0f / (float)ncones;
            ind.x *= inv_n;
            ind.y *= inv_n;
            ind.z *= inv_n;
            ind.w *= inv_n;
            ind.w *= P.ao_factor;
            ind.x *= I.dc[0] * P.diffuse_intensity;
            ind.y *= I.dc[1] * P.diffuse_intensity;
            ind.z *= I.dc[2] * P.diffuse_intensity;
            if (!kSplitSpec && wants_specular(P, I)) sO = specular_cone<kHalf>(vol, P, I, steps);
            float direct[3] = {0.0f, 0.0f, 0.0f};
            const bool want_direct = (P.mode == 4u || P.mode == 5u || P.mode == 8u);
            if (want_direct) {
                if (I.has_emission) {
                    for (int k = 0; k < 3; ++k) direct[k] = I.ef[k];
                } else {
                    float b[3];
                    brdf(I.normal, I.view, P.light_dir, I.dc, I.sc, I.rough, b);
                    const float visv = trace_visibility(P, I.world);
                    for (int k = 0; k < 3; ++k) direct[k] = b[k] * visv;
                }
            }
            switch (P.mode) {
            case 0: dO = make_float4(I.dc[0], I.dc[1], I.dc[2], 1.0f); break;
            case 1: dO = make_float4(I.sc[0], I.sc[1], I.sc[2], 1.0f); break;
            case 2:
                dO = make_float4(I.normal[0] * 0.5f + 0.5f, I.normal[1] * 0.5f + 0.5f, I.normal[2] * 0.5f + 0.5f, 1.0f);
                break;
            case 3: {
                const float colors[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
                int lower = (int)floorf(I.min_level);
                lower = lower < 0 ? 0 : (lower > 5 ? 5 : lower);
                const float fr = I.min_level - floorf(I.min_level);
                dO = make_float4((colors[lower][0] * (1.0f - fr) + colors[lower + 1][0] * fr) * 0.5f,
                                 (colors[lower][1] * (1.0f - fr) + colors[lower + 1][1] * fr) * 0.5f,
                                 (colors[lower][2] * (1.0f - fr) + colors[lower + 1][2] * fr) * 0.5f, 0.5f);
                break;
            }
            case 4: dO = make_float4(direct[0] * ind.w, direct[1] * ind.w, direct[2] * ind.w, 1.0f); break;
            case 5:
            case 8:
                dO = make_float4(direct[0] * ind.w + ind.x, direct[1] * ind.w + ind.y, direct[2] * ind.w + ind.z, 1.0f);
                break;
            case 7: dO = make_float4(ind.w, ind.w, ind.w, 1.0f); break;
            default: break;
            }
        }
        out_d[idx] = dO;
        if (!kSplitSpec) out_s[idx] = sO;
    }
    add_samples(samples, steps);
}

// Specular work ordering. The specular cone's step count is set by the
// roughness byte alone (aperture max(r, 0.05), step factor = voxel size, 63 to
// 336 steps over r in [0.1, 0.8]), so lanes of a wave that mixes roughness
// idle until its smoothest lane finishes. Each 64x64 pixel tile is
// counting-sorted by roughness in LDS and the permutation written to `perm`
// (tile t owns perm[t*4096, (t+1)*4096)); pixels without a specular cone go
// last, unused slots hold 0xffffffff. The order inside one roughness bin
// depends on LDS atomic order, which changes no pixel's value.
// Dispatch order: a tile's 16 chunks of 256 entries (one workgroup each) get
// a cost key, the bin of the chunk's first entry (its lowest roughness, the
// longest march), and within each XCD's share of the dispatch the chunks run
// longest first (k_spec_order), so no XCD ends the pass on a lone long
// chunk: 4.17 -> 3.55 ms for the 1080p trace pass, every pixel unchanged.
// The work list is perm[tiles * 4096], then keys[tiles * 16], then
// order[tiles * 16].
constexpr int kSpecTile = 64;

// Tile-local pixel of index i in Morton (Z) order, so that pixels of one bin
// fill 8 x 8 blocks in order and a wave's specular cones start close together.
__device__ __forceinline__ int morton_x(int i)
{
    return (i & 1) | ((i >> 1) & 2) | ((i >> 2) & 4) | ((i >> 3) & 8) | ((i >> 4) & 16) | ((i >> 5) & 32);
}
__device__ __forceinline__ int morton_y(int i) { return morton_x(i >> 1); }
constexpr int kSpecTilePixels = kSpecTile * kSpecTile;
constexpr int kSpecBins = 257;

__global__ __launch_bounds__(256) void k_spec_bin(TraceParams P, const uint32_t* __restrict__ g_diffuse,
                                                  const uint32_t* __restrict__ g_specular,
                                                  const float* __restrict__ g_depth, int tiles_x,
                                                  uint32_t* __restrict__ perm, uint32_t* __restrict__ keys)
{
    __shared__ uint32_t hist[kSpecBins];
    __shared__ uint32_t cursor[kSpecBins];
    const int tid = threadIdx.x;
    const int tile = blockIdx.x;
    const int tx0 = (tile % tiles_x) * kSpecTile, ty0 = (tile / tiles_x) * kSpecTile;
    for (int b = tid; b < kSpecBins; b += 256) hist[b] = 0u;
    __syncthreads();
    uint16_t key[kSpecTilePixels / 256];
#pragma unroll
    for (int j = 0; j < kSpecTilePixels / 256; ++j) {
        const int local = j * 256 + tid;
        const int px = tx0 + morton_x(local), py = ty0 + morton_y(local);
        uint16_t k = 0xffffu;
        if (px < P.width && py < P.height) {
            const size_t idx = (size_t)py * (size_t)P.width + (size_t)px;
            k = kSpecBins - 1;
            if (g_depth[idx] != 1.0f) {
                const uint32_t st = g_specular[idx];
                const bool spec = ((st & 0xffu) || ((st >> 8) & 0xffu) || ((st >> 16) & 0xffu)) && (st >> 24);
                if (spec && (P.mode == 6u || P.mode == 8u)) k = (uint16_t)(g_diffuse[idx] >> 24);
            }
            atomicAdd(&hist[k], 1u);
        }
        key[j] = k;
    }
    __syncthreads();
    // Bin starts: an exclusive scan of the 257 bins, thread t taking bin t
    // (a shuffle scan per wave, then the waves' sums) and thread 255 the
    // last bin as well.
    {
        __shared__ uint32_t wsum[4];
        const int lane = tid & 63, wave = tid >> 6;
        const uint32_t h = hist[tid];
        uint32_t v = h;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        if (lane == 63) wsum[wave] = v;
        __syncthreads();
        for (int w = 0; w < wave; ++w) v += wsum[w];
        cursor[tid] = v - h;
        if (tid == 255) cursor[kSpecBins - 1] = v;
    }
    __syncthreads();
    // Each chunk's cost key: the bin of its first entry (the lowest roughness
    // in the chunk, the longest march), written by the thread of that bin;
    // chunks past the tile's pixels get kSpecBins, the cheapest.
    {
        uint32_t* ck = keys + (size_t)tile * (kSpecTilePixels / 256);
        constexpr uint32_t kChunks = kSpecTilePixels / 256;
        for (int b = tid; b < kSpecBins; b += 256) {
            const uint32_t lo = cursor[b], hi = lo + hist[b];
            for (uint32_t c = (lo + 255u) / 256u; c < kChunks && c * 256u < hi; ++c) ck[c] = (uint32_t)b;
        }
        if (tid == 0) {
            const uint32_t used = cursor[kSpecBins - 1] + hist[kSpecBins - 1];
            for (uint32_t c = (used + 255u) / 256u; c < kChunks; ++c) ck[c] = (uint32_t)kSpecBins;
        }
    }
    __syncthreads();  // the keys read the bin starts before the scatter advances them
    uint32_t* out = perm + (size_t)tile * kSpecTilePixels;
#pragma unroll
    for (int j = 0; j < kSpecTilePixels / 256; ++j) {
        const int local = j * 256 + tid;
        if (key[j] == 0xffffu) continue;
        const int px = tx0 + morton_x(local), py = ty0 + morton_y(local);
        const uint32_t pos = atomicAdd(&cursor[key[j]], 1u);
        out[pos] = (uint32_t)py * (uint32_t)P.width + (uint32_t)px;
    }
    // unused slots of an edge tile: after the scatter the last bin's cursor
    // is the number of pixels the tile holds
    __syncthreads();
    const uint32_t used = cursor[kSpecBins - 1];
    for (uint32_t i = used + tid; i < (uint32_t)kSpecTilePixels; i += 256) out[i] = 0xffffffffu;
}

// Dispatch order of the specular chunks: XCD x runs positions
// [x per + min(x, rem), ...) of xcd_remap; within each XCD's range the chunks
// are counting-sorted by key, longest marches first.
__global__ __launch_bounds__(256) void k_spec_order(const uint32_t* __restrict__ keys, uint32_t n,
                                                    uint32_t* __restrict__ order)
{
    __shared__ uint32_t hist[kSpecBins + 1];
    const uint32_t x = blockIdx.x, per = n >> 3, rem = n & 7u;
    const uint32_t start = x * per + (x < rem ? x : rem), len = per + (x < rem ? 1u : 0u);
    for (uint32_t b = threadIdx.x; b <= (uint32_t)kSpecBins; b += 256) hist[b] = 0u;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < len; i += 256) atomicAdd(&hist[keys[start + i]], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 0; b <= kSpecBins; ++b) {
            const uint32_t h = hist[b];
            hist[b] = run;
            run += h;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < len; i += 256) order[start + atomicAdd(&hist[keys[start + i]], 1u)] = start + i;
}

// One lane per entry of `perm`: the specular cone of that pixel (or the
// reference's (0,0,0,1) where none is traced).
template <bool kHalf>
__global__ __launch_bounds__(256) void k_trace_specular(const typename Vol<kHalf>::T* __restrict__ vol, TraceParams P,
                                                        const uint32_t* __restrict__ perm,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ g_diffuse,
                                                        const uint16_t* __restrict__ g_normal,
                                                        const uint32_t* __restrict__ g_specular,
                                                        const uint16_t* __restrict__ g_emission,
                                                        const float* __restrict__ g_depth, float4* __restrict__ out_s,
                                                        unsigned long long* __restrict__ samples)
{
    const uint32_t pi = perm[(size_t)order[xcd_remap(blockIdx.x, gridDim.x)] * 256 + threadIdx.x];
    uint32_t steps = 0;
    if (pi != 0xffffffffu) {
        const int px = (int)(pi % (uint32_t)P.width), py = (int)(pi / (uint32_t)P.width);
        float4 sO = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        const float depth = g_depth[pi];
        if (depth != 1.0f) {
            PixelIn I;
            pixel_in(P, pi, px, py, depth, g_diffuse, g_normal, g_specular, g_emission, I);
            if (wants_specular(P, I)) sO = specular_cone<kHalf>(vol, P, I, steps);
        }
        out_s[pi] = sO;
    }
    add_samples(samples, steps);
}

}  // namespace vct


namespace vct {

struct SvoTraceParams {
    float center[3];  // scene box centre (voxelizer.vert:42-47)
    float ev;         // largest half extent of the scene box
};

// voxelConeTracing_Octree.frag:319-346 sampleSVO (top-down walk to the target
// resolution; uvec3() of a negative value saturates to 0).
__device__ __forceinline__ f4 sample_svo(const uint2* __restrict__ pool, const TraceParams& P, const SvoTraceParams& S,
                                         const float pos[3], uint32_t level)
{
    uint32_t res = (uint32_t)P.dim;
    uint32_t fp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float s = (pos[k] - S.center[k]) / S.ev;
        const float f = ((s + 1.0f) * 0.5f) * (float)res;
        const uint32_t u = f > 0.0f ? (f < 4294967040.0f ? (uint32_t)f : 0xffffffffu) : 0u;
        fp[k] = u > res ? res : u;
    }
    const uint32_t target = 1u << level;
    uint32_t idx = 0, cur = 0;
    do {
        res >>= 1;
        const uint32_t cx = fp[0] >= res, cy = fp[1] >= res, cz = fp[2] >= res;
        idx = cur + (cz | (cx << 1) | (cy << 2));
        cur = pool[idx].x & 0x7fffffffu;
        fp[0] -= cx * res;
        fp[1] -= cy * res;
        fp[2] -= cz * res;
    } while (cur != 0u && res > target);
    const uint32_t t = pool[idx].y;
    const float sc = 1.0f / 255.0f;
    return {(float)(t & 0xffu) * sc, (float)((t >> 8) & 0xffu) * sc, (float)((t >> 16) & 0xffu) * sc, (float)(t >> 24) * sc};
}

// voxelConeTracing_Octree.frag:359-409 traceCone
__device__ f4 svo_trace_cone(const uint2* __restrict__ pool, const TraceParams& P, const SvoTraceParams& S,
                             const float start_in[3], const float dir[3], float aperture, float max_dist,
                             float start_level, float step_factor, uint32_t& steps)
{
    f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
    const float cone_coef = 2.0f * tanf(aperture * 0.5f);
    float voxel_size = P.voxel_size * exp2f(start_level);
    float start[3];
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * voxel_size * P.start_offset * 0.5f;
    float step = 0.0f;
    float diameter = fmaxf(step * cone_coef, P.voxel_size);
    float occlusion = 0.0f;
    float seg = voxel_size;
    const float min_radius = P.voxel_size * P.dim * 0.5f;
    const float max_level = (float)(P.levels - 1);
    while (step < max_dist && occlusion < 1.0f) {
        float pos[3];
        for (int k = 0; k < 3; ++k) pos[k] = start[k] + dir[k] * step;
        const float dx = P.center[0] - pos[0], dy = P.center[1] - pos[1], dz = P.center[2] - pos[2];
        const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
        // hardware log2 / exp2 / pow as in the clipmap march (same results
        // for these arguments, see cone_level and cone_accumulate)
        const float min_level = ceilf(__builtin_amdgcn_logf(dist / min_radius));
        float cur = __builtin_amdgcn_logf(diameter / P.voxel_size);
        cur = fminf(fmaxf(fmaxf(start_level, cur), min_level), max_level);
        const f4 lo = sample_svo(pool, P, S, pos, (uint32_t)floorf(cur));
        const f4 hi = sample_svo(pool, P, S, pos, (uint32_t)ceilf(cur));
        const float fl = cur - floorf(cur);
        const f4 s = {lo.x * (1.0f - fl) + hi.x * fl, lo.y * (1.0f - fl) + hi.y * fl, lo.z * (1.0f - fl) + hi.z * fl,
                      lo.w * (1.0f - fl) + hi.w * fl};
        voxel_size = P.voxel_size * __builtin_amdgcn_exp2f(cur);
        const float correction = seg / voxel_size;
        const float opacity = fminf(
            fmaxf(1.0f - __builtin_amdgcn_exp2f(correction * __builtin_amdgcn_logf(fmaxf(1.0f - s.w, 0.0f))), 0.0f),
            1.0f);
        const float blend = fminf(fmaxf(1.0f - result.w, 0.0f), 1.0f);
        result.x += blend * (s.x * correction);
        result.y += blend * (s.y * correction);
        result.z += blend * (s.z * correction);
        result.w += blend * opacity;
        occlusion += (1.0f - occlusion) * opacity / (1.0f + (step + voxel_size) * P.decay);
        const float prev = step;
        step += fmaxf(diameter, P.voxel_size) * step_factor;
        seg = step - prev;
        diameter = step * cone_coef;
        ++steps;
    }
    return {fminf(fmaxf(result.x, 0.0f), 1.0f), fminf(fmaxf(result.y, 0.0f), 1.0f), fminf(fmaxf(result.z, 0.0f), 1.0f),
            fminf(fmaxf(1.0f - occlusion, 0.0f), 1.0f)};
}

__device__ __forceinline__ float clamp01(float x)
{
    return fminf(fmaxf(x, 0.0f), 1.0f);
}

// voxelConeTracing_Octree.frag:148-300 main
__global__ __launch_bounds__(256) void k_svo_trace(const uint2* __restrict__ pool, TraceParams P, SvoTraceParams S,
                                                   const uint32_t* __restrict__ g_diffuse,
                                                   const uint16_t* __restrict__ g_normal,
                                                   const uint32_t* __restrict__ g_specular,
                                                   const uint16_t* __restrict__ g_emission,
                                                   const float* __restrict__ g_depth, float4* __restrict__ out_d,
                                                   float4* __restrict__ out_s, unsigned long long* __restrict__ samples)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned tiles_x = ((unsigned)P.width + 15u) / 16u;
    const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    uint32_t steps = 0;
    if (px < P.width && py < P.height) {
        const size_t idx = (size_t)py * (size_t)P.width + (size_t)px;
        float4 dO = make_float4(0.0f, 0.0f, 0.0f, 1.0f), sO = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        const float depth = g_depth[idx];
        if (depth != 1.0f) {
            const float tcx = ((float)px + 0.5f) / (float)P.width, tcy = ((float)py + 0.5f) / (float)P.height;
            const float q[4] = {tcx * 2.0f - 1.0f, tcy * 2.0f - 1.0f, depth, 1.0f};
            float wp[4];
            for (int k = 0; k < 4; ++k)
                wp[k] = P.vp_inv[k] * q[0] + P.vp_inv[4 + k] * q[1] + P.vp_inv[8 + k] * q[2] + P.vp_inv[12 + k] * q[3];
            const float world[3] = {wp[0] / wp[3], wp[1] / wp[3], wp[2] / wp[3]};
            float view[3] = {P.eye[0] - world[0], P.eye[1] - world[1], P.eye[2] - world[2]};
            {
                const float l = sqrtf(view[0] * view[0] + view[1] * view[1] + view[2] * view[2]);
                for (int k = 0; k < 3; ++k) view[k] /= l;
            }
            const uint32_t dt = g_diffuse[idx], st = g_specular[idx];
            const float dc[3] = {u8_to_f(dt & 0xffu), u8_to_f((dt >> 8) & 0xffu), u8_to_f((dt >> 16) & 0xffu)};
            const float sc[3] = {u8_to_f(st & 0xffu), u8_to_f((st >> 8) & 0xffu), u8_to_f((st >> 16) & 0xffu)};
            const float rough = u8_to_f(dt >> 24), metallic = u8_to_f(st >> 24);
            const f4 nf = unpack_half(*(const uint2*)(g_normal + 4 * idx));
            const f4 ef = unpack_half(*(const uint2*)(g_emission + 4 * idx));
            float normal[3] = {nf.x * 2.0f - 1.0f, nf.y * 2.0f - 1.0f, nf.z * 2.0f - 1.0f};
            {
                const float l = sqrtf(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
                for (int k = 0; k < 3; ++k) normal[k] /= l;
            }
            const bool has_emission = ef.x > 0.0f || ef.y > 0.0f || ef.z > 0.0f;
            float min_level;
            {
                const float dx = P.center[0] - world[0], dy = P.center[1] - world[1], dz = P.center[2] - world[2];
                const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                const float min_radius = P.voxel_size * P.dim * 0.5f;
                const float ml = fmaxf(log2f(dist / min_radius), 0.0f);
                const float radius = min_radius * exp2f(ceilf(ml));
                const float f = dist / radius;
                min_level = f > 0.5f ? ceilf(ml) + (f - 0.5f) * 2.0f : ceilf(ml);
            }
            const float vs = P.voxel_size * exp2f(min_level);
            float start[3];
            for (int k = 0; k < 3; ++k) start[k] = world[k] + normal[k] * vs * P.start_offset;
            f4 ind = {0.0f, 0.0f, 0.0f, 1.0f};
            float valid = 0.0f;
            const int ncones = P.cones32 ? 32 : 16;
            const float aperture = P.cones32 ? 0.628319f : 0.872665f;
            const bool want_diffuse = (P.mode == 4u || P.mode == 5u || P.mode == 7u || P.mode == 8u);
            for (int i = 0; want_diffuse && i < ncones; ++i) {
                const float* d = P.cones32 ? c_dirs32[i] : c_dirs16[i];
                const float ct = normal[0] * d[0] + normal[1] * d[1] + normal[2] * d[2];
                if (ct < 0.0f) continue;
                const f4 c = svo_trace_cone(pool, P, S, start, d, aperture, 30.0f, min_level, P.step_factor, steps);
                ind.x += c.x * ct;
                ind.y += c.y * ct;
                ind.z += c.z * ct;
                ind.w += c.w * ct;
                valid += ct;
            }
            ind.x /= valid;
            ind.y /= valid;
            ind.z /= valid;
            ind.w /= valid;
            ind.w *= P.ao_factor;
            ind.x *= dc[0] * P.diffuse_intensity;
            ind.y *= dc[1] * P.diffuse_intensity;
            ind.z *= dc[2] * P.diffuse_intensity;
            ind = {clamp01(ind.x), clamp01(ind.y), clamp01(ind.z), clamp01(ind.w)};
            float spec[3] = {0.0f, 0.0f, 0.0f};
            if ((P.mode == 6u || P.mode == 8u) && (sc[0] > 1e-6f || sc[1] > 1e-6f || sc[2] > 1e-6f) && metallic > 1e-6f) {
                const float I[3] = {-view[0], -view[1], -view[2]};
                const float dni = normal[0] * I[0] + normal[1] * I[1] + normal[2] * I[2];
                float dir[3] = {I[0] - 2.0f * dni * normal[0], I[1] - 2.0f * dni * normal[1], I[2] - 2.0f * dni * normal[2]};
                const float l = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
                for (int k = 0; k < 3; ++k) dir[k] /= l;
                const f4 c = svo_trace_cone(pool, P, S, start, dir, fmaxf(rough, 0.05f), 30.0f, min_level,
                                            fmaxf(0.2f, P.voxel_size), steps);
                spec[0] = c.x * sc[0] * P.specular_intensity;
                spec[1] = c.y * sc[1] * P.specular_intensity;
                spec[2] = c.z * sc[2] * P.specular_intensity;
            }
            float direct[3] = {0.0f, 0.0f, 0.0f};
            if (P.mode == 4u || P.mode == 8u) {
                if (has_emission) {
                    direct[0] = ef.x;
                    direct[1] = ef.y;
                    direct[2] = ef.z;
                } else {
                    float b[3];
                    brdf(normal, view, P.light_dir, dc, sc, rough, b);
                    const float visv = trace_visibility(P, world);
                    for (int k = 0; k < 3; ++k) direct[k] = b[k] * visv;
                }
                for (int k = 0; k < 3; ++k) direct[k] = clamp01(direct[k]);
            }
            switch (P.mode) {
            case 0: dO = make_float4(dc[0], dc[1], dc[2], 1.0f); break;
            case 1: dO = make_float4(sc[0], sc[1], sc[2], 1.0f); break;
            case 2: dO = make_float4(normal[0] * 0.5f + 0.5f, normal[1] * 0.5f + 0.5f, normal[2] * 0.5f + 0.5f, 1.0f); break;
            case 3: {
                const float colors[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
                float o4[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                if (min_level < 6.0f) {
                    int lower = (int)floorf(min_level);
                    lower = lower < 0 ? 0 : lower;
                    const float fr = min_level - floorf(min_level);
                    for (int k = 0; k < 3; ++k) o4[k] = colors[lower][k] * (1.0f - fr) + colors[lower + 1][k] * fr;
                    o4[3] = 1.0f;
                }
                dO = make_float4(o4[0] * 0.5f, o4[1] * 0.5f, o4[2] * 0.5f, o4[3] * 0.5f);
                break;
            }
            case 4: dO = make_float4(direct[0] * ind.w, direct[1] * ind.w, direct[2] * ind.w, 1.0f); break;
            case 5: dO = make_float4(ind.x, ind.y, ind.z, 1.0f); break;
            case 6: sO = make_float4(spec[0], spec[1], spec[2], 1.0f); break;
            case 7: dO = make_float4(ind.w, ind.w, ind.w, 1.0f); break;
            case 8:
                dO = make_float4(direct[0] * ind.w + ind.x, direct[1] * ind.w + ind.y, direct[2] * ind.w + ind.z, 1.0f);
                sO = make_float4(spec[0], spec[1], spec[2], 1.0f);
                break;
            default: break;
            }
        }
        out_d[idx] = dO;
        out_s[idx] = sO;
    }
    if (samples) {
        unsigned long long v = steps;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0 && v) atomicAdd(samples, v);
    }
}

// ---- occupancy mask of the radiance volume (the trace's empty-space skip)

__device__ __forceinline__ bool texel_nonzero(uint32_t t) { return t != 0u; }
__device__ __forceinline__ bool texel_nonzero(uint2 t) { return (t.x | t.y) != 0u; }

// Bit i of level blockIdx.y: texel position i = (z*D + y)*D + x of the level
// holds a nonzero texel in any of the 6 faces (one lane per position, the
// wave's ballot stored as two words; positions past D^3 are zero).
template <typename T>
__global__ __launch_bounds__(256) void k_occ_bits(const T* __restrict__ vol, uint32_t D3, uint32_t words,
                                                  uint32_t raw_stride, uint32_t* __restrict__ raw)
{
    const uint32_t level = blockIdx.y;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    bool nz = false;
    if (i < D3) {
        const size_t base = (size_t)level * kFaces * D3 + i;
#pragma unroll
        // a plain OR, not ||: the six loads are independent and all in flight
        // (a short-circuit chain would wait for each face before the next)
        for (int f = 0; f < kFaces; ++f) nz = nz | texel_nonzero(vol[base + (size_t)f * D3]);
    }
    const unsigned long long b = __ballot(nz);
    const uint32_t lane = threadIdx.x & 63u, w = i >> 5;
    if ((lane == 0u || lane == 32u) && w < words)
        raw[(size_t)level * raw_stride + w] = lane == 0u ? (uint32_t)b : (uint32_t)(b >> 32);
}

// 32 bits of a bit array starting at bit b.
__device__ __forceinline__ uint32_t bits_at(const uint32_t* __restrict__ r, uint32_t b)
{
    const uint32_t w = b >> 5, sh = b & 31u;
    const uint32_t lo = r[w] >> sh;
    return sh ? (lo | (r[w + 1] << (32u - sh))) : lo;
}

// mask bit i = OR of the raw bits at i + {0, 1} + {0, D} + {0, D^2}: the
// 2x2x2 footprint with corner i. Reads stay within the level's words plus its
// zero padding words.
__global__ __launch_bounds__(256) void k_occ_dilate(const uint32_t* __restrict__ raw, uint32_t D, uint32_t words,
                                                    uint32_t raw_stride, uint32_t* __restrict__ mask)
{
    const uint32_t level = blockIdx.y;
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= words) return;
    const uint32_t* r = raw + (size_t)level * raw_stride;
    const uint32_t b = w * 32u, D2 = D * D;
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) m |= bits_at(r, b + ((k & 4) ? D2 : 0u) + ((k & 2) ? D : 0u) + (uint32_t)(k & 1));
    mask[(size_t)level * words + w] = m;
}

// Rebuilds the occupancy mask from the radiance volume on `s`.
static vct_status build_occupancy(vct_ctx* c, hipStream_t s)
{
    const uint32_t D = (uint32_t)c->D, D3 = D * D * D;
    const uint32_t raw_stride = c->occ_words + c->occ_pad;
    uint32_t* raw = (uint32_t*)c->occ_raw.ptr;
    if (c->cfg.radiance_format == VCT_FORMAT_RGBA16F)
        hipLaunchKernelGGL(k_occ_bits<uint2>, dim3((D3 + 255u) / 256u, (unsigned)c->levels), dim3(256), 0, s,
                           (const uint2*)c->radiance.ptr, D3, c->occ_words, raw_stride, raw);
    else
        hipLaunchKernelGGL(k_occ_bits<uint32_t>, dim3((D3 + 255u) / 256u, (unsigned)c->levels), dim3(256), 0, s,
                           (const uint32_t*)c->radiance.ptr, D3, c->occ_words, raw_stride, raw);
    hipLaunchKernelGGL(k_occ_dilate, dim3((c->occ_words + 255u) / 256u, (unsigned)c->levels), dim3(256), 0, s,
                       (const uint32_t*)raw, D, c->occ_words, raw_stride, (uint32_t*)c->occ_mask.ptr);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // namespace vct

using namespace vct;

static vct_status make_trace_params(vct_ctx* c, const vct_gbuffer* g, const vct_camera_ubo* cam,
                                    const vct_cone_tracing_desc* d, const vct_shadow_map* shadow,
                                    const vct_directional_light_desc* light,
                                    const vct_directional_light_shadow_desc* light_shadow, float* diffuse_out,
                                    float* specular_out, int levels, TraceParams* out)
{
    if (!c || !g || !cam || !d || !diffuse_out || !specular_out) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!g->diffuse || !g->normal || !g->specular || !g->emission || !g->depth || !g->width || !g->height)
        return fail(VCT_ERR_INVALID_ARGUMENT, "incomplete G-buffer");
    if (d->rendering_mode > 8u) return fail(VCT_ERR_INVALID_ARGUMENT, "rendering_mode");
    if (!(d->voxel_size > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "voxel_size");
    const bool need_light = d->rendering_mode == 4u || d->rendering_mode == 5u || d->rendering_mode == 8u;
    if (need_light && (!shadow || !shadow->depth || !shadow->width || !shadow->height || !light || !light_shadow))
        return fail(VCT_ERR_INVALID_ARGUMENT, "light, light_shadow and shadow map are required for this mode");
    TraceParams P{};
    for (int k = 0; k < 3; ++k) P.center[k] = d->volume_center[k];
    P.mode = d->rendering_mode;
    P.voxel_size = d->voxel_size;
    P.dim = d->volume_dimension;
    P.start_offset = d->trace_start_offset;
    P.diffuse_intensity = d->indirect_diffuse_intensity;
    P.ao_factor = d->ambient_occlusion_factor;
    P.step_factor = d->min_trace_step_factor > 0.2f ? d->min_trace_step_factor : 0.2f;
    P.specular_intensity = d->indirect_specular_intensity;
    P.decay = d->occlusion_decay;
    P.cones32 = d->enable_32_cones != 0;
    P.levels = levels;
    P.res = c->res;
    P.D = c->D;
    P.inv_extent0 = 1.0f / (d->voxel_size * d->volume_dimension);
    P.inv_voxel_size = 1.0f / d->voxel_size;
    P.inv_min_radius = 1.0f / (d->voxel_size * d->volume_dimension * 0.5f);
    {
        const size_t vb = texels_per_volume(c) * texel_bytes(c, VCT_VOLUME_RADIANCE);
        if (vb > 0xffffffffull) return fail(VCT_ERR_UNSUPPORTED, "radiance volume above 4 GiB");
        P.vol_bytes = (uint32_t)vb;
    }
    P.occ = (const uint32_t*)c->occ_mask.ptr;
    P.occ_words = c->occ_words;
    for (int k = 0; k < 16; ++k) P.vp_inv[k] = cam->view_proj_inv[k];
    for (int k = 0; k < 3; ++k) P.eye[k] = cam->eye_pos[k];
    P.width = (int)g->width;
    P.height = (int)g->height;
    P.quirks = c->cfg.quirks;
    if (light) {
        float nd[3] = {-light->direction[0], -light->direction[1], -light->direction[2]};
        float len = std::sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
        for (int k = 0; k < 3; ++k) P.light_dir[k] = len > 0.0f ? nd[k] / len : 0.0f;
    }
    if (light_shadow) {
        for (int k = 0; k < 16; ++k) {
            P.light_view[k] = light_shadow->view[k];
            P.light_proj[k] = light_shadow->proj[k];
        }
        P.z_near = light_shadow->z_near;
        P.z_far = light_shadow->z_far;
    }
    if (shadow) {
        P.shadow = shadow->depth;
        P.sw = (int)shadow->width;
        P.sh = (int)shadow->height;
    }
    *out = P;
    return VCT_OK;
}

static bool traces_specular(const TraceParams& P) { return P.mode == 6u || P.mode == 8u; }

// 64x64-pixel tiles of the specular work list of a width x height image.
static size_t spec_tiles(const vct_gbuffer* g)
{
    return (size_t)((g->width + kSpecTile - 1) / kSpecTile) * (size_t)((g->height + kSpecTile - 1) / kSpecTile);
}

// The specular cone runs in its own roughness-binned pass whenever the mode
// traces it (modes 6 and 8); the diffuse pass then leaves out_s alone.
// `perm` holds spec_tiles(g) * (4096 + 32) words when the mode traces it: the
// permutation, the chunk keys and the chunk order.
template <bool kHalf>
static vct_status launch_trace(hipStream_t s, const typename Vol<kHalf>::T* vol, const TraceParams& P,
                               const vct_gbuffer* g, float4* out_d, float4* out_s, unsigned long long* samples,
                               uint32_t* perm)
{
    const dim3 grid(((g->width + 15) / 16) * ((g->height + 15) / 16));
    if (!traces_specular(P)) {
        hipLaunchKernelGGL((k_trace<kHalf, false>), grid, dim3(256), 0, s, vol, P, g->diffuse, g->normal, g->specular,
                           g->emission, g->depth, out_d, out_s, samples);
        VCT_HIP(hipGetLastError());
        return VCT_OK;
    }
    const int tiles_x = (int)((g->width + kSpecTile - 1) / kSpecTile);
    const size_t tiles = spec_tiles(g);
    const uint32_t nchunks = (uint32_t)(tiles * (kSpecTilePixels / 256));
    uint32_t* keys = perm + tiles * kSpecTilePixels;
    uint32_t* order = keys + nchunks;
    hipLaunchKernelGGL(k_spec_bin, dim3((unsigned)tiles), dim3(256), 0, s, P, g->diffuse, g->specular, g->depth,
                       tiles_x, perm, keys);
    hipLaunchKernelGGL(k_spec_order, dim3(8), dim3(256), 0, s, (const uint32_t*)keys, nchunks, order);
    hipLaunchKernelGGL((k_trace<kHalf, true>), grid, dim3(256), 0, s, vol, P, g->diffuse, g->normal, g->specular,
                       g->emission, g->depth, out_d, out_s, samples);
    hipLaunchKernelGGL(k_trace_specular<kHalf>, dim3((unsigned)(tiles * (kSpecTilePixels / 256))), dim3(256), 0, s,
                       vol, P, (const uint32_t*)perm, (const uint32_t*)order, g->diffuse, g->normal, g->specular,
                       g->emission, g->depth, out_s, samples);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

extern "C" vct_status vct_trace(vct_ctx* c, vct_stream stream, const vct_gbuffer* g, const vct_camera_ubo* cam,
                                const vct_cone_tracing_desc* d, const vct_shadow_map* shadow,
                                const vct_directional_light_desc* light,
                                const vct_directional_light_shadow_desc* light_shadow, float* diffuse_out,
                                float* specular_out, unsigned long long* samples_out)
{
    TraceParams P;
    VCT_TRY(make_trace_params(c, g, cam, d, shadow, light, light_shadow, diffuse_out, specular_out,
                              c ? c->levels : 0, &P));
    if (d->volume_dimension != (float)c->res)
        return fail(VCT_ERR_INVALID_ARGUMENT, "volume_dimension must equal the resolution");
    uint32_t* perm = nullptr;
    if (traces_specular(P)) {
        VCT_TRY(ensure(c->spec_perm, 4u * spec_tiles(g) * (kSpecTilePixels + 2 * (kSpecTilePixels / 256))));
        perm = (uint32_t*)c->spec_perm.ptr;
    }
    if (c->mask_version != c->radiance_version) {
        VCT_TRY(build_occupancy(c, (hipStream_t)stream));
        c->mask_version = c->radiance_version;
    }
    if (c->cfg.radiance_format == VCT_FORMAT_RGBA16F)
        return launch_trace<true>((hipStream_t)stream, (const uint2*)c->radiance.ptr, P, g, (float4*)diffuse_out,
                                  (float4*)specular_out, samples_out, perm);
    return launch_trace<false>((hipStream_t)stream, (const uint32_t*)c->radiance.ptr, P, g, (float4*)diffuse_out,
                               (float4*)specular_out, samples_out, perm);
}

// ---------------------------------------------------------------------------
// Allocation-free trace. The specular work list lives in `scratch`, a static
// device array of kNoAllocTiles 64x64-pixel tiles (32.25 MiB: up to 2048 tiles,
// e.g. 3840x2160); every GPU the library runs on holds its own copy until the
// process ends. Its device address is looked up once per device (at context
// creation) and kept in g_scratch_ptr; the array's name is never used as a
// device pointer on the host. Calls that overlap on one GPU share the array
// and race; the caller serializes them (one stream, or events).
namespace vct {

constexpr size_t kNoAllocTiles = 2048;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kNoAllocTiles * (kSpecTilePixels + 2 * (kSpecTilePixels / 256))];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

vct_status trace_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(VCT_ERR_UNSUPPORTED, "device ordinal %d", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return VCT_OK;
    int cur = 0;
    VCT_HIP(hipGetDevice(&cur));
    if (cur != device) VCT_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr)
        return fail(VCT_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return VCT_OK;
}

}  // namespace vct

extern "C" vct_status vct_trace_noalloc(vct_ctx* c, vct_stream stream, const vct_gbuffer* g,
                                        const vct_camera_ubo* cam, const vct_cone_tracing_desc* d,
                                        const vct_shadow_map* shadow, const vct_directional_light_desc* light,
                                        const vct_directional_light_shadow_desc* light_shadow, float* diffuse_out,
                                        float* specular_out, unsigned long long* samples_out)
{
    TraceParams P;
    VCT_TRY(make_trace_params(c, g, cam, d, shadow, light, light_shadow, diffuse_out, specular_out,
                              c ? c->levels : 0, &P));
    if (d->volume_dimension != (float)c->res)
        return fail(VCT_ERR_INVALID_ARGUMENT, "volume_dimension must equal the resolution");
    uint32_t* perm = nullptr;
    if (traces_specular(P)) {
        const size_t tiles = spec_tiles(g);
        if (tiles > kNoAllocTiles)
            return fail(VCT_ERR_CAPACITY,
                        "a %ux%u image needs %zu specular tiles of 64x64 pixels; the allocation-free trace holds %zu",
                        g->width, g->height, tiles, kNoAllocTiles);
        VCT_TRY(trace_scratch_init(c->device));  // set at vct_create; a no-op here
        perm = (uint32_t*)g_scratch_ptr[c->device];
    }
    // rebuilt on every call: the host-side version check vct_trace uses would
    // not run again when a captured call is replayed
    VCT_TRY(build_occupancy(c, (hipStream_t)stream));
    c->mask_version = c->radiance_version;
    if (c->cfg.radiance_format == VCT_FORMAT_RGBA16F)
        return launch_trace<true>((hipStream_t)stream, (const uint2*)c->radiance.ptr, P, g, (float4*)diffuse_out,
                                  (float4*)specular_out, samples_out, perm);
    return launch_trace<false>((hipStream_t)stream, (const uint32_t*)c->radiance.ptr, P, g, (float4*)diffuse_out,
                               (float4*)specular_out, samples_out, perm);
}

// OctreeVoxelConeTracing::onUpdate (OctreeVoxelConeTracing.cpp:74-104). The
// scene box is a parameter (the shader hard-codes Sponza's,
// voxelConeTracing_Octree.frag:321-322); volume_dimension is the octree
// resolution the walk starts from (the reference pushes 256, :93).
extern "C" vct_status vct_svo_trace(vct_ctx* c, vct_stream stream, const vct_gbuffer* g, const vct_camera_ubo* cam,
                                    const vct_cone_tracing_desc* d, const vct_shadow_map* shadow,
                                    const vct_directional_light_desc* light,
                                    const vct_directional_light_shadow_desc* light_shadow,
                                    const float scene_bb_min[3], const float scene_bb_max[3], float* diffuse_out,
                                    float* specular_out, unsigned long long* samples_out)
{
    TraceParams P;
    VCT_TRY(make_trace_params(c, g, cam, d, shadow, light, light_shadow, diffuse_out, specular_out, 6, &P));
    if (!scene_bb_min || !scene_bb_max) return fail(VCT_ERR_INVALID_ARGUMENT, "null scene box");
    if (!c->svo.built) return fail(VCT_ERR_NOT_READY, "vct_svo_build has not run");
    const float dim = d->volume_dimension;
    if (!(dim >= 2.0f) || dim > 4096.0f || (((uint32_t)dim) & ((uint32_t)dim - 1)) || (float)(uint32_t)dim != dim)
        return fail(VCT_ERR_INVALID_ARGUMENT, "volume_dimension must be a power of two");
    SvoTraceParams S;
    float ext[3];
    for (int k = 0; k < 3; ++k) {
        ext[k] = scene_bb_max[k] - scene_bb_min[k];
        S.center[k] = (scene_bb_min[k] + scene_bb_max[k]) * 0.5f;
    }
    S.ev = std::fmax(ext[0], std::fmax(ext[1], ext[2])) * 0.5f;
    if (!(S.ev > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "empty scene box");
    dim3 grid(((g->width + 15) / 16) * ((g->height + 15) / 16));
    hipLaunchKernelGGL(k_svo_trace, grid, dim3(256), 0, (hipStream_t)stream, (const uint2*)c->svo.pool.ptr, P, S,
                       g->diffuse, g->normal, g->specular, g->emission, g->depth, (float4*)diffuse_out,
                       (float4*)specular_out, samples_out);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}
