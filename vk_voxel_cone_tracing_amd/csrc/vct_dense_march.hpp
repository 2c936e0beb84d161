// vct_dense_march.hpp - the cone march over the dense mip pyramid, shared by
// the forward trace and the adjoint (which recomputes the march per cone).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "vct_dense.hpp"
#include "vct_texel.hpp"

namespace vct {
namespace dense {

constexpr int kMaxLevels = 11;

struct Pyramid {
    unsigned long long offset[kMaxLevels];  // texel offset of each level
    unsigned long long block[kMaxLevels];   // texels per face block
    float inv_vs[kMaxLevels];               // n_m / extent
    int n[kMaxLevels];
    int nlevels;
    float vs;  // level-0 voxel size
};

struct March {
    float coef;  // 2 tan(aperture / 2)
    float max_dist;
    float step_factor;
    float start_offset;
    float decay;
    float ao_factor;
    float diffuse_intensity;
    int ncones;
    int cones32;
    float vp_inv[16];
    float grid_min[3];
};

__constant__ float c_dense_dirs16[16][3] = {
    {0.57735f, 0.57735f, 0.57735f},       {0.57735f, -0.57735f, -0.57735f},
    {-0.57735f, 0.57735f, -0.57735f},     {-0.57735f, -0.57735f, 0.57735f},
    {-0.903007f, -0.182696f, -0.388844f}, {-0.903007f, 0.182696f, 0.388844f},
    {0.903007f, -0.182696f, 0.388844f},   {0.903007f, 0.182696f, -0.388844f},
    {-0.388844f, -0.903007f, -0.182696f}, {0.388844f, -0.903007f, 0.182696f},
    {0.388844f, 0.903007f, -0.182696f},   {-0.388844f, 0.903007f, 0.182696f},
    {-0.182696f, -0.388844f, -0.903007f}, {0.182696f, 0.388844f, -0.903007f},
    {-0.182696f, 0.388844f, 0.903007f},   {0.182696f, -0.388844f, 0.903007f}};

__constant__ float c_dense_dirs32[32][3] = {
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

__device__ __forceinline__ const float* cone_dir(const March& M, int i)
{
    return M.cones32 ? c_dense_dirs32[i] : c_dense_dirs16[i];
}

// A texel's channels: the floats of an RGBA32F grid, or the bytes (0..255,
// not yet divided by 255) of an RGBA8 grid; trilinear() applies the 1/255
// once to its result instead of to each of the 8 texels. (Bytes passed as
// half-float denormals to v_fma_mix_f32 take fewer instructions but ran 5%
// slower: the byte conversions issue beside the FMAs, the mixed FMAs do not.)
template <bool kF32>
__device__ __forceinline__ f4 texel(const void* __restrict__ p)
{
    if (kF32) {
        const float4 t = *(const float4*)p;
        return {t.x, t.y, t.z, t.w};
    }
    const uint32_t t = *(const uint32_t*)p;
    return {(float)(t & 0xffu), (float)((t >> 8) & 0xffu), (float)((t >> 16) & 0xffu), (float)(t >> 24)};
}

// Footprint of one trilinear fetch of level m (bordered indices). Texel
// indices are 32-bit: a pyramid of at most 1024^3 holds about 2.0e9 texels
// (make_params refuses more than 2^32).
struct Footprint {
    uint32_t base;  // texel index of the (0,0,0) corner inside face 0
    uint32_t dy, dz;
    float fx, fy, fz;
    bool valid;
};

__device__ __forceinline__ Footprint footprint(const Pyramid& P, int m, const float p[3])
{
    Footprint F;
    const int n = P.n[m];
    const uint32_t D = (uint32_t)n + 2u;
    uint32_t i0[3];
    float fr[3];
    F.valid = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float u = p[k] * P.inv_vs[m] - 0.5f;
        const float f = floorf(u);
        fr[k] = u - f;
        if (!(f >= -1.0f) || f > (float)(n - 1)) F.valid = false;
        // f + 1 lies in [0, n] where the footprint is valid; the median only
        // keeps the conversion defined for the others, whose texels are not read
        i0[k] = (uint32_t)((int)__builtin_amdgcn_fmed3f(f, -1.0f, (float)(n - 1)) + 1);
    }
    // (i2 D + i1) < 2^24 and D < 2^24: two 24-bit multiply-adds
    F.base = (uint32_t)P.offset[m] + __umul24(__umul24(i0[2], D) + i0[1], D) + i0[0];
    F.dy = D;
    F.dz = D * D;
    F.fx = fr[0];
    F.fy = fr[1];
    F.fz = fr[2];
    return F;
}

template <bool kF32>
__device__ __forceinline__ f4 trilinear(const void* __restrict__ data, const Footprint& F, uint32_t face_off)
{
    using T = typename std::conditional<kF32, float4, uint32_t>::type;
    const T* r0 = (const T*)data + (F.base + face_off);  // rows (y, z), (y+1, z), (y, z+1), (y+1, z+1)
    const T* r1 = r0 + F.dy;
    const T* r2 = r0 + F.dz;
    const T* r3 = r2 + F.dy;
    const f4 t000 = texel<kF32>(r0), t100 = texel<kF32>(r0 + 1);
    const f4 t010 = texel<kF32>(r1), t110 = texel<kF32>(r1 + 1);
    const f4 t001 = texel<kF32>(r2), t101 = texel<kF32>(r2 + 1);
    const f4 t011 = texel<kF32>(r3), t111 = texel<kF32>(r3 + 1);
    const float ax = F.fx, ay = F.fy, az = F.fz;
    const float w000 = (1.0f - ax) * (1.0f - ay) * (1.0f - az), w100 = ax * (1.0f - ay) * (1.0f - az);
    const float w010 = (1.0f - ax) * ay * (1.0f - az), w110 = ax * ay * (1.0f - az);
    const float w001 = (1.0f - ax) * (1.0f - ay) * az, w101 = ax * (1.0f - ay) * az;
    const float w011 = (1.0f - ax) * ay * az, w111 = ax * ay * az;
    f4 r;
#define VCT_T(c) r.c = w000 * t000.c + w100 * t100.c + w010 * t010.c + w110 * t110.c + w001 * t001.c + w101 * t101.c + \
                       w011 * t011.c + w111 * t111.c
    VCT_T(x);
    VCT_T(y);
    VCT_T(z);
    VCT_T(w);
#undef VCT_T
    if (!kF32) {
        const float s = 1.0f / 255.0f;
        r.x *= s;
        r.y *= s;
        r.z *= s;
        r.w *= s;
    }
    return r;
}

// The three-face sample of level m (one face at level 0).
template <bool kF32>
__device__ __forceinline__ f4 sample(const void* __restrict__ data, const Pyramid& P, int m, const float p[3],
                                     const int face[3], const float w[3])
{
    const Footprint F = footprint(P, m, p);
    if (!F.valid) return {0.0f, 0.0f, 0.0f, 0.0f};
    if (m == 0) {
        const f4 t = trilinear<kF32>(data, F, 0);
        const float ws = w[0] + w[1] + w[2];
        return {t.x * ws, t.y * ws, t.z * ws, t.w * ws};
    }
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const f4 t = trilinear<kF32>(data, F, (uint32_t)face[a] * (uint32_t)P.block[m]);
        acc.x += t.x * w[a];
        acc.y += t.y * w[a];
        acc.z += t.z * w[a];
        acc.w += t.w * w[a];
    }
    return acc;
}

// One sample of the march: the level's three-face trilinear sample, lerped
// with the next level when the level is fractional.
template <bool kF32>
__device__ __forceinline__ f4 level_sample(const void* __restrict__ data, const Pyramid& P, const float pos[3],
                                           float lvl, const int face[3], const float w[3])
{
    const int lo = (int)floorf(lvl), hi = (int)ceilf(lvl);
    const float fl = lvl - floorf(lvl);
    f4 s = sample<kF32>(data, P, lo, pos, face, w);
    if (hi != lo) {
        const f4 b = sample<kF32>(data, P, hi, pos, face, w);
        s.x = s.x * (1.0f - fl) + b.x * fl;
        s.y = s.y * (1.0f - fl) + b.y * fl;
        s.z = s.z * (1.0f - fl) + b.z * fl;
        s.w = s.w * (1.0f - fl) + b.w * fl;
    }
    return s;
}

// The march's transcendental steps on the hardware units (v_log_f32,
// v_exp_f32): log2f and exp2f are these instructions for every argument but
// denormal inputs or results, which the march's level and voxel-size terms
// never produce; (1 - a)^c is exp2(c log2(1 - a)) as in the clipmap trace
// (1 - a = 0 gives log2 -inf and exp2 0, the power's value at 0). The adjoint
// replays the march with the same functions.
__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float hw_pow(float base, float c) { return hw_exp2(c * hw_log2(base)); }

__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// The step's constants of accumulate(): the opacity correction seg / vs_l as
// seg x (1 / vs) x 2^-l, and the reciprocal of the occlusion divisor
// 1 + (t + vs_l) decay (hardware reciprocal, 1 ulp). The adjoint's replay
// uses this function too, so forward and replay agree bit for bit.
__device__ __forceinline__ void step_consts(float lvl, float step, float seg, float vs, float inv_vs, float decay,
                                            float& corr, float& iq)
{
    const float vsl = vs * hw_exp2(lvl);
    corr = seg * (inv_vs * hw_exp2(-lvl));
    iq = hw_rcp(1.0f + (step + vsl) * decay);
}

// Front-to-back accumulation of one sample (voxelConeTracing.frag:370-386)
// with the step's constants; returns (1 - s.a)^corr for the adjoint.
__device__ __forceinline__ float accumulate(f4& result, float& occlusion, const f4& s, float corr, float iq)
{
    const float pc = hw_pow(fmaxf(1.0f - s.w, 0.0f), corr);  // 1 - a >= 0
    const float op = fminf(fmaxf(1.0f - pc, 0.0f), 1.0f);
    const float blend = fminf(fmaxf(1.0f - result.w, 0.0f), 1.0f);
    result.x += blend * (s.x * corr);
    result.y += blend * (s.y * corr);
    result.z += blend * (s.z * corr);
    result.w += blend * op;
    occlusion += (1.0f - occlusion) * op * iq;
    return pc;
}

__device__ __forceinline__ void accumulate(f4& result, float& occlusion, const f4& s, float lvl, float step, float seg,
                                           float vs, float inv_vs, float decay)
{
    float corr, iq;
    step_consts(lvl, step, seg, vs, inv_vs, decay, corr, iq);
    accumulate(result, occlusion, s, corr, iq);
}

// A cone's slot of the adjoint tape: the sample of accumulated step k goes
// to at[k * 64] (the lane's column of the wave's records), for k < cap;
// `count` ends as the number of accumulated steps.
struct TapeSlot {
    float4* at;
    uint32_t cap;
    uint32_t count;
};

__device__ __forceinline__ void tape_put(TapeSlot* t, const f4& s)
{
    if (t->count < t->cap) t->at[(size_t)t->count * 64u] = make_float4(s.x, s.y, s.z, s.w);
    ++t->count;
}

// traceCone (voxelConeTracing.frag:341-392) with the dense level rule.
// The step sequence t_k does not depend on the data (only the loop exit on
// full occlusion does), so with kAhead the samples of steps k and k+1 are
// fetched together and step k+1 is accumulated only if the loop would have
// reached it: the same arithmetic in the same order, with twice the loads in
// flight per lane. With a tape, every accumulated sample is also written to
// the cone's tape slot.
template <bool kF32, bool kAhead = false>
__device__ f4 march(const void* __restrict__ data, const Pyramid& P, const March& M, const float start_in[3],
                    const float dir[3], uint32_t& steps, TapeSlot* tape = nullptr)
{
    const float vs = P.vs;
    float start[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * vs * M.start_offset * 0.5f;
    float step = 0.0f, diameter = vs, occlusion = 0.0f, seg = vs;
    f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
    const int face[3] = {dir[0] > 0.0f ? 0 : 1, dir[1] > 0.0f ? 2 : 3, dir[2] > 0.0f ? 4 : 5};
    const float w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
    const float max_level = (float)(P.nlevels - 1);
    const float inv_vs = 1.0f / vs;
    while (step < M.max_dist && occlusion < 1.0f) {
        float pos[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pos[k] = start[k] + dir[k] * step;
        const float lvl = fminf(fmaxf(hw_log2(diameter * inv_vs), 0.0f), max_level);
        if (!kAhead) {
            const f4 s = level_sample<kF32>(data, P, pos, lvl, face, w);
            accumulate(result, occlusion, s, lvl, step, seg, vs, inv_vs, M.decay);
            if (tape) tape_put(tape, s);
            const float prev = step;
            step += fmaxf(diameter, vs) * M.step_factor;
            seg = step - prev;
            diameter = step * M.coef;
            ++steps;
        } else {
            const float step_b = step + fmaxf(diameter, vs) * M.step_factor;
            const float seg_b = step_b - step;
            const float diam_b = step_b * M.coef;
            float pos_b[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) pos_b[k] = start[k] + dir[k] * step_b;
            const float lvl_b = fminf(fmaxf(hw_log2(diam_b * inv_vs), 0.0f), max_level);
            const f4 sa = level_sample<kF32>(data, P, pos, lvl, face, w);
            const f4 sb = level_sample<kF32>(data, P, pos_b, lvl_b, face, w);
            accumulate(result, occlusion, sa, lvl, step, seg, vs, inv_vs, M.decay);
            if (tape) tape_put(tape, sa);
            ++steps;
            step = step_b;
            seg = seg_b;
            diameter = diam_b;
            if (!(step < M.max_dist && occlusion < 1.0f)) break;
            accumulate(result, occlusion, sb, lvl_b, step, seg, vs, inv_vs, M.decay);
            if (tape) tape_put(tape, sb);
            ++steps;
            const float prev = step;
            step += fmaxf(diameter, vs) * M.step_factor;
            seg = step - prev;
            diameter = step * M.coef;
        }
    }
    return {result.x, result.y, result.z, 1.0f - occlusion};
}

__device__ __forceinline__ void pixel_setup(const Pyramid& P, const March& M, int px, int py, int width, int height,
                                            float depth, const uint16_t* __restrict__ nrm, uint32_t diffuse,
                                            float start[3], float normal[3], float albedo[3])
{
    const float q[4] = {(((float)px + 0.5f) / (float)width) * 2.0f - 1.0f,
                        (((float)py + 0.5f) / (float)height) * 2.0f - 1.0f, depth, 1.0f};
    float wp[4];
    for (int k = 0; k < 4; ++k)
        wp[k] = M.vp_inv[k] * q[0] + M.vp_inv[4 + k] * q[1] + M.vp_inv[8 + k] * q[2] + M.vp_inv[12 + k] * q[3];
    const f4 nf = unpack_half(*(const uint2*)nrm);
    normal[0] = nf.x * 2.0f - 1.0f;
    normal[1] = nf.y * 2.0f - 1.0f;
    normal[2] = nf.z * 2.0f - 1.0f;
    const float l = sqrtf(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    for (int k = 0; k < 3; ++k) {
        normal[k] /= l;
        const float world = wp[k] / wp[3] - M.grid_min[k];
        start[k] = world + normal[k] * P.vs * M.start_offset;
    }
    albedo[0] = u8_to_f(diffuse & 0xffu);
    albedo[1] = u8_to_f((diffuse >> 8) & 0xffu);
    albedo[2] = u8_to_f((diffuse >> 16) & 0xffu);
}

// A lane's part of the adjoint tape: cone i's samples from col + i * stride
// (float4 units, one column of 64), its accumulated step count at
// counts[i * 64] (kTapeOverflow when the march took more than cap steps).
struct LaneTape {
    float4* col;
    size_t stride;
    uint8_t* counts;
    uint32_t cap;
};

constexpr uint8_t kTapeOverflow = 255;

// One pixel of the dense trace: 16 (or 32) cones weighted by cos, output
// (indirect diffuse x albedo x intensity, ambient occlusion). With a tape,
// the counts of the cones it marches are written (the caller zeroes them).
template <bool kF32, bool kAhead>
__device__ __forceinline__ float4 dense_pixel(const void* __restrict__ data, const Pyramid& P, const March& M, int px,
                                              int py, int width, int height, const uint32_t* __restrict__ g_diffuse,
                                              const uint16_t* __restrict__ g_normal,
                                              const float* __restrict__ g_depth, uint32_t& steps,
                                              const LaneTape* lt = nullptr)
{
    const size_t idx = (size_t)py * (size_t)width + (size_t)px;
    const float depth = g_depth[idx];
    if (depth == 1.0f) return make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float start[3], normal[3], albedo[3];
    pixel_setup(P, M, px, py, width, height, depth, g_normal + 4 * idx, g_diffuse[idx], start, normal, albedo);
    f4 ind = {0.0f, 0.0f, 0.0f, 1.0f};
    for (int i = 0; i < M.ncones; ++i) {
        const float* d = cone_dir(M, i);
        const float ct = normal[0] * d[0] + normal[1] * d[1] + normal[2] * d[2];
        if (ct < 0.0f) continue;
        f4 c;
        if (lt) {
            TapeSlot t{lt->col + (size_t)i * lt->stride, lt->cap, 0u};
            c = march<kF32, kAhead>(data, P, M, start, d, steps, &t);
            lt->counts[i * 64] = t.count > lt->cap ? kTapeOverflow : (uint8_t)t.count;
        } else {
            c = march<kF32, kAhead>(data, P, M, start, d, steps);
        }
        ind.x += c.x * ct;
        ind.y += c.y * ct;
        ind.z += c.z * ct;
        ind.w += c.w * ct;
    }
    const float inv = 1.0f / (float)M.ncones;
    return make_float4(ind.x * inv * albedo[0] * M.diffuse_intensity, ind.y * inv * albedo[1] * M.diffuse_intensity,
                       ind.z * inv * albedo[2] * M.diffuse_intensity, ind.w * inv * M.ao_factor);
}

// Longest march the adjoint keeps per-step state for (per-lane arrays).
constexpr int kAdjMaxSteps = 40;

// Bytes of the adjoint sweep's per-(wave, cone) header.
constexpr size_t kSweepHeaderBytes = 72;

// Step count of the march for these parameters, on the host: the step
// sequence does not depend on the data, only the exit on full occlusion does.
inline int march_step_count(const Pyramid& P, const March& M)
{
    float step = 0.0f, diameter = P.vs;
    int n = 0;
    while (step < M.max_dist && n <= kAdjMaxSteps + 1) {
        ++n;
        step += std::fmax(diameter, P.vs) * M.step_factor;
        diameter = step * M.coef;
    }
    return n;
}

// The adjoint tape of one view (vct_dense_trace_tape): per slot (a wave of
// 8x8 pixels of a 16x16 tile, wave-major within the tile, and a cone), `cap`
// steps x 64 lanes of float4 samples, slot after slot; then one step-count
// byte per (slot, lane). cap is the host's step count plus two steps of
// margin (device rounding may differ in the last bit). Behind them, the
// backward's own work area for the view: the sweep's headers (one per slot
// at most), the pixels' cone start points and the sweep / splat counters, so
// that backward calls of different views share nothing but the grid.
struct TapeLayout {
    size_t slots;
    uint32_t cap;
    size_t sample_bytes, count_bytes;
    size_t header_off, start_off, ctr_off;
    size_t bytes;
};

inline TapeLayout tape_layout(const Pyramid& P, const March& M, uint32_t width, uint32_t height,
                              size_t header_bytes)
{
    auto up = [](size_t v) { return (v + 255u) / 256u * 256u; };
    TapeLayout T;
    const size_t tiles = (size_t)((width + 15u) / 16u) * (size_t)((height + 15u) / 16u);
    T.slots = tiles * 4u * (size_t)M.ncones;
    T.cap = (uint32_t)march_step_count(P, M) + 2u;
    T.sample_bytes = T.slots * T.cap * 64u * sizeof(float4);
    T.count_bytes = T.slots * 64u;
    T.header_off = up(T.sample_bytes + T.count_bytes);
    T.start_off = up(T.header_off + T.slots * header_bytes);
    T.ctr_off = up(T.start_off + (size_t)width * height * sizeof(float4));
    T.bytes = T.ctr_off + 16u;
    return T;
}

inline vct_status make_params(const vct_dense* g, const vct_camera_ubo* cam, const vct_dense_trace_desc* d,
                              Pyramid* P, March* M)
{
    if (g->nlevels > kMaxLevels) return fail(VCT_ERR_UNSUPPORTED, "too many levels");
    const int last = g->nlevels - 1;
    if (last < 0 || (unsigned long long)g->offset[last] + (last ? 6ull : 1ull) * g->block[last] > 0xffffffffull)
        return fail(VCT_ERR_UNSUPPORTED, "the march indexes texels in 32 bits; this pyramid holds more");
    if (!(d->aperture > 0.0f) || !(d->max_distance > 0.0f))
        return fail(VCT_ERR_INVALID_ARGUMENT, "aperture and max_distance must be positive");
    for (int m = 0; m < g->nlevels; ++m) {
        P->offset[m] = g->offset[m];
        P->block[m] = g->block[m];
        P->n[m] = g->res >> m;
        P->inv_vs[m] = (float)(g->res >> m) / g->desc.extent;
    }
    P->nlevels = g->nlevels;
    P->vs = g->desc.extent / (float)g->res;
    M->coef = 2.0f * std::tan(d->aperture * 0.5f);
    M->max_dist = d->max_distance;
    M->step_factor = d->min_trace_step_factor > 0.2f ? d->min_trace_step_factor : 0.2f;
    M->start_offset = d->trace_start_offset;
    M->decay = d->occlusion_decay;
    M->ao_factor = d->ambient_occlusion_factor;
    M->diffuse_intensity = d->indirect_diffuse_intensity;
    M->cones32 = d->enable_32_cones != 0;
    M->ncones = M->cones32 ? 32 : 16;
    for (int k = 0; k < 16; ++k) M->vp_inv[k] = cam->view_proj_inv[k];
    for (int k = 0; k < 3; ++k) M->grid_min[k] = g->desc.grid_min[k];
    return VCT_OK;
}

}  // namespace dense
}  // namespace vct
