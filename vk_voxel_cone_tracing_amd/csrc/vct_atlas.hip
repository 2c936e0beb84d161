// vct_atlas.hip - clipmap maintenance kernels for gfx950: region clear (K3),
// anisotropic opacity/radiance downsample (K4/K5), copy-alpha (K6) and
// toroidal border wrap (K7).
//
// All of these are HBM-streaming kernels. Threads are laid out x-fastest so
// that a wavefront touches contiguous rows of one (level, face) block; the
// 2x2x2 downsample footprint of a thread pair is 16 contiguous bytes per row.
// Arithmetic is FP32 without contraction, in the reference's operation order,
// so the UNORM8 results are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include "vct_internal.hpp"
#include "vct_texel.hpp"

namespace vct {

template <bool kHalf>
struct Texel;

template <>
struct Texel<false> {
    using T = uint32_t;
    __device__ static f4 load(const T* v, size_t i) { return unpack_unorm(v[i]); }
    __device__ static void store(T* v, size_t i, f4 x) { v[i] = pack_unorm(x); }
};

template <>
struct Texel<true> {
    using T = uint2;
    __device__ static f4 load(const T* v, size_t i) { return unpack_half(v[i]); }
    __device__ static void store(T* v, size_t i, f4 x) { v[i] = pack_half(x); }
};

// ---------------------------------------------------------------- K3
template <bool kHalf>
__global__ void k_clear_region(typename Texel<kHalf>::T* __restrict__ vol, int D, int res, int level, int3 minc,
                               uint3 ext, int faces)
{
    uint32_t gx = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gy = blockIdx.y * blockDim.y + threadIdx.y;
    uint32_t gz = blockIdx.z * blockDim.z + threadIdx.z;
    if (gx >= ext.x || gy >= ext.y || gz >= ext.z) return;
    // clipmapCleaning.comp:22-30 with a non-negative (toroidal) modulo: res is a
    // power of two, so & (res - 1) wraps negative region corners too (GLSL %
    // leaves negative operands undefined; the voxelizer's fract() mapping wraps
    // them this way)
    int x = (((int)gx + minc.x) & (res - 1)) + 1;
    int y = (((int)gy + minc.y) & (res - 1)) + 1;
    int z = (((int)gz + minc.z) & (res - 1)) + 1;
    for (int f = 0; f < faces; ++f) {
        size_t i = vol_index(D, level, f, x, y, z);
        if (kHalf) ((uint2*)vol)[i] = make_uint2(0u, 0u);
        else ((uint32_t*)vol)[i] = 0u;
    }
}

// ---------------------------------------------------------------- K6
template <bool kHalf>
__global__ void k_copy_alpha(typename Texel<kHalf>::T* __restrict__ rad, const uint32_t* __restrict__ opa, int D,
                             int res, int level, int faces)
{
    uint32_t gx = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gy = blockIdx.y * blockDim.y + threadIdx.y;
    uint32_t gz = blockIdx.z * blockDim.z + threadIdx.z;
    if (gx >= (uint32_t)res || gy >= (uint32_t)res || gz >= (uint32_t)res) return;
    for (int f = 0; f < faces; ++f) {
        size_t i = vol_index(D, level, f, (int)gx + 1, (int)gy + 1, (int)gz + 1);
        uint32_t src = opa[i];
        if (kHalf) {
            uint2 t = ((uint2*)rad)[i];
            f4 v = unpack_half(t);
            v.w = u8_to_f(src >> 24);
            ((uint2*)rad)[i] = pack_half(v);
        } else {
            // UNORM8 -> float -> UNORM8 reproduces every byte, so the alpha
            // byte moves unchanged (copyAlphaImage.comp:25-28).
            uint32_t* r = (uint32_t*)rad;
            r[i] = (r[i] & 0x00ffffffu) | (src & 0xff000000u);
        }
    }
}

// K3 + K6 for a half-float radiance level that is about to be injected:
// every interior texel becomes (0, 0, 0, opacity alpha), which is what the
// full-level clear followed (after injection) by copy-alpha leaves on the
// texels no fragment reaches; injected texels keep this alpha.
__global__ void k_begin_radiance(uint2* __restrict__ rad, const uint32_t* __restrict__ opa, int D, int res,
                                 int level)
{
    uint32_t gx = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t gy = blockIdx.y * blockDim.y + threadIdx.y;
    uint32_t gz = blockIdx.z * blockDim.z + threadIdx.z;
    if (gx >= (uint32_t)res || gy >= (uint32_t)res || gz >= (uint32_t)res) return;
    for (int f = 0; f < kFaces; ++f) {
        size_t i = vol_index(D, level, f, (int)gx + 1, (int)gy + 1, (int)gz + 1);
        rad[i] = pack_half(f4{0.0f, 0.0f, 0.0f, u8_to_f(opa[i] >> 24)});
    }
}

// ---------------------------------------------------------------- K7
// One thread per texel of one border plane of one (level, face): the border
// texels are those of [0, n)^3 with a coordinate outside [1, res], i.e. on a
// plane at 0 or res + 1; each takes the interior texel its coordinates wrap to
// (borderWrapping.comp:20-36), so no copy reads a texel another one writes.
// A texel on two planes gets the same value twice.
template <typename T>
__global__ __launch_bounds__(256) void k_wrap_border(T* __restrict__ vol, int D, int res, int n, int faces)
{
    const int u = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int v = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int plane = (int)blockIdx.z % 6, lf = (int)blockIdx.z / 6;
    const int l = lf / faces, f = lf % faces;
    const int axis = plane >> 1, c = (plane & 1) ? res + 1 : 0;
    if (u >= n || v >= n || c >= n) return;
    int g[3];
    g[axis] = c;
    g[axis == 0 ? 1 : 0] = u;
    g[axis == 2 ? 1 : 2] = v;
    int r[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = ((g[k] + res - 1) & (res - 1)) + 1;
    vol[vol_index(D, l, f, g[0], g[1], g[2])] = vol[vol_index(D, l, f, r[0], r[1], r[2])];
}

// ---------------------------------------------------------------- K4 / K5

__device__ __forceinline__ int image_coord(int p, int res)
{
    int q = p / res;
    if (q < 0) q = -q;
    return (p + res * (q + 1)) & (res - 1);
}

__device__ __forceinline__ float mixf(float x, float y, float a)
{
#pragma clang fp contract(off)
    return x * (1.0f - a) + y * a;
}

template <bool kHalf, bool kRadiance>
__global__ __launch_bounds__(256) void k_downsample(typename Texel<kHalf>::T* __restrict__ vol, int D, int res,
                                                    int level, int3 prev_min, int region_size)
{
#pragma clang fp contract(off)
    using TX = Texel<kHalf>;
    const int half = res >> 1;
    int gx = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int gy = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    int gz = (int)(blockIdx.z * blockDim.z + threadIdx.z);
    if (gx >= half || gy >= half || gz >= half) return;
    const int cur[3] = {(prev_min.x >> 1) + gx, (prev_min.y >> 1) + gy, (prev_min.z >> 1) + gz};
    const int pm[3] = {prev_min.x, prev_min.y, prev_min.z};
    int wp[3], ps[3];
    for (int k = 0; k < 3; ++k) {
        wp[k] = image_coord(cur[k], res) + 1;
        ps[k] = image_coord(cur[k] << 1, res) + 1;
    }
    // transition band (opacityDownSample.comp:55-64)
    float dist[3];
    for (int k = 0; k < 3; ++k) {
        float center = (float)(pm[k] >> 1) + (float)(half >> 1);
        dist[k] = fabsf((float)cur[k] + 0.5f - center) - 0.5f;
    }
    float lerp = 0.0f;
    const float inv = 1.0f / ((float)region_size + 1.0f);
    const float thr = (float)((half >> 1) - region_size);
    if (dist[0] >= thr || dist[1] >= thr || dist[2] >= thr) {
        float m = fmaxf(dist[0], fmaxf(dist[1], dist[2]));
        lerp = m - (float)((half >> 1) - region_size) + 1.0f;
        lerp = lerp * inv;
    }
    // front/back pairs per face: opacityDownSample.comp:94-138 (face_pair_child)
#pragma unroll
    for (int f = 0; f < kFaces; ++f) {
        f4 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            v[i] = TX::load(vol, vol_index(D, level - 1, f, ps[0] + (i & 1), ps[1] + ((i >> 1) & 1), ps[2] + (i >> 2)));
        const size_t wi = vol_index(D, level, f, wp[0], wp[1], wp[2]);
        f4 self = TX::load(vol, wi);
        if (!kRadiance) {
            float s = 0.0f;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const f4 a = v[face_pair_child(f, p, 0)];
                const f4 b = v[face_pair_child(f, p, 1)];
                s = s + a.w;
                s = s + (1.0f - a.w) * b.w;
            }
            float down = s * 0.25f;
            self.w = mixf(down, self.x, lerp);
            self.y = self.w;
            TX::store(vol, wi, self);
        } else {
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
            f4 out = {mixf(acc[0] * 0.25f, self.x, lerp), mixf(acc[1] * 0.25f, self.y, lerp),
                      mixf(acc[2] * 0.25f, self.z, lerp), mixf(acc[3] * 0.25f, self.w, lerp)};
            TX::store(vol, wi, out);
        }
    }
}

}  // namespace vct

using namespace vct;

static vct_status check_volume(vct_ctx* c, vct_volume v)
{
    if (!c) return fail(VCT_ERR_INVALID_ARGUMENT, "null context");
    if (v != VCT_VOLUME_OPACITY && v != VCT_VOLUME_RADIANCE) return fail(VCT_ERR_INVALID_ARGUMENT, "bad volume");
    return VCT_OK;
}

static bool is_half(const vct_ctx* c, vct_volume v)
{
    return v == VCT_VOLUME_RADIANCE && c->cfg.radiance_format == VCT_FORMAT_RGBA16F;
}

extern "C" {

vct_status vct_clear_region(vct_ctx* c, vct_stream stream, vct_volume v, const vct_image_cleaning_desc* d)
{
    VCT_TRY(check_volume(c, v));
    if (!d) return fail(VCT_ERR_INVALID_ARGUMENT, "null desc");
    if (d->clip_level < 0 || d->clip_level >= c->levels) return fail(VCT_ERR_INVALID_ARGUMENT, "clip_level");
    if (d->clipmap_resolution != c->res) return fail(VCT_ERR_INVALID_ARGUMENT, "clipmap_resolution");
    if (d->face_count > (uint32_t)kFaces) return fail(VCT_ERR_INVALID_ARGUMENT, "face_count");
    for (int k = 0; k < 3; ++k) {
        if (d->region_min_corner[k] < -(1 << 28) || d->region_min_corner[k] > (1 << 28))
            return fail(VCT_ERR_INVALID_ARGUMENT, "region_min_corner out of range");
        if (d->clip_max_extent[k] > (uint32_t)c->res) return fail(VCT_ERR_INVALID_ARGUMENT, "extent > resolution");
    }
    if (!d->clip_max_extent[0] || !d->clip_max_extent[1] || !d->clip_max_extent[2] || !d->face_count) return VCT_OK;
    if (v == VCT_VOLUME_RADIANCE) radiance_written(c);
    dim3 block(64, 2, 2);
    dim3 grid((d->clip_max_extent[0] + 63) / 64, (d->clip_max_extent[1] + 1) / 2, (d->clip_max_extent[2] + 1) / 2);
    int3 minc = make_int3(d->region_min_corner[0], d->region_min_corner[1], d->region_min_corner[2]);
    uint3 ext = make_uint3(d->clip_max_extent[0], d->clip_max_extent[1], d->clip_max_extent[2]);
    hipStream_t s = (hipStream_t)stream;
    if (is_half(c, v))
        hipLaunchKernelGGL(k_clear_region<true>, grid, block, 0, s, (uint2*)c->radiance.ptr, c->D, c->res,
                           d->clip_level, minc, ext, (int)d->face_count);
    else
        hipLaunchKernelGGL(k_clear_region<false>, grid, block, 0, s,
                           (uint32_t*)(v == VCT_VOLUME_OPACITY ? c->opacity.ptr : c->radiance.ptr), c->D, c->res,
                           d->clip_level, minc, ext, (int)d->face_count);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_begin_radiance_level(vct_ctx* c, vct_stream stream, uint32_t level)
{
    if (!c) return fail(VCT_ERR_INVALID_ARGUMENT, "null context");
    if (level >= (uint32_t)c->levels) return fail(VCT_ERR_INVALID_ARGUMENT, "level out of range");
    if (c->cfg.radiance_format != VCT_FORMAT_RGBA16F)
        return fail(VCT_ERR_UNSUPPORTED, "the fused clear / copy-alpha is for RGBA16F radiance");
    hipStream_t s = (hipStream_t)stream;
    dim3 block(64, 2, 2);
    dim3 grid((c->res + 63) / 64, (c->res + 1) / 2, (c->res + 1) / 2);
    radiance_written(c);
    hipLaunchKernelGGL(k_begin_radiance, grid, block, 0, s, (uint2*)c->radiance.ptr, (const uint32_t*)c->opacity.ptr,
                       c->D, c->res, (int)level);
    VCT_HIP(hipGetLastError());
    c->keep_alpha[level] = true;
    return VCT_OK;
}

vct_status vct_copy_alpha(vct_ctx* c, vct_stream stream, const vct_copy_alpha_desc* d)
{
    if (!c || !d) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (d->clip_level < 0 || d->clip_level >= c->levels) return fail(VCT_ERR_INVALID_ARGUMENT, "clip_level");
    if (d->clipmap_resolution != c->res) return fail(VCT_ERR_INVALID_ARGUMENT, "clipmap_resolution");
    if (d->face_count < 0 || d->face_count > kFaces) return fail(VCT_ERR_INVALID_ARGUMENT, "face_count");
    dim3 block(64, 2, 2);
    dim3 grid((c->res + 63) / 64, (c->res + 1) / 2, (c->res + 1) / 2);
    hipStream_t s = (hipStream_t)stream;
    radiance_written(c);
    if (is_half(c, VCT_VOLUME_RADIANCE))
        hipLaunchKernelGGL(k_copy_alpha<true>, grid, block, 0, s, (uint2*)c->radiance.ptr,
                           (const uint32_t*)c->opacity.ptr, c->D, c->res, d->clip_level, d->face_count);
    else
        hipLaunchKernelGGL(k_copy_alpha<false>, grid, block, 0, s, (uint32_t*)c->radiance.ptr,
                           (const uint32_t*)c->opacity.ptr, c->D, c->res, d->clip_level, d->face_count);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_wrap_border(vct_ctx* c, vct_stream stream, vct_volume v, const vct_border_wrapping_desc* d)
{
    VCT_TRY(check_volume(c, v));
    if (!d) return fail(VCT_ERR_INVALID_ARGUMENT, "null desc");
    if (d->clipmap_resolution != (uint32_t)c->res || d->clip_border_width != 2u)
        return fail(VCT_ERR_INVALID_ARGUMENT, "desc must be {resolution, 2, faces, levels}");
    if (d->face_count > (uint32_t)kFaces || d->clip_region_count > (uint32_t)c->levels)
        return fail(VCT_ERR_INVALID_ARGUMENT, "face_count / clip_region_count out of range");
    const int rwb = c->res + 2;
    // BorderWrapper.cpp:139-140 dispatches ((res + border) >> 3) groups of 8.
    const int n = (c->cfg.quirks & VCT_QUIRK_BORDER_SKIP_FAR) ? 8 * (rwb >> 3) : rwb;
    if (d->clip_region_count == 0 || d->face_count == 0) return VCT_OK;
    if (v == VCT_VOLUME_RADIANCE) radiance_written(c);
    dim3 block(64, 4, 1);
    dim3 grid((n + 63) / 64, (n + 3) / 4, 6u * d->clip_region_count * d->face_count);
    hipStream_t s = (hipStream_t)stream;
    if (is_half(c, v))
        hipLaunchKernelGGL(k_wrap_border<uint2>, grid, block, 0, s, (uint2*)c->radiance.ptr, c->D, c->res, n,
                           (int)d->face_count);
    else
        hipLaunchKernelGGL(k_wrap_border<uint32_t>, grid, block, 0, s,
                           (uint32_t*)(v == VCT_VOLUME_OPACITY ? c->opacity.ptr : c->radiance.ptr), c->D, c->res, n,
                           (int)d->face_count);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_downsample(vct_ctx* c, vct_stream stream, vct_volume v, const vct_downsample_desc* d)
{
    VCT_TRY(check_volume(c, v));
    if (!d) return fail(VCT_ERR_INVALID_ARGUMENT, "null desc");
    if (d->clip_level < 1 || d->clip_level >= c->levels)
        return fail(VCT_ERR_INVALID_ARGUMENT, "clip_level must be in [1, levels)");
    if (d->clipmap_resolution != c->res) return fail(VCT_ERR_INVALID_ARGUMENT, "clipmap_resolution");
    if (d->downsample_region_size < 0 || d->downsample_region_size >= c->res / 4)
        return fail(VCT_ERR_INVALID_ARGUMENT, "downsample_region_size");
    const int half = c->res >> 1;
    dim3 block(half < 64 ? half : 64, 2, 2);
    dim3 grid((half + block.x - 1) / block.x, (half + 1) / 2, (half + 1) / 2);
    int3 pm = make_int3(d->prev_region_min_corner[0], d->prev_region_min_corner[1], d->prev_region_min_corner[2]);
    hipStream_t s = (hipStream_t)stream;
    if (v == VCT_VOLUME_RADIANCE) radiance_written(c);
    if (v == VCT_VOLUME_OPACITY)
        hipLaunchKernelGGL((k_downsample<false, false>), grid, block, 0, s, (uint32_t*)c->opacity.ptr, c->D, c->res,
                           d->clip_level, pm, d->downsample_region_size);
    else if (is_half(c, v))
        hipLaunchKernelGGL((k_downsample<true, true>), grid, block, 0, s, (uint2*)c->radiance.ptr, c->D, c->res,
                           d->clip_level, pm, d->downsample_region_size);
    else
        hipLaunchKernelGGL((k_downsample<false, true>), grid, block, 0, s, (uint32_t*)c->radiance.ptr, c->D, c->res,
                           d->clip_level, pm, d->downsample_region_size);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // extern "C"
