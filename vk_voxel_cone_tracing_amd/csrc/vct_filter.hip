// vct_filter.hip - the step after the cone trace: SpecularFilterPass
// (specularFilter.frag:25-53). Per pixel: the traced diffuse contribution plus
// the specular contribution filtered by the 32-direction x 8-ring blur
// (filter.glsl:9-24) or the 15x15 bilateral filter (filter.glsl:26-63), then
// the optional Uncharted2 tonemap (tonemapping.glsl:9-26). Inputs are sampled
// like the reference's "VCTSampler": bilinear, clamp to edge. One lane per
// pixel, 16x16 tiles in XCD order; the inputs are the two RGBA32F images the
// trace wrote, so the taps of a tile stay in L2.
#include <hip/hip_runtime.h>

#include "vct_internal.hpp"

static_assert(sizeof(vct_specular_filter_desc) == 16, "SpecularFilterPassDesc layout");

namespace vct {

__device__ __forceinline__ float4 tex_clamp_linear(const float4* __restrict__ img, int W, int H, float s, float t)
{
#pragma clang fp contract(off)
    const float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int i0 = (int)fu, j0 = (int)fv;
    float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            int i = i0 + dx, j = j0 + dy;
            i = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
            j = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
            const float w = (dx ? a : 1.0f - a) * (dy ? b : 1.0f - b);
            const float4 p = img[(size_t)j * (size_t)W + (size_t)i];
            r.x += w * p.x;
            r.y += w * p.y;
            r.z += w * p.z;
            r.w += w * p.w;
        }
    return r;
}

__constant__ float c_kernel15[15] = {0.031225216f, 0.033322271f, 0.035206333f, 0.036826804f, 0.038138565f,
                                     0.039104044f, 0.039695028f, 0.039894000f, 0.039695028f, 0.039104044f,
                                     0.038138565f, 0.036826804f, 0.035206333f, 0.033322271f, 0.031225216f};

__device__ __forceinline__ float uncharted2(float x)
{
#pragma clang fp contract(off)
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

__global__ __launch_bounds__(256) void k_specular_filter(const float4* __restrict__ diffuse,
                                                         const float4* __restrict__ specular, int W, int H,
                                                         vct_specular_filter_desc d, float4* __restrict__ out)
{
#pragma clang fp contract(off)
    const unsigned tiles_x = ((unsigned)W + 15u) / 16u;
    const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
    const int px = (int)(tile % tiles_x) * 16 + (int)(threadIdx.x & 15);
    const int py = (int)(tile / tiles_x) * 16 + (int)(threadIdx.x >> 4);
    if (px >= W || py >= H) return;
    const float s = ((float)px + 0.5f) / (float)W, t = ((float)py + 0.5f) / (float)H;
    float4 col = tex_clamp_linear(diffuse, W, H, s, t);
    float4 sc = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (d.filter_method == 1) {
        // filter.glsl:9-24: the direction loop runs on a float accumulator
        const float two_pi = 6.28318530718f;
        const float rx = 0.01f / (float)W, ry = 0.01f / (float)H;
        float4 c = tex_clamp_linear(specular, W, H, s, t);
        for (float a = 0.0f; a < two_pi; a += two_pi / 32.0f) {
            const float ca = cosf(a), sa = sinf(a);
            for (float i = 1.0f / 8.0f; i <= 1.0f; i += 1.0f / 8.0f) {
                const float4 q = tex_clamp_linear(specular, W, H, s + ca * rx * i, t + sa * ry * i);
                c.x += q.x;
                c.y += q.y;
                c.z += q.z;
                c.w += q.w;
            }
        }
        const float n = 8.0f * 32.0f - 15.0f;
        sc = make_float4(c.x / n, c.y / n, c.z / n, c.w / n);
    } else {
        // filter.glsl:26-63
        const float bsigma = 10.0f;
        const float4 center = tex_clamp_linear(specular, W, H, s, t);
        const float bZ = 1.0f / (0.39894f * expf(-0.5f * 0.0f * 0.0f / (bsigma * bsigma)) / bsigma);
        const float ix = 1.0f / (float)W, iy = 1.0f / (float)H;
        float Z = 0.0f, ax = 0.0f, ay = 0.0f, az = 0.0f;
        for (int i = -7; i <= 7; ++i)
            for (int j = -7; j <= 7; ++j) {
                const float4 q = tex_clamp_linear(specular, W, H, s + (float)i * ix, t + (float)j * iy);
                const float dx = q.x - center.x, dy = q.y - center.y, dz = q.z - center.z;
                const float dd = dx * dx + dy * dy + dz * dz;
                const float f = (0.39894f * expf(-0.5f * dd / (bsigma * bsigma)) / bsigma) * bZ * c_kernel15[7 + j] *
                                c_kernel15[7 + i];
                Z += f;
                ax += f * q.x;
                ay += f * q.y;
                az += f * q.z;
            }
        sc.x += ax / Z;
        sc.y += ay / Z;
        sc.z += az / Z;
    }
    col.x += sc.x;
    col.y += sc.y;
    col.z += sc.z;
    if (d.tonemap_enable == 1) {
        const float white = 1.0f / uncharted2(11.2f);
        const float g = 1.0f / d.tonemap_gamma;
        col = make_float4(powf(uncharted2(col.x * d.tonemap_exposure) * white, g),
                          powf(uncharted2(col.y * d.tonemap_exposure) * white, g),
                          powf(uncharted2(col.z * d.tonemap_exposure) * white, g), col.w);
    }
    out[(size_t)py * (size_t)W + (size_t)px] = col;
}

}  // namespace vct

using namespace vct;

extern "C" vct_status vct_specular_filter(vct_stream stream, const float* diffuse, const float* specular, uint32_t width,
                                          uint32_t height, const vct_specular_filter_desc* desc, float* out)
{
    if (!diffuse || !specular || !desc || !out || !width || !height)
        return fail(VCT_ERR_INVALID_ARGUMENT, "null argument or empty image");
    if (out == diffuse || out == specular) return fail(VCT_ERR_INVALID_ARGUMENT, "out must not alias an input");
    const unsigned tiles = ((width + 15) / 16) * ((height + 15) / 16);
    hipLaunchKernelGGL(k_specular_filter, dim3(tiles), dim3(256), 0, (hipStream_t)stream, (const float4*)diffuse,
                       (const float4*)specular, (int)width, (int)height, *desc, (float4*)out);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}
