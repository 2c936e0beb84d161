// vct_dense_trace.hip - cone tracing over the dense mip pyramid (gfx950).
//
// Per pixel: the reference's 16 (or 32) diffuse cones (voxelConeTracing.frag
// :166-203) marched through the dense pyramid; output RGBA32F =
// (indirect diffuse x albedo x intensity, ambient occlusion). Level 0 has one
// face, so its sample is one trilinear fetch scaled by the weight sum; levels
// >= 1 blend the three direction-selected faces. Every face block has a zero
// border, so only positions more than one texel outside the grid need a test.
#include <hip/hip_runtime.h>

#include "vct_dense.hpp"
#include "vct_dense_march.hpp"

namespace vct {
namespace dense {

// Register budget of four waves per SIMD: a 512x512 view is 1024
// workgroups, one round of four waves per SIMD, so a five-wave budget buys no
// occupancy and only makes the scheduler keep fewer texel loads in flight
// (64 views of 512x512: 28.8 -> 22.6 ms).
template <bool kF32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_dense_trace(const void* __restrict__ data, Pyramid P, March M,
                                                     const uint32_t* __restrict__ g_diffuse,
                                                     const uint16_t* __restrict__ g_normal,
                                                     const float* __restrict__ g_depth, int width, int height,
                                                     float4* __restrict__ out, unsigned long long* __restrict__ samples)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned tiles_x = ((unsigned)width + 15u) / 16u;
    const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    uint32_t steps = 0;
    if (px < width && py < height)
        out[(size_t)py * (size_t)width + (size_t)px] =
            dense_pixel<kF32, false>(data, P, M, px, py, width, height, g_diffuse, g_normal, g_depth, steps);
    if (samples) {
        unsigned long long v = steps;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0 && v) atomicAdd(samples, v);
    }
}

// The forward trace that also writes the adjoint tape (RGBA32F grids): the
// same pixels, plus every accumulated sample of every cone in the
// (wave, cone) slot layout of TapeLayout and each lane's step count per cone.
__global__ __launch_bounds__(256) void k_dense_trace_tape(const void* __restrict__ data, Pyramid P, March M,
                                                          const uint32_t* __restrict__ g_diffuse,
                                                          const uint16_t* __restrict__ g_normal,
                                                          const float* __restrict__ g_depth, int width, int height,
                                                          float4* __restrict__ out,
                                                          unsigned long long* __restrict__ samples,
                                                          float4* __restrict__ tape, uint8_t* __restrict__ counts,
                                                          uint32_t cap)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned tiles_x = ((unsigned)width + 15u) / 16u;
    const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    const size_t slot0 = ((size_t)tile * 4u + (size_t)wave) * (size_t)M.ncones;
    LaneTape lt{tape + slot0 * cap * 64u + lane, (size_t)cap * 64u, counts + slot0 * 64u + lane, cap};
    for (int i = 0; i < M.ncones; ++i) lt.counts[i * 64] = 0;
    uint32_t steps = 0;
    if (px < width && py < height)
        out[(size_t)py * (size_t)width + (size_t)px] =
            dense_pixel<true, false>(data, P, M, px, py, width, height, g_diffuse, g_normal, g_depth, steps, &lt);
    if (samples) {
        unsigned long long v = steps;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0 && v) atomicAdd(samples, v);
    }
}

}  // namespace dense
}  // namespace vct

using namespace vct;

extern "C" vct_status vct_dense_trace(vct_dense* g, vct_stream stream, const vct_gbuffer* gb,
                                      const vct_camera_ubo* cam, const vct_dense_trace_desc* d, float* out,
                                      unsigned long long* samples_out)
{
    if (!g || !gb || !cam || !d || !out) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!gb->diffuse || !gb->normal || !gb->depth || !gb->width || !gb->height)
        return fail(VCT_ERR_INVALID_ARGUMENT, "G-buffer needs diffuse, normal and depth");
    dense::Pyramid P;
    dense::March M;
    VCT_TRY(dense::make_params(g, cam, d, &P, &M));
    dim3 grid(((gb->width + 15) / 16) * ((gb->height + 15) / 16));
    hipStream_t s = (hipStream_t)stream;
    if (g->texel_bytes == 16)
        hipLaunchKernelGGL(dense::k_dense_trace<true>, grid, dim3(256), 0, s, g->data.ptr, P, M, gb->diffuse,
                           gb->normal, gb->depth, (int)gb->width, (int)gb->height, (float4*)out, samples_out);
    else
        hipLaunchKernelGGL(dense::k_dense_trace<false>, grid, dim3(256), 0, s, g->data.ptr, P, M, gb->diffuse,
                           gb->normal, gb->depth, (int)gb->width, (int)gb->height, (float4*)out, samples_out);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

extern "C" vct_status vct_dense_tape_bytes(vct_dense* g, const vct_gbuffer* gb, const vct_camera_ubo* cam,
                                           const vct_dense_trace_desc* d, size_t* bytes)
{
    if (!g || !gb || !cam || !d || !bytes) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!gb->width || !gb->height) return fail(VCT_ERR_INVALID_ARGUMENT, "empty G-buffer");
    dense::Pyramid P;
    dense::March M;
    VCT_TRY(dense::make_params(g, cam, d, &P, &M));
    const dense::TapeLayout T = dense::tape_layout(P, M, gb->width, gb->height, dense::kSweepHeaderBytes);
    if (T.cap > (uint32_t)dense::kAdjMaxSteps)
        return fail(VCT_ERR_UNSUPPORTED, "the march takes more than %d steps", dense::kAdjMaxSteps - 2);
    *bytes = T.bytes;
    return VCT_OK;
}

extern "C" vct_status vct_dense_trace_tape(vct_dense* g, vct_stream stream, const vct_gbuffer* gb,
                                           const vct_camera_ubo* cam, const vct_dense_trace_desc* d, float* out,
                                           unsigned long long* samples_out, void* tape, size_t tape_bytes)
{
    if (!g || !gb || !cam || !d || !out || !tape) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!gb->diffuse || !gb->normal || !gb->depth || !gb->width || !gb->height)
        return fail(VCT_ERR_INVALID_ARGUMENT, "G-buffer needs diffuse, normal and depth");
    if (g->texel_bytes != 16) return fail(VCT_ERR_UNSUPPORTED, "the adjoint tape needs an RGBA32F grid");
    dense::Pyramid P;
    dense::March M;
    VCT_TRY(dense::make_params(g, cam, d, &P, &M));
    const dense::TapeLayout T = dense::tape_layout(P, M, gb->width, gb->height, dense::kSweepHeaderBytes);
    if (T.cap > (uint32_t)dense::kAdjMaxSteps)
        return fail(VCT_ERR_UNSUPPORTED, "the march takes more than %d steps", dense::kAdjMaxSteps - 2);
    if (tape_bytes < T.bytes)
        return fail(VCT_ERR_INVALID_ARGUMENT, "the tape holds %zu bytes; this view's takes %zu", tape_bytes, T.bytes);
    if ((T.sample_bytes / sizeof(float4)) >> 32) return fail(VCT_ERR_UNSUPPORTED, "tape above 2^32 samples");
    if ((uintptr_t)tape % 16u) return fail(VCT_ERR_INVALID_ARGUMENT, "the tape must be 16-byte aligned");
    dim3 grid(((gb->width + 15) / 16) * ((gb->height + 15) / 16));
    hipLaunchKernelGGL(dense::k_dense_trace_tape, grid, dim3(256), 0, (hipStream_t)stream, g->data.ptr, P, M,
                       gb->diffuse, gb->normal, gb->depth, (int)gb->width, (int)gb->height, (float4*)out, samples_out,
                       (float4*)tape, (uint8_t*)tape + T.sample_bytes, T.cap);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}
