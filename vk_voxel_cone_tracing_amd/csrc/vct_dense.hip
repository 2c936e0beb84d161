// vct_dense.hip - the "plain" dense-grid method on gfx950: one res^3 grid
// voxelized from triangles, a 6-face anisotropic mip pyramid down to 1^3, and
// the cone march for ambient occlusion + indirect diffuse.
//
// This applies the reference's kernels (dominant-axis voxelization,
// radianceDownSample's front-to-back pairs, traceCone) to a single dense grid
// instead of a clipmap; the level rule is clamp(log2(diameter / voxelSize),
// 0, levels - 1). Storage per level m: level 0 keeps one face (all six faces
// are equal after voxelization), levels >= 1 keep six faces; every face block
// carries a one-texel zero border so the trilinear footprint never needs a
// bounds test inside the grid. Texels are RGBA8 or RGBA32F.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <new>

#include "vct_dense.hpp"
#include "vct_dense_mip.hpp"
#include "vct_internal.hpp"
#include "vct_raster.hpp"
#include "vct_texel.hpp"

static_assert(sizeof(vct_dense_desc) == 32, "vct_dense_desc layout");
static_assert(sizeof(vct_dense_trace_desc) == 32, "vct_dense_trace_desc layout");

namespace vct {

vct_status dense_alloc(vct_dense* g)
{
    size_t total = 0;
    for (int m = 0; m < g->nlevels; ++m) {
        g->offset[m] = total;
        const int n = g->res >> m;
        g->block[m] = (size_t)(n + 2) * (n + 2) * (n + 2);
        total += g->block[m] * (m == 0 ? 1u : 6u);
    }
    g->texels = total;
    return ensure(g->data, total * g->texel_bytes);
}

__global__ void k_dense_translate(const float* __restrict__ pos, uint32_t n9, float mx, float my, float mz,
                                  float* __restrict__ out)
{
#pragma clang fp contract(off)
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n9) return;
    const uint32_t k = i % 3u;
    out[i] = pos[i] - (k == 0 ? mx : (k == 1 ? my : mz));
}

struct DenseRaster {
    float vs;
    int lo[3], hi[3];
    int samples;
    int res;
};

__global__ void k_dense_count(const float* __restrict__ tpos, uint32_t ntris, DenseRaster R,
                              unsigned long long* __restrict__ counts)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    float p[9];
    for (int k = 0; k < 9; ++k) p[k] = tpos[9u * t + k];
    TriSetup s;
    unsigned long long n = 0;
    if (tri_setup(p, R.vs, R.lo, R.hi, s)) n = (unsigned long long)s.nu * s.nv * (unsigned long long)R.samples;
    counts[t] = n;
}

__global__ void k_dense_total(const unsigned long long* __restrict__ counts,
                              const unsigned long long* __restrict__ offsets, uint32_t n,
                              unsigned long long* __restrict__ total)
{
    total[0] = n ? offsets[n - 1] + counts[n - 1] : 0ull;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Every covered sample proposes its triangle for its voxel; the smallest
// triangle index (the first fragment in rasterization order) wins.
__global__ __launch_bounds__(256) void k_dense_raster(const float* __restrict__ tpos,
                                                      const unsigned long long* __restrict__ offsets, uint32_t ntris,
                                                      const unsigned long long* __restrict__ total, DenseRaster R,
                                                      uint32_t* __restrict__ winner)
{
#pragma clang fp contract(off)
    const unsigned long long n = total[0];
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long wi = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; wi < n; wi += stride) {
        uint32_t lo = 0, hi = ntris;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (offsets[mid] <= wi) lo = mid; else hi = mid;
        }
        const uint32_t t = lo;
        float p[9];
        for (int k = 0; k < 9; ++k) p[k] = tpos[9u * t + k];
        TriSetup s;
        if (!tri_setup(p, R.vs, R.lo, R.hi, s)) continue;
        float fp[3];
        double lam[3];
        if (!tri_sample(p, s, R.vs, R.samples, (uint32_t)(wi - offsets[t]), fp, lam)) continue;
        int ic[3];
        bool inside = true;
        for (int k = 0; k < 3; ++k) {
            const float f = floorf(fp[k] / R.vs);
            if (!(f >= 0.0f) || f >= (float)R.res) inside = false;
            ic[k] = (int)f;
        }
        if (!inside) continue;
        atomicMin(&winner[((size_t)ic[2] * R.res + ic[1]) * R.res + ic[0]], t);
    }
}

template <bool kF32>
__global__ void k_dense_resolve(const uint32_t* __restrict__ winner, const uint32_t* __restrict__ mat,
                                const vct_gltf_material* __restrict__ mats, int res, void* __restrict__ level0)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (size_t)res * res * res) return;
    const uint32_t t = winner[v];
    if (t == 0xffffffffu) return;
    const int x = (int)(v % res), y = (int)((v / res) % res), z = (int)(v / ((size_t)res * res));
    const size_t D = (size_t)res + 2;
    const size_t i = ((size_t)(z + 1) * D + (size_t)(y + 1)) * D + (size_t)(x + 1);
    const float* bc = mats[mat[t]].base_color_factor;
    if (kF32) ((float4*)level0)[i] = make_float4(bc[0], bc[1], bc[2], 1.0f);
    else ((uint32_t*)level0)[i] = pack_unorm(f4{bc[0], bc[1], bc[2], 1.0f});
}

// radianceDownSample.comp:95-139 over the whole grid, one level.
template <bool kF32>
__global__ __launch_bounds__(256) void k_dense_mip(const void* __restrict__ src, void* __restrict__ dst, int n,
                                                   bool src_single_face)
{
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    const int z = (int)(blockIdx.z * blockDim.z + threadIdx.z);
    if (x >= n || y >= n || z >= n) return;
    dense_mip_texel<kF32>(src, dst, n, x, y, z, src_single_face);
}

}  // namespace vct

using namespace vct;

extern "C" {

vct_status vct_dense_create(vct_dense** out, const vct_dense_desc* d)
{
    if (!out || !d) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint32_t r = d->resolution;
    if (r < 2 || r > 1024 || (r & (r - 1))) return fail(VCT_ERR_INVALID_ARGUMENT, "resolution must be a power of two in [2, 1024]");
    if (d->format != VCT_FORMAT_RGBA8 && d->format != VCT_FORMAT_RGBA32F)
        return fail(VCT_ERR_INVALID_ARGUMENT, "format must be RGBA8 or RGBA32F");
    if (!(d->extent > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "extent must be positive");
    if (d->raster_samples != 1 && d->raster_samples != 4) return fail(VCT_ERR_INVALID_ARGUMENT, "raster_samples");
    vct_dense* g = new (std::nothrow) vct_dense();
    if (!g) return fail(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
    g->desc = *d;
    g->res = (int)r;
    int n = 0;
    while ((1u << n) < r) ++n;
    g->nlevels = n + 1;
    g->texel_bytes = d->format == VCT_FORMAT_RGBA32F ? 16u : 4u;
    vct_config cfg{8, 0, VCT_FORMAT_RGBA8, 0, d->raster_samples, 0, 1.0f, 0};
    vct_status s = vct_create(&g->scratch, &cfg);
    if (s == VCT_OK) s = dense_alloc(g);
    if (s == VCT_OK) {
        hipError_t e = hipMemset(g->data.ptr, 0, g->texels * g->texel_bytes);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) s = fail(VCT_ERR_HIP, "clear failed: %s", hipGetErrorString(e));
    }
    if (s != VCT_OK) {
        vct_dense_destroy(g);
        return s;
    }
    *out = g;
    return VCT_OK;
}

vct_status vct_dense_destroy(vct_dense* g)
{
    if (!g) return VCT_OK;
    (void)hipDeviceSynchronize();
    release(g->data);
    release(g->winner);
    release(g->bwd_rec);
    release(g->bwd_hdr);
    release(g->bwd_start);
    release(g->bwd_ctr);
    if (g->scratch) vct_destroy(g->scratch);
    delete g;
    return VCT_OK;
}

vct_status vct_dense_clear(vct_dense* g, vct_stream stream)
{
    if (!g) return fail(VCT_ERR_INVALID_ARGUMENT, "null grid");
    VCT_HIP(hipMemsetAsync(g->data.ptr, 0, g->texels * g->texel_bytes, (hipStream_t)stream));
    return VCT_OK;
}

vct_status vct_dense_level(vct_dense* g, uint32_t level, void** ptr, size_t* bytes, uint32_t* faces)
{
    if (!g || !ptr) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (level >= (uint32_t)g->nlevels) return fail(VCT_ERR_INVALID_ARGUMENT, "level out of range");
    *ptr = (char*)g->data.ptr + g->offset[level] * g->texel_bytes;
    if (bytes) *bytes = g->block[level] * (level == 0 ? 1u : 6u) * g->texel_bytes;
    if (faces) *faces = level == 0 ? 1u : 6u;
    return VCT_OK;
}

// Voxelizes a mesh into level 0: opacity 1 and the base colour of the first
// triangle (in draw order) that covers the voxel.
vct_status vct_dense_voxelize(vct_dense* g, vct_stream stream, const vct_mesh* mesh)
{
    if (!g) return fail(VCT_ERR_INVALID_ARGUMENT, "null grid");
    hipStream_t s = (hipStream_t)stream;
    vct_ctx* c = g->scratch;
    TriSoup soup;
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup, TextureUse::kRefuse));
    if (!soup.count) return VCT_OK;
    const uint32_t n9 = soup.count * 9u;
    VCT_TRY(ensure(c->rec_val, 4u * (size_t)n9));
    float* tpos = (float*)c->rec_val.ptr;
    hipLaunchKernelGGL(k_dense_translate, dim3((n9 + 255) / 256), dim3(256), 0, s, soup.pos, n9, g->desc.grid_min[0],
                       g->desc.grid_min[1], g->desc.grid_min[2], tpos);
    DenseRaster R;
    R.vs = g->desc.extent / (float)g->res;
    for (int k = 0; k < 3; ++k) {
        R.lo[k] = 0;
        R.hi[k] = g->res - 1;
    }
    R.samples = (int)g->desc.raster_samples;
    R.res = g->res;
    VCT_TRY(ensure(c->counts, 8u * (size_t)soup.count));
    VCT_TRY(ensure(c->offsets, 8u * (size_t)soup.count));
    VCT_TRY(ensure(c->totals, 64));
    hipLaunchKernelGGL(k_dense_count, dim3((soup.count + 255) / 256), dim3(256), 0, s, tpos, soup.count, R,
                       (unsigned long long*)c->counts.ptr);
    size_t tmp = 0;
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (unsigned long long*)c->counts.ptr,
                                             (unsigned long long*)c->offsets.ptr, (int)soup.count, s));
    VCT_TRY(ensure(c->scan_tmp, tmp));
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, (unsigned long long*)c->counts.ptr,
                                             (unsigned long long*)c->offsets.ptr, (int)soup.count, s));
    hipLaunchKernelGGL(k_dense_total, dim3(1), dim3(1), 0, s, (const unsigned long long*)c->counts.ptr,
                       (const unsigned long long*)c->offsets.ptr, soup.count, (unsigned long long*)c->totals.ptr);
    const size_t nv = (size_t)g->res * g->res * g->res;
    VCT_TRY(ensure(g->winner, 4u * nv));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, (uint32_t*)g->winner.ptr, nv,
                       0xffffffffu);
    hipLaunchKernelGGL(k_dense_raster, dim3(8192), dim3(256), 0, s, tpos, (const unsigned long long*)c->offsets.ptr,
                       soup.count, (const unsigned long long*)c->totals.ptr, R, (uint32_t*)g->winner.ptr);
    if (g->texel_bytes == 16)
        hipLaunchKernelGGL(k_dense_resolve<true>, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s,
                           (const uint32_t*)g->winner.ptr, soup.mat, mesh->materials, g->res, g->data.ptr);
    else
        hipLaunchKernelGGL(k_dense_resolve<false>, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s,
                           (const uint32_t*)g->winner.ptr, soup.mat, mesh->materials, g->res, g->data.ptr);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

// Builds levels 1..log2(res) from level 0 (DownSampler applied per level).
vct_status vct_dense_mipmap(vct_dense* g, vct_stream stream)
{
    if (!g) return fail(VCT_ERR_INVALID_ARGUMENT, "null grid");
    hipStream_t s = (hipStream_t)stream;
    for (int m = 1; m < g->nlevels; ++m) {
        const int n = g->res >> m;
        dim3 block(n < 64 ? n : 64, n < 2 ? 1 : 2, n < 2 ? 1 : 2);
        dim3 grid((n + block.x - 1) / block.x, (n + block.y - 1) / block.y, (n + block.z - 1) / block.z);
        const void* src = (const char*)g->data.ptr + g->offset[m - 1] * g->texel_bytes;
        void* dst = (char*)g->data.ptr + g->offset[m] * g->texel_bytes;
        if (g->texel_bytes == 16)
            hipLaunchKernelGGL(k_dense_mip<true>, grid, block, 0, s, src, dst, n, m == 1);
        else
            hipLaunchKernelGGL(k_dense_mip<false>, grid, block, 0, s, src, dst, n, m == 1);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_dense_read_level(vct_dense* g, vct_stream stream, uint32_t level, void* dst, size_t bytes)
{
    void* p;
    size_t b;
    VCT_TRY(vct_dense_level(g, level, &p, &b, nullptr));
    if (!dst || bytes != b) return fail(VCT_ERR_INVALID_ARGUMENT, "buffer is %zu bytes, level is %zu", bytes, b);
    VCT_HIP(hipMemcpyAsync(dst, p, b, hipMemcpyDeviceToHost, (hipStream_t)stream));
    VCT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return VCT_OK;
}

vct_status vct_dense_write_level(vct_dense* g, vct_stream stream, uint32_t level, const void* src, size_t bytes)
{
    void* p;
    size_t b;
    VCT_TRY(vct_dense_level(g, level, &p, &b, nullptr));
    if (!src || bytes != b) return fail(VCT_ERR_INVALID_ARGUMENT, "buffer is %zu bytes, level is %zu", bytes, b);
    VCT_HIP(hipMemcpyAsync(p, src, b, hipMemcpyHostToDevice, (hipStream_t)stream));
    VCT_HIP(hipStreamSynchronize((hipStream_t)stream));
    return VCT_OK;
}

}  // extern "C"
