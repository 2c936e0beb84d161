// vct_ctx.hip - context lifetime, clip regions, volume access and the mesh
// vertex stage of libvct (gfx950).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "vct_internal.hpp"
#include "vct_textures.hpp"

static_assert(sizeof(vct_clipmap_region) == 28, "ClipmapRegion layout");
static_assert(sizeof(vct_voxelization_desc) == 64, "VoxelizationDesc std140 layout");
static_assert(sizeof(vct_downsample_desc) == 24, "DownSampleDesc layout");
static_assert(sizeof(vct_image_cleaning_desc) == 36, "ImageCleaningDesc layout");
static_assert(sizeof(vct_border_wrapping_desc) == 16, "BorderWrappingDesc layout");
static_assert(sizeof(vct_copy_alpha_desc) == 12, "CopyAlphaDesc layout");
static_assert(sizeof(vct_cone_tracing_desc) == 52, "VoxelConeTracingDesc layout");
static_assert(sizeof(vct_directional_light_desc) == 32, "DirectionalLightDesc std140 layout");
static_assert(sizeof(vct_directional_light_shadow_desc) == 144, "DirectionalLightShadowDesc std140 layout");
static_assert(sizeof(vct_camera_ubo) == 144, "CameraUBO layout");
static_assert(sizeof(vct_gltf_material) == 80, "GltfShadeMaterial std430 layout");
static_assert(sizeof(vct_mesh) == 120, "vct_mesh layout");
static_assert(sizeof(vct_texture) == 16, "vct_texture layout");
static_assert(offsetof(vct_gltf_material, emissive_factor) == 48, "GltfShadeMaterial emissiveFactor offset");
static_assert(offsetof(vct_voxelization_desc, clip_max_extent) == 28, "VoxelizationDesc clipMaxExtent offset");
static_assert(offsetof(vct_voxelization_desc, clipmap_resolution) == 60, "VoxelizationDesc clipmapResolution offset");
static_assert(offsetof(vct_directional_light_shadow_desc, z_near) == 128, "DirectionalLightShadowDesc zNear offset");

namespace vct {

static thread_local char g_last_error[512] = "";

vct_status fail(vct_status s, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return s;
}

vct_status ensure(DeviceBuffer& b, size_t bytes)
{
    if (bytes <= b.bytes && b.ptr) return VCT_OK;
    if (b.ptr) {
        (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        (void)hipGetLastError();
        return fail(VCT_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.bytes = want;
    return VCT_OK;
}

void release(DeviceBuffer& b)
{
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

vct_status upload_small(vct_ctx* ctx, hipStream_t s, DeviceBuffer& dst, const void* src, size_t bytes)
{
    if (bytes == 0) return VCT_OK;
    // The pinned block keeps the bytes of the last finished upload: when dst
    // already holds exactly these bytes (an unchanged draw list, the common
    // case from one call to the next) nothing is copied and nothing waits.
    if (ctx->pinned_dst == &dst && ctx->pinned_valid == bytes && dst.ptr && dst.bytes >= bytes &&
        std::memcmp(ctx->pinned, src, bytes) == 0)
        return VCT_OK;
    ctx->pinned_dst = nullptr;
    ctx->pinned_valid = 0;
    VCT_TRY(ensure(dst, bytes));
    if (bytes > ctx->pinned_bytes) {
        if (ctx->pinned) hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        VCT_HIP(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
        ctx->pinned_bytes = bytes;
    }
    // Every upload waits for its copy before returning, so no copy out of the
    // pinned block is still in flight here.
    std::memcpy(ctx->pinned, src, bytes);
    VCT_HIP(hipMemcpyAsync(dst.ptr, ctx->pinned, bytes, hipMemcpyHostToDevice, s));
    VCT_HIP(hipStreamSynchronize(s));
    ctx->pinned_dst = &dst;
    ctx->pinned_valid = bytes;
    return VCT_OK;
}

// ---------------------------------------------------------------------------
// Vertex stage: one thread per triangle of the flattened draw list.

struct DevPrim {
    uint32_t first_index;
    uint32_t tri_count;
    int32_t vertex_offset;
    uint32_t instance_index;
    uint32_t material_index;
    uint32_t first_tri;
};

__global__ void k_vertex_stage(const float* __restrict__ positions, const float* __restrict__ normals,
                               const float* __restrict__ texcoords, const uint32_t* __restrict__ indices,
                               const float* __restrict__ nodes, const DevPrim* __restrict__ prims, uint32_t nprims,
                               uint32_t ntris, float* __restrict__ out_pos, float* __restrict__ out_nrm,
                               uint32_t* __restrict__ out_mat, float* __restrict__ out_uv)
{
#pragma clang fp contract(off)
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    // upper_bound on first_tri
    uint32_t lo = 0, hi = nprims;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (prims[mid].first_tri <= t) lo = mid; else hi = mid;
    }
    const DevPrim pr = prims[lo];
    const uint32_t local = t - pr.first_tri;
    const float* m = nodes + 32u * pr.instance_index;
    const float* it = m + 16;
    for (int k = 0; k < 3; ++k) {
        uint32_t vi = (uint32_t)((int64_t)indices[pr.first_index + 3u * local + (uint32_t)k] + pr.vertex_offset);
        float px = positions[3u * vi], py = positions[3u * vi + 1], pz = positions[3u * vi + 2];
        float nx = normals[3u * vi], ny = normals[3u * vi + 1], nz = normals[3u * vi + 2];
        for (int r = 0; r < 3; ++r) {
            out_pos[9u * t + 3 * k + r] = ((m[r] * px + m[4 + r] * py) + m[8 + r] * pz) + m[12 + r];
            out_nrm[9u * t + 3 * k + r] = (it[r] * nx + it[4 + r] * ny) + it[8 + r] * nz;
        }
        out_uv[6u * t + 2 * k] = texcoords ? texcoords[2u * vi] : 0.0f;
        out_uv[6u * t + 2 * k + 1] = texcoords ? texcoords[2u * vi + 1] : 0.0f;
    }
    out_mat[t] = pr.material_index;
}

vct_status read_texture_indices(hipStream_t s, const vct_gltf_material* mats, uint32_t count, int32_t* idx)
{
    if (!count) return VCT_OK;
    vct_gltf_material* host = new (std::nothrow) vct_gltf_material[count];
    if (!host) return fail(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
    hipError_t e = hipMemcpyAsync(host, mats, sizeof(vct_gltf_material) * count, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        delete[] host;
        return fail(VCT_ERR_HIP, "reading the material table failed: %s", hipGetErrorString(e));
    }
    // msaaVoxelizer.frag:64, msaaInjectRadiance.frag:73-82,132-135 and
    // voxelizer.frag:52-76 sample these three textures
    for (uint32_t i = 0; i < count; ++i) {
        idx[3 * i + 0] = host[i].base_color_texture;
        idx[3 * i + 1] = host[i].emissive_texture;
        idx[3 * i + 2] = host[i].occlusion_texture;
    }
    delete[] host;
    return VCT_OK;
}

// The drawn materials' texture indices against `use` and the mesh's texture
// table. The material table is read back when its pointer, count or the
// mesh's material_version differs from the last one read.
static vct_status check_textures(vct_ctx* ctx, hipStream_t s, const vct_mesh* mesh, TextureUse use)
{
    if (use == TextureUse::kIgnore) return VCT_OK;
    if (ctx->mat_table != (const void*)mesh->materials || ctx->mat_count != mesh->material_count ||
        ctx->mat_version != mesh->material_version) {
        delete[] ctx->mat_textures;
        ctx->mat_textures = nullptr;
        ctx->mat_table = nullptr;
        ctx->mat_count = 0;
        int32_t* idx = new (std::nothrow) int32_t[3u * (mesh->material_count ? mesh->material_count : 1u)];
        if (!idx) return fail(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
        vct_status st = read_texture_indices(s, mesh->materials, mesh->material_count, idx);
        if (st != VCT_OK) {
            delete[] idx;
            return st;
        }
        ctx->mat_textures = idx;
        ctx->mat_table = mesh->materials;
        ctx->mat_count = mesh->material_count;
        ctx->mat_version = mesh->material_version;
    }
    for (uint32_t i = 0; i < mesh->primitive_count; ++i) {
        const vct_primitive& p = mesh->primitives[i];
        if (!p.index_count) continue;
        for (int k = 0; k < 3; ++k) {
            const int32_t t = ctx->mat_textures[3u * p.material_index + (uint32_t)k];
            if (t < 0) continue;
            if (use == TextureUse::kRefuse)
                return fail(VCT_ERR_UNSUPPORTED,
                            "primitive %u: material %u references a base-colour, emissive or occlusion texture; this "
                            "pass uses material factors only",
                            i, p.material_index);
            if ((uint32_t)t >= mesh->texture_count || !mesh->textures)
                return fail(VCT_ERR_INVALID_ARGUMENT, "primitive %u: material %u references texture %d of %u", i,
                            p.material_index, t, mesh->texture_count);
            if (!mesh->texcoords)
                return fail(VCT_ERR_INVALID_ARGUMENT, "primitive %u: a textured material needs texcoords", i);
        }
    }
    return VCT_OK;
}

vct_status run_vertex_stage(vct_ctx* ctx, hipStream_t s, const vct_mesh* mesh, TriSoup* out, TextureUse use)
{
    if (!mesh || !mesh->positions || !mesh->normals || !mesh->indices || !mesh->node_matrices || !mesh->materials)
        return fail(VCT_ERR_INVALID_ARGUMENT, "mesh buffers must be non-null");
    if (!mesh->primitives && mesh->primitive_count)
        return fail(VCT_ERR_INVALID_ARGUMENT, "mesh primitive list is null");
    if (!mesh->textures && mesh->texture_count)
        return fail(VCT_ERR_INVALID_ARGUMENT, "mesh texture list is null");
    for (uint32_t i = 0; i < mesh->primitive_count; ++i)
        if (mesh->primitives[i].material_index >= mesh->material_count)
            return fail(VCT_ERR_INVALID_ARGUMENT, "primitive %u material out of range", i);
    for (uint32_t i = 0; i < mesh->texture_count; ++i)
        if (!mesh->textures[i].texels || !mesh->textures[i].width || !mesh->textures[i].height ||
            mesh->textures[i].width > 65536u || mesh->textures[i].height > 65536u)
            return fail(VCT_ERR_INVALID_ARGUMENT, "texture %u: texels and a size in [1, 65536] are required", i);
    VCT_TRY(check_textures(ctx, s, mesh, use));
    // the kernels take the texture table whenever the mesh has one
    const bool sampled = use == TextureUse::kSample && mesh->texture_count > 0;
    // host side of GLTFScene::cmdDraw: flatten the draws; the texture table
    // follows the primitive table in the same upload
    static thread_local DevPrim* hp = nullptr;
    static thread_local size_t hp_cap = 0;
    const size_t tex_slots = ((size_t)mesh->texture_count * sizeof(DevTexture) + sizeof(DevPrim) - 1) / sizeof(DevPrim);
    if (mesh->primitive_count + 1 + tex_slots > hp_cap) {
        delete[] hp;
        hp_cap = mesh->primitive_count + 1 + tex_slots;
        hp = new (std::nothrow) DevPrim[hp_cap];
        if (!hp) {
            hp_cap = 0;
            return fail(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
        }
    }
    uint64_t ntris = 0;
    uint32_t n = 0;
    for (uint32_t i = 0; i < mesh->primitive_count; ++i) {
        const vct_primitive& p = mesh->primitives[i];
        if (p.index_count % 3u) return fail(VCT_ERR_INVALID_ARGUMENT, "primitive %u index_count not a multiple of 3", i);
        if ((uint64_t)p.first_index + p.index_count > mesh->index_count)
            return fail(VCT_ERR_INVALID_ARGUMENT, "primitive %u indices out of range", i);
        if (p.instance_index >= mesh->instance_count)
            return fail(VCT_ERR_INVALID_ARGUMENT, "primitive %u instance out of range", i);
        if (p.index_count == 0) continue;
        hp[n].first_index = p.first_index;
        hp[n].tri_count = p.index_count / 3u;
        hp[n].vertex_offset = p.vertex_offset;
        hp[n].instance_index = p.instance_index;
        hp[n].material_index = p.material_index;
        hp[n].first_tri = (uint32_t)ntris;
        ntris += p.index_count / 3u;
        ++n;
    }
    if (ntris > 0xffffffffull / 9u) return fail(VCT_ERR_CAPACITY, "too many triangles");
    DevTexture* ht = (DevTexture*)(hp + n);
    for (uint32_t i = 0; sampled && i < mesh->texture_count; ++i)
        ht[i] = DevTexture{mesh->textures[i].texels, mesh->textures[i].width, mesh->textures[i].height};
    out->count = (uint32_t)ntris;
    VCT_TRY(ensure(ctx->soup_pos, 36u * (ntris ? ntris : 1)));
    VCT_TRY(ensure(ctx->soup_nrm, 36u * (ntris ? ntris : 1)));
    VCT_TRY(ensure(ctx->soup_mat, 4u * (ntris ? ntris : 1)));
    VCT_TRY(ensure(ctx->soup_uv, 24u * (ntris ? ntris : 1)));
    out->pos = (const float*)ctx->soup_pos.ptr;
    out->nrm = (const float*)ctx->soup_nrm.ptr;
    out->mat = (const uint32_t*)ctx->soup_mat.ptr;
    out->uv = (const float*)ctx->soup_uv.ptr;
    out->tex = nullptr;
    out->tex_count = 0;
    if (ntris == 0) return VCT_OK;
    const uint32_t ntex = sampled ? mesh->texture_count : 0u;
    VCT_TRY(upload_small(ctx, s, ctx->prims, hp, sizeof(DevPrim) * n + sizeof(DevTexture) * ntex));
    // null without a texture table: the kernels skip the lookups
    out->tex = ntex ? (const DevTexture*)((const DevPrim*)ctx->prims.ptr + n) : nullptr;
    out->tex_count = ntex;
    uint32_t threads = 256;
    uint32_t blocks = (uint32_t)((ntris + threads - 1) / threads);
    hipLaunchKernelGGL(k_vertex_stage, dim3(blocks), dim3(threads), 0, s, mesh->positions, mesh->normals,
                       mesh->texcoords, mesh->indices, mesh->node_matrices, (const DevPrim*)ctx->prims.ptr, n,
                       (uint32_t)ntris, (float*)ctx->soup_pos.ptr, (float*)ctx->soup_nrm.ptr,
                       (uint32_t*)ctx->soup_mat.ptr, (float*)ctx->soup_uv.ptr);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

// ---------------------------------------------------------------------------
// Reference-layout conversion for vct_read_volume / vct_write_volume.

template <typename T>
__global__ void k_to_atlas(const T* __restrict__ src, T* __restrict__ dst, int D, int levels, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t W = (size_t)D * kFaces, H = (size_t)D * levels;
    size_t x = i % W, y = (i / W) % H, z = i / (W * H);
    int face = (int)(x / D), level = (int)(y / D);
    dst[i] = src[vol_index(D, level, face, (int)(x % D), (int)(y % D), (int)z)];
}

template <typename T>
__global__ void k_from_atlas(const T* __restrict__ src, T* __restrict__ dst, int D, int levels, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t W = (size_t)D * kFaces, H = (size_t)D * levels;
    size_t x = i % W, y = (i / W) % H, z = i / (W * H);
    int face = (int)(x / D), level = (int)(y / D);
    dst[vol_index(D, level, face, (int)(x % D), (int)(y % D), (int)z)] = src[i];
}

}  // namespace vct

using namespace vct;

extern "C" {

const char* vct_version(void)
{
    return "vct-mi355x 0.2.0 (abi 2, gfx950)";
}

const char* vct_status_string(vct_status s)
{
    switch (s) {
    case VCT_OK: return "VCT_OK";
    case VCT_ERR_INVALID_ARGUMENT: return "VCT_ERR_INVALID_ARGUMENT";
    case VCT_ERR_OUT_OF_MEMORY: return "VCT_ERR_OUT_OF_MEMORY";
    case VCT_ERR_HIP: return "VCT_ERR_HIP";
    case VCT_ERR_CAPACITY: return "VCT_ERR_CAPACITY";
    case VCT_ERR_UNSUPPORTED: return "VCT_ERR_UNSUPPORTED";
    case VCT_ERR_NOT_READY: return "VCT_ERR_NOT_READY";
    }
    return "VCT_ERR_UNKNOWN";
}

const char* vct_last_error(void)
{
    return g_last_error;
}

vct_status vct_create(vct_ctx** out_ctx, const vct_config* cfg)
{
    if (!out_ctx || !cfg) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    *out_ctx = nullptr;
    const uint32_t r = cfg->voxel_resolution;
    if (r < 8 || r > 256 || (r & (r - 1)))
        return fail(VCT_ERR_INVALID_ARGUMENT, "voxel_resolution must be a power of two in [8, 256]");
    if (cfg->clip_levels > (uint32_t)kMaxLevels)
        return fail(VCT_ERR_INVALID_ARGUMENT, "clip_levels must be in [0, %d]", kMaxLevels);
    if (cfg->radiance_format > VCT_FORMAT_RGBA16F) return fail(VCT_ERR_INVALID_ARGUMENT, "bad radiance_format");
    if (cfg->raster_samples != 1 && cfg->raster_samples != 4)
        return fail(VCT_ERR_INVALID_ARGUMENT, "raster_samples must be 1 or 4");
    if (cfg->clip_levels && cfg->downsample_region_size >= r / 4)
        return fail(VCT_ERR_INVALID_ARGUMENT, "downsample_region_size must be below resolution / 4");
    if (!(cfg->extent_l0 > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "extent_l0 must be positive");
    int dev = 0;
    VCT_HIP(hipGetDevice(&dev));
    vct_ctx* c = new (std::nothrow) vct_ctx();
    if (!c) return fail(VCT_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->cfg = *cfg;
    c->res = (int)r;
    c->D = (int)r + 2;
    c->levels = (int)cfg->clip_levels;
    c->device = dev;
    vct_status s = VCT_OK;
    if (c->levels) s = ensure(c->opacity, texels_per_volume(c) * 4u);
    if (s == VCT_OK && c->levels) s = ensure(c->radiance, texels_per_volume(c) * texel_bytes(c, VCT_VOLUME_RADIANCE));
    if (s == VCT_OK && c->levels) {
        const uint32_t D = (uint32_t)c->D, D3 = D * D * D;
        c->occ_words = ((D3 + 63u) / 64u) * 2u;
        c->occ_pad = (D * D + D + 1u) / 32u + 3u;
        s = ensure(c->occ_raw, 4u * (size_t)c->levels * (c->occ_words + c->occ_pad));
        if (s == VCT_OK) s = ensure(c->occ_mask, 4u * (size_t)c->levels * c->occ_words);
    }
    if (s == VCT_OK && c->levels) {
        hipError_t e = hipMemset(c->occ_raw.ptr, 0, c->occ_raw.bytes);
        if (e == hipSuccess) e = hipMemset(c->opacity.ptr, 0, texels_per_volume(c) * 4u);
        if (e == hipSuccess) e = hipMemset(c->radiance.ptr, 0, texels_per_volume(c) * texel_bytes(c, VCT_VOLUME_RADIANCE));
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) s = fail(VCT_ERR_HIP, "initial clear failed: %s", hipGetErrorString(e));
    }
    // device address of the allocation-free trace's static work list on this device
    if (s == VCT_OK) s = trace_scratch_init(dev);
    if (s != VCT_OK) {
        vct_destroy(c);
        return s;
    }
    // VoxelizationPass::createVoxelClipmap (VoxelizationPass.cpp:335-345)
    for (int i = 0; i < c->levels; ++i) {
        for (int k = 0; k < 3; ++k) {
            c->regions[i].min_corner[k] = -(int32_t)(r >> 1);
            c->regions[i].extent[k] = r;
        }
        c->regions[i].voxel_size = (cfg->extent_l0 * (float)(1u << i)) / (float)r;
    }
    *out_ctx = c;
    return VCT_OK;
}

vct_status vct_destroy(vct_ctx* c)
{
    if (!c) return VCT_OK;
    (void)hipDeviceSynchronize();
    release(c->opacity);
    release(c->radiance);
    release(c->occ_raw);
    release(c->occ_mask);
    DeviceBuffer* bufs[] = {&c->soup_pos, &c->soup_nrm, &c->soup_mat, &c->prims, &c->prim_first, &c->counts,
                            &c->offsets, &c->scan_tmp, &c->totals, &c->rec_key[0], &c->rec_key[1], &c->rec_idx[0],
                            &c->rec_idx[1], &c->rec_val, &c->sort_tmp, &c->staging, &c->materials_flat,
                            &c->svo.frags, &c->svo.pool, &c->svo.counters, &c->spec_perm, &c->svo.cells,
                            &c->svo.grids, &c->svo.bricks, &c->soup_uv};
    for (DeviceBuffer* b : bufs) release(*b);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->pinned_word) (void)hipHostFree(c->pinned_word);
    delete[] c->mat_textures;
    delete c;
    return VCT_OK;
}

vct_status vct_set_clip_regions(vct_ctx* c, const vct_clipmap_region* regions, uint32_t count)
{
    if (!c || !regions) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (count != (uint32_t)c->levels) return fail(VCT_ERR_INVALID_ARGUMENT, "expected %d regions", c->levels);
    for (uint32_t i = 0; i < count; ++i) {
        for (int k = 0; k < 3; ++k)
            if (regions[i].extent[k] != (uint32_t)c->res)
                return fail(VCT_ERR_INVALID_ARGUMENT, "region %u extent must equal the resolution", i);
        if (!(regions[i].voxel_size > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "region %u voxel_size", i);
        c->regions[i] = regions[i];
    }
    c->regions_set = true;
    return VCT_OK;
}

vct_status vct_get_clip_regions(const vct_ctx* c, vct_clipmap_region* regions, uint32_t count)
{
    if (!c || !regions) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (count != (uint32_t)c->levels) return fail(VCT_ERR_INVALID_ARGUMENT, "expected %d regions", c->levels);
    for (uint32_t i = 0; i < count; ++i) regions[i] = c->regions[i];
    return VCT_OK;
}

vct_status vct_compute_clip_regions(const vct_config* cfg, const float cam[3], vct_clipmap_region* regions,
                                    uint32_t count)
{
    if (!cfg || !cam || !regions) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    const uint32_t r = cfg->voxel_resolution;
    if (r < 8 || r > 256 || (r & (r - 1))) return fail(VCT_ERR_INVALID_ARGUMENT, "voxel_resolution");
    if (count != cfg->clip_levels || count > (uint32_t)kMaxLevels)
        return fail(VCT_ERR_INVALID_ARGUMENT, "expected %u regions", cfg->clip_levels);
    if (!(cfg->extent_l0 > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "extent_l0");
    // Application::updateClipRegionBoundingBox (Application.cpp:116-128) and
    // VoxelizationPass::calculateChangeDelta (VoxelizationPass.cpp:438-449)
    // applied to the initial regions of createVoxelClipmap (:335-357).
    static const int32_t kMinChange[8] = {2, 2, 2, 2, 2, 1, 1, 1}; // VoxelizationPass.h:57
    for (uint32_t i = 0; i < count; ++i) {
        vct_clipmap_region reg;
        float vs = (cfg->extent_l0 * (float)(1u << i)) / (float)r;
        float half = (float)((((uint32_t)cfg->extent_l0) >> 1) * (1u << i));
        for (int k = 0; k < 3; ++k) {
            reg.min_corner[k] = -(int32_t)(r >> 1);
            reg.extent[k] = r;
            float bb_min = cam[k] - half;
            float delta_w = bb_min - (float)reg.min_corner[k] * vs;
            float min_change = vs * (float)kMinChange[i];
            int32_t delta = (int32_t)std::trunc(delta_w / min_change) * kMinChange[i];
            reg.min_corner[k] += delta;
        }
        reg.voxel_size = vs;
        regions[i] = reg;
    }
    return VCT_OK;
}

vct_status vct_update_clip_regions(vct_ctx* c, const float cam[3])
{
    if (!c || !cam) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    vct_config cfg = c->cfg;
    cfg.clip_levels = (uint32_t)c->levels;
    VCT_TRY(vct_compute_clip_regions(&cfg, cam, c->regions, (uint32_t)c->levels));
    c->regions_set = true;
    return VCT_OK;
}

vct_status vct_move_clip_regions(vct_ctx* c, const float cam[3], vct_clipmap_region* revox, uint32_t* counts)
{
    if (!c || !cam || !revox || !counts) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->regions_set) return fail(VCT_ERR_NOT_READY, "the regions have not been set");
    static const int32_t kMinChange[8] = {2, 2, 2, 2, 2, 1, 1, 1};  // VoxelizationPass.h:57
    int32_t old_min[kMaxLevels][3];
    for (int i = 0; i < c->levels; ++i)
        for (int k = 0; k < 3; ++k) old_min[i][k] = c->regions[i].min_corner[k];
    for (int i = 0; i < c->levels; ++i) {
        vct_clipmap_region& reg = c->regions[i];
        const float vs = reg.voxel_size;
        // camera box of level i (Application.cpp:116-128)
        const float half = (float)((((uint32_t)c->cfg.extent_l0) >> 1) * (1u << i));
        int32_t delta[3];
        bool full = false;
        for (int k = 0; k < 3; ++k) {
            const float delta_w = (cam[k] - half) - (float)reg.min_corner[k] * vs;
            const float min_change = vs * (float)kMinChange[i];
            delta[k] = (int32_t)std::trunc(delta_w / min_change) * kMinChange[i];
            reg.min_corner[k] += delta[k];
            if ((uint32_t)std::abs(delta[k]) >= reg.extent[k]) full = true;
        }
        vct_clipmap_region* out = revox + VCT_REVOX_PER_LEVEL * i;
        uint32_t n = 0;
        if (full) {
            out[n++] = reg;
        } else {
            for (int k = 0; k < 3; ++k) {
                const int32_t ad = std::abs(delta[k]);
                if (ad < kMinChange[i]) continue;
                vct_clipmap_region slab = reg;
                slab.extent[k] = (uint32_t)ad;
                slab.min_corner[k] = delta[k] > 0 ? reg.min_corner[k] + (int32_t)reg.extent[k] - ad : reg.min_corner[k];
                out[n++] = slab;
            }
            // The downsample of level i overwrites the mip channels of the cells
            // under level i-1 ((prev_min >> 1) + [0, res/2), opacityDownSample.comp);
            // when level i-1 moved, the cells it left still hold the old mips.
            // Revoxelize the old inner box (intersected with the new region).
            if (i > 0) {
                const vct_clipmap_region& prev = c->regions[i - 1];
                bool moved = false;
                vct_clipmap_region box = reg;
                for (int k = 0; k < 3; ++k) {
                    moved |= (prev.min_corner[k] >> 1) != (old_min[i - 1][k] >> 1);
                    const int32_t lo = std::max(old_min[i - 1][k] >> 1, reg.min_corner[k]);
                    const int32_t hi = std::min((old_min[i - 1][k] >> 1) + (int32_t)(prev.extent[k] >> 1),
                                                reg.min_corner[k] + (int32_t)reg.extent[k]);
                    box.min_corner[k] = lo;
                    box.extent[k] = hi > lo ? (uint32_t)(hi - lo) : 0u;
                }
                if (moved && box.extent[0] && box.extent[1] && box.extent[2]) out[n++] = box;
            }
        }
        counts[i] = n;
    }
    return VCT_OK;
}

vct_status vct_make_voxelization_desc(const vct_ctx* c, uint32_t level, const vct_clipmap_region* region,
                                      vct_voxelization_desc* d)
{
    if (!c || !d) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (level >= (uint32_t)c->levels) return fail(VCT_ERR_INVALID_ARGUMENT, "level out of range");
    const vct_clipmap_region& target = c->regions[level];
    const vct_clipmap_region& reg = region ? *region : target;
    // Voxelizer::cmdVoxelize (Voxelizer.cpp:225-262)
    for (int k = 0; k < 3; ++k) {
        d->region_min_corner[k] = (float)reg.min_corner[k] * reg.voxel_size - 1e-6f;
        d->region_max_corner[k] = (float)(reg.min_corner[k] + (int32_t)reg.extent[k]) * reg.voxel_size + 1e-6f;
        d->prev_region_min_corner[k] = 0.0f;
        d->prev_region_max_corner[k] = 0.0f;
    }
    d->clip_level = level;
    d->clip_max_extent = (float)target.extent[0] * target.voxel_size;
    d->voxel_size = target.voxel_size;
    if (level > 0) {
        const vct_clipmap_region& prev = c->regions[level - 1];
        for (int k = 0; k < 3; ++k) {
            d->prev_region_min_corner[k] = (float)prev.min_corner[k] * prev.voxel_size;
            d->prev_region_max_corner[k] = (float)(prev.min_corner[k] + (int32_t)prev.extent[k]) * prev.voxel_size;
        }
    }
    d->clipmap_resolution = c->res;
    return VCT_OK;
}

vct_status vct_clear_volume(vct_ctx* c, vct_stream stream, vct_volume v)
{
    if (!c || (v != VCT_VOLUME_OPACITY && v != VCT_VOLUME_RADIANCE))
        return fail(VCT_ERR_INVALID_ARGUMENT, "bad argument");
    DeviceBuffer& b = v == VCT_VOLUME_OPACITY ? c->opacity : c->radiance;
    VCT_HIP(hipMemsetAsync(b.ptr, 0, texels_per_volume(c) * texel_bytes(c, v), (hipStream_t)stream));
    if (v == VCT_VOLUME_RADIANCE) radiance_written(c);
    return VCT_OK;
}

vct_status vct_volume_bytes(const vct_ctx* c, vct_volume v, size_t* out)
{
    if (!c || !out || (v != VCT_VOLUME_OPACITY && v != VCT_VOLUME_RADIANCE))
        return fail(VCT_ERR_INVALID_ARGUMENT, "bad argument");
    *out = texels_per_volume(c) * texel_bytes(c, v);
    return VCT_OK;
}

static vct_status convert(vct_ctx* c, hipStream_t s, vct_volume v, void* host, size_t bytes, bool read)
{
    if (!c || !host || (v != VCT_VOLUME_OPACITY && v != VCT_VOLUME_RADIANCE))
        return fail(VCT_ERR_INVALID_ARGUMENT, "bad argument");
    const size_t n = texels_per_volume(c);
    const size_t tb = texel_bytes(c, v);
    if (bytes != n * tb) return fail(VCT_ERR_INVALID_ARGUMENT, "buffer is %zu bytes, volume is %zu", bytes, n * tb);
    DeviceBuffer& vol = v == VCT_VOLUME_OPACITY ? c->opacity : c->radiance;
    VCT_TRY(ensure(c->staging, n * tb));
    const unsigned threads = 256;
    const unsigned blocks = (unsigned)((n + threads - 1) / threads);
    if (read) {
        if (tb == 4)
            hipLaunchKernelGGL(k_to_atlas<uint32_t>, dim3(blocks), dim3(threads), 0, s, (const uint32_t*)vol.ptr,
                               (uint32_t*)c->staging.ptr, c->D, c->levels, n);
        else
            hipLaunchKernelGGL(k_to_atlas<uint2>, dim3(blocks), dim3(threads), 0, s, (const uint2*)vol.ptr,
                               (uint2*)c->staging.ptr, c->D, c->levels, n);
        VCT_HIP(hipGetLastError());
        VCT_HIP(hipMemcpyAsync(host, c->staging.ptr, n * tb, hipMemcpyDeviceToHost, s));
        VCT_HIP(hipStreamSynchronize(s));
    } else {
        if (v == VCT_VOLUME_RADIANCE) radiance_written(c);
        VCT_HIP(hipMemcpyAsync(c->staging.ptr, host, n * tb, hipMemcpyHostToDevice, s));
        if (tb == 4)
            hipLaunchKernelGGL(k_from_atlas<uint32_t>, dim3(blocks), dim3(threads), 0, s,
                               (const uint32_t*)c->staging.ptr, (uint32_t*)vol.ptr, c->D, c->levels, n);
        else
            hipLaunchKernelGGL(k_from_atlas<uint2>, dim3(blocks), dim3(threads), 0, s, (const uint2*)c->staging.ptr,
                               (uint2*)vol.ptr, c->D, c->levels, n);
        VCT_HIP(hipGetLastError());
        VCT_HIP(hipStreamSynchronize(s));
    }
    return VCT_OK;
}

vct_status vct_read_volume(vct_ctx* c, vct_stream stream, vct_volume v, void* dst, size_t bytes)
{
    return convert(c, (hipStream_t)stream, v, dst, bytes, true);
}

vct_status vct_write_volume(vct_ctx* c, vct_stream stream, vct_volume v, const void* src, size_t bytes)
{
    return convert(c, (hipStream_t)stream, v, const_cast<void*>(src), bytes, false);
}

}  // extern "C"
