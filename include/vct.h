/*
 * vct.h - C ABI of the MI355X voxel cone tracing hot path.
 *
 * One shared library (libvct.so, HIP for gfx950) exports the per-pass entry
 * points that the reference records as Vulkan command-buffer work. Each entry
 * point replaces one reference pass helper; the citation beside it names the
 * reference interface (path:line, relative to the reference repository).
 *
 * Conventions
 *  - Every call returns vct_status; nothing aborts.
 *  - Calls that take a vct_stream enqueue on that HIP stream (hipStream_t,
 *    NULL = legacy default stream) and are ordered only by that stream.
 *  - A vct_ctx owns the voxel volumes, the SVO pool and all scratch memory.
 *    The caller owns meshes, G-buffers, shadow maps and output images
 *    (device pointers). A vct_ctx is not thread-safe.
 *  - Descriptor structs are byte-compatible with the reference's std140 /
 *    push-constant structs (sizes are checked with static asserts in the
 *    library).
 */
#ifndef VCT_H
#define VCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCT_ABI_VERSION 2u

typedef struct vct_ctx vct_ctx;
typedef void* vct_stream; /* hipStream_t */

typedef enum vct_status {
    VCT_OK = 0,
    VCT_ERR_INVALID_ARGUMENT = 1,
    VCT_ERR_OUT_OF_MEMORY = 2,
    VCT_ERR_HIP = 3,
    VCT_ERR_CAPACITY = 4,     /* a fixed-size pool would overflow */
    VCT_ERR_UNSUPPORTED = 5,  /* e.g. textured materials on the dense method */
    VCT_ERR_NOT_READY = 6     /* a prerequisite pass has not run */
} vct_status;

typedef enum vct_volume {
    VCT_VOLUME_OPACITY = 0,  /* "VoxelOpacity"  (Voxelizer.cpp:262-264) */
    VCT_VOLUME_RADIANCE = 1  /* "VoxelRadiance" (Voxelizer.cpp:268-270) */
} vct_volume;

typedef enum vct_format {
    VCT_FORMAT_RGBA8 = 0,   /* VK_FORMAT_R8G8B8A8_UNORM, the reference's only format */
    VCT_FORMAT_RGBA16F = 1, /* half-float radiance (deterministic FP32 accumulate) */
    VCT_FORMAT_RGBA32F = 2  /* float grids of the dense method (adjoint) */
} vct_format;

/* Reference behaviours that are defects or driver-defined. Each bit selects
 * the reference's behaviour; a clear bit selects the evident intent. */
enum {
    VCT_QUIRK_CLAMP_X_RANGE = 1u << 0,      /* msaaVoxelizer.frag:43-60 closed region, float(vec3) clamp; off: half-open voxel cells */
    VCT_QUIRK_BORDER_SKIP_FAR = 1u << 1,    /* BorderWrapper.cpp:139-140 writes 0..127 only */
    VCT_QUIRK_SHADOW_RAW_DEPTH = 1u << 2,   /* shadow.glsl:8-36 averages raw depth */
    VCT_QUIRK_SVO_HALF_COORDS = 1u << 3,    /* octreeNodeFlag.comp:28-32 fragPos >> 1 */
    VCT_QUIRK_SVO_MIP_LEVEL_TEST = 1u << 4, /* octreeNodeMipmapWrite.comp:50-52 */
    VCT_QUIRK_SVO_UNTEXTURED_BLACK = 1u << 5, /* voxelizer.frag:72-77 */
    VCT_QUIRK_ALL = 0x3fu
};

/* Rendering modes, VoxelConeTracingPass.h:17-28. */
enum {
    VCT_MODE_DIFFUSE = 0,
    VCT_MODE_SPECULAR = 1,
    VCT_MODE_NORMAL = 2,
    VCT_MODE_MIN_LEVEL = 3,
    VCT_MODE_DIRECT = 4,
    VCT_MODE_INDIRECT_DIFFUSE = 5,
    VCT_MODE_INDIRECT_SPECULAR = 6,
    VCT_MODE_AMBIENT_OCCLUSION = 7,
    VCT_MODE_COMBINED_GI = 8
};

typedef struct vct_config {
    uint32_t voxel_resolution;       /* DEFAULT_VOXEL_RESOLUTION, EngineConfig.h:28 (power of two, 8..256) */
    uint32_t clip_levels;            /* DEFAULT_CLIP_REGION_COUNT, EngineConfig.h:26 (1..8; 0 = scratch only) */
    uint32_t radiance_format;        /* vct_format */
    uint32_t quirks;                 /* VCT_QUIRK_* bits */
    uint32_t raster_samples;         /* samples per voxel cell for triangle coverage: 1 or 4 */
    uint32_t downsample_region_size; /* DEFAULT_DOWNSAMPLE_REGION_SIZE, EngineConfig.h:29 */
    float extent_l0;                 /* DEFAULT_VOXEL_EXTENT_L0, EngineConfig.h:30 */
    uint32_t reserved;
} vct_config;

/* ClipmapRegion.h:10-19 */
typedef struct vct_clipmap_region {
    int32_t min_corner[3];
    uint32_t extent[3];
    float voxel_size;
} vct_clipmap_region;

/* Voxelizer.h:69-78 (std140 UBO) */
typedef struct vct_voxelization_desc {
    float region_min_corner[3];
    uint32_t clip_level;
    float region_max_corner[3];
    float clip_max_extent;
    float prev_region_min_corner[3];
    float voxel_size;
    float prev_region_max_corner[3];
    int32_t clipmap_resolution;
} vct_voxelization_desc;

/* DownSampler.h:49-55 */
typedef struct vct_downsample_desc {
    int32_t prev_region_min_corner[3];
    int32_t clip_level;
    int32_t clipmap_resolution;
    int32_t downsample_region_size;
} vct_downsample_desc;

/* ClipmapCleaner.h:40-47 (push constant) */
typedef struct vct_image_cleaning_desc {
    int32_t region_min_corner[3];
    int32_t clip_level;
    uint32_t clip_max_extent[3];
    int32_t clipmap_resolution;
    uint32_t face_count;
} vct_image_cleaning_desc;

/* BorderWrapper.h BorderWrappingDesc, values at BorderWrapper.cpp:69-73 */
typedef struct vct_border_wrapping_desc {
    uint32_t clipmap_resolution;
    uint32_t clip_border_width;
    uint32_t face_count;
    uint32_t clip_region_count;
} vct_border_wrapping_desc;

/* CopyAlpha.h CopyAlphaDesc, values at CopyAlpha.cpp:98-103 */
typedef struct vct_copy_alpha_desc {
    int32_t clip_level;
    int32_t clipmap_resolution;
    int32_t face_count;
} vct_copy_alpha_desc;

/* VoxelConeTracingPass.h:46-59 (push constant) */
typedef struct vct_cone_tracing_desc {
    float volume_center[3];
    uint32_t rendering_mode;
    float voxel_size;
    float volume_dimension;
    float trace_start_offset;
    float indirect_diffuse_intensity;
    float ambient_occlusion_factor;
    float min_trace_step_factor;
    float indirect_specular_intensity;
    float occlusion_decay;
    int32_t enable_32_cones;
} vct_cone_tracing_desc;

/* light.glsl:8-13 */
typedef struct vct_directional_light_desc {
    float direction[3];
    float intensity;
    float color[3];
    int32_t padding;
} vct_directional_light_desc;

/* light.glsl:15-20 (std140: 136 bytes of members, 144 with tail padding) */
typedef struct vct_directional_light_shadow_desc {
    float view[16]; /* column-major, glm layout */
    float proj[16];
    float z_near;
    float z_far;
    float padding[2];
} vct_directional_light_shadow_desc;

/* Camera.h:37-42 CameraUBO */
typedef struct vct_camera_ubo {
    float view_proj[16];
    float view_proj_inv[16];
    float eye_pos[3];
    int32_t padding;
} vct_camera_ubo;

/* gltf.glsl:8-26 GltfShadeMaterial (std430, 80 bytes) */
typedef struct vct_gltf_material {
    float base_color_factor[4];
    int32_t base_color_texture;
    float metallic_factor;
    float roughness_factor;
    int32_t metallic_roughness_texture;
    int32_t emissive_texture;
    int32_t alpha_mode;
    float alpha_cutoff;
    int32_t double_sided;
    float emissive_factor[3];
    int32_t normal_texture;
    float normal_texture_scale;
    int32_t occlusion_texture;
    float occlusion_texture_strength;
    int32_t padding;
} vct_gltf_material;

/* A scene texture (GLTFScene::uploadImage, GLTFScene.cpp:192-340): level 0 of
 * an RGBA8 UNORM image, row-major, texel (x, y) at texels[y * width + x],
 * texture coordinate (0, 0) at the corner of texel (0, 0). The voxel passes
 * sample it as the reference's samplers do: REPEAT addressing, LINEAR
 * filtering and maxLod 0 (GLTFScene.cpp:339, Sampler.cpp:29-50), so texture()
 * and textureLod() both read level 0; no mip levels are needed. */
typedef struct vct_texture {
    const uint32_t* texels; /* device */
    uint32_t width;
    uint32_t height;
} vct_texture;

/* One drawIndexed call of GLTFScene::cmdDraw (GLTFScene.cpp:457-489) with its
 * push constants {instanceIndex, materialIndex} (msaaVoxelizer.vert:25-29). */
typedef struct vct_primitive {
    uint32_t first_index;
    uint32_t index_count;
    int32_t vertex_offset;
    uint32_t instance_index;
    uint32_t material_index;
} vct_primitive;

/* The triangle stream of SceneManager (SceneManager.cpp:124-204): device
 * vertex/index/matrix/material buffers plus the host list of draws. */
typedef struct vct_mesh {
    const float* positions;     /* device, 3 floats per vertex */
    const float* normals;       /* device, 3 floats per vertex */
    const float* texcoords;     /* device, 2 floats per vertex, may be NULL */
    const uint32_t* indices;    /* device */
    const float* node_matrices; /* device, 32 floats per instance: model, itModel (NodeMatrix) */
    const vct_gltf_material* materials; /* device */
    uint32_t vertex_count;
    uint32_t index_count;
    uint32_t instance_count;
    uint32_t material_count;
    const vct_primitive* primitives; /* host */
    uint32_t primitive_count;
    float scene_bb_min[3]; /* SceneManager::getSceneBoundingBox, used by the SVO voxelizer */
    float scene_bb_max[3];
    /* uTextures (msaaVoxelizer.frag:18): host array of texture_count entries
     * that the materials' texture indices refer to; may be NULL when no drawn
     * material references a texture. The opacity voxelizer, the radiance
     * injector and the SVO voxelizer sample the occlusion, emissive and
     * base-colour textures as their shaders do (texcoords must then be set);
     * the dense method and the G-buffer producer refuse textured materials
     * (VCT_ERR_UNSUPPORTED). Texture indices are also bounds-checked on the
     * device: an index outside the table is never read. */
    const vct_texture* textures;
    uint32_t texture_count;
    /* The validation above reads the material table's texture indices back
     * to the host once per (materials, material_count, material_version);
     * a caller that rewrites a material table in place, or frees one and may
     * get a new one at the same address, gives it a new version. */
    uint32_t material_version;
} vct_mesh;

/* A D32 shadow map (DirectionalLight::createShadowMap, DirectionalLight.cpp:59-73). */
typedef struct vct_shadow_map {
    const float* depth; /* device, width*height floats, row-major */
    uint32_t width;
    uint32_t height;
} vct_shadow_map;

/* G-buffer targets read by the trace (GBufferPass.cpp:177-194). */
typedef struct vct_gbuffer {
    const uint32_t* diffuse;  /* RGBA8: rgb base colour, a perceptual roughness */
    const uint16_t* normal;   /* RGBA16F: rgb = n * 0.5 + 0.5 */
    const uint32_t* specular; /* RGBA8: rgb specular colour, a metallic */
    const uint16_t* emission; /* RGBA16F */
    const float* depth;       /* D32, 1.0 = background */
    uint32_t width;
    uint32_t height;
} vct_gbuffer;

/* ------------------------------------------------------------------------ */

/* Library version string. */
const char* vct_version(void);
/* Human-readable name of a status code. */
const char* vct_status_string(vct_status status);
/* Last error detail recorded by the calling thread (empty if none). */
const char* vct_last_error(void);

/* Voxelizer::initializeVoxelizer + createVoxelClipmap (Voxelizer.cpp:143-153,
 * 249-299): allocates both atlases and scratch. */
vct_status vct_create(vct_ctx** out_ctx, const vct_config* config);
vct_status vct_destroy(vct_ctx* ctx);

/* The "ClipmapRegions" resource (RenderPassManager put/get,
 * VoxelizationPass.cpp:335-357). */
vct_status vct_set_clip_regions(vct_ctx* ctx, const vct_clipmap_region* regions, uint32_t count);
vct_status vct_get_clip_regions(const vct_ctx* ctx, vct_clipmap_region* regions, uint32_t count);
/* Application::updateClipRegionBoundingBox (Application.cpp:116-128) followed by
 * VoxelizationPass::calculateChangeDelta (VoxelizationPass.cpp:438-449) from the
 * initial regions of createVoxelClipmap. Writes the regions into ctx. */
vct_status vct_update_clip_regions(vct_ctx* ctx, const float camera_pos[3]);
/* The same regions computed on the host without a context (no device call):
 * config->voxel_resolution, clip_levels and extent_l0 are read; `count` must
 * equal clip_levels. */
vct_status vct_compute_clip_regions(const vct_config* config, const float camera_pos[3],
                                    vct_clipmap_region* regions, uint32_t count);
/* Incremental (toroidal) update for a moving camera:
 * VoxelizationPass::fillRevoxelizationRegions (VoxelizationPass.cpp:450-494)
 * with calculateChangeDelta (:438-449) taken from the CURRENT regions. Moves
 * every level's region in ctx and writes the regions that must be cleared and
 * voxelized again: up to 3 slabs per level (one per axis that moved by at
 * least the level's minimum change) or the whole region when a move reaches
 * the extent. For a level >= 1 whose downsampled inner box (the cells under
 * the level below) moved, one more region: the old inner box within the new
 * region, whose cells still hold the previous mips (the reference omits it).
 * revox[level * VCT_REVOX_PER_LEVEL + i], i < counts[level]; revox holds
 * levels * VCT_REVOX_PER_LEVEL entries, counts `levels`. */
#define VCT_REVOX_PER_LEVEL 4
vct_status vct_move_clip_regions(vct_ctx* ctx, const float camera_pos[3], vct_clipmap_region* revox,
                                 uint32_t* counts);
/* Host half of Voxelizer::cmdVoxelize (Voxelizer.cpp:225-262). */
vct_status vct_make_voxelization_desc(const vct_ctx* ctx, uint32_t level,
                                      const vct_clipmap_region* region,
                                      vct_voxelization_desc* out_desc);

/* vkCmdClearColorImage of a whole atlas (VoxelizationPass.cpp:104-125). */
vct_status vct_clear_volume(vct_ctx* ctx, vct_stream stream, vct_volume volume);

/* K1: opacity voxelization of one clip level (VoxelizationPass::onUpdate,
 * VoxelizationPass.cpp:190-212; Voxelizer::cmdVoxelize, Voxelizer.h:36).
 * Stream-ordered; the host waits only when the draw list (mesh->primitives)
 * differs from the one the context uploaded last, for its upload. */
vct_status vct_voxelize_opacity(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh,
                                const vct_voxelization_desc* desc, uint32_t level);

/* K2: radiance injection of one clip level (RadianceInjectionPass::onUpdate,
 * RadianceInjectionPass.cpp:134-159). The host waits once per call for the
 * record count that sizes the sort (and for a changed draw list, as in
 * vct_voxelize_opacity); while the context's record buffer grows, the raster
 * runs a second time. */
vct_status vct_inject_radiance(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh,
                               const vct_shadow_map* shadow,
                               const vct_directional_light_desc* light,
                               const vct_directional_light_shadow_desc* light_shadow,
                               const vct_voxelization_desc* desc, uint32_t level);

/* VoxelizationPass::onUpdate's and RadianceInjectionPass::onUpdate's loops
 * over the clip levels (VoxelizationPass.cpp:190-212,
 * RadianceInjectionPass.cpp:134-159) as one call each (no counterpart as one
 * call): descs[i] (i < count <= 8, distinct clip levels) are voxelized or
 * injected exactly as `count` vct_voxelize_opacity / vct_inject_radiance
 * calls in that order would, bit for bit, with the levels' work in shared
 * launches; the injection waits on the host once for all levels (the record
 * count) and sorts once. */
vct_status vct_voxelize_opacity_levels(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh,
                                       const vct_voxelization_desc* descs, uint32_t count);
vct_status vct_inject_radiance_levels(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh,
                                      const vct_shadow_map* shadow, const vct_directional_light_desc* light,
                                      const vct_directional_light_shadow_desc* light_shadow,
                                      const vct_voxelization_desc* descs, uint32_t count);

/* K3: ClipmapCleaner::cmdClearOpacityClipRegion / cmdClearRadianceClipRegion
 * (ClipmapCleaner.h:27-37). */
vct_status vct_clear_region(vct_ctx* ctx, vct_stream stream, vct_volume volume,
                            const vct_image_cleaning_desc* desc);

/* K4/K5: DownSampler::cmdDownSampleOpacity / cmdDownSampleRadiance
 * (DownSampler.h:29-40). */
vct_status vct_downsample(vct_ctx* ctx, vct_stream stream, vct_volume volume,
                          const vct_downsample_desc* desc);

/* K6: CopyAlpha::cmdImageCopyAlpha (CopyAlpha.h:26): radiance.a <- opacity.a. */
/* K3 + K6 fused for an RGBA16F radiance level that is about to be injected
 * (no reference counterpart as one call): writes (0, 0, 0, opacity alpha)
 * to every interior texel of the level, and the next vct_inject_radiance of
 * that level keeps the stored alpha. begin + inject leaves the level exactly
 * as a full-level vct_clear_region, vct_inject_radiance and vct_copy_alpha do
 * (RadianceInjectionPass.cpp:64-159) with one pass over the level instead of
 * two. RGBA8 radiance keeps its count in alpha while injecting, so it is
 * refused (VCT_ERR_UNSUPPORTED). */
vct_status vct_begin_radiance_level(vct_ctx* ctx, vct_stream stream, uint32_t level);
vct_status vct_copy_alpha(vct_ctx* ctx, vct_stream stream, const vct_copy_alpha_desc* desc);

/* K7: BorderWrapper::cmdWrappingOpacityBorder / cmdWrappingRadianceBorder
 * (BorderWrapper.h:28-36). */
vct_status vct_wrap_border(vct_ctx* ctx, vct_stream stream, vct_volume volume,
                           const vct_border_wrapping_desc* desc);

/* K8: VoxelConeTracingPass::onUpdate (VoxelConeTracingPass.cpp:75-106).
 * diffuse_out / specular_out: device RGBA32F images (width*height*4 floats).
 * samples_out: optional device counter, incremented by the executed cone
 * samples. */
vct_status vct_trace(vct_ctx* ctx, vct_stream stream, const vct_gbuffer* gbuffer,
                     const vct_camera_ubo* camera, const vct_cone_tracing_desc* desc,
                     const vct_shadow_map* shadow,
                     const vct_directional_light_desc* light,
                     const vct_directional_light_shadow_desc* light_shadow,
                     float* diffuse_out, float* specular_out,
                     unsigned long long* samples_out);

/* No counterpart in the reference. vct_trace for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the same arguments, kernels and results, with the
 * specular work list in a static device array of the library instead of
 * context scratch, and the radiance occupancy mask rebuilt on every call
 * (vct_trace rebuilds it only when the host saw the radiance volume change,
 * which a replayed graph would not). It makes no allocation; an image needing
 * more than 2048 specular tiles of 64x64 pixels (32.25 MiB; 3840x2160 needs 2040) returns
 * VCT_ERR_CAPACITY. Every GPU holds one such array for the whole process and
 * all calls on that GPU share it: calls that overlap on one GPU race, so the
 * caller serializes them (one stream, or events between streams). */
vct_status vct_trace_noalloc(vct_ctx* ctx, vct_stream stream, const vct_gbuffer* gbuffer,
                             const vct_camera_ubo* camera, const vct_cone_tracing_desc* desc,
                             const vct_shadow_map* shadow, const vct_directional_light_desc* light,
                             const vct_directional_light_shadow_desc* light_shadow, float* diffuse_out,
                             float* specular_out, unsigned long long* samples_out);

/* Volume access in the reference atlas layout: a 3-D image of
 * ((res+2)*6) x ((res+2)*levels) x (res+2) texels, x fastest
 * (Voxelizer.h:40-52). RGBA8 texels are 4 bytes, RGBA16F 8 bytes.
 * dst/src are HOST pointers; the call synchronises the stream. */
vct_status vct_volume_bytes(const vct_ctx* ctx, vct_volume volume, size_t* out_bytes);
vct_status vct_read_volume(vct_ctx* ctx, vct_stream stream, vct_volume volume,
                           void* dst, size_t dst_bytes);
vct_status vct_write_volume(vct_ctx* ctx, vct_stream stream, vct_volume volume,
                            const void* src, size_t src_bytes);

/* ---- sparse voxel octree path ------------------------------------------ */

/* K9: SparseVoxelizer::preVoxelize + cmdVoxelize (SparseVoxelizer.h:34-53,
 * SparseVoxelizer.cpp:248-326): fragment list of uvec2 entries
 * (x | y<<12 | (z&0xff)<<24, (z>>8)<<28 | RGBA8 & 0x0fffffff) at resolution
 * 2^octree_level, normalised by mesh->scene_bb_min/max. *nfrag_out receives
 * getFragmentCount(). */
vct_status vct_svo_voxelize(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh, const vct_shadow_map* shadow,
                            const vct_directional_light_desc* light,
                            const vct_directional_light_shadow_desc* light_shadow, uint32_t octree_level,
                            uint32_t* nfrag_out);
/* K10-K15: OctreeBuilder::cmdBuild (OctreeBuilder.h:22-41, OctreeBuilder.cpp:205-345). */
vct_status vct_svo_build(vct_ctx* ctx, vct_stream stream);
/* OctreeBuilder::readOctreeNodes (OctreeBuilder.cpp:359-385): host copy of the
 * allocated nodes (2 uint32 per node); with out == NULL only *n_nodes is set. */
vct_status vct_svo_read_nodes(vct_ctx* ctx, vct_stream stream, uint32_t* out, size_t cap_nodes, size_t* n_nodes);
/* SparseVoxelizer::readVoxelFragments (SparseVoxelizer.cpp:328-353). */
vct_status vct_svo_read_fragments(vct_ctx* ctx, vct_stream stream, uint32_t* out, size_t cap_fragments,
                                  size_t* n_fragments);
/* 3^3 RGBA8 bricks of every node of the built octree (no reference shader
 * fills bricks; EngineConfig.h:24-25 only declares their counts). Texel
 * (i, j, k), i, j, k in {0, 1, 2}, of the brick of the node at depth d and
 * cell c (resolution 2^d) samples the node's depth at lattice point
 * 2c + (i, j, k) in half-cell units (corners, edge midpoints, face centres,
 * centre): per channel the mean of the colours of the depth-d nodes whose
 * cells touch the point (1, 2, 4 or 8 cells; a cell without a node, or
 * outside the grid, counts as zero), truncated to an integer. A trilinear
 * fetch anywhere in a node's cell then reads only its brick and is
 * continuous across nodes. Brick n belongs to node n (27 uint32, x fastest);
 * octree levels up to 10. */
vct_status vct_svo_build_bricks(vct_ctx* ctx, vct_stream stream);
/* Host copy of the bricks (27 uint32 each); with out == NULL only *n_bricks. */
vct_status vct_svo_read_bricks(vct_ctx* ctx, vct_stream stream, uint32_t* out, size_t cap_bricks,
                               size_t* n_bricks);
/* OctreeBuilder::getOctreeBuffer: device pointer to the node pool. */
vct_status vct_svo_pool(vct_ctx* ctx, const uint32_t** pool, uint32_t* node_count);
/* K16: OctreeVoxelConeTracing::onUpdate (OctreeVoxelConeTracing.cpp:74-104)
 * over the octree built by vct_svo_build. */
vct_status vct_svo_trace(vct_ctx* ctx, vct_stream stream, const vct_gbuffer* gbuffer, const vct_camera_ubo* camera,
                         const vct_cone_tracing_desc* desc, const vct_shadow_map* shadow,
                         const vct_directional_light_desc* light,
                         const vct_directional_light_shadow_desc* light_shadow, const float scene_bb_min[3],
                         const float scene_bb_max[3], float* diffuse_out, float* specular_out,
                         unsigned long long* samples_out);

/* ---- dense "plain" grid ------------------------------------------------
 * No reference counterpart as a whole: the reference's voxelization,
 * anisotropic downsample (radianceDownSample.comp:95-139) and cone march
 * (voxelConeTracing.frag:341-392) applied to one res^3 grid with a 6-face mip
 * pyramid down to 1^3 instead of a clipmap (level rule
 * clamp(log2(diameter / voxelSize), 0, levels - 1), zero outside the grid). */
typedef struct vct_dense vct_dense;

typedef struct vct_dense_desc {
    float grid_min[3];       /* world position of the grid's minimum corner */
    float extent;            /* world edge length of the cube */
    uint32_t resolution;     /* power of two */
    uint32_t format;         /* VCT_FORMAT_RGBA8 or VCT_FORMAT_RGBA32F */
    uint32_t raster_samples; /* 1 or 4 */
    uint32_t reserved;
} vct_dense_desc;

typedef struct vct_dense_trace_desc {
    float aperture;                   /* diffuse cone aperture (reference 0.872665 for 16 cones) */
    float max_distance;               /* MAX_TRACE_DISTANCE */
    float min_trace_step_factor;      /* max(0.2, .) is used, voxelConeTracing.frag:182 */
    float trace_start_offset;
    float occlusion_decay;
    float ambient_occlusion_factor;
    float indirect_diffuse_intensity;
    int32_t enable_32_cones;
} vct_dense_trace_desc;

vct_status vct_dense_create(vct_dense** out, const vct_dense_desc* desc);
vct_status vct_dense_destroy(vct_dense* grid);
vct_status vct_dense_clear(vct_dense* grid, vct_stream stream);
/* Device pointer, byte size and face count (1 for level 0, else 6) of one
 * level: faces of (n+2)^3 texels with a zero border, x fastest. */
vct_status vct_dense_level(vct_dense* grid, uint32_t level, void** ptr, size_t* bytes, uint32_t* faces);
vct_status vct_dense_read_level(vct_dense* grid, vct_stream stream, uint32_t level, void* dst, size_t bytes);
vct_status vct_dense_write_level(vct_dense* grid, vct_stream stream, uint32_t level, const void* src, size_t bytes);
/* K1 on the dense grid: level 0 texels covered by a triangle get
 * (base colour, 1) of the first covering triangle in draw order. */
vct_status vct_dense_voxelize(vct_dense* grid, vct_stream stream, const vct_mesh* mesh);
/* K5 applied level by level: levels 1..log2(res) from level 0. */
vct_status vct_dense_mipmap(vct_dense* grid, vct_stream stream);
/* K8's cone march over the dense pyramid: per pixel RGBA32F
 * (indirect diffuse rgb x albedo x intensity, ambient occlusion). */
vct_status vct_dense_trace(vct_dense* grid, vct_stream stream, const vct_gbuffer* gbuffer,
                           const vct_camera_ubo* camera, const vct_dense_trace_desc* desc, float* out,
                           unsigned long long* samples_out);

/* ---- adjoint of the dense method ------------------------------------------
 * No reference counterpart: gradients of an image loss with respect to the
 * level-0 albedo (rgb) and opacity (a) of an RGBA32F dense grid, through
 * vct_dense_mipmap and vct_dense_trace. The gradient pyramid has the grid's
 * layout (vct_dense_texels texels of 4 floats); the caller zeroes it, the
 * trace adjoint accumulates into it (float atomics, so the last bits depend
 * on summation order), the mip adjoint carries levels >= 1 down to level 0.
 * Typical use: set_level0, mipmap, trace (forward); zero grad; per view
 * trace_backward; mipmap_backward; level0_grad. */
enum {
    VCT_ADJOINT_RECOMPUTE = 1u << 0 /* fetch samples again in the reverse sweep instead of keeping them */
};
vct_status vct_dense_texels(vct_dense* grid, size_t* texels);
/* albedo: res^3 x 3 floats, opacity: res^3 floats (device, x fastest). */
vct_status vct_dense_set_level0(vct_dense* grid, vct_stream stream, const float* albedo, const float* opacity);
/* grad_out: d loss / d out of vct_dense_trace (RGBA32F per pixel). `status`
 * (device uint32, zeroed by the caller) gets bit 0 if a march outgrew the
 * per-lane state (only possible for step counts the call rejects) and bit 1
 * if the sweep's gradient records would have outgrown their buffer (never
 * with the library's sizing); the gradient is then incomplete. */
vct_status vct_dense_trace_backward(vct_dense* grid, vct_stream stream, const vct_gbuffer* gbuffer,
                                    const vct_camera_ubo* camera, const vct_dense_trace_desc* desc,
                                    const float* grad_out, float* grad_pyramid, uint32_t flags, uint32_t* status);
/* The adjoint tape (samples stored by the forward instead of recomputed by
 * the backward): vct_dense_trace_tape gives vct_dense_trace's output and
 * also writes every accumulated sample of every cone, with each pixel's step
 * count per cone, to `tape` (16-byte aligned device memory of at least
 * vct_dense_tape_bytes bytes: 16 bytes x cones x (march steps + 2) plus one
 * byte per cone for each pixel of the image rounded up to 16x16 tiles, then
 * the backward's work area, 72 bytes per 64 pixels and cone plus 16 per
 * pixel). vct_dense_trace_backward_tape then runs the reverse sweep on the
 * stored samples instead of marching again, with vct_dense_trace_backward's
 * result and status bits; it overwrites the tape with the sample gradients,
 * so one tape serves one backward. Backward calls of different views, each
 * with its own tape, touch nothing shared but the grid's pyramid (read) and
 * grad_pyramid (atomic adds): they may run concurrently on several streams. */
vct_status vct_dense_tape_bytes(vct_dense* grid, const vct_gbuffer* gbuffer, const vct_camera_ubo* camera,
                                const vct_dense_trace_desc* desc, size_t* bytes);
vct_status vct_dense_trace_tape(vct_dense* grid, vct_stream stream, const vct_gbuffer* gbuffer,
                                const vct_camera_ubo* camera, const vct_dense_trace_desc* desc, float* out,
                                unsigned long long* samples_out, void* tape, size_t tape_bytes);
vct_status vct_dense_trace_backward_tape(vct_dense* grid, vct_stream stream, const vct_gbuffer* gbuffer,
                                         const vct_camera_ubo* camera, const vct_dense_trace_desc* desc,
                                         const float* grad_out, void* tape, size_t tape_bytes, float* grad_pyramid,
                                         uint32_t* status);
/* vct_dense_trace_backward_tape over `views` views of one image size in one
 * call (host arrays of views G-buffers and cameras, and of device pointers to
 * each view's output gradient and tape, each tape of at least tape_bytes):
 * the views share the launches (one per 16 views), so no launch's tail
 * leaves the chip idle between views. */
vct_status vct_dense_trace_backward_tapes(vct_dense* grid, vct_stream stream, uint32_t views,
                                          const vct_gbuffer* gbuffers, const vct_camera_ubo* cameras,
                                          const vct_dense_trace_desc* desc, const float* const* grad_outs,
                                          void* const* tapes, size_t tape_bytes, float* grad_pyramid,
                                          uint32_t* status);
vct_status vct_dense_mipmap_backward(vct_dense* grid, vct_stream stream, float* grad_pyramid);
vct_status vct_dense_level0_grad(vct_dense* grid, vct_stream stream, const float* grad_pyramid, float* grad_albedo,
                                 float* grad_opacity);

/* ---- batched dense method ------------------------------------------------
 * No reference counterpart (the reference renders one scene): many scenes of
 * uneven size, each voxelized into its own dense grid (vct_dense_voxelize's
 * rule), mipped (vct_dense_mipmap) and traced (vct_dense_trace), captured
 * once in a hipGraph and replayed by vct_batch_run. One task table (vertex
 * chunks, raster chunks, mip bricks, pixel tiles); both schedules run the
 * vertex chunks, then the raster of every scene's triangles as one
 * sample-parallel pass; then, chosen by vct_batch_tuning.workgroups_per_cu,
 * 0 (the default) launches each mip level and the pixel tiles over every
 * scene as one grid each, and 1, 2 or 4 run ONE persistent launch whose
 * workgroups pull the mip bricks and pixel tiles from a queue ordered by
 * phase, a task waiting on its scene's completion counter of the previous
 * mip level, so small scenes trace while large ones still mip. Results are
 * bit-identical to the per-scene vct_dense_* calls under every schedule. */
typedef struct vct_batch vct_batch;

typedef struct vct_batch_scene {
    vct_mesh mesh;          /* device buffers + host primitive list */
    vct_dense_desc grid;    /* format must be VCT_FORMAT_RGBA8 */
    vct_gbuffer gbuffer;    /* diffuse, normal and depth are read */
    vct_camera_ubo camera;
    float* out;             /* device RGBA32F, gbuffer.width x gbuffer.height */
} vct_batch_scene;

/* Schedule knobs picked by vct_batch_autotune. */
typedef struct vct_batch_tuning {
    uint32_t raster_chunk;      /* triangles per raster task of the task table: 64, 128 or
                                   256 (vct_batch_plan); the launches raster sample-parallel
                                   whatever it is */
    uint32_t mip_brick;         /* destination brick edge of a mip task: 4 or 8 */
    uint32_t trace_schedule;    /* 0: one cone step per loop trip; 1: two-step lookahead
                                   (the samples of steps k and k+1 fetched together,
                                   step k+1 accumulated only if the march reaches it) */
    uint32_t workgroups_per_cu; /* 0: one flat launch per phase (vertex, raster, each mip
                                   level, trace), kernel boundaries as dependencies;
                                   1, 2 or 4: vertex and raster launches, then one
                                   persistent launch of CUs x this workgroups pulling the
                                   mip bricks and pixel tiles from a queue */
} vct_batch_tuning;

vct_status vct_batch_create(vct_batch** out, const vct_batch_scene* scenes, uint32_t count,
                            const vct_dense_trace_desc* trace);
vct_status vct_batch_destroy(vct_batch* batch);
vct_status vct_batch_set_tuning(vct_batch* batch, const vct_batch_tuning* tuning);
vct_status vct_batch_get_tuning(const vct_batch* batch, vct_batch_tuning* tuning);
/* Replays the captured launch on `stream`; `samples_out` (device, may be
 * NULL) accumulates executed cone samples and is not cleared here. */
vct_status vct_batch_run(vct_batch* batch, vct_stream stream, unsigned long long* samples_out);
/* Synchronizes and reports whether any task of the previous runs gave up
 * waiting on a dependency (VCT_ERR_HIP with a message), then clears it. */
vct_status vct_batch_status(vct_batch* batch);
/* Times candidate schedules `reps` times on `stream` (coordinate descent over
 * raster chunk, mip brick and trace schedule, phased launches only), keeps
 * the fastest and returns it (and its median milliseconds). */
vct_status vct_batch_autotune(vct_batch* batch, vct_stream stream, uint32_t reps, vct_batch_tuning* best,
                              float* best_ms);
/* Level `level` of scene `scene` in the vct_dense_level layout. */
vct_status vct_batch_level(vct_batch* batch, uint32_t scene, uint32_t level, void** ptr, size_t* bytes,
                           uint32_t* faces);
vct_status vct_batch_read_level(vct_batch* batch, vct_stream stream, uint32_t scene, uint32_t level, void* dst,
                                size_t bytes);
/* The task queue both schedules run, computed on the host without a batch
 * (no device call): for scenes of these shapes and tuning.raster_chunk /
 * tuning.mip_brick, *n_tasks tasks of 4 uint32 each {type << 24 | scene, a,
 * b, level} (type 0 vertex chunk [a, b) of triangles, 1 raster chunk [a, b),
 * 2 mip brick a of level `level`, 3 pixel tile a) in queue order, and
 * need[scene * VCT_BATCH_PHASES + phase] = the task count of each phase
 * (0 vertex, 1 raster, 1 + m mip level m); a raster task waits for phase 0 of
 * its scene, a level-m brick for phase m, a tile for the scene's last level.
 * With tasks == NULL only *n_tasks (and need, if non-NULL) are written. */
#define VCT_BATCH_PHASES 12
typedef struct vct_batch_scene_shape {
    uint32_t triangles;
    uint32_t resolution;
    uint32_t width;
    uint32_t height;
} vct_batch_scene_shape;
vct_status vct_batch_plan(const vct_batch_scene_shape* scenes, uint32_t count, const vct_batch_tuning* tuning,
                          uint32_t* tasks, size_t cap_tasks, size_t* n_tasks, uint32_t* need);

/* ---- the step after the trace ---------------------------------------------
 * SpecularFilterPass (SpecularFilterPass.cpp:75-90, specularFilter.frag:25-53):
 * out = diffuse + filtered specular (filter_method 1: the 32-direction x
 * 8-ring blur of filter.glsl:9-24; otherwise the 15x15 bilateral filter of
 * filter.glsl:26-63), then the Uncharted2 tonemap (tonemapping.glsl:9-26)
 * when tonemap_enable == 1. Images are RGBA32F width x height (device).
 * Defaults of the reference: gamma 2.2, exposure 0.1, tonemap off, filter 1
 * (SpecularFilterPass.h:48-51). */
typedef struct vct_specular_filter_desc { /* SpecularFilterPass.h:31-37 (push constant) */
    float tonemap_gamma;
    float tonemap_exposure;
    int32_t tonemap_enable;
    int32_t filter_method;
} vct_specular_filter_desc;
vct_status vct_specular_filter(vct_stream stream, const float* diffuse, const float* specular, uint32_t width,
                               uint32_t height, const vct_specular_filter_desc* desc, float* out);

/* ---- input producers (no reference counterpart on the hot path) ----------
 * The reference rasterizes its G-buffer (GBufferPass.cpp:170-206) and light
 * depth map (ReflectiveShadowMapPass.cpp:38-88) before the GI passes run;
 * these ray-cast equivalents produce the same channel layouts on the device
 * so that a frame can be driven without a rasterizer. `surface` holds five
 * floats per material: perceptual roughness, metallic, specular rgb. */
vct_status vct_raycast_gbuffer(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh, const float* surface,
                               const vct_camera_ubo* camera, uint32_t* diffuse, uint16_t* normal,
                               uint32_t* specular, uint16_t* emission, float* depth, uint32_t width,
                               uint32_t height);
vct_status vct_raycast_shadow(vct_ctx* ctx, vct_stream stream, const vct_mesh* mesh,
                              const vct_directional_light_shadow_desc* light_shadow, float* depth,
                              uint32_t width, uint32_t height);

#ifdef __cplusplus
}
#endif

#endif /* VCT_H */
