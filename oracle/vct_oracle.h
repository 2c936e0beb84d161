/*
 * vct_oracle.h - CPU restatement of the reference voxel-GI shaders.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Every function restates one GLSL program or helper of the reference from
 * its text (citations are path:line relative to the reference repository).
 * Volumes use the reference atlas layout: a 3-D image of
 * W = (res+2)*6, H = (res+2)*levels, D = res+2 texels with x fastest
 * (Voxelizer.h:40-52, Voxelizer.cpp:249-270). RGBA8 texels are uint32 with R
 * in the low byte; RGBA16F texels are 4 x uint16 half floats.
 *
 * Arithmetic: FP32 with IEEE basic operations and no contraction
 * (gcc -ffp-contract=off), so integer/byte results (occupancy, UNORM8 mips,
 * RGBA8 CAS averages, node pools) are reproducible bit for bit. The cone
 * trace also has an FP64 twin (orc_trace_f64) used to state error.
 *
 * Order-dependent reference operations (imageAtomicCompSwap running averages,
 * atomicAdd node allocation) are serialised in a fixed order: fragments in
 * (triangle, cell row, cell column, sample) order, flagged nodes in pool-index
 * order.
 */
#ifndef VCT_ORACLE_H
#define VCT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_FMT_RGBA8 = 0, ORC_FMT_RGBA16F = 1 };

enum {
    ORC_QUIRK_CLAMP_X_RANGE = 1u << 0,
    ORC_QUIRK_BORDER_SKIP_FAR = 1u << 1,
    ORC_QUIRK_SHADOW_RAW_DEPTH = 1u << 2,
    ORC_QUIRK_SVO_HALF_COORDS = 1u << 3,
    ORC_QUIRK_SVO_MIP_LEVEL_TEST = 1u << 4,
    ORC_QUIRK_SVO_UNTEXTURED_BLACK = 1u << 5
};

typedef struct orc_volume {
    void* texels;   /* uint32_t* (RGBA8) or uint16_t* (RGBA16F, 4 per texel) */
    int fmt;
    int res;        /* interior resolution (128) */
    int levels;     /* clip levels (6) */
} orc_volume;

/* An RGBA8 texture (R in the low byte), rows of `width` texels. */
typedef struct orc_texture {
    const uint32_t* texels;
    uint32_t width, height;
} orc_texture;

/* World-space triangle soup after the vertex stage (msaaVoxelizer.vert:31-37). */
typedef struct orc_tris {
    const float* pos;      /* 9 floats per triangle */
    const float* nrm;      /* 9 floats per triangle (itModel * normal, not normalised) */
    const uint32_t* mat;   /* material index per triangle */
    uint32_t count;
    const float* uv;       /* 6 floats per triangle (the vertices' texcoords) */
    const orc_texture* textures; /* NULL: no texture lookups at all */
    uint32_t texture_count;
} orc_tris;

/* Texture indices are -1 for none; they are only read when the soup has a
 * texture table. */
typedef struct orc_material {
    float base_color[4];
    float emissive[3];
    int32_t base_color_texture, emissive_texture, occlusion_texture;
} orc_material;

typedef struct orc_light {
    float direction[3];
    float intensity;
    float color[3];
    float view[16];
    float proj[16];
    float z_near;
    float z_far;
} orc_light;

typedef struct orc_shadow {
    const float* depth;
    int width;
    int height;
} orc_shadow;

typedef struct orc_vox_desc {
    int region_cell_min[3];    /* ClipmapRegion.minCorner of the level (voxels) */
    int region_cell_extent[3]; /* ClipmapRegion.extent */
    float region_min[3];
    float region_max[3];
    float clip_max_extent;
    float voxel_size;
    int clip_level;
    int clipmap_resolution;
} orc_vox_desc;

typedef struct orc_trace_params {
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
    int enable_32_cones;
    int clip_levels;
    float view_proj_inv[16];
    float eye_pos[3];
} orc_trace_params;

typedef struct orc_gbuffer {
    const uint32_t* diffuse;
    const uint16_t* normal;
    const uint32_t* specular;
    const uint16_t* emission;
    const float* depth;
    int width;
    int height;
} orc_gbuffer;

/* ---- numeric helpers (atomic.glsl:4-22, Vulkan UNORM / half conversions) ---- */
void orc_unpack_rgba8(uint32_t v, float out[4]);          /* convRGBA8ToVec4 */
uint32_t orc_pack_rgba8(const float v[4]);                /* convVec4ToRGBA8 */
uint32_t orc_float_to_unorm8_rgba(const float v[4]);      /* imageStore to RGBA8_UNORM */
uint16_t orc_float_to_half(float f);
float orc_half_to_float(uint16_t h);
uint32_t orc_rgba8_avg(uint32_t stored, const float value[4]); /* one uncontended CAS-loop step */

/* ---- atlas kernels ---- */
size_t orc_volume_texels(int res, int levels);
void orc_clear_region(orc_volume* vol, const int min_corner[3], const int extent[3],
                      int clip_level, int face_count);                 /* clipmapCleaning.comp */
void orc_copy_alpha(orc_volume* radiance, const orc_volume* opacity, int clip_level,
                    int face_count);                                   /* copyAlphaImage.comp */
void orc_wrap_border(orc_volume* vol, int border_width_total, int face_count,
                     int region_count, int skip_far);                  /* borderWrapping.comp */
void orc_downsample_opacity(orc_volume* vol, const int prev_min_corner[3], int clip_level,
                            int region_size);                          /* opacityDownSample.comp */
void orc_downsample_radiance(orc_volume* vol, const int prev_min_corner[3], int clip_level,
                             int region_size);                         /* radianceDownSample.comp */

/* ---- triangle voxelization (K1) and radiance injection (K2) ---- */
int orc_dominant_axis(const float* p);                     /* msaaVoxelizer.geom:27-32 */
void orc_voxelize_opacity(orc_volume* vol, const orc_tris* tris, const orc_vox_desc* d,
                          int samples, uint32_t quirks);
/* The same with the materials' occlusion textures (msaaVoxelizer.frag:64). */
void orc_voxelize_opacity_textured(orc_volume* vol, const orc_tris* tris, const orc_material* mats,
                                   const orc_vox_desc* d, int samples, uint32_t quirks);
/* Bilinear level-0 REPEAT sample of a texture (the reference's samplers:
 * GLTFScene.cpp:339, Sampler.cpp:29-50) and the texcoord of a fragment. */
void orc_sample_texture(const orc_texture* t, const float uv[2], float out[4]);
void orc_fragment_uv(const orc_tris* tris, uint32_t tri, const double lam[3], float uv[2]);
float orc_visibility(const orc_shadow* sm, const orc_light* light, const float pos[3],
                     uint32_t quirks);                     /* shadow.glsl:28-36 */
void orc_inject_radiance(orc_volume* vol, const orc_tris* tris, const orc_material* mats,
                         const orc_shadow* sm, const orc_light* light, const orc_vox_desc* d,
                         int samples, uint32_t quirks);

/* ---- cone tracing (K8) ---- */
float orc_calc_min_level(const orc_trace_params* p, const float pos[3]);
void orc_trace_cone(const orc_volume* rad, const orc_trace_params* p, const float start[3],
                    const float dir[3], float aperture, float max_dist, float start_level,
                    float step_factor, float out[4], uint32_t* steps);
void orc_trace(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
               const orc_shadow* sm, const orc_light* light, uint32_t quirks,
               int row_begin, int row_end, float* diffuse_out, float* specular_out,
               uint64_t* samples);
/* orc_trace over an arbitrary list of rows (rows run in parallel, one per thread at a time) */
void orc_trace_rowlist(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
                       const orc_shadow* sm, const orc_light* light, uint32_t quirks, const int* rows,
                       int nrows, float* diffuse_out, float* specular_out, uint64_t* samples);
void orc_trace_rowlist_f64(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
                           const orc_shadow* sm, const orc_light* light, uint32_t quirks, const int* rows,
                           int nrows, double* diffuse_out, double* specular_out, uint64_t* samples);
void orc_trace_f64(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
                   const orc_shadow* sm, const orc_light* light, uint32_t quirks,
                   int row_begin, int row_end, double* diffuse_out, double* specular_out,
                   uint64_t* samples);
void orc_brdf(const float n[3], const float v[3], const float l[3], const float diffuse_color[3],
              const float specular_color[3], float perceptual_roughness, float metallic,
              float out[3]);                                /* brdf.glsl:65-78 */

/* test hooks */
uint32_t orc_raster_list(const float* pos, uint32_t count, float voxel_size, const int lo[3], const int hi[3],
                         int samples, float* out, uint32_t cap);
void orc_sample_clipmap(const orc_volume* vol, const orc_trace_params* p, const float pos[3], int level,
                        const float dir[3], float out[4]);

/* ---- G-buffer / shadow producers (ray cast) ---- */
void orc_raycast_gbuffer(const orc_tris* tris, const orc_material* mats,
                         const float* roughness_metal, /* per material: perceptual roughness, metallic, spec rgb */
                         const float view_proj_inv[16], const float view_proj[16],
                         int width, int height, uint32_t* diffuse, uint16_t* normal,
                         uint32_t* specular, uint16_t* emission, float* depth);
void orc_raycast_shadow(const orc_tris* tris, const float view[16], const float proj[16],
                        int width, int height, float* depth);

/* ---- sparse voxel octree (K9-K16) ---- */
uint32_t orc_svo_fragments(const orc_tris* tris, const orc_material* mats, const orc_shadow* sm,
                           const orc_light* light, const float bb_min[3], const float bb_max[3],
                           int resolution, int samples, uint32_t quirks, uint32_t* frags, uint32_t cap);
int orc_svo_build(const uint32_t* frags, uint32_t nfrag, int level, uint32_t quirks,
                  uint32_t* pool, uint32_t pool_cap, uint32_t* node_count);
/* 3^3 RGBA8 bricks of every node of a built pool (build-defined): texel
 * (i, j, k) of node n at depth d and cell c is the truncated mean of the
 * depth-d node colours of the cells touching lattice point 2c + (i, j, k)
 * (half-cell units; outside cells count as zero). bricks: 27 per node,
 * x fastest. Returns 0, or -1 for a bad pool. */
int orc_svo_bricks(const uint32_t* pool, uint32_t node_count, int level, uint32_t* bricks);
/* test hooks: fragment words (voxelizer.frag:99-100) and their position decode
 * (octreeNodeFlag.comp:28-32, >> 1 with ORC_QUIRK_SVO_HALF_COORDS); one sampleSVO */
void orc_svo_pack_fragments(const uint32_t* xyz, const uint32_t* rgba8, uint32_t n, uint32_t* out);
void orc_svo_unpack_fragments(const uint32_t* frags, uint32_t n, uint32_t quirks, uint32_t* xyz);
void orc_sample_svo(const uint32_t* pool, const orc_trace_params* p, const float bb_min[3], const float bb_max[3],
                    const float pos[3], uint32_t level, float out[4]);
void orc_svo_trace(const uint32_t* pool, const orc_gbuffer* g, const orc_trace_params* p,
                   const float bb_min[3], const float bb_max[3], const orc_shadow* sm,
                   const orc_light* light, uint32_t quirks, int row_begin, int row_end,
                   float* diffuse_out, float* specular_out, uint64_t* samples);

/* ---- dense "plain" grid with a full mip pyramid (orc_dense.c) ---- */
typedef struct orc_dense {
    void* levels[16]; /* level 0: 1 face, level m >= 1: 6 faces; (n+2)^3 texels per face, zero border */
    int fmt;          /* ORC_FMT_RGBA8 or ORC_FMT_RGBA32F */
    int res;
    int nlevels;      /* log2(res) + 1 */
    float grid_min[3];
    float extent;
} orc_dense;

typedef struct orc_dense_params {
    float aperture;
    float max_dist;
    float step_factor;
    float start_offset;
    float decay;
    float ao_factor;
    float diffuse_intensity;
    int cones32;
    float view_proj_inv[16];
} orc_dense_params;

enum { ORC_FMT_RGBA32F = 2 };

size_t orc_dense_level_texels(int res, int level);
void orc_dense_voxelize(orc_dense* g, const orc_tris* tris, const orc_material* mats, int samples);
void orc_dense_mip(orc_dense* g);
void orc_dense_trace(const orc_dense* g, const orc_gbuffer* gb, const orc_dense_params* p, int row_begin,
                     int row_end, double* out, uint64_t* samples);

/* ---- FP64 adjoint of the dense method (orc_adjoint.c) ---- */
typedef struct orc_adj orc_adj;
orc_adj* orc_adj_create(int res, float extent, const float grid_min[3]);
void orc_adj_free(orc_adj* a);
void orc_adj_set_level0(orc_adj* a, const double* rgba);
void orc_adj_mip(orc_adj* a);
void orc_adj_render(orc_adj* a, const orc_gbuffer* gb, const orc_dense_params* P, double* out);
void orc_adj_backward(orc_adj* a, const orc_gbuffer* gb, const orc_dense_params* P, const double* grad_out,
                      double* grad_level0);

/* ---- SpecularFilterPass (orc_filter.c) ---- */
void orc_specular_filter(const float* diffuse, const float* specular, int W, int H, float gamma, float exposure,
                         int tonemap_enable, int filter_method, float* out);

#ifdef __cplusplus
}
#endif

#endif
