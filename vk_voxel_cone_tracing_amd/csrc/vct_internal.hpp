// vct_internal.hpp - context, memory layout and device helpers shared by the
// gfx950 kernels of libvct.
//
// Volume layout in HBM. Each volume (opacity, radiance) is a dense array of
// (levels x 6 faces) blocks of D^3 texels, D = res + 2 (one border texel per
// side), x fastest inside a block:
//     index(level, face, x, y, z) = ((level*6 + face)*D + z)*D*D + y*D + x
// This is the reference atlas (Voxelizer.h:40-52: x = face*D + x, y = level*D
// + y) with the face/level blocks made contiguous, so one trilinear footprint
// stays inside one 2.2 MB block and the per-level kernels stream contiguous
// memory. RGBA8 texels are uint32 (R in the low byte); RGBA16F texels are
// 4 halves (8 bytes).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/vct.h"

namespace vct {

constexpr int kFaces = 6;
constexpr int kMaxLevels = 8;
constexpr size_t kPinnedWordBytes = 512;  // vct_ctx::pinned_word

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Grow-only device scratch allocation.
vct_status ensure(DeviceBuffer& b, size_t bytes);
void release(DeviceBuffer& b);

// Per-thread error detail for vct_last_error().
vct_status fail(vct_status s, const char* fmt, ...);

#define VCT_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return ::vct::fail(VCT_ERR_HIP, "%s failed: %s (%s:%d)", #expr,              \
                               hipGetErrorString(_e), __FILE__, __LINE__);               \
    } while (0)

#define VCT_TRY(expr)                           \
    do {                                        \
        vct_status _s = (expr);                 \
        if (_s != VCT_OK) return _s;            \
    } while (0)

struct DevTexture;

// World-space triangle soup produced by the vertex stage.
struct TriSoup {
    const float* pos;     // 9 floats per triangle
    const float* nrm;     // 9 floats per triangle
    const uint32_t* mat;  // material index per triangle
    const float* uv;      // 6 floats per triangle (zero without texcoords)
    const DevTexture* tex;  // the mesh's textures on the device (null without a table)
    uint32_t tex_count;
    uint32_t count;
};

// What a pass does with materials that reference textures.
enum class TextureUse { kSample, kRefuse, kIgnore };

struct SvoState {
    DeviceBuffer frags;      // uvec2 per fragment
    DeviceBuffer pool;       // uvec2 per node
    DeviceBuffer counters;   // small device counters
    uint32_t nfrag = 0;
    uint32_t level = 0;
    uint32_t pool_nodes = 0; // capacity
    uint32_t node_count = 0; // 8 + 8 * allocated blocks, known after build
    bool built = false;
    // pool range [begin, begin + num) of the nodes of each depth 1..L
    uint32_t level_begin[12] = {}, level_num[12] = {};
    uint32_t depths = 0;     // depths with nodes
    DeviceBuffer cells;      // packed cell of every node (brick builder)
    DeviceBuffer grids;      // dense node-colour grid of every depth (brick builder)
    DeviceBuffer bricks;     // 27 RGBA8 per node
    bool bricks_built = false;
};

}  // namespace vct

struct vct_ctx {
    vct_config cfg{};
    int res = 0;
    int D = 0;
    int levels = 0;
    int device = 0;
    vct_clipmap_region regions[vct::kMaxLevels]{};
    bool regions_set = false;

    vct::DeviceBuffer opacity;   // uint32 texels
    vct::DeviceBuffer radiance;  // uint32 or uint2 texels

    // scratch
    vct::DeviceBuffer soup_pos, soup_nrm, soup_mat;
    vct::DeviceBuffer prims, prim_first;
    vct::DeviceBuffer counts, offsets, scan_tmp, totals;
    vct::DeviceBuffer rec_key[2], rec_idx[2], rec_val, sort_tmp;
    vct::DeviceBuffer staging;
    vct::DeviceBuffer materials_flat;
    vct::DeviceBuffer spec_perm;         // roughness-sorted pixel order of the specular cones
    bool keep_alpha[vct::kMaxLevels]{};  // set by vct_begin_radiance_level, used by the next injection
    uint32_t inject_cap = 0;             // injection record capacity (grow-only)
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    const vct::DeviceBuffer* pinned_dst = nullptr;  // buffer that holds the pinned block's first pinned_valid bytes
    size_t pinned_valid = 0;
    void* pinned_word = nullptr;  // pinned landing slot of small device-to-host reads (kPinnedWordBytes)

    vct::SvoState svo;

    // Occupancy of the radiance volume for the trace's empty-space skip:
    // occ_raw holds one bit per texel position of a level (any face nonzero),
    // occ_words words per level followed by occ_pad zero words; occ_mask the
    // same bits dilated to the 2x2x2 footprint (occ_words per level). Every
    // entry point that writes radiance bumps radiance_version; the trace
    // rebuilds the mask when mask_version differs.
    vct::DeviceBuffer occ_raw, occ_mask;
    uint32_t occ_words = 0, occ_pad = 0;
    uint64_t radiance_version = 1, mask_version = 0;

    // The texture indices of the last material table read back, keyed by
    // the table's pointer, count and the mesh's material_version.
    const void* mat_table = nullptr;
    uint32_t mat_count = 0, mat_version = 0;
    int32_t* mat_textures = nullptr;  // base-colour, emissive, occlusion texture index per material
    vct::DeviceBuffer soup_uv;
};

namespace vct {

// Records a write to the radiance volume (the occupancy mask is stale).
inline void radiance_written(vct_ctx* c) { ++c->radiance_version; }

inline size_t texels_per_volume(const vct_ctx* c)
{
    return (size_t)c->D * c->D * c->D * kFaces * (size_t)c->levels;
}
inline size_t texel_bytes(const vct_ctx* c, vct_volume v)
{
    return (v == VCT_VOLUME_RADIANCE && c->cfg.radiance_format == VCT_FORMAT_RGBA16F) ? 8u : 4u;
}

__host__ __device__ inline size_t vol_index(int D, int level, int face, int x, int y, int z)
{
    return ((((size_t)level * kFaces + (size_t)face) * (size_t)D + (size_t)z) * (size_t)D + (size_t)y) * (size_t)D +
           (size_t)x;
}

// Children are indexed i = x | y << 1 | z << 2. Face f looks along axis
// f >> 1 (x, y, z), towards + for even f; its 4 front/back pairs are the child
// pairs that differ only in that axis' bit, listed by increasing index of the
// child with the bit clear, the front child being the one with the bit clear
// for + faces and set for - faces (radianceDownSample.comp:95-139).
__host__ __device__ constexpr int face_pair_child(int f, int p, int back)
{
    const int axis = f >> 1;
    const int c = ((p >> axis) << (axis + 1)) | (p & ((1 << axis) - 1));
    const int set = (f & 1) ? !back : back;
    return set ? c | (1 << axis) : c;
}

// Workgroups of a launch are dispatched to the device's 8 XCDs round-robin
// (workgroup b runs on XCD b % 8) and every XCD has its own L2. This maps the
// dispatch index b of n workgroups to a work index such that each XCD walks
// one contiguous eighth of the work: neighbouring pixel tiles, whose cones
// fetch overlapping voxels, then share an L2. It is a bijection on [0, n).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n)
{
    const unsigned xcd = b & 7u, k = b >> 3, per = n >> 3, rem = n & 7u;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// Mesh vertex stage (msaaVoxelizer.vert:31-37): world-space triangles, their
// texture coordinates and the mesh's texture table on the device. A draw
// list whose materials reference a base-colour, emissive or occlusion
// texture is refused (VCT_ERR_UNSUPPORTED) with TextureUse::kRefuse; with
// kSample every such index must name one of the mesh's textures.
vct_status run_vertex_stage(vct_ctx* ctx, hipStream_t s, const vct_mesh* mesh, TriSoup* out,
                            TextureUse use = TextureUse::kSample);

// The base-colour, emissive and occlusion texture indices of every material
// of a device material table (3 per material), read back to the host.
vct_status read_texture_indices(hipStream_t s, const vct_gltf_material* mats, uint32_t count, int32_t* idx);

// Looks up, once per device, the device address of the allocation-free
// trace's static work list (vct_trace_noalloc); called by vct_create.
vct_status trace_scratch_init(int device);

// Pinned staging copy of a small host array to a device buffer (stream ordered).
vct_status upload_small(vct_ctx* ctx, hipStream_t s, DeviceBuffer& dst, const void* src, size_t bytes);

}  // namespace vct
