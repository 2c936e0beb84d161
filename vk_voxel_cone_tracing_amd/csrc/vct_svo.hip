// vct_svo.hip - sparse voxel octree path for gfx950: fragment-list
// voxelization (K9), the per-level node init/flag/alloc build (K10-K13), leaf
// and mipmap colour writes (K14/K15) and octree cone tracing (K16).
//
// Determinism. The reference hands out fragment-list slots and child blocks
// with atomicAdd, so its layout depends on invocation order. Here:
//  * every covered raster sample gets its slot from an exclusive scan over the
//    (triangle, cell, sample) work items, i.e. rasterization order;
//  * child blocks are handed to the flagged nodes of a level in pool-index
//    order with a wavefront ballot/popcount prefix scan (three passes: wave
//    counts, block offsets, assignment), which is the reference's outcome
//    when invocations run in index order;
//  * the order-dependent leaf running average is applied per leaf in
//    fragment order after a radix sort keyed by (leaf, fragment).
// The node pool is therefore a pure function of the fragment list.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "vct_internal.hpp"
#include "vct_raster.hpp"
#include "vct_shadow.hpp"
#include "vct_texel.hpp"
#include "vct_textures.hpp"

namespace vct {

struct SvoRaster {
    float vs;  // 2 / R: one raster cell per voxel of the normalised [-1, 1] box
    int lo[3], hi[3];
    int samples;
    int res;
    uint32_t quirks;
};

__global__ void k_svo_normalize(const float* __restrict__ pos, uint32_t n9, float cx, float cy, float cz, float ev,
                                float* __restrict__ out)
{
#pragma clang fp contract(off)
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n9) return;
    const uint32_t k = i % 3u;
    const float c = k == 0 ? cx : (k == 1 ? cy : cz);
    out[i] = (pos[i] - c) / ev;  // voxelizer.vert:42-47
}

__global__ void k_svo_count(const float* __restrict__ npos, uint32_t ntris, SvoRaster R,
                            unsigned long long* __restrict__ counts)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    float p[9];
    for (int k = 0; k < 9; ++k) p[k] = npos[9u * t + k];
    TriSetup s;
    unsigned long long n = 0;
    if (tri_setup(p, R.vs, R.lo, R.hi, s)) n = (unsigned long long)s.nu * s.nv * (unsigned long long)R.samples;
    counts[t] = n;
}

// The total of an exclusive scan: its last element plus the last input.
template <typename T>
__global__ void k_scan_total(const T* __restrict__ scan, const T* __restrict__ in, uint32_t n,
                             unsigned long long* __restrict__ out)
{
    out[0] = (unsigned long long)scan[n - 1] + (unsigned long long)in[n - 1];
}

// covered[wi] = 1 for every covered sample: each is one fragment invocation
// and consumes one fragment-list slot (voxelizer.frag:55).
__global__ __launch_bounds__(256) void k_svo_cover(const float* __restrict__ npos,
                                                   const unsigned long long* __restrict__ offsets, uint32_t ntris,
                                                   uint32_t nwork, SvoRaster R, TexInputs T,
                                                   uint32_t* __restrict__ covered)
{
    const uint32_t base = blockIdx.x * blockDim.x;
    const TriRange tr = block_triangles(offsets, ntris, base, min(base + blockDim.x, nwork) - 1u);
    const uint32_t wi = base + threadIdx.x;
    if (wi >= nwork) return;
    const uint32_t t = work_triangle(offsets, tr.lo, tr.hi, wi);
    float p[9];
    for (int k = 0; k < 9; ++k) p[k] = npos[9u * t + k];
    TriSetup s;
    uint32_t c = 0;
    if (tri_setup(p, R.vs, R.lo, R.hi, s)) {
        float fp[3];
        double lam[3];
        c = tri_sample(p, s, R.vs, R.samples, (uint32_t)(wi - offsets[t]), fp, lam) ? 1u : 0u;
        if (c && T.tex) {
            // voxelizer.frag:52-53: occlusion-texture discard, before the slot
            float uv[2];
            interp_uv(T.uv + 6u * t, lam, uv);
            if (occluded(T, T.mats[T.mat[t]], uv)) c = 0u;
        }
    }
    covered[wi] = c;
}

// The work index of every fragment: covered sample wi is fragment slot[wi].
__global__ void k_svo_compact(const uint32_t* __restrict__ covered, const uint32_t* __restrict__ slot, uint32_t nwork,
                              uint32_t* __restrict__ work_of)
{
    const uint32_t wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi < nwork && covered[wi]) work_of[slot[wi]] = wi;
}

// voxelizer.frag:48-100, one thread per fragment (so every lane of a wave
// has a fragment to light), in fragment order.
__global__ __launch_bounds__(256) void k_svo_write(const float* __restrict__ npos, const float* __restrict__ nrm,
                                                   TexInputs T, const unsigned long long* __restrict__ offsets,
                                                   uint32_t ntris,
                                                   uint32_t nfrag, SvoRaster R, LightParams L,
                                                   const uint32_t* __restrict__ work_of, uint2* __restrict__ frags)
{
#pragma clang fp contract(off)
    const uint32_t f0 = blockIdx.x * blockDim.x;
    const TriRange tr = block_triangles(offsets, ntris, work_of[f0], work_of[min(f0 + blockDim.x, nfrag) - 1u]);
    const uint32_t cur = f0 + threadIdx.x;
    if (cur >= nfrag) return;
    const uint32_t wi = work_of[cur];
    const uint32_t t = work_triangle(offsets, tr.lo, tr.hi, wi);
    float p[9];
    for (int k = 0; k < 9; ++k) p[k] = npos[9u * t + k];
    TriSetup s;
    if (!tri_setup(p, R.vs, R.lo, R.hi, s)) return;
    float fp[3];
    double lam[3];
    if (!tri_sample(p, s, R.vs, R.samples, (uint32_t)(wi - offsets[t]), fp, lam)) return;
    float q[3];
    for (int k = 0; k < 3; ++k) q[k] = (fp[k] + 1.0f) * 0.5f;  // biasAndScale (voxelizer.geom:27-30)
    const vct_gltf_material& m = T.mats[T.mat[t]];
    float uv[2] = {0.0f, 0.0f};
    if (T.tex) interp_uv(T.uv + 6u * t, lam, uv);
    float rad[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (m.emissive_factor[0] > 0.0f || m.emissive_factor[1] > 0.0f || m.emissive_factor[2] > 0.0f) {
        // voxelizer.frag:60-69; the emissive texture adds to the factor
        float em[3] = {m.emissive_factor[0], m.emissive_factor[1], m.emissive_factor[2]};
        if (has_texture(T, m.emissive_texture)) {
            const f4 te = sample_texture(T.tex[m.emissive_texture], uv);
            em[0] = em[0] + te.x;
            em[1] = em[1] + te.y;
            em[2] = em[2] + te.z;
        }
        for (int k = 0; k < 3; ++k) rad[k] = fminf(fmaxf(em[k], 0.0f), 1.0f);
    } else {
        // voxelizer.frag:72-77: texture x factor (black without a texture
        // under VCT_QUIRK_SVO_UNTEXTURED_BLACK)
        float color[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        if (!(R.quirks & VCT_QUIRK_SVO_UNTEXTURED_BLACK) || m.base_color_texture > -1)
            for (int k = 0; k < 4; ++k) color[k] = m.base_color_factor[k];
        if (has_texture(T, m.base_color_texture)) {
            const f4 tb = sample_texture(T.tex[m.base_color_texture], uv);
            color[0] = tb.x * color[0];
            color[1] = tb.y * color[1];
            color[2] = tb.z * color[2];
            color[3] = tb.w * color[3];
        }
        const float* n0 = nrm + 9u * t;
        float ni[3];
        for (int k = 0; k < 3; ++k)
            ni[k] = (float)(lam[0] * (double)n0[k] + lam[1] * (double)n0[3 + k] + lam[2] * (double)n0[6 + k]);
        const float len = sqrtf(ni[0] * ni[0] + ni[1] * ni[1] + ni[2] * ni[2]);
        const float nn[3] = {ni[0] / len, ni[1] / len, ni[2] / len};
        float ndl = nn[0] * L.dir[0] + nn[1] * L.dir[1] + nn[2] * L.dir[2];
        ndl = fminf(fmaxf(ndl, 0.001f), 1.0f);
        // calcVisibility at fs_in.position, the normalised [0,1] position (voxelizer.frag:86)
        const float vis = visibility(L, q, R.quirks);
        float lc[3];
        for (int k = 0; k < 3; ++k) lc[k] = ndl * vis * L.color[k] * L.intensity;
        if (lc[0] == 0.0f && lc[1] == 0.0f && lc[2] == 0.0f) return;  // discard: the slot stays zero
        for (int k = 0; k < 3; ++k) rad[k] = fminf(fmaxf(lc[k] * color[k] * color[3], 0.0f), 1.0f);
        rad[3] = 1.0f;
    }
    uint32_t vp[3];
    for (int k = 0; k < 3; ++k) {
        const float f = q[k] * (float)R.res;
        uint32_t u = f > 0.0f ? (uint32_t)f : 0u;
        vp[k] = u > (uint32_t)R.res ? (uint32_t)R.res : u;
    }
    const f4 v255 = {rad[0] * 255.0f, rad[1] * 255.0f, rad[2] * 255.0f, rad[3] * 255.0f};
    frags[cur] = make_uint2(vp[0] | (vp[1] << 12u) | ((vp[2] & 0xffu) << 24u),
                            ((vp[2] >> 8u) << 28u) | (pack_raw(v255) & 0x0fffffffu));
}

__global__ void k_fill_u2(uint2* __restrict__ p, uint32_t n, uint2 v)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ void frag_pos(uint2 f, bool half, uint32_t p[3])
{
    p[0] = f.x & 0xfffu;
    p[1] = (f.x >> 12u) & 0xfffu;
    p[2] = ((f.x >> 24u) & 0xffu) | ((f.y >> 20u) & 0xf00u);
    if (half) {
        p[0] >>= 1;
        p[1] >>= 1;
        p[2] >>= 1;
    }
}

// The descent shared by octreeNodeFlag/LeafWrite/MipmapWrite.
__device__ __forceinline__ uint32_t walk(const uint2* __restrict__ pool, uint32_t p[3], uint32_t res, uint32_t stop,
                                         uint32_t& res_out)
{
    uint32_t idx = 0, cur = 0;
    do {
        res >>= 1;
        const uint32_t cx = p[0] >= res, cy = p[1] >= res, cz = p[2] >= res;
        idx = cur + (cz | (cx << 1) | (cy << 2));
        cur = pool[idx].x & 0x7fffffffu;
        p[0] -= cx * res;
        p[1] -= cy * res;
        p[2] -= cz * res;
    } while (cur != 0u && res > stop);
    res_out = res;
    return idx;
}

// octreeNodeFlag.comp: mark the deepest existing node of every fragment.
// The pass at `depth` would walk from the root along the previous pass's
// path and one step further (the node a fragment reached then was flagged,
// so it has children now): node_of keeps each fragment's node between passes
// and the pass takes that one step. The child slot at depth d is bit
// L - 1 - d of the position (z | x << 1 | y << 2, as in walk()); a coordinate
// of 2^L takes walk()'s path for 2^L - 1.
__global__ __launch_bounds__(256) void k_svo_flag(uint2* __restrict__ pool, const uint2* __restrict__ frags,
                                                  uint32_t nfrag, uint32_t L, uint32_t depth, bool half,
                                                  uint32_t* __restrict__ node_of)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfrag) return;
    uint32_t p[3];
    frag_pos(frags[f], half, p);
    const uint32_t top = (1u << L) - 1u, b = L - 1u - depth;
    const uint32_t x = min(p[0], top), y = min(p[1], top), z = min(p[2], top);
    const uint32_t slot = ((z >> b) & 1u) | (((x >> b) & 1u) << 1) | (((y >> b) & 1u) << 2);
    uint32_t idx = slot;
    if (depth > 0) {
        const uint32_t parent = node_of[f];
        const uint32_t cur = pool[parent].x & 0x7fffffffu;
        idx = cur ? cur + slot : parent;  // walk() stops at a node without children
    }
    node_of[f] = idx;
    // Setting a set bit changes nothing, so the pool ends the same when the
    // lanes of a wave that reach one node leave it to one of them, and when a
    // flag already visible is not set again. Near the root the fragments of a
    // wave reach a few nodes (nearly all fragments reach one node at the
    // first depth), and atomics on one address serialize: up to four of a
    // wave's nodes are flagged by one lane each, the rest lane by lane.
    const uint32_t lane = __lane_id();
    bool todo = true;
#pragma unroll
    for (int round = 0; round < 4; ++round) {
        const unsigned long long left = __ballot(todo);
        if (left == 0ull) break;
        const int leader = __ffsll((long long)left) - 1;
        const uint32_t node = __builtin_amdgcn_readlane(idx, leader);
        if (todo && idx == node) {
            todo = false;
            if ((int)lane == leader && !(pool[idx].x & 0x80000000u)) atomicOr(&pool[idx].x, 0x80000000u);
        }
    }
    if (todo && !(pool[idx].x & 0x80000000u)) atomicOr(&pool[idx].x, 0x80000000u);
}

// octreeNodeAlloc.comp with ordered allocation. Pass 1: per-wave flagged counts.
__global__ __launch_bounds__(256) void k_svo_alloc_count(const uint2* __restrict__ pool, uint32_t begin, uint32_t num,
                                                         uint32_t* __restrict__ wave_counts)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool flagged = i < num && (pool[begin + i].x & 0x80000000u);
    const unsigned long long ballot = __ballot(flagged);
    if ((threadIdx.x & 63) == 0) wave_counts[i >> 6] = (uint32_t)__popcll(ballot);
}

// Pass 3: rank = wave offset + popcount of lower lanes; child block
// (base + rank + 1) << 3 with the subdivide flag kept.
__global__ __launch_bounds__(256) void k_svo_alloc_assign(uint2* __restrict__ pool, uint32_t begin, uint32_t num,
                                                          const uint32_t* __restrict__ wave_offsets, uint32_t base)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool flagged = i < num && (pool[begin + i].x & 0x80000000u);
    const unsigned long long ballot = __ballot(flagged);
    if (!flagged) return;
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long lower = lane ? (ballot & (~0ull >> (64 - lane))) : 0ull;
    const uint32_t rank = wave_offsets[i >> 6] + (uint32_t)__popcll(lower);
    pool[begin + i] = make_uint2(((base + rank + 1u) << 3u) | 0x80000000u, 0u);
}

// octreeNodeModifyArg.comp: the child blocks handed out so far, on the device
// (the last wave's offset plus its count), read back once per level.
__global__ void k_svo_advance(const uint32_t* __restrict__ wave_offsets, const uint32_t* __restrict__ wave_counts,
                              uint32_t waves, uint32_t* __restrict__ counter)
{
    counter[0] += wave_offsets[waves - 1] + wave_counts[waves - 1];
}

// octreeNodeLeafWrite.comp, ordered: one key (leaf << fbits | fragment) per
// fragment that reaches a leaf; fragments that do not take `none`, above
// every such key, and sort last.
__global__ __launch_bounds__(256) void k_svo_leaf_keys(const uint2* __restrict__ pool, const uint2* __restrict__ frags,
                                                       uint32_t nfrag, uint32_t R, bool half, uint32_t fbits,
                                                       unsigned long long none, unsigned long long* __restrict__ keys)
{
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfrag) return;
    uint32_t p[3], r;
    frag_pos(frags[f], half, p);
    const uint32_t idx = walk(pool, p, R, 1u, r);
    keys[f] = r == 1u ? (((unsigned long long)idx << fbits) | f) : none;
}

__global__ __launch_bounds__(256) void k_svo_leaf_apply(uint2* __restrict__ pool, const uint2* __restrict__ frags,
                                                        const unsigned long long* __restrict__ keys, uint32_t n,
                                                        uint32_t fbits, unsigned long long none)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || keys[i] == none) return;
    const unsigned long long leaf = keys[i] >> fbits;
    if (i > 0 && (keys[i - 1] >> fbits) == leaf) return;
    const unsigned long long fmask = (1ull << fbits) - 1ull;
    uint32_t t = pool[leaf].y;
    for (uint32_t j = i; j < n && keys[j] != none && (keys[j] >> fbits) == leaf; ++j)
        t = rgba8_avg(t, unpack_raw(frags[(uint32_t)(keys[j] & fmask)].y & 0x0fffffffu));
    pool[leaf].y = t;
}

// octreeNodeMipmapWrite.comp evaluated per node: every flagged node of the
// depth range gets the truncating average of its eight children. The
// reference reaches exactly these nodes (the ancestors of fragments) once per
// fragment and writes the same value each time.
__global__ __launch_bounds__(256) void k_svo_mip(uint2* __restrict__ pool, uint32_t begin, uint32_t num)
{
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const uint2 node = pool[begin + i];
    if (!(node.x & 0x80000000u)) return;
    const uint32_t child = node.x & 0x7fffffffu;
    if (child == 0u) return;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t j = 0; j < 8; ++j) {
        const f4 v = unpack_raw(pool[child + j].y);
        s[0] = s[0] + v.x / 8.0f;
        s[1] = s[1] + v.y / 8.0f;
        s[2] = s[2] + v.z / 8.0f;
        s[3] = s[3] + v.w / 8.0f;
    }
    pool[begin + i].y = pack_raw(f4{s[0], s[1], s[2], s[3]});
}

static LightParams make_light(const vct_shadow_map* shadow, const vct_directional_light_desc* light,
                              const vct_directional_light_shadow_desc* ls)
{
    LightParams L{};
    float nd[3] = {-light->direction[0], -light->direction[1], -light->direction[2]};
    float len = std::sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
    for (int k = 0; k < 3; ++k) {
        L.dir[k] = nd[k] / len;
        L.color[k] = light->color[k];
    }
    L.intensity = light->intensity;
    for (int k = 0; k < 16; ++k) {
        L.view[k] = ls->view[k];
        L.proj[k] = ls->proj[k];
    }
    L.z_near = ls->z_near;
    L.z_far = ls->z_far;
    L.depth = shadow->depth;
    L.width = (int)shadow->width;
    L.height = (int)shadow->height;
    return L;
}

static unsigned blocks_for(uint64_t n, unsigned threads = 256)
{
    return (unsigned)((n + threads - 1) / threads);
}

}  // namespace vct

using namespace vct;

extern "C" {

// SparseVoxelizer::preVoxelize + cmdVoxelize (SparseVoxelizer.cpp:248-326).
vct_status vct_svo_voxelize(vct_ctx* c, vct_stream stream, const vct_mesh* mesh, const vct_shadow_map* shadow,
                            const vct_directional_light_desc* light,
                            const vct_directional_light_shadow_desc* light_shadow, uint32_t octree_level,
                            uint32_t* nfrag_out)
{
    if (!c || !mesh || !shadow || !shadow->depth || !light || !light_shadow)
        return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (octree_level < 1 || octree_level > 11)
        return fail(VCT_ERR_INVALID_ARGUMENT, "octree_level must be in [1, 11] (EngineConfig.h:20-21)");
    hipStream_t s = (hipStream_t)stream;
    TriSoup soup;
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup));
    const uint32_t R = 1u << octree_level;
    c->svo.level = octree_level;
    c->svo.built = false;
    c->svo.nfrag = 0;
    if (nfrag_out) *nfrag_out = 0;
    if (!soup.count) return VCT_OK;
    // voxelizer.vert:42-47 normalisation
    float ext[3], center[3];
    for (int k = 0; k < 3; ++k) {
        ext[k] = mesh->scene_bb_max[k] - mesh->scene_bb_min[k];
        center[k] = (mesh->scene_bb_min[k] + mesh->scene_bb_max[k]) * 0.5f;
    }
    const float ev = std::fmax(ext[0], std::fmax(ext[1], ext[2])) * 0.5f;
    if (!(ev > 0.0f)) return fail(VCT_ERR_INVALID_ARGUMENT, "empty scene bounding box");
    const uint32_t n9 = soup.count * 9u;
    VCT_TRY(ensure(c->rec_val, 4u * (size_t)n9));
    float* npos = (float*)c->rec_val.ptr;
    hipLaunchKernelGGL(k_svo_normalize, dim3(blocks_for(n9)), dim3(256), 0, s, soup.pos, n9, center[0], center[1],
                       center[2], ev, npos);
    SvoRaster Rr;
    Rr.vs = 2.0f / (float)R;
    for (int k = 0; k < 3; ++k) {
        Rr.lo[k] = -(int)(R / 2);
        Rr.hi[k] = (int)(R / 2) - 1;
    }
    Rr.samples = (int)c->cfg.raster_samples;
    Rr.res = (int)R;
    Rr.quirks = c->cfg.quirks;
    VCT_TRY(ensure(c->counts, 8u * (size_t)soup.count));
    VCT_TRY(ensure(c->offsets, 8u * (size_t)soup.count));
    VCT_TRY(ensure(c->totals, 64));
    hipLaunchKernelGGL(k_svo_count, dim3(blocks_for(soup.count)), dim3(256), 0, s, npos, soup.count, Rr,
                       (unsigned long long*)c->counts.ptr);
    size_t tmp = 0;
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (unsigned long long*)c->counts.ptr,
                                             (unsigned long long*)c->offsets.ptr, (int)soup.count, s));
    VCT_TRY(ensure(c->scan_tmp, tmp));
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, (unsigned long long*)c->counts.ptr,
                                             (unsigned long long*)c->offsets.ptr, (int)soup.count, s));
    if (!c->pinned_word) VCT_HIP(hipHostMalloc(&c->pinned_word, kPinnedWordBytes, hipHostMallocDefault));
    volatile unsigned long long* landed = (volatile unsigned long long*)c->pinned_word;
    hipLaunchKernelGGL(k_scan_total<unsigned long long>, dim3(1), dim3(1), 0, s,
                       (const unsigned long long*)c->offsets.ptr, (const unsigned long long*)c->counts.ptr, soup.count,
                       (unsigned long long*)c->totals.ptr);
    VCT_HIP(hipMemcpyAsync((void*)landed, c->totals.ptr, 8, hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    const unsigned long long nwork64 = landed[0];
    if (nwork64 >= 0xffffffffull) return fail(VCT_ERR_CAPACITY, "raster work %llu exceeds 2^32", nwork64);
    const uint32_t nwork = (uint32_t)nwork64;
    if (!nwork) return VCT_OK;
    VCT_TRY(ensure(c->rec_idx[0], 4u * (size_t)nwork));
    VCT_TRY(ensure(c->rec_idx[1], 4u * (size_t)nwork));
    uint32_t* covered = (uint32_t*)c->rec_idx[0].ptr;
    uint32_t* slot = (uint32_t*)c->rec_idx[1].ptr;
    const TexInputs T{soup.uv, soup.mat, mesh->materials, soup.tex, soup.tex_count};
    hipLaunchKernelGGL(k_svo_cover, dim3(blocks_for(nwork)), dim3(256), 0, s, npos,
                       (const unsigned long long*)c->offsets.ptr, soup.count, nwork, Rr, T, covered);
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, covered, slot, (int)nwork, s));
    VCT_TRY(ensure(c->scan_tmp, tmp));
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, covered, slot, (int)nwork, s));
    // the count pass's host read-back (SparseVoxelizer.cpp:279)
    hipLaunchKernelGGL(k_scan_total<uint32_t>, dim3(1), dim3(1), 0, s, (const uint32_t*)slot, (const uint32_t*)covered,
                       nwork, (unsigned long long*)c->totals.ptr);
    VCT_HIP(hipMemcpyAsync((void*)landed, c->totals.ptr, 8, hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    const uint32_t nfrag = (uint32_t)landed[0];
    VCT_TRY(ensure(c->svo.frags, 8u * (size_t)(nfrag ? nfrag : 1)));
    if (nfrag) {
        hipLaunchKernelGGL(k_fill_u2, dim3(blocks_for(nfrag)), dim3(256), 0, s, (uint2*)c->svo.frags.ptr, nfrag,
                           make_uint2(0u, 0u));
        const LightParams L = make_light(shadow, light, light_shadow);
        VCT_TRY(ensure(c->rec_key[0], 4u * (size_t)nfrag));
        uint32_t* work_of = (uint32_t*)c->rec_key[0].ptr;
        hipLaunchKernelGGL(k_svo_compact, dim3(blocks_for(nwork)), dim3(256), 0, s, covered, slot, nwork, work_of);
        hipLaunchKernelGGL(k_svo_write, dim3(blocks_for(nfrag)), dim3(256), 0, s, npos, soup.nrm, T,
                           (const unsigned long long*)c->offsets.ptr, soup.count, nfrag, Rr, L,
                           work_of, (uint2*)c->svo.frags.ptr);
        VCT_HIP(hipGetLastError());
    }
    c->svo.nfrag = nfrag;
    if (nfrag_out) *nfrag_out = nfrag;
    return VCT_OK;
}

// ---- 3^3 RGBA8 bricks (build-defined; EngineConfig.h:24-25 declares brick
// counts, no reference shader fills bricks). Texel (i, j, k) of the brick of
// node n at depth d, cell c (resolution 2^d) is the lattice point
// 2c + (i, j, k) in half-cell units: per channel the mean over the depth-d
// cells touching it ({c-1, c}, {c} or {c, c+1} per axis for i = 0, 1, 2;
// cells outside the grid count as zero) of their node colours, truncated.
// Cells are packed x | y << 10 | z << 20 (depth <= 10).

__device__ __forceinline__ uint32_t pack_cell(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 10) | (z << 20); }

// Root children: slot s = z | x << 1 | y << 2 (octreeNodeFlag.comp:40).
__global__ void k_brick_root_cells(uint32_t* __restrict__ cells)
{
    const uint32_t s = threadIdx.x;
    if (s < 8u) cells[s] = pack_cell((s >> 1) & 1u, (s >> 2) & 1u, s & 1u);
}

// The 8 children of every subdivided node of one depth range.
__global__ __launch_bounds__(256) void k_brick_child_cells(const uint2* __restrict__ pool, uint32_t begin,
                                                           uint32_t num, uint32_t node_count,
                                                           uint32_t* __restrict__ cells)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const uint32_t n = begin + i;
    const uint32_t child = pool[n].x & 0x7fffffffu;
    if (!child || child + 8u > node_count) return;
    const uint32_t cc = cells[n];
    const uint32_t x = 2u * (cc & 1023u), y = 2u * ((cc >> 10) & 1023u), z = 2u * (cc >> 20);
#pragma unroll
    for (uint32_t s = 0; s < 8u; ++s) cells[child + s] = pack_cell(x + ((s >> 1) & 1u), y + ((s >> 2) & 1u), z + (s & 1u));
}

// The depth's node colours into its dense grid (2^d cells per axis).
__global__ __launch_bounds__(256) void k_brick_scatter(const uint2* __restrict__ pool, const uint32_t* __restrict__ cells,
                                                       uint32_t begin, uint32_t num, uint32_t depth,
                                                       uint32_t* __restrict__ grid)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const uint32_t n = begin + i, cc = cells[n], r = 1u << depth;
    const size_t at = ((size_t)(cc >> 20) * r + ((cc >> 10) & 1023u)) * r + (cc & 1023u);
    grid[at] = pool[n].y;
}

// RGBA8 -> four 16-bit lanes of a 64-bit word (sums of up to 8 bytes fit a lane).
__device__ __forceinline__ unsigned long long widen(uint32_t v)
{
    return (unsigned long long)(v & 0xffu) | ((unsigned long long)((v >> 8) & 0xffu) << 16) |
           ((unsigned long long)((v >> 16) & 0xffu) << 32) | ((unsigned long long)(v >> 24) << 48);
}

// Per-lane floor(lane / 2^sh), repacked to RGBA8. Every lane holds a sum of
// at most 2^sh bytes, so the quotient fits 8 bits; the mask keeps the next
// lane's low bits (shifted down into bits 16 - sh and up) out of it.
__device__ __forceinline__ uint32_t narrow(unsigned long long v, uint32_t sh)
{
    const unsigned long long m = 0xffull;
    return (uint32_t)(((v >> sh) & m) | (((v >> (16 + sh)) & m) << 8) | (((v >> (32 + sh)) & m) << 16) |
                      (((v >> (48 + sh)) & m) << 24));
}

__global__ __launch_bounds__(256) void k_brick_fill(const uint32_t* __restrict__ cells, uint32_t begin, uint32_t num,
                                                    uint32_t depth, const uint32_t* __restrict__ grid,
                                                    uint32_t* __restrict__ bricks)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const uint32_t n = begin + i, cc = cells[n];
    const int r = 1 << depth;
    const int c[3] = {(int)(cc & 1023u), (int)((cc >> 10) & 1023u), (int)(cc >> 20)};
    // the 3x3x3 cells around c, widened, x fastest: v[dz][dy][dx] for d in {-1, 0, 1}
    unsigned long long v[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int x = c[0] + t % 3 - 1, y = c[1] + (t / 3) % 3 - 1, z = c[2] + t / 9 - 1;
        const bool in = x >= 0 && y >= 0 && z >= 0 && x < r && y < r && z < r;
        v[t] = in ? widen(grid[((size_t)z * (size_t)r + (size_t)y) * (size_t)r + (size_t)x]) : 0ull;
    }
    // separable sums: per axis texel 0 takes cells {-1, 0}, 1 takes {0}, 2 takes {0, +1}
    unsigned long long sx[27], sy[27];
#pragma unroll
    for (int zy = 0; zy < 9; ++zy) {
        sx[zy * 3 + 0] = v[zy * 3 + 0] + v[zy * 3 + 1];
        sx[zy * 3 + 1] = v[zy * 3 + 1];
        sx[zy * 3 + 2] = v[zy * 3 + 1] + v[zy * 3 + 2];
    }
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            sy[(z * 3 + 0) * 3 + x] = sx[(z * 3 + 0) * 3 + x] + sx[(z * 3 + 1) * 3 + x];
            sy[(z * 3 + 1) * 3 + x] = sx[(z * 3 + 1) * 3 + x];
            sy[(z * 3 + 2) * 3 + x] = sx[(z * 3 + 1) * 3 + x] + sx[(z * 3 + 2) * 3 + x];
        }
    uint32_t* out = bricks + 27u * (size_t)n;
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const unsigned long long a = sy[(0 * 3 + y) * 3 + x], b = sy[(1 * 3 + y) * 3 + x], d = sy[(2 * 3 + y) * 3 + x];
            const uint32_t shxy = (uint32_t)(x != 1) + (uint32_t)(y != 1);
            out[(0 * 3 + y) * 3 + x] = narrow(a + b, shxy + 1u);
            out[(1 * 3 + y) * 3 + x] = narrow(b, shxy);
            out[(2 * 3 + y) * 3 + x] = narrow(b + d, shxy + 1u);
        }
}

// OctreeBuilder::cmdBuild (OctreeBuilder.cpp:205-345).
vct_status vct_svo_build(vct_ctx* c, vct_stream stream)
{
    if (!c) return fail(VCT_ERR_INVALID_ARGUMENT, "null context");
    if (!c->svo.level) return fail(VCT_ERR_NOT_READY, "vct_svo_voxelize has not run");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t L = c->svo.level, R = 1u << L, nfrag = c->svo.nfrag;
    const bool half = (c->cfg.quirks & VCT_QUIRK_SVO_HALF_COORDS) != 0;
    // OctreeBuilder.cpp:110-113: clamp(8 * fragments, 1e6, 5e8) nodes
    uint64_t cap = (uint64_t)nfrag << 3;
    if (cap < 1000000ull) cap = 1000000ull;
    if (cap > 500000000ull) cap = 500000000ull;
    VCT_TRY(ensure(c->svo.pool, 8u * (size_t)cap));
    c->svo.pool_nodes = (uint32_t)cap;
    uint2* pool = (uint2*)c->svo.pool.ptr;
    const uint2* frags = (const uint2*)c->svo.frags.ptr;
    uint32_t begin = 0, num = 8, counter = 0;
    std::vector<uint32_t> level_begin, level_num;
    VCT_TRY(ensure(c->counts, 8u * ((cap >> 6) + 8)));
    VCT_TRY(ensure(c->offsets, 8u * ((cap >> 6) + 8)));
    VCT_TRY(ensure(c->rec_idx[1], 4u * (size_t)(nfrag ? nfrag : 1)));
    uint32_t* node_of = (uint32_t*)c->rec_idx[1].ptr;  // each fragment's node between flag passes
    VCT_TRY(ensure(c->svo.counters, 16));
    uint32_t* dev_counter = (uint32_t*)c->svo.counters.ptr;
    VCT_HIP(hipMemsetAsync(dev_counter, 0, 4, s));
    if (!c->pinned_word) VCT_HIP(hipHostMalloc(&c->pinned_word, kPinnedWordBytes, hipHostMallocDefault));
    volatile uint32_t* host_counter = (volatile uint32_t*)c->pinned_word;
    for (uint32_t i = 1; i <= L; ++i) {
        if (num == 0) break;  // no node was flagged: the tree ends here (the reference's zero-group dispatches)
        if ((uint64_t)begin + num > cap) return fail(VCT_ERR_CAPACITY, "octree node pool overflow (%llu nodes)", (unsigned long long)cap);
        level_begin.push_back(begin);
        level_num.push_back(num);
        // octreeNodeInit.comp
        hipLaunchKernelGGL(k_fill_u2, dim3(blocks_for(num)), dim3(256), 0, s, pool + begin, num, make_uint2(0u, 0u));
        // octreeNodeFlag.comp
        if (nfrag)
            hipLaunchKernelGGL(k_svo_flag, dim3(blocks_for(nfrag)), dim3(256), 0, s, pool, frags, nfrag, L, i - 1u, half,
                               node_of);
        if (i == L) break;
        // octreeNodeAlloc.comp + octreeNodeModifyArg.comp
        const uint32_t waves = (num + 63) / 64;
        uint32_t* wave_counts = (uint32_t*)c->counts.ptr;
        uint32_t* wave_offsets = (uint32_t*)c->offsets.ptr;
        hipLaunchKernelGGL(k_svo_alloc_count, dim3(blocks_for(num)), dim3(256), 0, s, pool, begin, num, wave_counts);
        size_t tmp = 0;
        VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, wave_counts, wave_offsets, (int)waves, s));
        VCT_TRY(ensure(c->scan_tmp, tmp));
        VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, wave_counts, wave_offsets, (int)waves, s));
        hipLaunchKernelGGL(k_svo_alloc_assign, dim3(blocks_for(num)), dim3(256), 0, s, pool, begin, num, wave_offsets,
                           counter);
        hipLaunchKernelGGL(k_svo_advance, dim3(1), dim3(1), 0, s, wave_offsets, wave_counts, waves, dev_counter);
        VCT_HIP(hipMemcpyAsync((void*)host_counter, dev_counter, 4, hipMemcpyDeviceToHost, s));
        VCT_HIP(hipStreamSynchronize(s));
        counter = *host_counter;
        begin += num;
        num = (counter << 3) + 8u - begin;
    }
    VCT_HIP(hipGetLastError());
    const uint64_t total = ((uint64_t)counter << 3) + 8u;
    if (total > cap) return fail(VCT_ERR_CAPACITY, "octree node pool overflow");
    c->svo.node_count = (uint32_t)total;
    // octreeNodeLeafWrite.comp (fragment-ordered running average)
    if (nfrag) {
        for (int b = 0; b < 2; ++b) {
            VCT_TRY(ensure(c->rec_key[b], 8u * (size_t)nfrag));
        }
        // key bits: fragment index below, leaf index above; a leaf index is
        // below total - 1, so the all-ones `none` is above every real key
        auto bits = [](unsigned long long v) {
            int b = 0;
            while (b < 64 && (v >> b) != 0ull) ++b;
            return b;
        };
        const uint32_t fbits = (uint32_t)std::max(1, bits(nfrag - 1u));
        const int end_bit = (int)fbits + bits(total);
        const unsigned long long none = end_bit >= 64 ? ~0ull : (1ull << end_bit) - 1ull;
        unsigned long long* k0 = (unsigned long long*)c->rec_key[0].ptr;
        unsigned long long* k1 = (unsigned long long*)c->rec_key[1].ptr;
        hipLaunchKernelGGL(k_svo_leaf_keys, dim3(blocks_for(nfrag)), dim3(256), 0, s, pool, frags, nfrag, R, half,
                           fbits, none, k0);
        size_t tmp = 0;
        VCT_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, k0, k1, (int)nfrag, 0, end_bit, s));
        VCT_TRY(ensure(c->sort_tmp, tmp));
        VCT_HIP(hipcub::DeviceRadixSort::SortKeys(c->sort_tmp.ptr, tmp, k0, k1, (int)nfrag, 0, end_bit, s));
        hipLaunchKernelGGL(k_svo_leaf_apply, dim3(blocks_for(nfrag)), dim3(256), 0, s, pool, frags, k1, nfrag, fbits,
                           none);
    }
    // octreeNodeMipmapWrite.comp for i = 2..L (OctreeBuilder.cpp:328-344). The
    // target of pass i is the node whose descent resolution equals i (quirk:
    // only i in {2, 4, 8} exist) or 2^(i-1); that node sits at depth
    // L - log2(target), whose pool range is level_begin/num of that depth.
    const bool quirk = (c->cfg.quirks & VCT_QUIRK_SVO_MIP_LEVEL_TEST) != 0;
    for (uint32_t i = 2; i <= L; ++i) {
        uint32_t target = quirk ? i : (1u << (i - 1));
        if (target & (target - 1)) continue;  // not a power of two: no node has that resolution
        uint32_t lg = 0;
        while ((1u << lg) < target) ++lg;
        if (lg >= L) continue;
        const uint32_t depth = L - lg;  // 1-based depth of the target nodes
        if (depth < 1 || depth > level_begin.size()) continue;
        hipLaunchKernelGGL(k_svo_mip, dim3(blocks_for(level_num[depth - 1])), dim3(256), 0, s, pool,
                           level_begin[depth - 1], level_num[depth - 1]);
    }
    VCT_HIP(hipGetLastError());
    c->svo.depths = (uint32_t)level_begin.size();
    for (uint32_t d = 0; d < c->svo.depths; ++d) {
        c->svo.level_begin[d] = level_begin[d];
        c->svo.level_num[d] = level_num[d];
    }
    c->svo.built = true;
    c->svo.bricks_built = false;
    return VCT_OK;
}

vct_status vct_svo_build_bricks(vct_ctx* c, vct_stream stream)
{
    if (!c) return fail(VCT_ERR_INVALID_ARGUMENT, "null context");
    if (!c->svo.built) return fail(VCT_ERR_NOT_READY, "vct_svo_build has not run");
    const uint32_t L = c->svo.level;
    if (L > 10) return fail(VCT_ERR_UNSUPPORTED, "bricks are built for octree levels up to 10 (level %u)", L);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t n = c->svo.node_count;
    size_t grid_total = 0;
    unsigned long long off[12] = {};
    for (uint32_t d = 1; d <= c->svo.depths; ++d) {
        off[d] = grid_total;
        grid_total += (size_t)1 << (3 * d);
    }
    VCT_TRY(ensure(c->svo.cells, 4u * (size_t)n));
    VCT_TRY(ensure(c->svo.grids, 4u * (grid_total ? grid_total : 1)));
    VCT_TRY(ensure(c->svo.bricks, 4u * 27u * (size_t)n));
    const uint2* pool = (const uint2*)c->svo.pool.ptr;
    uint32_t* cells = (uint32_t*)c->svo.cells.ptr;
    uint32_t* grids = (uint32_t*)c->svo.grids.ptr;
    VCT_HIP(hipMemsetAsync(grids, 0, 4u * grid_total, s));
    hipLaunchKernelGGL(k_brick_root_cells, dim3(1), dim3(64), 0, s, cells);
    for (uint32_t d = 1; d <= c->svo.depths; ++d) {
        const uint32_t b = c->svo.level_begin[d - 1], num = c->svo.level_num[d - 1];
        if (!num) continue;
        if (d < c->svo.depths)
            hipLaunchKernelGGL(k_brick_child_cells, dim3(blocks_for(num)), dim3(256), 0, s, pool, b, num, n, cells);
        hipLaunchKernelGGL(k_brick_scatter, dim3(blocks_for(num)), dim3(256), 0, s, pool, cells, b, num, d,
                           grids + off[d]);
    }
    for (uint32_t d = 1; d <= c->svo.depths; ++d) {
        const uint32_t b = c->svo.level_begin[d - 1], num = c->svo.level_num[d - 1];
        if (!num) continue;
        hipLaunchKernelGGL(k_brick_fill, dim3(blocks_for(num)), dim3(256), 0, s, (const uint32_t*)cells, b, num, d,
                           (const uint32_t*)(grids + off[d]), (uint32_t*)c->svo.bricks.ptr);
    }
    VCT_HIP(hipGetLastError());
    c->svo.bricks_built = true;
    return VCT_OK;
}

vct_status vct_svo_read_bricks(vct_ctx* c, vct_stream stream, uint32_t* out, size_t cap_bricks, size_t* n_bricks)
{
    if (!c || !n_bricks) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->svo.bricks_built) return fail(VCT_ERR_NOT_READY, "vct_svo_build_bricks has not run");
    *n_bricks = c->svo.node_count;
    if (!out) return VCT_OK;
    if (cap_bricks < c->svo.node_count)
        return fail(VCT_ERR_CAPACITY, "output holds %zu bricks, the octree has %u", cap_bricks, c->svo.node_count);
    hipStream_t s = (hipStream_t)stream;
    VCT_HIP(hipMemcpyAsync(out, c->svo.bricks.ptr, 4u * 27u * (size_t)c->svo.node_count, hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    return VCT_OK;
}

// OctreeBuilder::readOctreeNodes (OctreeBuilder.cpp:359-385): host copy of the
// allocated part of the pool (uvec2 per node).
vct_status vct_svo_read_nodes(vct_ctx* c, vct_stream stream, uint32_t* out, size_t cap_nodes, size_t* n_nodes)
{
    if (!c || !n_nodes) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->svo.built) return fail(VCT_ERR_NOT_READY, "vct_svo_build has not run");
    *n_nodes = c->svo.node_count;
    if (!out) return VCT_OK;
    if (cap_nodes < c->svo.node_count) return fail(VCT_ERR_CAPACITY, "output holds %zu nodes, pool has %u", cap_nodes, c->svo.node_count);
    hipStream_t s = (hipStream_t)stream;
    VCT_HIP(hipMemcpyAsync(out, c->svo.pool.ptr, 8u * (size_t)c->svo.node_count, hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    return VCT_OK;
}

// SparseVoxelizer::readVoxelFragments (SparseVoxelizer.cpp:328-353).
vct_status vct_svo_read_fragments(vct_ctx* c, vct_stream stream, uint32_t* out, size_t cap, size_t* n)
{
    if (!c || !n) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    *n = c->svo.nfrag;
    if (!out || !c->svo.nfrag) return VCT_OK;
    if (cap < c->svo.nfrag) return fail(VCT_ERR_CAPACITY, "output holds %zu fragments, list has %u", cap, c->svo.nfrag);
    hipStream_t s = (hipStream_t)stream;
    VCT_HIP(hipMemcpyAsync(out, c->svo.frags.ptr, 8u * (size_t)c->svo.nfrag, hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    return VCT_OK;
}

// Device pointer to the node pool ("OctreeBuilder::getOctreeBuffer").
vct_status vct_svo_pool(vct_ctx* c, const uint32_t** pool, uint32_t* node_count)
{
    if (!c || !pool || !node_count) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!c->svo.built) return fail(VCT_ERR_NOT_READY, "vct_svo_build has not run");
    *pool = (const uint32_t*)c->svo.pool.ptr;
    *node_count = c->svo.node_count;
    return VCT_OK;
}

}  // extern "C"
