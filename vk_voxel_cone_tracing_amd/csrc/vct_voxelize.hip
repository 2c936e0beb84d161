// vct_voxelize.hip - clip-level triangle voxelization (K1) and radiance
// injection (K2) for gfx950.
//
// Both passes take one or several clip levels per call and expand
// (level, triangle) pairs into (level, triangle, cell, sample) work items
// with one exclusive scan over the per-pair cell counts, so every wavefront
// gets an even share of raster work regardless of triangle size and level,
// and all levels of a pass share the launches (and, for K2, the one host
// wait for the record count and the one sort). K1 writes opacity
// directly (the stores are idempotent). K2 reproduces the reference's
// order-dependent imageAtomicCompSwap running average deterministically: each
// fragment emits one record per written face texel keyed by
// (texel address in the volume, work index), a radix sort groups the records
// by texel in fragment order (a texel belongs to one level, whose work items
// keep their order), and one lane per texel applies the running average (RGBA8)
// or the FP32 sum/count mean (RGBA16F) in that order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>

#include "vct_internal.hpp"
#include "vct_raster.hpp"
#include "vct_shadow.hpp"
#include "vct_texel.hpp"
#include "vct_textures.hpp"

namespace vct {

struct VoxParams {
    float region_min[3];
    float region_max[3];
    float clip_max_extent;
    float voxel_size;
    int clip_level;
    int res;
    int D;
    int lo[3];
    int hi[3];
    int samples;
    uint32_t quirks;
};

// Voxel cell of a coordinate on one axis: floor(p / voxel_size).
__device__ __forceinline__ int voxel_cell(const VoxParams& P, float p)
{
    return (int)floorf(p / P.voxel_size);
}

// msaaVoxelizer.frag:43-55 / msaaInjectRadiance.frag:162-174. With
// VCT_QUIRK_CLAMP_X_RANGE, the reference's calculateImageCoords: clamp into the
// closed region (with the x range on every axis), then fract(p / extent).
// Without it, half-open voxel cells: texel = cell mod res with the cell clamped
// into the region's cells, so a point's texel does not depend on where the
// region sits and a sub-region voxelization writes exactly the texels a full
// one writes inside it.
__device__ __forceinline__ void image_coords(const VoxParams& P, const float pos[3], int ic[3])
{
#pragma clang fp contract(off)
    if (P.quirks & VCT_QUIRK_CLAMP_X_RANGE) {
        const float c = P.voxel_size * 0.25f;
        for (int k = 0; k < 3; ++k) {
            const float lo = P.region_min[0] + c;
            const float hi = P.region_max[0] - c;
            const float w = fminf(fmaxf(pos[k], lo), hi);
            const float q = w / P.clip_max_extent;
            const float fr = q - floorf(q);
            ic[k] = ((int)(fr * (float)P.res)) % P.res + 1;
        }
        return;
    }
    for (int k = 0; k < 3; ++k) {
        const int g = min(max(voxel_cell(P, pos[k]), P.lo[k] + 1), P.hi[k] - 1);
        ic[k] = (g & (P.res - 1)) + 1;
    }
}

// msaaVoxelizer.frag:59-60: the reference keeps fragments in the closed region;
// without VCT_QUIRK_CLAMP_X_RANGE the region is its half-open cells.
__device__ __forceinline__ bool in_region(const VoxParams& P, const float fp[3])
{
    if (P.quirks & VCT_QUIRK_CLAMP_X_RANGE)
        return !(fp[0] < P.region_min[0] || fp[1] < P.region_min[1] || fp[2] < P.region_min[2] ||
                 fp[0] > P.region_max[0] || fp[1] > P.region_max[1] || fp[2] > P.region_max[2]);
    for (int k = 0; k < 3; ++k) {
        const int g = voxel_cell(P, fp[k]);
        if (g <= P.lo[k] || g >= P.hi[k]) return false;
    }
    return true;
}

// The clip levels of one call: their raster parameters and, for the
// injection, the depth-clip box of each. Passed by value (kernel arguments);
// a work item's slot is found from the slot starts the expansion wrote.
struct LevelSet {
    VoxParams P[kMaxLevels];
    float box_lo[kMaxLevels][3], box_hi[kMaxLevels][3];
    int n;               // slots in use
    uint32_t keep_mask;  // bit l: clip level l keeps its stored alpha (RGBA16F after vct_begin_radiance_level)
};

// counts[slot * ntris + t]: the work items (cells x samples) of triangle t in
// the slot's level (grid: triangles x slots).
__global__ void k_count(const float* __restrict__ pos, uint32_t ntris, LevelSet S,
                        unsigned long long* __restrict__ counts)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = (int)blockIdx.y;
    if (t >= ntris) return;
    const VoxParams& P = S.P[slot];
    float p[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) p[k] = pos[9u * t + k];
    TriSetup s;
    unsigned long long n = 0;
    if (tri_setup(p, P.voxel_size, P.lo, P.hi, s)) n = (unsigned long long)s.nu * s.nv * (unsigned long long)P.samples;
    counts[(size_t)slot * ntris + t] = n;
}

// The work items of the slots are numbered consecutively, each slot starting
// at a multiple of 64 (kWaveAlign), so that every wave of a grid-stride loop
// (64 consecutive items from a multiple of 64) lies in one slot: the slot's
// parameters are then wave-uniform (scalar registers). totals layout:
// [0] the numbered items (padding included), [1] the injection's record
// counter (0), [2 + k] slot k's first item, [10 + k] its first item in the
// unpadded scan (offsets), [18 + k] its item count. The record counter
// (kTotRec) sits 256 bytes past the start, in a cache line of its own: every
// work item reads the slot table, and sharing a line with the counter that
// every fragment increments atomically would make each of those reads wait.
constexpr unsigned long long kWaveAlign = 64;
constexpr int kTotA = 2, kTotU = 2 + kMaxLevels, kTotC = 2 + 2 * kMaxLevels;
constexpr int kTotRec = 32, kTotWords = kTotRec + 1;
static_assert(kTotC + kMaxLevels <= kTotRec, "slot table overlaps the record counter");
static_assert(8u * kTotWords <= kPinnedWordBytes, "totals read back into the pinned landing slot");

__global__ void k_total(const unsigned long long* __restrict__ counts, const unsigned long long* __restrict__ offsets,
                        uint32_t ntris, int nslots, unsigned long long* __restrict__ totals)
{
    const size_t n = (size_t)ntris * (size_t)nslots;
    const unsigned long long all = n ? offsets[n - 1] + counts[n - 1] : 0ull;
    unsigned long long a = 0;
    for (int k = 0; k < nslots; ++k) {
        const unsigned long long u = offsets[(size_t)k * ntris];
        const unsigned long long next = k + 1 < nslots ? offsets[(size_t)(k + 1) * ntris] : all;
        totals[kTotA + k] = a;
        totals[kTotU + k] = u;
        totals[kTotC + k] = next - u;
        a += (next - u + kWaveAlign - 1) / kWaveAlign * kWaveAlign;
    }
    totals[0] = a;
    totals[kTotRec] = 0ull;
}

__global__ void k_reset_records(unsigned long long* __restrict__ total) { total[kTotRec] = 0ull; }

// Bits of the work index in an injection record key: enough for n - 1 (n
// work items). The key is (texel index in the volume) << bits | index.
__host__ __device__ inline uint32_t work_bits(unsigned long long n)
{
    uint32_t b = 1;
    while (n > 1ull && b < 32 && ((n - 1ull) >> b) != 0ull) ++b;  // n <= 2^32 (checked)
    return b;
}

// The slot of work item wi (wave-uniform, see k_total), its position in the
// unpadded numbering (the work index the injection's record keys carry) and
// its triangle (a search within the slot's triangles: with tens of thousands
// of triangles a full search per lane is cheaper than narrowing it with
// block-uniform searches first). valid is false for a slot's padding.
struct WorkItem {
    int slot;
    bool valid;
    uint32_t t;       // triangle
    uint32_t sample;  // index of the (cell, sample) within the triangle
};

__device__ __forceinline__ WorkItem work_item(const unsigned long long* __restrict__ offsets,
                                              const unsigned long long* __restrict__ totals, uint32_t ntris, int nslots,
                                              unsigned long long wi)
{
    int slot = 0;
    for (int k = 1; k < nslots; ++k) slot += wi >= totals[kTotA + k] ? 1 : 0;
    slot = __builtin_amdgcn_readfirstlane(slot);  // one slot per wave
    WorkItem it;
    it.slot = slot;
    const unsigned long long local = wi - totals[kTotA + slot];
    it.valid = local < totals[kTotC + slot];
    it.t = 0;
    it.sample = 0;
    if (it.valid) {
        const unsigned long long index = totals[kTotU + slot] + local;
        const uint32_t base = (uint32_t)slot * ntris;
        const uint32_t j = work_triangle(offsets, base, base + ntris - 1u, index);
        it.t = j - base;
        it.sample = (uint32_t)(index - offsets[j]);
    }
    return it;
}

// The unpadded work index of item wi of `slot`.
__device__ __forceinline__ unsigned long long work_index(const unsigned long long* __restrict__ totals, int slot,
                                                         unsigned long long wi)
{
    return wi - totals[kTotA + slot] + totals[kTotU + slot];
}

__global__ __launch_bounds__(256) void k_raster_opacity(const float* __restrict__ pos,
                                                        const unsigned long long* __restrict__ offsets, uint32_t ntris,
                                                        const unsigned long long* __restrict__ totals, LevelSet S,
                                                        TexInputs T, uint32_t* __restrict__ vol)
{
    const unsigned long long n = totals[0];
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long wi = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; wi < n; wi += stride) {
        const WorkItem it = work_item(offsets, totals, ntris, S.n, wi);
        if (!it.valid) continue;
        const VoxParams P = S.P[it.slot];
        const uint32_t t = it.t;
        float p[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) p[k] = pos[9u * t + k];
        TriSetup s;
        if (!tri_setup(p, P.voxel_size, P.lo, P.hi, s)) continue;
        float fp[3];
        double lam[3];
        if (!tri_sample(p, s, P.voxel_size, P.samples, it.sample, fp, lam)) continue;
        if (!in_region(P, fp)) continue;
        if (T.tex) {
            // msaaVoxelizer.frag:64: occlusion-texture discard
            float uv[2];
            interp_uv(T.uv + 6u * t, lam, uv);
            if (occluded(T, T.mats[T.mat[t]], uv)) continue;
        }
        int ic[3];
        image_coords(P, fp, ic);
        // imageStore(vec4(1.0)) into the six faces (msaaVoxelizer.frag:69-73)
        for (int f = 0; f < kFaces; ++f) vol[vol_index(P.D, P.clip_level, f, ic[0], ic[1], ic[2])] = 0xffffffffu;
    }
}

struct Record {
    float r, g, b;
};

__global__ __launch_bounds__(256) void k_raster_inject(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                       TexInputs T, const unsigned long long* __restrict__ offsets,
                                                       uint32_t ntris, unsigned long long* __restrict__ totals,
                                                       LevelSet S, LightParams L,
                                                       unsigned long long* __restrict__ keys,
                                                       uint32_t* __restrict__ idx, Record* __restrict__ vals,
                                                       uint32_t cap)
{
#pragma clang fp contract(off)
    const unsigned long long n = totals[0];
    if (n > 0xffffffffull) return;  // record keys hold 32 bits of work index; the host reports it
    const uint32_t wbits = work_bits(n);
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long wi = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; wi < n; wi += stride) {
        const WorkItem it = work_item(offsets, totals, ntris, S.n, wi);
        if (!it.valid) continue;
        const VoxParams P = S.P[it.slot];
        const float box_lo[3] = {S.box_lo[it.slot][0], S.box_lo[it.slot][1], S.box_lo[it.slot][2]};
        const float box_hi[3] = {S.box_hi[it.slot][0], S.box_hi[it.slot][1], S.box_hi[it.slot][2]};
        const uint32_t t = it.t;
        float p[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) p[k] = pos[9u * t + k];
        TriSetup s;
        if (!tri_setup(p, P.voxel_size, P.lo, P.hi, s)) continue;
        float fp[3];
        double lam[3];
        if (!tri_sample(p, s, P.voxel_size, P.samples, it.sample, fp, lam)) continue;
        if (fp[0] < box_lo[0] || fp[1] < box_lo[1] || fp[2] < box_lo[2] || fp[0] > box_hi[0] || fp[1] > box_hi[1] ||
            fp[2] > box_hi[2])
            continue;
        const vct_gltf_material& m = T.mats[T.mat[t]];
        float uv[2] = {0.0f, 0.0f};
        if (T.tex) {
            // msaaInjectRadiance.frag:73-74: occlusion-texture discard
            interp_uv(T.uv + 6u * t, lam, uv);
            if (occluded(T, m, uv)) continue;
        }
        int ic[3];
        image_coords(P, fp, ic);
        Record out[6];
        int faces[6];
        int nout = 0;
        if (m.emissive_factor[0] > 0.0f || m.emissive_factor[1] > 0.0f || m.emissive_factor[2] > 0.0f) {
            // msaaInjectRadiance.frag:76-86; the emissive texture adds to the factor
            float em[3] = {m.emissive_factor[0], m.emissive_factor[1], m.emissive_factor[2]};
            if (has_texture(T, m.emissive_texture)) {
                const f4 te = sample_texture(T.tex[m.emissive_texture], uv);
                em[0] = em[0] + te.x;
                em[1] = em[1] + te.y;
                em[2] = em[2] + te.z;
            }
            Record e = {fminf(fmaxf(em[0], 0.0f), 1.0f), fminf(fmaxf(em[1], 0.0f), 1.0f),
                        fminf(fmaxf(em[2], 0.0f), 1.0f)};
            for (int f = 0; f < 6; ++f) {
                out[f] = e;
                faces[f] = f;
            }
            nout = 6;
        } else {
            // msaaInjectRadiance.frag:131-153
            const float* n0 = nrm + 9u * t;
            float ni[3];
            for (int k = 0; k < 3; ++k)
                ni[k] = (float)(lam[0] * (double)n0[k] + lam[1] * (double)n0[3 + k] + lam[2] * (double)n0[6 + k]);
            const float len = sqrtf(ni[0] * ni[0] + ni[1] * ni[1] + ni[2] * ni[2]);
            const float nn[3] = {ni[0] / len, ni[1] / len, ni[2] / len};
            float ndl = nn[0] * L.dir[0] + nn[1] * L.dir[1] + nn[2] * L.dir[2];
            ndl = fminf(fmaxf(ndl, 0.001f), 1.0f);
            const float vis = visibility(L, fp, P.quirks);
            float lc[3];
            for (int k = 0; k < 3; ++k) lc[k] = ndl * vis * L.color[k] * L.intensity;
            if (lc[0] == 0.0f && lc[1] == 0.0f && lc[2] == 0.0f) continue;
            // msaaInjectRadiance.frag:131-136: the base-colour texture scales the factor
            float col[4] = {m.base_color_factor[0], m.base_color_factor[1], m.base_color_factor[2],
                            m.base_color_factor[3]};
            if (has_texture(T, m.base_color_texture)) {
                const f4 tb = sample_texture(T.tex[m.base_color_texture], uv);
                col[0] = col[0] * tb.x;
                col[1] = col[1] * tb.y;
                col[2] = col[2] * tb.z;
                col[3] = col[3] * tb.w;
            }
            float rad[3];
            for (int k = 0; k < 3; ++k) rad[k] = fminf(fmaxf(lc[k] * col[k] * col[3], 0.0f), 1.0f);
            const int face[3] = {(-nn[0] > 0.0f) ? 0 : 1, (-nn[1] > 0.0f) ? 2 : 3, (-nn[2] > 0.0f) ? 4 : 5};
            for (int k = 0; k < 3; ++k) {
                const float w = fabsf(nn[k]);
                out[k] = {rad[0] * w, rad[1] * w, rad[2] * w};
                faces[k] = face[k];
            }
            nout = 3;
        }
        const unsigned long long first = atomicAdd(&totals[kTotRec], (unsigned long long)nout);
        const unsigned long long index = work_index(totals, it.slot, wi);
        for (int j = 0; j < nout; ++j) {
            const unsigned long long r = first + (unsigned long long)j;
            if (r >= cap) break;  // capacity is checked on the host from the counter
            const size_t addr = vol_index(P.D, P.clip_level, faces[j], ic[0], ic[1], ic[2]);
            keys[r] = ((unsigned long long)addr << wbits) | index;
            idx[r] = (uint32_t)r;
            vals[r] = out[j];
        }
    }
}

// One lane per run head of the sorted records: the texel's records in
// fragment order, applied as the reference's running average (RGBA8) or the
// FP32 sum / count mean (RGBA16F).
template <bool kHalf>
__global__ __launch_bounds__(256) void k_apply(const unsigned long long* __restrict__ keys,
                                               const uint32_t* __restrict__ idx, const Record* __restrict__ vals,
                                               uint32_t n, void* __restrict__ vol, uint32_t keep_mask, uint32_t wbits,
                                               size_t level_texels)
{
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long addr = keys[i] >> wbits;
    if (i > 0 && (keys[i - 1] >> wbits) == addr) return;  // not the head of its texel's run
    if (!kHalf) {
        uint32_t t = ((uint32_t*)vol)[addr];
        for (uint32_t j = i; j < n && (keys[j] >> wbits) == addr; ++j) {
            const Record v = vals[idx[j]];
            t = rgba8_avg(t, f4{v.r * 255.0f, v.g * 255.0f, v.b * 255.0f, 1.0f});
        }
        ((uint32_t*)vol)[addr] = t;
    } else {
        float sr = 0.0f, sg = 0.0f, sb = 0.0f, cnt = 0.0f;
        for (uint32_t j = i; j < n && (keys[j] >> wbits) == addr; ++j) {
            const Record v = vals[idx[j]];
            sr = sr + v.r;
            sg = sg + v.g;
            sb = sb + v.b;
            cnt = cnt + 1.0f;
        }
        const bool keep = (keep_mask >> (uint32_t)(addr / level_texels)) & 1u;
        const float a = keep ? unpack_half(((const uint2*)vol)[addr]).w : 1.0f;
        ((uint2*)vol)[addr] = pack_half(f4{sr / cnt, sg / cnt, sb / cnt, a});
    }
}

static vct_status make_params(vct_ctx* c, const vct_voxelization_desc* d, uint32_t level, VoxParams* P)
{
    if (!d) return fail(VCT_ERR_INVALID_ARGUMENT, "null desc");
    if (level >= (uint32_t)c->levels || d->clip_level != level)
        return fail(VCT_ERR_INVALID_ARGUMENT, "level mismatch (desc %u, call %u)", d->clip_level, level);
    if (d->clipmap_resolution != c->res) return fail(VCT_ERR_INVALID_ARGUMENT, "clipmap_resolution");
    if (!(d->voxel_size > 0.0f) || !(d->clip_max_extent > 0.0f))
        return fail(VCT_ERR_INVALID_ARGUMENT, "voxel_size / clip_max_extent must be positive");
    for (int k = 0; k < 3; ++k) {
        P->region_min[k] = d->region_min_corner[k];
        P->region_max[k] = d->region_max_corner[k];
        // integer region recovered from the corners written by Voxelizer::cmdVoxelize
        const float lo = std::rint((d->region_min_corner[k] + 1e-6f) / d->voxel_size);
        const float hi = std::rint((d->region_max_corner[k] - 1e-6f) / d->voxel_size);
        if (!(std::fabs(lo) < 1e8f) || !(std::fabs(hi) < 1e8f) || hi < lo)
            return fail(VCT_ERR_INVALID_ARGUMENT, "region corners");
        P->lo[k] = (int)lo - 1;
        P->hi[k] = (int)hi;
    }
    P->clip_max_extent = d->clip_max_extent;
    P->voxel_size = d->voxel_size;
    P->clip_level = (int)level;
    P->res = c->res;
    P->D = c->D;
    P->samples = (int)c->cfg.raster_samples;
    P->quirks = c->cfg.quirks;
    return VCT_OK;
}

// The level set of `count` descriptors (their clip levels distinct), with
// each level's depth-clip box for the injection.
static vct_status make_level_set(vct_ctx* c, const vct_voxelization_desc* descs, uint32_t count, LevelSet* S)
{
    if (!descs || count == 0 || count > (uint32_t)kMaxLevels)
        return fail(VCT_ERR_INVALID_ARGUMENT, "1 to %d voxelization descriptors are required", kMaxLevels);
    std::memset(S, 0, sizeof(*S));
    uint32_t seen = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t level = descs[i].clip_level;
        if (level < 32u && (seen >> level) & 1u)
            return fail(VCT_ERR_INVALID_ARGUMENT, "clip level %u appears twice", level);
        VCT_TRY(make_params(c, &descs[i], level, &S->P[i]));
        seen |= 1u << level;
        for (int k = 0; k < 3; ++k) {
            S->box_lo[i][k] = (float)S->P[i].lo[k] * S->P[i].voxel_size;
            S->box_hi[i][k] = (float)(S->P[i].hi[k] + 1) * S->P[i].voxel_size;
        }
    }
    S->n = (int)count;
    return VCT_OK;
}

// Per-(level, triangle) counts, their exclusive scan and the totals.
static vct_status expand(vct_ctx* c, hipStream_t s, const TriSoup& soup, const LevelSet& S)
{
    const size_t n = (size_t)soup.count * (size_t)S.n;
    if (n > 0x7fffffffull) return fail(VCT_ERR_CAPACITY, "%zu triangle-level pairs exceed 2^31", n);
    VCT_TRY(ensure(c->counts, 8u * (n ? n : 1)));
    VCT_TRY(ensure(c->offsets, 8u * (n ? n : 1)));
    VCT_TRY(ensure(c->totals, 8u * kTotWords));
    if (n) {
        hipLaunchKernelGGL(k_count, dim3((soup.count + 255) / 256, (unsigned)S.n), dim3(256), 0, s, soup.pos,
                           soup.count, S, (unsigned long long*)c->counts.ptr);
        VCT_HIP(hipGetLastError());
        size_t tmp = 0;
        VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, (unsigned long long*)c->counts.ptr,
                                                 (unsigned long long*)c->offsets.ptr, (int)n, s));
        VCT_TRY(ensure(c->scan_tmp, tmp));
        VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, (unsigned long long*)c->counts.ptr,
                                                 (unsigned long long*)c->offsets.ptr, (int)n, s));
    }
    hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, s, (const unsigned long long*)c->counts.ptr,
                       (const unsigned long long*)c->offsets.ptr, soup.count, S.n,
                       (unsigned long long*)c->totals.ptr);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

static vct_status voxelize_levels(vct_ctx* c, hipStream_t s, const vct_mesh* mesh, const vct_voxelization_desc* descs,
                                  uint32_t count)
{
    if (!c) return fail(VCT_ERR_INVALID_ARGUMENT, "null context");
    LevelSet S;
    VCT_TRY(make_level_set(c, descs, count, &S));
    TriSoup soup;
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup));
    if (!soup.count) return VCT_OK;
    VCT_TRY(expand(c, s, soup, S));
    // grid-stride over a device-side total: no host synchronisation
    const TexInputs T{soup.uv, soup.mat, mesh->materials, soup.tex, soup.tex_count};
    hipLaunchKernelGGL(k_raster_opacity, dim3(8192), dim3(256), 0, s, soup.pos,
                       (const unsigned long long*)c->offsets.ptr, soup.count,
                       (const unsigned long long*)c->totals.ptr, S, T, (uint32_t*)c->opacity.ptr);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

static vct_status inject_levels(vct_ctx* c, hipStream_t s, const vct_mesh* mesh, const vct_shadow_map* shadow,
                                const vct_directional_light_desc* light,
                                const vct_directional_light_shadow_desc* light_shadow,
                                const vct_voxelization_desc* descs, uint32_t count)
{
    if (!c) return fail(VCT_ERR_INVALID_ARGUMENT, "null context");
    if (!shadow || !shadow->depth || !shadow->width || !shadow->height || !light || !light_shadow)
        return fail(VCT_ERR_INVALID_ARGUMENT, "light, light_shadow and shadow map are required");
    LevelSet S;
    VCT_TRY(make_level_set(c, descs, count, &S));
    radiance_written(c);
    // after vct_begin_radiance_level the texels already carry the opacity alpha
    for (int i = 0; i < S.n; ++i) {
        const int level = S.P[i].clip_level;
        if (c->keep_alpha[level]) S.keep_mask |= 1u << level;
        c->keep_alpha[level] = false;
    }
    TriSoup soup;
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup));
    if (!soup.count) return VCT_OK;
    VCT_TRY(expand(c, s, soup, S));

    LightParams L;
    {
        // normalize(-uDirectionalLight.direction) with IEEE sqrt/div
        float nd[3] = {-light->direction[0], -light->direction[1], -light->direction[2]};
        float len = std::sqrt(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
        for (int k = 0; k < 3; ++k) L.dir[k] = nd[k] / len;
        for (int k = 0; k < 3; ++k) L.color[k] = light->color[k];
        L.intensity = light->intensity;
        for (int k = 0; k < 16; ++k) {
            L.view[k] = light_shadow->view[k];
            L.proj[k] = light_shadow->proj[k];
        }
        L.z_near = light_shadow->z_near;
        L.z_far = light_shadow->z_far;
        L.depth = shadow->depth;
        L.width = (int)shadow->width;
        L.height = (int)shadow->height;
    }
    // The raster writes into a grow-only record buffer and counts every record
    // (those past the capacity are counted, not stored). One host read-back
    // of the count sizes the sort; a count above the capacity grows the
    // buffer and runs the raster again, which happens while the capacity
    // warms up. Record keys carry the work index, so the sorted order does
    // not depend on which run stored them.
    if (!c->pinned_word) VCT_HIP(hipHostMalloc(&c->pinned_word, kPinnedWordBytes, hipHostMallocDefault));
    volatile unsigned long long* landed = (volatile unsigned long long*)c->pinned_word;
    unsigned long long tot[2] = {0, 0};
    const TexInputs T{soup.uv, soup.mat, mesh->materials, soup.tex, soup.tex_count};
    for (int attempt = 0;; ++attempt) {
        const uint32_t cap = c->inject_cap;
        hipLaunchKernelGGL(k_raster_inject, dim3(8192), dim3(256), 0, s, soup.pos, soup.nrm, T,
                           (const unsigned long long*)c->offsets.ptr, soup.count, (unsigned long long*)c->totals.ptr,
                           S, L, (unsigned long long*)c->rec_key[0].ptr, (uint32_t*)c->rec_idx[0].ptr,
                           (Record*)c->rec_val.ptr, cap);
        VCT_HIP(hipGetLastError());
        VCT_HIP(hipMemcpyAsync((void*)landed, c->totals.ptr, 8u * kTotWords, hipMemcpyDeviceToHost, s));
        VCT_HIP(hipStreamSynchronize(s));
        tot[0] = landed[0];
        tot[1] = landed[kTotRec];
        if (tot[0] > 0xffffffffull) return fail(VCT_ERR_CAPACITY, "raster work %llu exceeds 2^32", tot[0]);
        if (tot[1] <= cap) break;
        if (tot[1] > 0xfffffff0ull || attempt > 0)
            return fail(VCT_ERR_CAPACITY, "injection records overflow (%llu > %u)", tot[1], cap);
        const unsigned long long want = tot[1] + tot[1] / 4 + 1024;
        const uint32_t grown = want > 0xfffffff0ull ? 0xfffffff0u : (uint32_t)want;
        for (int b = 0; b < 2; ++b) {
            VCT_TRY(ensure(c->rec_key[b], 8u * (size_t)grown));
            VCT_TRY(ensure(c->rec_idx[b], 4u * (size_t)grown));
        }
        VCT_TRY(ensure(c->rec_val, sizeof(Record) * (size_t)grown));
        c->inject_cap = grown;
        hipLaunchKernelGGL(k_reset_records, dim3(1), dim3(1), 0, s, (unsigned long long*)c->totals.ptr);
        VCT_HIP(hipGetLastError());
    }
    const uint32_t nrec = (uint32_t)tot[1];
    if (!nrec) return VCT_OK;
    // key bits: the texel index in the volume above the work index
    const uint32_t wbits = work_bits(tot[0]);
    const size_t level_texels = (size_t)kFaces * (size_t)c->D * (size_t)c->D * (size_t)c->D;
    const size_t volume_texels = level_texels * (size_t)c->levels;
    int end_bit = (int)wbits;
    while (end_bit < 64 && ((volume_texels - 1) >> (end_bit - (int)wbits)) != 0) ++end_bit;
    size_t tmp = 0;
    VCT_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (unsigned long long*)c->rec_key[0].ptr,
                                               (unsigned long long*)c->rec_key[1].ptr, (uint32_t*)c->rec_idx[0].ptr,
                                               (uint32_t*)c->rec_idx[1].ptr, (int)nrec, 0, end_bit, s));
    VCT_TRY(ensure(c->sort_tmp, tmp));
    VCT_HIP(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp.ptr, tmp, (unsigned long long*)c->rec_key[0].ptr,
                                               (unsigned long long*)c->rec_key[1].ptr, (uint32_t*)c->rec_idx[0].ptr,
                                               (uint32_t*)c->rec_idx[1].ptr, (int)nrec, 0, end_bit, s));
    const bool half = c->cfg.radiance_format == VCT_FORMAT_RGBA16F;
    if (half)
        hipLaunchKernelGGL(k_apply<true>, dim3((nrec + 255) / 256), dim3(256), 0, s,
                           (const unsigned long long*)c->rec_key[1].ptr, (const uint32_t*)c->rec_idx[1].ptr,
                           (const Record*)c->rec_val.ptr, nrec, c->radiance.ptr, S.keep_mask, wbits, level_texels);
    else
        hipLaunchKernelGGL(k_apply<false>, dim3((nrec + 255) / 256), dim3(256), 0, s,
                           (const unsigned long long*)c->rec_key[1].ptr, (const uint32_t*)c->rec_idx[1].ptr,
                           (const Record*)c->rec_val.ptr, nrec, c->radiance.ptr, 0u, wbits, level_texels);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // namespace vct

using namespace vct;

extern "C" {

vct_status vct_voxelize_opacity(vct_ctx* c, vct_stream stream, const vct_mesh* mesh, const vct_voxelization_desc* d,
                                uint32_t level)
{
    if (d && d->clip_level != level)
        return fail(VCT_ERR_INVALID_ARGUMENT, "level mismatch (desc %u, call %u)", d->clip_level, level);
    return voxelize_levels(c, (hipStream_t)stream, mesh, d, 1);
}

vct_status vct_voxelize_opacity_levels(vct_ctx* c, vct_stream stream, const vct_mesh* mesh,
                                       const vct_voxelization_desc* descs, uint32_t count)
{
    return voxelize_levels(c, (hipStream_t)stream, mesh, descs, count);
}

vct_status vct_inject_radiance(vct_ctx* c, vct_stream stream, const vct_mesh* mesh, const vct_shadow_map* shadow,
                               const vct_directional_light_desc* light,
                               const vct_directional_light_shadow_desc* light_shadow,
                               const vct_voxelization_desc* d, uint32_t level)
{
    if (d && d->clip_level != level)
        return fail(VCT_ERR_INVALID_ARGUMENT, "level mismatch (desc %u, call %u)", d->clip_level, level);
    return inject_levels(c, (hipStream_t)stream, mesh, shadow, light, light_shadow, d, 1);
}

vct_status vct_inject_radiance_levels(vct_ctx* c, vct_stream stream, const vct_mesh* mesh,
                                      const vct_shadow_map* shadow, const vct_directional_light_desc* light,
                                      const vct_directional_light_shadow_desc* light_shadow,
                                      const vct_voxelization_desc* descs, uint32_t count)
{
    return inject_levels(c, (hipStream_t)stream, mesh, shadow, light, light_shadow, descs, count);
}

}  // extern "C"
