// vct_producers.hip - G-buffer and shadow-map producers for gfx950.
//
// The reference rasterizes these in GBufferPass (GBufferPass.cpp:170-206) and
// ReflectiveShadowMapPass (ReflectiveShadowMapPass.cpp:38-88); they feed the
// hot path (cone trace and radiance injection) but are not part of it. Here
// they are ray cast: one lane per pixel and one 256-lane workgroup per 16x16
// pixel tile, which tests only the triangles binned to its tile (their
// projected boxes), streamed through LDS 256 at a time.
// Outputs use the reference's G-buffer channel layout.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "vct_internal.hpp"
#include "vct_texel.hpp"

namespace vct {

struct Mat16 {
    float m[16];
};

struct RayTri {
    float p0[3], e1[3], e2[3];
};

constexpr int kTile = 256;

// Moller-Trumbore; returns t (or a huge value) and the barycentrics of v1, v2.
__device__ __forceinline__ float intersect(const float o[3], const float d[3], const RayTri& tr, float& bu, float& bv)
{
    const float px = d[1] * tr.e2[2] - d[2] * tr.e2[1];
    const float py = d[2] * tr.e2[0] - d[0] * tr.e2[2];
    const float pz = d[0] * tr.e2[1] - d[1] * tr.e2[0];
    const float det = tr.e1[0] * px + tr.e1[1] * py + tr.e1[2] * pz;
    if (fabsf(det) < 1e-30f) return 3.0e38f;
    const float inv = 1.0f / det;
    const float tx = o[0] - tr.p0[0], ty = o[1] - tr.p0[1], tz = o[2] - tr.p0[2];
    const float u = (tx * px + ty * py + tz * pz) * inv;
    const float qx = ty * tr.e1[2] - tz * tr.e1[1];
    const float qy = tz * tr.e1[0] - tx * tr.e1[2];
    const float qz = tx * tr.e1[1] - ty * tr.e1[0];
    const float v = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
    const float t = (tr.e2[0] * qx + tr.e2[1] * qy + tr.e2[2] * qz) * inv;
    if (u < 0.0f || v < 0.0f || u + v > 1.0f) return 3.0e38f;
    bu = u;
    bv = v;
    return t;
}

__device__ __forceinline__ void mat_vec(const float* m, const float v[4], float out[4])
{
    for (int k = 0; k < 4; ++k) out[k] = m[k] * v[0] + m[4 + k] * v[1] + m[8 + k] * v[2] + m[12 + k] * v[3];
}

// ---- screen-tile binning: a ray is only tested against the triangles whose
// projected box covers its 16x16 pixel tile. The nearest hit is the least
// (t, triangle index) pair, so it does not depend on the order in which a
// tile's list was filled, and equals the hit of a test against every
// triangle.
constexpr int kBinSide = 16;  // pixels per tile side; a tile is one 256-lane workgroup

struct BinGrid {
    Mat16 vp;  // world -> clip, the matrix the rays are cast with the inverse of
    int w, h, tiles_x, tiles_y;
};

// The tiles [tx0, tx1] x [ty0, ty1] that triangle p may cover (false: none).
// Rays start on the near plane (clip z = 0) and run forward, so only the
// part of the triangle with clip z >= 0 can be hit: the triangle is clipped
// to it (at most 4 vertices, all with w > 0 there) and the clipped polygon's
// projected box taken. The box is widened by a pixel each way so that
// rounding between the projection and the rays' inverse cannot drop a pixel
// whose ray hits the triangle.
__device__ __forceinline__ bool tile_box(const float* __restrict__ p, const BinGrid& G, int& tx0, int& ty0, int& tx1,
                                         int& ty1)
{
    float c[3][4];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float v[4] = {p[3 * k], p[3 * k + 1], p[3 * k + 2], 1.0f};
        mat_vec(G.vp.m, v, c[k]);
    }
    float xmin = 3.0e38f, ymin = 3.0e38f, xmax = -3.0e38f, ymax = -3.0e38f;
    int n = 0;
    bool any_w = true;
    auto add = [&](const float q[4]) {
        ++n;
        if (!(q[3] > 0.0f)) {
            any_w = false;
            return;
        }
        const float x = q[0] / q[3], y = q[1] / q[3];
        xmin = fminf(xmin, x);
        xmax = fmaxf(xmax, x);
        ymin = fminf(ymin, y);
        ymax = fmaxf(ymax, y);
    };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float* a = c[k];
        const float* b = c[(k + 1) % 3];
        const bool ina = a[2] >= 0.0f, inb = b[2] >= 0.0f;
        if (ina) add(a);
        if (ina != inb) {
            const float sc = a[2] / (a[2] - b[2]);
            const float q[4] = {a[0] + sc * (b[0] - a[0]), a[1] + sc * (b[1] - a[1]), 0.0f, a[3] + sc * (b[3] - a[3])};
            add(q);
        }
    }
    if (n == 0) return false;  // wholly behind the near plane
    if (!any_w) {
        // a clipped vertex without a projection (a degenerate projection):
        // every tile, conservatively
        tx0 = 0;
        ty0 = 0;
        tx1 = G.tiles_x - 1;
        ty1 = G.tiles_y - 1;
        return true;
    }
    // pixel px's ray is at NDC x = (px + 0.5) / w * 2 - 1
    const float px0 = floorf((xmin + 1.0f) * 0.5f * (float)G.w - 0.5f) - 1.0f;
    const float px1 = ceilf((xmax + 1.0f) * 0.5f * (float)G.w - 0.5f) + 1.0f;
    const float py0 = floorf((ymin + 1.0f) * 0.5f * (float)G.h - 0.5f) - 1.0f;
    const float py1 = ceilf((ymax + 1.0f) * 0.5f * (float)G.h - 0.5f) + 1.0f;
    if (!(px1 >= 0.0f) || !(py1 >= 0.0f) || !(px0 <= (float)(G.w - 1)) || !(py0 <= (float)(G.h - 1))) return false;
    tx0 = (int)fmaxf(px0, 0.0f) / kBinSide;
    ty0 = (int)fmaxf(py0, 0.0f) / kBinSide;
    tx1 = (int)fminf(px1, (float)(G.w - 1)) / kBinSide;
    ty1 = (int)fminf(py1, (float)(G.h - 1)) / kBinSide;
    return true;
}

// Per triangle: its tile box (int4: tx0, ty0, tx1, ty1) and tile count (0:
// culled). The (triangle, tile) pairs are then expanded one per lane, so a
// wall covering thousands of tiles does not hold one lane for all of them.
__global__ __launch_bounds__(256) void k_bin_boxes(const float* __restrict__ soup, uint32_t ntris, BinGrid G,
                                                   int4* __restrict__ boxes, uint32_t* __restrict__ ntiles)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntris) return;
    int tx0, ty0, tx1, ty1;
    if (!tile_box(soup + 9u * t, G, tx0, ty0, tx1, ty1)) {
        ntiles[t] = 0u;
        return;
    }
    boxes[t] = make_int4(tx0, ty0, tx1, ty1);
    ntiles[t] = (uint32_t)(tx1 - tx0 + 1) * (uint32_t)(ty1 - ty0 + 1);
}

// Pair p: its triangle (the last with toff[t] <= p) and tile.
__device__ __forceinline__ void pair_at(const uint32_t* __restrict__ toff, uint32_t ntris,
                                        const int4* __restrict__ boxes, int tiles_x, uint32_t p, uint32_t& t,
                                        uint32_t& tile)
{
    uint32_t lo = 0, hi = ntris;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (toff[mid] <= p) lo = mid; else hi = mid;
    }
    t = lo;
    const int4 b = boxes[t];
    const uint32_t i = p - toff[t], bw = (uint32_t)(b.z - b.x + 1);
    tile = (uint32_t)(b.y + (int)(i / bw)) * (uint32_t)tiles_x + (uint32_t)(b.x + (int)(i % bw));
}

__global__ __launch_bounds__(256) void k_bin_count(const uint32_t* __restrict__ toff, uint32_t ntris,
                                                   const int4* __restrict__ boxes, int tiles_x, uint32_t npairs,
                                                   uint32_t* __restrict__ counts)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    uint32_t t, tile;
    pair_at(toff, ntris, boxes, tiles_x, p, t, tile);
    atomicAdd(&counts[tile], 1u);
}

__global__ __launch_bounds__(256) void k_bin_fill(const uint32_t* __restrict__ toff, uint32_t ntris,
                                                  const int4* __restrict__ boxes, int tiles_x, uint32_t npairs,
                                                  const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursor,
                                                  uint32_t* __restrict__ lists)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    uint32_t t, tile;
    pair_at(toff, ntris, boxes, tiles_x, p, t, tile);
    lists[offsets[tile] + atomicAdd(&cursor[tile], 1u)] = t;
}

// Casts the ray through pixel (px, py) of a viewProjInv against the n
// triangles of this workgroup's tile list; returns the nearest hit (least t,
// then least triangle index).
__device__ void cast(const float* __restrict__ soup, const uint32_t* __restrict__ list, uint32_t n, const float* inv,
                     int px, int py, int w, int h, bool valid, float o[3], float d[3], float& best_t, int& best_tri,
                     float& bu, float& bv)
{
    __shared__ RayTri tile[kTile];
    __shared__ uint32_t ids[kTile];
    if (valid) {
        const float nx = (((float)px + 0.5f) / (float)w) * 2.0f - 1.0f;
        const float ny = (((float)py + 0.5f) / (float)h) * 2.0f - 1.0f;
        const float a[4] = {nx, ny, 0.0f, 1.0f}, b[4] = {nx, ny, 1.0f, 1.0f};
        float pa[4], pb[4];
        mat_vec(inv, a, pa);
        mat_vec(inv, b, pb);
        for (int k = 0; k < 3; ++k) {
            o[k] = pa[k] / pa[3];
            d[k] = pb[k] / pb[3] - o[k];
        }
    }
    best_t = 3.0e38f;
    best_tri = -1;
    for (uint32_t base = 0; base < n; base += kTile) {
        const uint32_t j = base + threadIdx.x;
        if (j < n) {
            const uint32_t t = list[j];
            const float* p = soup + 9u * t;
            RayTri r;
            for (int k = 0; k < 3; ++k) {
                r.p0[k] = p[k];
                r.e1[k] = p[3 + k] - p[k];
                r.e2[k] = p[6 + k] - p[k];
            }
            tile[threadIdx.x] = r;
            ids[threadIdx.x] = t;
        }
        __syncthreads();
        const uint32_t m = n - base < (uint32_t)kTile ? n - base : (uint32_t)kTile;
        if (valid) {
            for (uint32_t i = 0; i < m; ++i) {
                float u, v;
                const float th = intersect(o, d, tile[i], u, v);
                const int id = (int)ids[i];
                if (th > 0.0f && th <= 1.0f && (th < best_t || (th == best_t && id < best_tri))) {
                    best_t = th;
                    best_tri = id;
                    bu = u;
                    bv = v;
                }
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kTile) void k_gbuffer(const float* __restrict__ soup, const float* __restrict__ nrm,
                                                   const uint32_t* __restrict__ mat,
                                                   const vct_gltf_material* __restrict__ mats,
                                                   const float* __restrict__ surface,
                                                   const uint32_t* __restrict__ offsets,
                                                   const uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ lists, int w, int h, Mat16 inv_m,
                                                   Mat16 vp_m, uint32_t* __restrict__ g_diffuse,
                                                   uint2* __restrict__ g_normal, uint32_t* __restrict__ g_specular,
                                                   uint2* __restrict__ g_emission, float* __restrict__ g_depth)
{
    const float* inv = inv_m.m;
    const float* vp = vp_m.m;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool valid = px < w && py < h;
    const uint32_t tile = blockIdx.y * gridDim.x + blockIdx.x;
    float o[3] = {0, 0, 0}, d[3] = {0, 0, 1}, t, bu = 0.0f, bv = 0.0f;
    int tri;
    cast(soup, lists + offsets[tile], counts[tile], inv, px, py, w, h, valid, o, d, t, tri, bu, bv);
    if (!valid) return;
    const size_t idx = (size_t)py * (size_t)w + (size_t)px;
    if (tri < 0) {
        g_depth[idx] = 1.0f;
        g_diffuse[idx] = 0u;
        g_specular[idx] = 0u;
        g_normal[idx] = make_uint2(0u, 0u);
        g_emission[idx] = make_uint2(0u, 0u);
        return;
    }
    const float hp[4] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t, 1.0f};
    float c[4];
    mat_vec(vp, hp, c);
    g_depth[idx] = c[2] / c[3];
    const float* n = nrm + 9u * (uint32_t)tri;
    const float w0 = 1.0f - bu - bv;
    float nn[3];
    for (int k = 0; k < 3; ++k) nn[k] = w0 * n[k] + bu * n[3 + k] + bv * n[6 + k];
    const float len = sqrtf(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    g_normal[idx] = pack_half(f4{nn[0] / len * 0.5f + 0.5f, nn[1] / len * 0.5f + 0.5f, nn[2] / len * 0.5f + 0.5f, 1.0f});
    const uint32_t m = mat[tri];
    const vct_gltf_material& mt = mats[m];
    const float* s = surface + 5u * m;  // perceptual roughness, metallic, specular rgb
    g_diffuse[idx] = f_to_u8(mt.base_color_factor[0]) | (f_to_u8(mt.base_color_factor[1]) << 8) |
                     (f_to_u8(mt.base_color_factor[2]) << 16) | (f_to_u8(s[0]) << 24);
    g_specular[idx] = f_to_u8(s[2]) | (f_to_u8(s[3]) << 8) | (f_to_u8(s[4]) << 16) | (f_to_u8(s[1]) << 24);
    g_emission[idx] = pack_half(f4{mt.emissive_factor[0], mt.emissive_factor[1], mt.emissive_factor[2], 1.0f});
}

__global__ __launch_bounds__(kTile) void k_shadow(const float* __restrict__ soup, const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ counts,
                                                  const uint32_t* __restrict__ lists, int w, int h, Mat16 inv_m,
                                                  float* __restrict__ depth)
{
    const float* inv = inv_m.m;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool valid = px < w && py < h;
    const uint32_t tile = blockIdx.y * gridDim.x + blockIdx.x;
    float o[3] = {0, 0, 0}, d[3] = {0, 0, 1}, t, bu, bv;
    int tri;
    cast(soup, lists + offsets[tile], counts[tile], inv, px, py, w, h, valid, o, d, t, tri, bu, bv);
    if (!valid) return;
    // orthographic light: NDC depth is linear along the near-to-far ray
    depth[(size_t)py * (size_t)w + (size_t)px] = tri < 0 ? 1.0f : t;
}

__global__ void k_scan_total32(const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ counts, uint32_t n,
                               unsigned long long* __restrict__ total)
{
    total[0] = (unsigned long long)offsets[n - 1] + (unsigned long long)counts[n - 1];
}

// The tile lists of the soup for a w x h image seen through vp: per tile its
// triangle count, the offset of its list, and the lists (in context scratch;
// one host read-back, of the number of (triangle, tile) pairs).
struct Bins {
    const uint32_t *offsets, *counts, *lists;
};

static vct_status bin_triangles(vct_ctx* c, hipStream_t s, const TriSoup& soup, const Mat16& vp, int w, int h,
                                Bins* out)
{
    BinGrid G;
    G.vp = vp;
    G.w = w;
    G.h = h;
    G.tiles_x = (w + kBinSide - 1) / kBinSide;
    G.tiles_y = (h + kBinSide - 1) / kBinSide;
    const uint32_t ntiles = (uint32_t)G.tiles_x * (uint32_t)G.tiles_y;
    const uint32_t nt = soup.count ? soup.count : 1u;
    VCT_TRY(ensure(c->counts, 4u * (size_t)ntiles));
    VCT_TRY(ensure(c->offsets, 4u * (size_t)ntiles));
    VCT_TRY(ensure(c->rec_idx[1], 4u * (size_t)ntiles));
    VCT_TRY(ensure(c->rec_val, 16u * (size_t)nt));    // triangle boxes
    VCT_TRY(ensure(c->rec_key[0], 4u * (size_t)nt));  // triangle tile counts
    VCT_TRY(ensure(c->rec_key[1], 4u * (size_t)nt));  // their exclusive scan
    VCT_TRY(ensure(c->totals, 64));
    uint32_t* counts = (uint32_t*)c->counts.ptr;
    uint32_t* offsets = (uint32_t*)c->offsets.ptr;
    uint32_t* cursor = (uint32_t*)c->rec_idx[1].ptr;
    int4* boxes = (int4*)c->rec_val.ptr;
    uint32_t* tcount = (uint32_t*)c->rec_key[0].ptr;
    uint32_t* toff = (uint32_t*)c->rec_key[1].ptr;
    VCT_HIP(hipMemsetAsync(counts, 0, 4u * (size_t)ntiles, s));
    VCT_HIP(hipMemsetAsync(cursor, 0, 4u * (size_t)ntiles, s));
    unsigned long long npairs = 0;
    if (soup.count) {
        hipLaunchKernelGGL(k_bin_boxes, dim3((soup.count + 255) / 256), dim3(256), 0, s, soup.pos, soup.count, G,
                           boxes, tcount);
        size_t tmp = 0;
        VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, tcount, toff, (int)soup.count, s));
        VCT_TRY(ensure(c->scan_tmp, tmp));
        VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, tcount, toff, (int)soup.count, s));
        hipLaunchKernelGGL(k_scan_total32, dim3(1), dim3(1), 0, s, (const uint32_t*)toff, (const uint32_t*)tcount,
                           soup.count, (unsigned long long*)c->totals.ptr);
        if (!c->pinned_word) VCT_HIP(hipHostMalloc(&c->pinned_word, kPinnedWordBytes, hipHostMallocDefault));
        volatile unsigned long long* landed = (volatile unsigned long long*)c->pinned_word;
        VCT_HIP(hipMemcpyAsync((void*)landed, c->totals.ptr, 8, hipMemcpyDeviceToHost, s));
        VCT_HIP(hipStreamSynchronize(s));
        npairs = landed[0];
        // the per-triangle tile counts are exact, so the 32-bit scan is exact below 2^32
        if (npairs >= 0xffffffffull) return fail(VCT_ERR_CAPACITY, "%llu (triangle, tile) pairs exceed 2^32", npairs);
    }
    VCT_TRY(ensure(c->rec_idx[0], 4u * (size_t)(npairs ? npairs : 1)));
    uint32_t* lists = (uint32_t*)c->rec_idx[0].ptr;
    const uint32_t np = (uint32_t)npairs;
    if (np)
        hipLaunchKernelGGL(k_bin_count, dim3((np + 255) / 256), dim3(256), 0, s, (const uint32_t*)toff, soup.count,
                           (const int4*)boxes, G.tiles_x, np, counts);
    size_t tmp = 0;
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, counts, offsets, (int)ntiles, s));
    VCT_TRY(ensure(c->scan_tmp, tmp));
    VCT_HIP(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp.ptr, tmp, counts, offsets, (int)ntiles, s));
    if (np)
        hipLaunchKernelGGL(k_bin_fill, dim3((np + 255) / 256), dim3(256), 0, s, (const uint32_t*)toff, soup.count,
                           (const int4*)boxes, G.tiles_x, np, (const uint32_t*)offsets, cursor, lists);
    VCT_HIP(hipGetLastError());
    *out = Bins{offsets, counts, lists};
    return VCT_OK;
}

static void invert4(const double* m, double* o)
{
    o[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    o[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    o[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    o[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    o[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    o[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    o[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    o[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    o[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    o[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    o[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    o[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    o[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    o[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    o[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    o[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const double det = m[0] * o[0] + m[1] * o[4] + m[2] * o[8] + m[3] * o[12];
    for (int i = 0; i < 16; ++i) o[i] /= det;
}

}  // namespace vct

using namespace vct;

extern "C" {

vct_status vct_raycast_gbuffer(vct_ctx* c, vct_stream stream, const vct_mesh* mesh, const float* surface,
                               const vct_camera_ubo* cam, uint32_t* diffuse, uint16_t* normal, uint32_t* specular,
                               uint16_t* emission, float* depth, uint32_t width, uint32_t height)
{
    if (!c || !mesh || !surface || !cam || !diffuse || !normal || !specular || !emission || !depth || !width || !height)
        return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t s = (hipStream_t)stream;
    TriSoup soup;
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup, TextureUse::kRefuse));
    Mat16 inv, vp;
    for (int i = 0; i < 16; ++i) {
        inv.m[i] = cam->view_proj_inv[i];
        vp.m[i] = cam->view_proj[i];
    }
    Bins B;
    VCT_TRY(bin_triangles(c, s, soup, vp, (int)width, (int)height, &B));
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(k_gbuffer, grid, dim3(kTile), 0, s, soup.pos, soup.nrm, soup.mat, mesh->materials, surface,
                       B.offsets, B.counts, B.lists, (int)width, (int)height, inv, vp, diffuse, (uint2*)normal,
                       specular, (uint2*)emission, depth);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_raycast_shadow(vct_ctx* c, vct_stream stream, const vct_mesh* mesh,
                              const vct_directional_light_shadow_desc* ls, float* depth, uint32_t width,
                              uint32_t height)
{
    if (!c || !mesh || !ls || !depth || !width || !height) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t s = (hipStream_t)stream;
    TriSoup soup;
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup, TextureUse::kIgnore));
    double vp[16], inv[16];
    for (int col = 0; col < 4; ++col)
        for (int r = 0; r < 4; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += (double)ls->proj[k * 4 + r] * (double)ls->view[col * 4 + k];
            vp[col * 4 + r] = acc;
        }
    invert4(vp, inv);
    Mat16 invf, vpf;
    for (int i = 0; i < 16; ++i) {
        invf.m[i] = (float)inv[i];
        vpf.m[i] = (float)vp[i];
    }
    Bins B;
    VCT_TRY(bin_triangles(c, s, soup, vpf, (int)width, (int)height, &B));
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(k_shadow, grid, dim3(kTile), 0, s, soup.pos, B.offsets, B.counts, B.lists, (int)width,
                       (int)height, invf, depth);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // extern "C"
