// vct_producers.hip - G-buffer and shadow-map producers for gfx950.
//
// The reference rasterizes these in GBufferPass (GBufferPass.cpp:170-206) and
// ReflectiveShadowMapPass (ReflectiveShadowMapPass.cpp:38-88); they feed the
// hot path (cone trace and radiance injection) but are not part of it. Here
// they are ray cast: one lane per pixel, triangles streamed through LDS in
// tiles of 256 so that a workgroup's 256 rays test each triangle from LDS.
// Outputs use the reference's G-buffer channel layout.
#include <hip/hip_runtime.h>

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

// Casts the ray through pixel (px, py) of a viewProjInv; returns the nearest hit.
__device__ void cast(const float* __restrict__ soup, uint32_t ntris, const float* inv, int px, int py, int w, int h,
                     bool valid, float o[3], float d[3], float& best_t, int& best_tri, float& bu, float& bv)
{
    __shared__ RayTri tile[kTile];
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
    for (uint32_t base = 0; base < ntris; base += kTile) {
        const uint32_t t = base + threadIdx.x;
        if (t < ntris) {
            const float* p = soup + 9u * t;
            RayTri r;
            for (int k = 0; k < 3; ++k) {
                r.p0[k] = p[k];
                r.e1[k] = p[3 + k] - p[k];
                r.e2[k] = p[6 + k] - p[k];
            }
            tile[threadIdx.x] = r;
        }
        __syncthreads();
        const uint32_t n = ntris - base < (uint32_t)kTile ? ntris - base : (uint32_t)kTile;
        if (valid) {
            for (uint32_t i = 0; i < n; ++i) {
                float u, v;
                const float th = intersect(o, d, tile[i], u, v);
                if (th > 0.0f && th <= 1.0f && th < best_t) {
                    best_t = th;
                    best_tri = (int)(base + i);
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
                                                   const float* __restrict__ surface, uint32_t ntris, int w, int h,
                                                   Mat16 inv_m, Mat16 vp_m, uint32_t* __restrict__ g_diffuse,
                                                   uint2* __restrict__ g_normal, uint32_t* __restrict__ g_specular,
                                                   uint2* __restrict__ g_emission, float* __restrict__ g_depth)
{
    const float* inv = inv_m.m;
    const float* vp = vp_m.m;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool valid = px < w && py < h;
    float o[3] = {0, 0, 0}, d[3] = {0, 0, 1}, t, bu = 0.0f, bv = 0.0f;
    int tri;
    cast(soup, ntris, inv, px, py, w, h, valid, o, d, t, tri, bu, bv);
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

__global__ __launch_bounds__(kTile) void k_shadow(const float* __restrict__ soup, uint32_t ntris, int w, int h,
                                                  Mat16 inv_m, float* __restrict__ depth)
{
    const float* inv = inv_m.m;
    const int px = blockIdx.x * 16 + (threadIdx.x & 15);
    const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool valid = px < w && py < h;
    float o[3] = {0, 0, 0}, d[3] = {0, 0, 1}, t, bu, bv;
    int tri;
    cast(soup, ntris, inv, px, py, w, h, valid, o, d, t, tri, bu, bv);
    if (!valid) return;
    // orthographic light: NDC depth is linear along the near-to-far ray
    depth[(size_t)py * (size_t)w + (size_t)px] = tri < 0 ? 1.0f : t;
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
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup));
    Mat16 inv, vp;
    for (int i = 0; i < 16; ++i) {
        inv.m[i] = cam->view_proj_inv[i];
        vp.m[i] = cam->view_proj[i];
    }
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(k_gbuffer, grid, dim3(kTile), 0, s, soup.pos, soup.nrm, soup.mat, mesh->materials, surface,
                       soup.count, (int)width, (int)height, inv, vp, diffuse, (uint2*)normal, specular,
                       (uint2*)emission, depth);
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
    VCT_TRY(run_vertex_stage(c, s, mesh, &soup));
    double vp[16], inv[16];
    for (int col = 0; col < 4; ++col)
        for (int r = 0; r < 4; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += (double)ls->proj[k * 4 + r] * (double)ls->view[col * 4 + k];
            vp[col * 4 + r] = acc;
        }
    invert4(vp, inv);
    Mat16 invf;
    for (int i = 0; i < 16; ++i) invf.m[i] = (float)inv[i];
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(k_shadow, grid, dim3(kTile), 0, s, soup.pos, soup.count, (int)width, (int)height, invf, depth);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // extern "C"
