// vct_dense_grad.hip - the adjoint of the dense method: gradients of an
// image loss with respect to the level-0 albedo and opacity of an RGBA32F
// dense grid, through the cone march (vct_dense_trace) and the anisotropic
// mips (vct_dense_mipmap). No reference counterpart.
//
// Trace: per pixel and active cone the march is replayed forward keeping the
// per-step state (step length, segment, level, accumulated alpha, occlusion;
// with kStore also the fetched sample), then swept in reverse:
//   op = 1 - (1 - s.a)^c,  b = 1 - R.a,  q = 1 + (t + vs_l) decay
//   R.rgb += b c s.rgb,  R.a += b op,  O += (1 - O) op / q
// gives  ds.rgb = gR.rgb b c,  d op = gR.a b + gO (1 - O) / q,
//        ds.a = d op c (1 - s.a)^(c-1),
//        gR.a <- gR.a (1 - op) - c gR.rgb . s.rgb,  gO <- gO (1 - op / q).
// The sample gradient is spread over the texels of the sample's footprints
// (trilinear x face weight x level lerp) with hardware float atomics. Without
// kStore the samples are fetched again in the reverse sweep instead of kept,
// which halves the per-lane state.
//
// Mips: each child texel gathers from its parent (no atomics):
//   parent = mean over 4 pairs of a + (1 - a.a) b
//   => front child a: 0.25 g.rgb, 0.25 g.a - 0.25 g . b;  back child b: 0.25 (1 - a.a) g.
#include <hip/hip_runtime.h>

#include "vct_dense.hpp"
#include "vct_dense_march.hpp"
#include "vct_internal.hpp"

namespace vct {
namespace dense {

constexpr int kAdjMaxSteps = 40;

__global__ void k_set_level0(const float* __restrict__ albedo, const float* __restrict__ opacity, int res,
                             float4* __restrict__ level0)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n3 = (size_t)res * res * res;
    if (v >= n3) return;
    const int x = (int)(v % res), y = (int)((v / res) % res), z = (int)(v / ((size_t)res * res));
    const size_t D = (size_t)res + 2;
    level0[((size_t)(z + 1) * D + (size_t)(y + 1)) * D + (size_t)(x + 1)] =
        make_float4(albedo[3 * v], albedo[3 * v + 1], albedo[3 * v + 2], opacity[v]);
}

__global__ void k_level0_grad(const float4* __restrict__ grad0, int res, float* __restrict__ g_albedo,
                              float* __restrict__ g_opacity)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n3 = (size_t)res * res * res;
    if (v >= n3) return;
    const int x = (int)(v % res), y = (int)((v / res) % res), z = (int)(v / ((size_t)res * res));
    const size_t D = (size_t)res + 2;
    const float4 g = grad0[((size_t)(z + 1) * D + (size_t)(y + 1)) * D + (size_t)(x + 1)];
    g_albedo[3 * v] = g.x;
    g_albedo[3 * v + 1] = g.y;
    g_albedo[3 * v + 2] = g.z;
    g_opacity[v] = g.w;
}

// Adds scale * g spread over level m's footprint at p (the transpose of
// sample()).
__device__ __forceinline__ void scatter_level(float* __restrict__ grad, const Pyramid& P, int m, const float p[3],
                                              const int face[3], const float w[3], float scale, const f4& g)
{
    const Footprint F = footprint(P, m, p);
    if (!F.valid || scale == 0.0f) return;
    const float ax = F.fx, ay = F.fy, az = F.fz;
    const float cw[8] = {(1.0f - ax) * (1.0f - ay) * (1.0f - az), ax * (1.0f - ay) * (1.0f - az),
                         (1.0f - ax) * ay * (1.0f - az),          ax * ay * (1.0f - az),
                         (1.0f - ax) * (1.0f - ay) * az,          ax * (1.0f - ay) * az,
                         (1.0f - ax) * ay * az,                   ax * ay * az};
    const unsigned long long off[8] = {0, 1, F.dy, F.dy + 1, F.dz, F.dz + 1, F.dz + F.dy, F.dz + F.dy + 1};
    const int nf = m == 0 ? 1 : 3;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a >= nf) break;
        const float fw = scale * (m == 0 ? (w[0] + w[1] + w[2]) : w[a]);
        const unsigned long long fb = F.base + (m == 0 ? 0ull : (unsigned long long)face[a] * P.block[m]);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float k = fw * cw[c];
            float* t = grad + 4ull * (fb + off[c]);
            unsafeAtomicAdd(t + 0, k * g.x);
            unsafeAtomicAdd(t + 1, k * g.y);
            unsafeAtomicAdd(t + 2, k * g.z);
            unsafeAtomicAdd(t + 3, k * g.w);
        }
    }
}

struct StepState {
    float step, seg, lvl, rw, occ;
};

// One cone: replay forward, then the reverse sweep. Returns false if the
// march is longer than the state buffer.
template <bool kStore>
__device__ bool cone_backward(const void* __restrict__ data, float* __restrict__ grad, const Pyramid& P,
                              const March& M, const float start_in[3], const float dir[3], const f4& gC)
{
    const float vs = P.vs;
    float start[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * vs * M.start_offset * 0.5f;
    const int face[3] = {dir[0] > 0.0f ? 0 : 1, dir[1] > 0.0f ? 2 : 3, dir[2] > 0.0f ? 4 : 5};
    const float w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
    const float max_level = (float)(P.nlevels - 1);
    const float inv_vs = 1.0f / vs;
    StepState st[kAdjMaxSteps];
    f4 smp[kStore ? kAdjMaxSteps : 1];
    float step = 0.0f, diameter = vs, occlusion = 0.0f, seg = vs;
    f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
    int n = 0;
    while (step < M.max_dist && occlusion < 1.0f) {
        if (n == kAdjMaxSteps) return false;
        float pos[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pos[k] = start[k] + dir[k] * step;
        const float lvl = fminf(fmaxf(log2f(diameter * inv_vs), 0.0f), max_level);
        const f4 s = level_sample<true>(data, P, pos, lvl, face, w);
        st[n] = {step, seg, lvl, result.w, occlusion};
        if (kStore) smp[n] = s;
        ++n;
        accumulate(result, occlusion, s, lvl, step, seg, vs, M.decay);
        const float prev = step;
        step += fmaxf(diameter, vs) * M.step_factor;
        seg = step - prev;
        diameter = step * M.coef;
    }
    f4 gR = {gC.x, gC.y, gC.z, 0.0f};
    float gO = -gC.w;
    for (int k = n - 1; k >= 0; --k) {
        const StepState S = st[k];
        float pos[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) pos[i] = start[i] + dir[i] * S.step;
        const f4 s = kStore ? smp[kStore ? k : 0] : level_sample<true>(data, P, pos, S.lvl, face, w);
        const float vsl = vs * exp2f(S.lvl);
        const float c = S.seg / vsl;
        const float q = 1.0f + (S.step + vsl) * M.decay;
        const float base = 1.0f - s.w;
        const float op = 1.0f - powf(base, c);
        const float b = 1.0f - S.rw;
        const float g_op = gR.w * b + gO * (1.0f - S.occ) / q;
        f4 gs;
        gs.x = gR.x * b * c;
        gs.y = gR.y * b * c;
        gs.z = gR.z * b * c;
        gs.w = base > 0.0f ? g_op * c * powf(base, c - 1.0f) : 0.0f;
        const float dot = gR.x * s.x + gR.y * s.y + gR.z * s.z;
        gR.w = gR.w * (1.0f - op) - c * dot;
        gO = gO * (1.0f - op / q);
        const int lo = (int)floorf(S.lvl), hi = (int)ceilf(S.lvl);
        const float fl = S.lvl - floorf(S.lvl);
        if (hi != lo) {
            scatter_level(grad, P, lo, pos, face, w, 1.0f - fl, gs);
            scatter_level(grad, P, hi, pos, face, w, fl, gs);
        } else {
            scatter_level(grad, P, lo, pos, face, w, 1.0f, gs);
        }
    }
    return true;
}

template <bool kStore>
__global__ __launch_bounds__(256) void k_dense_trace_bwd(const void* __restrict__ data, Pyramid P, March M,
                                                         const uint32_t* __restrict__ g_diffuse,
                                                         const uint16_t* __restrict__ g_normal,
                                                         const float* __restrict__ g_depth, int width, int height,
                                                         const float4* __restrict__ grad_out,
                                                         float* __restrict__ grad, uint32_t* __restrict__ status)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned tiles_x = ((unsigned)width + 15u) / 16u;
    const unsigned tile = xcd_remap(blockIdx.x, gridDim.x);
    const int px = (int)(tile % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
    const int py = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
    if (px >= width || py >= height) return;
    const size_t idx = (size_t)py * (size_t)width + (size_t)px;
    const float depth = g_depth[idx];
    if (depth == 1.0f) return;
    float start[3], normal[3], albedo[3];
    pixel_setup(P, M, px, py, width, height, depth, g_normal + 4 * idx, g_diffuse[idx], start, normal, albedo);
    const float4 go = grad_out[idx];
    const float inv = 1.0f / (float)M.ncones;
    for (int i = 0; i < M.ncones; ++i) {
        const float* d = cone_dir(M, i);
        const float ct = normal[0] * d[0] + normal[1] * d[1] + normal[2] * d[2];
        if (ct < 0.0f) continue;
        const f4 gC = {go.x * inv * albedo[0] * M.diffuse_intensity * ct, go.y * inv * albedo[1] * M.diffuse_intensity * ct,
                       go.z * inv * albedo[2] * M.diffuse_intensity * ct, go.w * inv * M.ao_factor * ct};
        if (!cone_backward<kStore>(data, grad, P, M, start, d, gC)) atomicOr(status, 1u);
    }
}

// Gradient of child level c = m - 1 from parent level m (see the header).
__global__ __launch_bounds__(256) void k_dense_mip_bwd(const float4* __restrict__ child, const float4* __restrict__ gpar,
                                                       float4* __restrict__ gchild, int nc, bool single_face)
{
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n3 = (size_t)nc * nc * nc;
    if (v >= n3) return;
    const int x = (int)(v % nc), y = (int)((v / nc) % nc), z = (int)(v / ((size_t)nc * nc));
    const int np = nc >> 1;
    const size_t Dc = (size_t)nc + 2, Dp = (size_t)np + 2;
    const size_t cblock = Dc * Dc * Dc, pblock = Dp * Dp * Dp;
    const size_t pidx = ((size_t)((z >> 1) + 1) * Dp + (size_t)((y >> 1) + 1)) * Dp + (size_t)((x >> 1) + 1);
    const int i = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int f = 0; f < kFaces; ++f) {
        const int axis = f >> 1;
        const int bit = 1 << axis;
        const int j = i ^ bit;
        const bool front = ((i & bit) != 0) == ((f & 1) != 0);
        const int jx = x + ((j & 1) - (i & 1)), jy = y + (((j >> 1) & 1) - ((i >> 1) & 1)),
                  jz = z + (((j >> 2) & 1) - ((i >> 2) & 1));
        const size_t cf = single_face ? 0 : (size_t)f * cblock;
        const size_t sidx = cf + ((size_t)(z + 1) * Dc + (size_t)(y + 1)) * Dc + (size_t)(x + 1);
        const size_t tidx = cf + ((size_t)(jz + 1) * Dc + (size_t)(jy + 1)) * Dc + (size_t)(jx + 1);
        const float4 g = gpar[(size_t)f * pblock + pidx];
        const float4 partner = child[tidx];
        f4 d;
        if (front) {
            const float dot = g.x * partner.x + g.y * partner.y + g.z * partner.z + g.w * partner.w;
            d = {0.25f * g.x, 0.25f * g.y, 0.25f * g.z, 0.25f * g.w - 0.25f * dot};
        } else {
            const float t = 0.25f * (1.0f - partner.w);
            d = {t * g.x, t * g.y, t * g.z, t * g.w};
        }
        if (single_face) {
            acc.x += d.x;
            acc.y += d.y;
            acc.z += d.z;
            acc.w += d.w;
        } else {
            float4 o = gchild[sidx];
            gchild[sidx] = make_float4(o.x + d.x, o.y + d.y, o.z + d.z, o.w + d.w);
        }
    }
    if (single_face) {
        const size_t sidx = ((size_t)(z + 1) * Dc + (size_t)(y + 1)) * Dc + (size_t)(x + 1);
        float4 o = gchild[sidx];
        gchild[sidx] = make_float4(o.x + acc.x, o.y + acc.y, o.z + acc.z, o.w + acc.w);
    }
}

}  // namespace dense
}  // namespace vct

using namespace vct;

static vct_status need_f32(const vct_dense* g)
{
    if (!g) return fail(VCT_ERR_INVALID_ARGUMENT, "null grid");
    if (g->texel_bytes != 16) return fail(VCT_ERR_UNSUPPORTED, "the adjoint needs an RGBA32F grid");
    return VCT_OK;
}

// Step count of the march for these parameters: the step sequence does not
// depend on the data, only the exit on full occlusion does.
static int march_steps(const dense::Pyramid& P, const dense::March& M)
{
    float step = 0.0f, diameter = P.vs;
    int n = 0;
    while (step < M.max_dist && n <= dense::kAdjMaxSteps + 1) {
        ++n;
        step += std::fmax(diameter, P.vs) * M.step_factor;
        diameter = step * M.coef;
    }
    return n;
}

extern "C" {

vct_status vct_dense_texels(vct_dense* g, size_t* texels)
{
    if (!g || !texels) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    *texels = g->texels;
    return VCT_OK;
}

vct_status vct_dense_set_level0(vct_dense* g, vct_stream stream, const float* albedo, const float* opacity)
{
    VCT_TRY(need_f32(g));
    if (!albedo || !opacity) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n3 = (size_t)g->res * g->res * g->res;
    hipLaunchKernelGGL(dense::k_set_level0, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       albedo, opacity, g->res, (float4*)g->data.ptr);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_dense_trace_backward(vct_dense* g, vct_stream stream, const vct_gbuffer* gb,
                                    const vct_camera_ubo* cam, const vct_dense_trace_desc* d, const float* grad_out,
                                    float* grad_pyramid, uint32_t flags, uint32_t* status)
{
    VCT_TRY(need_f32(g));
    if (!gb || !cam || !d || !grad_out || !grad_pyramid || !status)
        return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (!gb->diffuse || !gb->normal || !gb->depth || !gb->width || !gb->height)
        return fail(VCT_ERR_INVALID_ARGUMENT, "G-buffer needs diffuse, normal and depth");
    dense::Pyramid P;
    dense::March M;
    VCT_TRY(dense::make_params(g, cam, d, &P, &M));
    if (march_steps(P, M) + 1 > dense::kAdjMaxSteps)
        return fail(VCT_ERR_UNSUPPORTED, "the march takes more than %d steps", dense::kAdjMaxSteps - 1);
    dim3 grid(((gb->width + 15) / 16) * ((gb->height + 15) / 16));
    hipStream_t s = (hipStream_t)stream;
    if (flags & VCT_ADJOINT_RECOMPUTE)
        hipLaunchKernelGGL(dense::k_dense_trace_bwd<false>, grid, dim3(256), 0, s, g->data.ptr, P, M, gb->diffuse,
                           gb->normal, gb->depth, (int)gb->width, (int)gb->height, (const float4*)grad_out,
                           grad_pyramid, status);
    else
        hipLaunchKernelGGL(dense::k_dense_trace_bwd<true>, grid, dim3(256), 0, s, g->data.ptr, P, M, gb->diffuse,
                           gb->normal, gb->depth, (int)gb->width, (int)gb->height, (const float4*)grad_out,
                           grad_pyramid, status);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_dense_mipmap_backward(vct_dense* g, vct_stream stream, float* grad_pyramid)
{
    VCT_TRY(need_f32(g));
    if (!grad_pyramid) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    const float4* data = (const float4*)g->data.ptr;
    float4* grad = (float4*)grad_pyramid;
    for (int m = g->nlevels - 1; m >= 1; --m) {
        const int nc = g->res >> (m - 1);
        const size_t n3 = (size_t)nc * nc * nc;
        hipLaunchKernelGGL(dense::k_dense_mip_bwd, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, data + g->offset[m - 1], (const float4*)(grad + g->offset[m]),
                           grad + g->offset[m - 1], nc, m == 1);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_dense_level0_grad(vct_dense* g, vct_stream stream, const float* grad_pyramid, float* grad_albedo,
                                 float* grad_opacity)
{
    VCT_TRY(need_f32(g));
    if (!grad_pyramid || !grad_albedo || !grad_opacity) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n3 = (size_t)g->res * g->res * g->res;
    hipLaunchKernelGGL(dense::k_level0_grad, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float4*)grad_pyramid, g->res, grad_albedo, grad_opacity);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // extern "C"
