// vct_dense_grad.hip - the adjoint of the dense method: gradients of an
// image loss with respect to the level-0 albedo and opacity of an RGBA32F
// dense grid, through the cone march (vct_dense_trace) and the anisotropic
// mips (vct_dense_mipmap). No reference counterpart.
//
// Trace: see k_dense_trace_bwd (a reverse sweep per cone, gradients
// gathered per step in LDS windows before they reach global memory).
//
// Mips: each child texel gathers from its parent (no atomics):
//   parent = mean over 4 pairs of a + (1 - a.a) b
//   => front child a: 0.25 g.rgb, 0.25 g.a - 0.25 g . b;  back child b: 0.25 (1 - a.a) g.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vct_dense.hpp"
#include "vct_dense_march.hpp"
#include "vct_internal.hpp"

namespace vct {
namespace dense {

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

// Bordered integer corner and fractions of level m's trilinear footprint at p:
// the arithmetic of footprint().
struct Corner {
    int i[3];
    float f[3];
    bool valid;
};

__device__ __forceinline__ Corner corner_of(const Pyramid& P, int m, const float p[3])
{
    Corner c;
    c.valid = true;
    const int n = P.n[m];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float u = p[k] * P.inv_vs[m] - 0.5f;
        const float f = floorf(u);
        c.f[k] = u - f;
        if (!(f >= -1.0f) || f > (float)(n - 1)) c.valid = false;
        c.i[k] = (int)fminf(fmaxf(f, -1.0f), (float)(n - 1)) + 1;
    }
    return c;
}

// LDS windows accumulate in 32-bit fixed point: on gfx950 ds_add_u32 costs
// about half of ds_add_f64 per wave instruction and ds_add_f32 is more than
// twenty times slower. Each (wave, cone) carries the largest magnitude of its
// sample gradients per channel, which bounds every window sum (see
// fixed_scale); the per-workgroup accumulators of the coarse levels, shared
// by many cones, stay in double.
constexpr int kPoolUnits = 1024;  // a wave's LDS pool for its two live level windows, in 16 B units (16 KB)
constexpr int kMaxCopies = 4;     // lane-striped copies of a small window
constexpr int kCoarseN = 4;       // levels with at most 4^3 interior cells stay in LDS per tile
constexpr int kCoarseFloats = 24 * (64 + 8 + 1);

// A value every lane holds alike, moved to a scalar register: what is
// derived from it (window boxes, scales, loop bounds) then stays in scalar
// registers instead of taking a vector register per value.
__device__ __forceinline__ float uniform(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return uniform(v);
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return uniform(v);
}

// N wave maxima at once: each butterfly step shuffles all N values, so the
// N chains' latencies overlap instead of adding up.
template <int N>
__device__ __forceinline__ void wave_max_n(float (&v)[N])
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        float t[N];
#pragma unroll
        for (int i = 0; i < N; ++i) t[i] = __shfl_xor(v[i], o, 64);
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = fmaxf(v[i], t[i]);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = uniform(v[i]);
}

// The interior cells of level m that a set of start points can reach while
// the level is in use: their box swept along the cone over the level's steps
// [ta, tb], clipped to the interior (border texels are constant zero and take
// no gradient). An empty set gives vol 0. A footprint with corner u covers
// cells floor(u) + 1 and + 2 (bordered), so the box is exact: a sample whose
// position rounds one cell past it (its position is computed in another
// order) takes splat_level's per-corner path, which adds what falls outside
// the window to global memory. (One cell of margin on each side, to keep such
// samples in the window, made the fine levels' windows 2.4x larger and the
// backward 1.3x slower, measured.)
struct Box {
    int lo[3], d[3];
    int vol, base;
};

__device__ __forceinline__ Box box_of(const Pyramid& P, int m, const float blo[3], const float bhi[3],
                                      const float dir[3], float ta, float tb)
{
    Box B;
    B.vol = 1;
    B.base = 0;
    const int n = P.n[m];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo_w = fminf(blo[k] + dir[k] * ta, blo[k] + dir[k] * tb);
        const float hi_w = fmaxf(bhi[k] + dir[k] * ta, bhi[k] + dir[k] * tb);
        const float a = fminf(fmaxf(lo_w * P.inv_vs[m] - 0.5f, -4.0f), (float)n + 4.0f);
        const float b = fminf(fmaxf(hi_w * P.inv_vs[m] - 0.5f, -4.0f), (float)n + 4.0f);
        const int lo = max((int)floorf(a) + 1, 1);
        const int hi = min((int)floorf(b) + 2, n);
        B.lo[k] = lo;
        B.d[k] = hi >= lo ? hi - lo + 1 : 0;
        B.vol *= B.d[k];
    }
    return B;
}

// Words one copy of a window part of `vol` texels takes: 4 channel planes
// and one word of padding, so that copy c of a texel lies c words further
// along the banks and the lanes of one instruction, which add to the same
// texel in different copies, do not meet in one bank.
__device__ __forceinline__ int copy_stride(int vol) { return vol ? 4 * vol + 1 : 0; }

// A wave's LDS window for one level: one box, or two when the wave's start
// points split across the middle of their longest axis (a depth edge) and the
// two boxes fit where the one does not. Layout in LDS: per part, `copies`
// blocks of 4 channel planes of `vol` 32-bit fixed-point texels from the
// part's base (in elements), copy_stride(vol) elements per copy; lds is
// false when nothing fits in `cap` 16-byte units; `used` is the units it
// takes; q[c] is channel c's fixed-point scale.
struct Window {
    Box part[2];
    int parts, axis, copies, used;
    float mid;
    bool lds;
    float q[4];
};

// Fixed-point scale for a window of a (wave, cone) whose sample gradients are
// at most gmax in magnitude, over nsteps steps: a texel element takes at most
// 64 lanes x nsteps contributions, each at most gmax (trilinear and level
// weights are at most 1), so the sum of the scaled contributions stays below
// 2^30 in magnitude. Contributions are truncated to integers: the error per
// contribution is below one unit, gmax * 64 * nsteps / 2^30.
__device__ __forceinline__ float fixed_scale(float gmax, int nsteps)
{
    return gmax > 1e-30f ? 1073741824.0f / (gmax * 64.0f * (float)nsteps) : 0.0f;
}

__device__ __forceinline__ Window window_of(const Pyramid& P, int m, const float start[3], bool lane_in,
                                            const float blo[3], const float bhi[3], const float dir[3], float ta,
                                            float tb, int base, int cap)
{
    Window W;
    W.parts = 1;
    W.axis = 0;
    W.mid = 0.0f;
    W.part[0] = box_of(P, m, blo, bhi, dir, ta, tb);
    W.part[1].vol = 0;
    if (2 * W.part[0].vol > cap) {
        int ax = 0;
        if (bhi[1] - blo[1] > bhi[ax] - blo[ax]) ax = 1;
        if (bhi[2] - blo[2] > bhi[ax] - blo[ax]) ax = 2;
        const float mid = 0.5f * (blo[ax] + bhi[ax]);
        const bool up = lane_in && start[ax] >= mid, down = lane_in && start[ax] < mid;
        float lo0[3], hi0[3], lo1[3], hi1[3];
        float r[12];  // the twelve reductions side by side (minima as maxima of negatives)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            r[k] = down ? -start[k] : -INFINITY;
            r[3 + k] = down ? start[k] : -INFINITY;
            r[6 + k] = up ? -start[k] : -INFINITY;
            r[9 + k] = up ? start[k] : -INFINITY;
        }
        wave_max_n<12>(r);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo0[k] = -r[k];
            hi0[k] = r[3 + k];
            lo1[k] = -r[6 + k];
            hi1[k] = r[9 + k];
        }
        const Box b0 = box_of(P, m, lo0, hi0, dir, ta, tb), b1 = box_of(P, m, lo1, hi1, dir, ta, tb);
        if (b0.vol + b1.vol < W.part[0].vol) {
            W.parts = 2;
            W.axis = ax;
            W.mid = mid;
            W.part[0] = b0;
            W.part[1] = b1;
        }
    }
    const int total = W.part[0].vol + W.part[1].vol;
    const int words = copy_stride(W.part[0].vol) + copy_stride(W.part[1].vol);  // one copy of both parts
    W.lds = total > 0 && words <= 4 * cap;
    int copies = kMaxCopies;
    while (copies > 1 && copies * words > 4 * cap) copies >>= 1;
    W.copies = copies;
    W.used = W.lds ? (copies * words + 3) / 4 : 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) W.q[c] = 0.0f;
    W.part[0].base = base;
    W.part[1].base = base + copies * copy_stride(W.part[0].vol);
    return W;
}

// Per-workgroup accumulators of the coarsest levels (n <= kCoarseN, at most
// three: n = 4, 2, 1): 6 faces x 4 channel planes of n^3 interior cells per
// level, level first + r from offset b[r].
struct Coarse {
    int first;  // first level held (nlevels when none)
    int b0, b1, b2;
};

__device__ __forceinline__ Coarse coarse_layout(const Pyramid& P)
{
    Coarse C;
    C.first = P.nlevels;
    for (int m = 1; m < P.nlevels; ++m)
        if (P.n[m] <= kCoarseN) {
            C.first = m;
            break;
        }
    C.b0 = 0;
    C.b1 = C.first < P.nlevels ? 24 * P.n[C.first] * P.n[C.first] * P.n[C.first] : 0;
    C.b2 = C.first + 1 < P.nlevels ? C.b1 + 24 * P.n[C.first + 1] * P.n[C.first + 1] * P.n[C.first + 1] : C.b1;
    return C;
}

__device__ __forceinline__ int coarse_base(const Coarse& C, int m)
{
    const int r = m - C.first;
    return r == 0 ? C.b0 : (r == 1 ? C.b1 : C.b2);
}

// Adds (face weights) x v to channel ch of interior texel (x, y, z) of level
// m: the tile's LDS accumulator for a coarse level, else global memory; one
// face at level 0 (weighted by w0 + w1 + w2, as sample() does), three above.
__device__ __forceinline__ void add_channel(float* __restrict__ grad, double* __restrict__ coarse, const Coarse& C,
                                            const Pyramid& P, int m, int x, int y, int z, int ch, const int face[3],
                                            const float w[3], float v)
{
    if (m >= C.first) {
        const int n = P.n[m], n3 = n * n * n;
        const int t = ((z - 1) * n + (y - 1)) * n + (x - 1);
        for (int a = 0; a < 3; ++a) atomicAdd(coarse + coarse_base(C, m) + (face[a] * 4 + ch) * n3 + t, (double)(w[a] * v));
        return;
    }
    const unsigned long long D = (unsigned long long)P.n[m] + 2;
    const unsigned long long idx = P.offset[m] + ((unsigned long long)z * D + (unsigned long long)y) * D + x;
    if (m == 0) {
        unsafeAtomicAdd(grad + 4ull * idx + ch, (w[0] + w[1] + w[2]) * v);
        return;
    }
    for (int a = 0; a < 3; ++a)
        unsafeAtomicAdd(grad + 4ull * (idx + (unsigned long long)face[a] * P.block[m]) + ch, w[a] * v);
}

// Adds v (in fixed-point units) to texel element t of a window whose channel
// planes are `plane` elements apart. Channel planes (not 4 channels per
// texel) let one LDS instruction's 64 lanes spread over all banks.
__device__ __forceinline__ void lds_add4(int* __restrict__ win, int t, int plane, const float v[4])
{
    atomicAdd(win + t, (int)v[0]);
    atomicAdd(win + t + plane, (int)v[1]);
    atomicAdd(win + t + 2 * plane, (int)v[2]);
    atomicAdd(win + t + 3 * plane, (int)v[3]);
}

// Gradient of one step's samples on level m: into the wave's LDS window when
// it has one, else straight to add_channel. A lane visits its 8 corners in
// the order j ^ r with r set by its pixel position in the 8 x 8 wave, so the
// lanes of one footprint rarely meet in the same cell in one LDS instruction.
__device__ __forceinline__ void splat_level(float* __restrict__ grad, int* __restrict__ win,
                                            double* __restrict__ coarse, const Coarse& C, const Pyramid& P, int m,
                                            const Window& W, float scale, const float start[3], const float pos[3],
                                            const f4& gs, const int face[3], const float w[3], int lane)
{
    const Corner c = corner_of(P, m, pos);
    if (!c.valid) return;
    const bool second = W.parts == 2 && start[W.axis] >= W.mid;
    Box B;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        B.lo[k] = second ? W.part[1].lo[k] : W.part[0].lo[k];
        B.d[k] = second ? W.part[1].d[k] : W.part[0].d[k];
    }
    B.vol = second ? W.part[1].vol : W.part[0].vol;
    B.base = second ? W.part[1].base : W.part[0].base;
    const float g[4] = {gs.x * scale, gs.y * scale, gs.z * scale, gs.w * scale};
    const float gq[4] = {g[0] * W.q[0], g[1] * W.q[1], g[2] * W.q[2], g[3] * W.q[3]};
    const int px = lane & 7, py = lane >> 3;
    const int r = (px & 1) | ((py & 1) << 1) | ((((px >> 1) ^ (py >> 1)) & 1) << 2);
    const int rx = c.i[0] - B.lo[0], ry = c.i[1] - B.lo[1], rz = c.i[2] - B.lo[2];
    if (W.lds && rx >= 0 && ry >= 0 && rz >= 0 && rx + 1 < B.d[0] && ry + 1 < B.d[1] && rz + 1 < B.d[2]) {
        // the whole footprint lies in the window (whose cells are all interior)
        const int d0 = B.d[0], d01 = B.d[0] * B.d[1];
        const int cp = (lane & (W.copies - 1)) * copy_stride(B.vol);
        const int bx = r & 1, by = (r >> 1) & 1, bz = r >> 2;
        const int t0 = B.base + cp + (rz + bz) * d01 + (ry + by) * d0 + rx + bx;
        const int sx = bx ? -1 : 1, sy = by ? -d0 : d0, sz = bz ? -d01 : d01;
        const float wx[2] = {bx ? c.f[0] : 1.0f - c.f[0], bx ? 1.0f - c.f[0] : c.f[0]};
        const float wy[2] = {by ? c.f[1] : 1.0f - c.f[1], by ? 1.0f - c.f[1] : c.f[1]};
        const float wz[2] = {bz ? c.f[2] : 1.0f - c.f[2], bz ? 1.0f - c.f[2] : c.f[2]};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int jx = j & 1, jy = (j >> 1) & 1, jz = j >> 2;
            const float k = wx[jx] * wy[jy] * wz[jz];
            const float v[4] = {k * gq[0], k * gq[1], k * gq[2], k * gq[3]};
            lds_add4(win, t0 + (jx ? sx : 0) + (jy ? sy : 0) + (jz ? sz : 0), B.vol, v);
        }
        return;
    }
    const int n = P.n[m];
    const int cp = W.lds ? (lane & (W.copies - 1)) * copy_stride(B.vol) : 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
        const int x = c.i[0] + qx, y = c.i[1] + qy, z = c.i[2] + qz;
        if (x < 1 || x > n || y < 1 || y > n || z < 1 || z > n) continue;
        const float k = (qx ? c.f[0] : 1.0f - c.f[0]) * (qy ? c.f[1] : 1.0f - c.f[1]) * (qz ? c.f[2] : 1.0f - c.f[2]);
        const int ux = x - B.lo[0], uy = y - B.lo[1], uz = z - B.lo[2];
        if (W.lds && (unsigned)ux < (unsigned)B.d[0] && (unsigned)uy < (unsigned)B.d[1] &&
            (unsigned)uz < (unsigned)B.d[2]) {
            const float v[4] = {k * gq[0], k * gq[1], k * gq[2], k * gq[3]};
            lds_add4(win, B.base + cp + (uz * B.d[1] + uy) * B.d[0] + ux, B.vol, v);
        } else {
            for (int ch = 0; ch < 4; ++ch) add_channel(grad, coarse, C, P, m, x, y, z, ch, face, w, k * g[ch]);
        }
    }
}

// Sums a window's copies into global memory (or the tile's coarse
// accumulator) and leaves the window zeroed. Lane l takes (texel, channel)
// entries l, l + 64, ...: consecutive lanes add to consecutive floats of a row
// of the gradient pyramid. Texel coordinates come from float reciprocals,
// exact here since a window holds at most 2^10 texels.
__device__ __forceinline__ void flush_window(float* __restrict__ grad, int* __restrict__ win,
                                             double* __restrict__ coarse, const Coarse& C, const Pyramid& P, int m,
                                             const Window& W, const int face[3], const float w[3], int lane)
{
    const int n = P.n[m], D = n + 2;
    const float ws = w[0] + w[1] + w[2];
    float inv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) inv[c] = W.q[c] > 0.0f ? 1.0f / W.q[c] : 0.0f;
#pragma unroll
    for (int pp = 0; pp < 2; ++pp) {
        if (pp >= W.parts) break;
        const Box& B = pp ? W.part[1] : W.part[0];
        const int d0 = B.d[0], d01 = B.d[0] * B.d[1];
        const float inv_d0 = 1.0f / (float)d0, inv_d01 = 1.0f / (float)d01;
        const unsigned long long org =
            P.offset[m] + ((unsigned long long)B.lo[2] * D + (unsigned long long)B.lo[1]) * D + B.lo[0];
        float* g0 = grad + 4ull * org;
        float* gf[3];
        for (int a = 0; a < 3; ++a) gf[a] = g0 + (m == 0 ? 0ull : 4ull * (unsigned long long)face[a] * P.block[m]);
        const int stride = copy_stride(B.vol);
        for (int e = lane; e < 4 * B.vol; e += 64) {
            int acc = 0;
            const int t = e >> 2, ch = e & 3;
            const int li0 = B.base + ch * B.vol + t;
            // up to kMaxCopies reads in flight, then the zeroing
            int part[kMaxCopies];
#pragma unroll
            for (int cc = 0; cc < kMaxCopies; ++cc) part[cc] = cc < W.copies ? win[li0 + cc * stride] : 0;
#pragma unroll
            for (int cc = 0; cc < kMaxCopies; ++cc) {
                acc += part[cc];
                if (cc < W.copies) win[li0 + cc * stride] = 0;
            }
            if (acc == 0) continue;
            const int z = (int)(((float)t + 0.5f) * inv_d01);
            const int rem = t - z * d01;
            const int y = (int)(((float)rem + 0.5f) * inv_d0);
            const int x = rem - y * d0;
            const float fv = (float)acc * (ch == 0 ? inv[0] : ch == 1 ? inv[1] : ch == 2 ? inv[2] : inv[3]);
            if (m >= C.first) {
                const int n3 = n * n * n;
                const int tc = ((B.lo[2] + z - 1) * n + (B.lo[1] + y - 1)) * n + (B.lo[0] + x - 1);
                for (int a = 0; a < 3; ++a) atomicAdd(coarse + coarse_base(C, m) + (face[a] * 4 + ch) * n3 + tc, (double)(w[a] * fv));
                continue;
            }
            const int off = 4 * ((z * D + y) * D + x) + ch;
            if (m == 0) {
                unsafeAtomicAdd(g0 + off, ws * fv);
            } else {
                for (int a = 0; a < 3; ++a) unsafeAtomicAdd(gf[a] + off, w[a] * fv);
            }
        }
    }
}

// Orders one wave's LDS accesses before and after (the wave's lanes write
// cells that other lanes of the same wave read next).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The march's step sequence (march() in vct_dense_march.hpp): it does not
// depend on the data, only the exit on full occlusion does. Returns the
// number of steps below max_dist, or -1 when there are more than kAdjMaxSteps.
__device__ __forceinline__ int step_table(const Pyramid& P, const March& M, float* step_t, float* seg_t, float* lvl_t)
{
    const float vs = P.vs, inv_vs = 1.0f / vs, max_level = (float)(P.nlevels - 1);
    float step = 0.0f, diameter = vs, seg = vs;
    int cnt = 0;
    while (step < M.max_dist && cnt < kAdjMaxSteps) {
        step_t[cnt] = step;
        seg_t[cnt] = seg;
        lvl_t[cnt] = fminf(fmaxf(hw_log2(diameter * inv_vs), 0.0f), max_level);
        ++cnt;
        const float prev = step;
        step += fmaxf(diameter, vs) * M.step_factor;
        seg = step - prev;
        diameter = step * M.coef;
    }
    return step < M.max_dist ? -1 : cnt;
}

// Per-step constants of accumulate() (vct_dense_march.hpp), from the same
// step_consts(): the opacity correction and the reciprocal of the occlusion
// divisor 1 + (t + vs_l) decay.
__device__ __forceinline__ void step_coefs(const Pyramid& P, const March& M, int cnt, const float* step_t,
                                           const float* seg_t, const float* lvl_t, float* corr_t, float* iq_t)
{
    const float inv_vs = 1.0f / P.vs;
    for (int k = 0; k < cnt; ++k) step_consts(lvl_t[k], step_t[k], seg_t[k], P.vs, inv_vs, M.decay, corr_t[k], iq_t[k]);
}

// accumulate() with the step's constants from the table; returns
// (1 - s.a)^corr for the reverse sweep.
__device__ __forceinline__ float replay_step(f4& result, float& occlusion, const f4& s, float corr, float iq)
{
    return accumulate(result, occlusion, s, corr, iq);
}

// One (wave, cone) of the reverse sweep, handed from the sweep kernel to the
// splat kernel: its sample gradients are records base + k * 64 + lane for
// steps k < kc; `mask` holds the lanes that trace the cone, blo / bhi bound
// their start points and gmax[c] bounds the magnitude of channel c of every
// record.
struct SweepHeader {
    uint32_t cone, kc, base;
    int px0, py0;  // the wave's 8 x 8 pixel block
    uint32_t pad;
    unsigned long long mask;
    float blo[3], bhi[3];
    float gmax[4];
};
static_assert(sizeof(SweepHeader) == kSweepHeaderBytes, "SweepHeader layout");

// Sweep: one 16 x 16 pixel tile and a group of cones per workgroup (small
// groups keep the grid many rounds deep, so waves with long marches do not
// leave the chip idle at the end), each wave (8 x 8 pixels) on its own, no
// LDS beyond the step table. Per cone every lane replays its
// march keeping the per-step state (two steps' samples fetched together, as
// in the forward), then the wave runs the reverse recurrence in lockstep
// (every lane's step k is at the same distance and level):
//   op = 1 - (1 - s.a)^c,  b = 1 - R.a,  q = 1 + (t + vs_l) decay
//   R.rgb += b c s.rgb,  R.a += b op,  O += (1 - O) op / q
// gives  ds.rgb = gR.rgb b c,  d op = gR.a b + gO (1 - O) / q,
//        ds.a = d op c (1 - s.a)^(c-1),
//        gR.a <- gR.a (1 - op) - c gR.rgb . s.rgb,  gO <- gO (1 - op / q).
// The sample gradients ds (zero where a lane's march has ended) go to one
// record block per (wave, cone). Where the samples come from:
//   kSweepRecompute  the replay fetches them, the reverse sweep again;
//   kSweepKeep       the replay fetches them and keeps them per lane;
//   kSweepTape       the forward wrote them to the tape (vct_dense_trace_tape)
//                    with each lane's step count: the replay only reads them,
//                    and the record block is the (wave, cone)'s tape slot,
//                    overwritten in place (rec is the tape's sample array).
//                    Marches of at most kTapeRegs steps skip this kernel: the
//                    fused splat runs their sweep (sweep_slot).
enum { kSweepRecompute = 0, kSweepKeep = 1, kSweepTape = 2 };
constexpr int kTapeRegs = 16;

// One step of the reverse recurrence above: returns the sample gradient ds
// and carries gR.a and gO back over the step. pc = (1 - s.a)^c from the
// replay, rw and occ the accumulated alpha and occlusion before the step.
__device__ __forceinline__ float4 reverse_step(f4& gR, float& gO, const f4& s, float c, float iq, float pc, float rw,
                                               float occ)
{
    const float base_a = fmaxf(1.0f - s.w, 0.0f);
    const float op = 1.0f - pc;
    const float b = 1.0f - rw;
    const float g_op = gR.w * b + gO * (1.0f - occ) * iq;
    float4 out;
    out.x = gR.x * b * c;
    out.y = gR.y * b * c;
    out.z = gR.z * b * c;
    out.w = base_a > 0.0f ? g_op * c * (pc / base_a) : 0.0f;  // d/da of 1 - base_a^c
    const float dot = gR.x * s.x + gR.y * s.y + gR.z * s.z;
    gR.w = gR.w * (1.0f - op) - c * dot;
    gO = gO * (1.0f - op * iq);
    return out;
}

template <int kMode>
__global__ __launch_bounds__(256) void k_dense_bwd_sweep(const void* __restrict__ data, Pyramid P, March M,
                                                         const uint32_t* __restrict__ g_diffuse,
                                                         const uint16_t* __restrict__ g_normal,
                                                         const float* __restrict__ g_depth, int width, int height,
                                                         const float4* __restrict__ grad_out, uint32_t tile0,
                                                         uint32_t rec_cap, float4* __restrict__ starts,
                                                         float4* __restrict__ rec,
                                                         SweepHeader* __restrict__ hdr, uint32_t* __restrict__ ctr,
                                                         uint32_t* __restrict__ status,
                                                         const uint8_t* __restrict__ tape_n, uint32_t tape_cap)
{
    constexpr bool kStore = kMode != kSweepRecompute;
    __shared__ float s_step[kAdjMaxSteps], s_seg[kAdjMaxSteps], s_lvl[kAdjMaxSteps];
    __shared__ float s_corr[kAdjMaxSteps], s_iq[kAdjMaxSteps];
    __shared__ int s_nsteps;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned tiles_x = ((unsigned)width + 15u) / 16u;
    const unsigned tile = tile0 + xcd_remap(blockIdx.x, gridDim.x);
    const int px0 = (int)(tile % tiles_x) * 16 + (wave & 1) * 8, py0 = (int)(tile / tiles_x) * 16 + (wave >> 1) * 8;
    const int px = px0 + (lane & 7), py = py0 + (lane >> 3);
    const float vs = P.vs;
    if (tid == 0) {
        const int cnt = step_table(P, M, s_step, s_seg, s_lvl);
        s_nsteps = cnt;
        step_coefs(P, M, cnt, s_step, s_seg, s_lvl, s_corr, s_iq);
    }
    bool on = px < width && py < height;
    size_t idx = 0;
    float start[3] = {0.0f, 0.0f, 0.0f}, normal[3] = {0.0f, 0.0f, 1.0f}, albedo[3] = {0.0f, 0.0f, 0.0f};
    float4 go = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (on) {
        idx = (size_t)py * (size_t)width + (size_t)px;
        const float depth = g_depth[idx];
        on = depth != 1.0f;
        if (on) {
            pixel_setup(P, M, px, py, width, height, depth, g_normal + 4 * idx, g_diffuse[idx], start, normal, albedo);
            go = grad_out[idx];
        }
        if (blockIdx.y == 0) starts[idx] = make_float4(start[0], start[1], start[2], 0.0f);
    }
    __syncthreads();
    const int nsteps = s_nsteps;
    if (nsteps < 0) {
        if (tid == 0) atomicOr(status, 1u);
        return;
    }
    if (!(wave_max(on ? 1.0f : 0.0f) > 0.0f)) return;  // no surface pixel in this wave
    const float inv = 1.0f / (float)M.ncones;
    // this workgroup's cones: gridDim.y groups of consecutive cones
    const int per_group = (M.ncones + (int)gridDim.y - 1) / (int)gridDim.y;
    const int cone0 = (int)blockIdx.y * per_group, cone1 = min(cone0 + per_group, M.ncones);
    for (int i = cone0; i < cone1; ++i) {
        const float* d = cone_dir(M, i);
        const float dir[3] = {d[0], d[1], d[2]};
        const int face[3] = {dir[0] > 0.0f ? 0 : 1, dir[1] > 0.0f ? 2 : 3, dir[2] > 0.0f ? 4 : 5};
        const float w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
        const float ct = normal[0] * dir[0] + normal[1] * dir[1] + normal[2] * dir[2];
        const bool lane_on = on && ct >= 0.0f;
        float cstart[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) cstart[k] = start[k] + dir[k] * vs * M.start_offset * 0.5f;
        // forward replay, two steps' samples in flight
        float st_rw[kAdjMaxSteps], st_occ[kAdjMaxSteps], st_pc[kAdjMaxSteps];
        f4 smp[kStore ? kAdjMaxSteps : 1];
        int n = 0;
        const size_t slot = ((size_t)tile * 4u + (size_t)wave) * (size_t)M.ncones + (size_t)i;
        if (kMode == kSweepTape) {
            int nt = lane_on ? (int)tape_n[slot * 64u + (size_t)lane] : 0;
            if (nt > nsteps) {  // kTapeOverflow, or a march longer than the step table
                atomicOr(status, 1u);
                nt = 0;
            }
            const float4* col = rec + slot * (size_t)tape_cap * 64u + (size_t)lane;
            for (int k = 0; k < nt; ++k) {
                const float4 t = col[(size_t)k * 64u];
                smp[kStore ? k : 0] = {t.x, t.y, t.z, t.w};
            }
            f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
            float occlusion = 0.0f;
            for (int k = 0; k < nt; ++k) {
                st_rw[k] = result.w;
                st_occ[k] = occlusion;
                st_pc[k] = replay_step(result, occlusion, smp[kStore ? k : 0], s_corr[k], s_iq[k]);
            }
            n = nt;
        } else if (lane_on) {
            f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
            float occlusion = 0.0f;
            for (int k = 0; k < nsteps && occlusion < 1.0f; k += 2) {
                float pa[3], pb[3];
                const bool two = k + 1 < nsteps;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    pa[a] = cstart[a] + dir[a] * s_step[k];
                    pb[a] = cstart[a] + dir[a] * s_step[two ? k + 1 : k];
                }
                const f4 sa = level_sample<true>(data, P, pa, s_lvl[k], face, w);
                const f4 sb = level_sample<true>(data, P, pb, s_lvl[two ? k + 1 : k], face, w);
                st_rw[k] = result.w;
                st_occ[k] = occlusion;
                if (kStore) smp[kStore ? k : 0] = sa;
                st_pc[k] = replay_step(result, occlusion, sa, s_corr[k], s_iq[k]);
                n = k + 1;
                if (!two || !(occlusion < 1.0f)) break;
                st_rw[k + 1] = result.w;
                st_occ[k + 1] = occlusion;
                if (kStore) smp[kStore ? k + 1 : 0] = sb;
                st_pc[k + 1] = replay_step(result, occlusion, sb, s_corr[k + 1], s_iq[k + 1]);
                n = k + 2;
            }
        }
        float blo[3], bhi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            blo[k] = wave_min(lane_on ? start[k] : INFINITY);
            bhi[k] = wave_max(lane_on ? start[k] : -INFINITY);
        }
        const int kc = (int)wave_max((float)n);
        if (kc == 0) continue;
        const unsigned long long mask = __ballot(lane_on);
        uint32_t base = 0, h = 0;
        if (kMode == kSweepTape) {
            base = (uint32_t)(slot * (size_t)tape_cap * 64u);  // the host keeps the tape below 2^32 float4
        } else {
            if (lane == 0) base = atomicAdd(&ctr[0], (uint32_t)kc * 64u);
            base = __shfl(base, 0, 64);
            if (base + (uint32_t)kc * 64u > rec_cap) {  // never with the host's sizing; checked all the same
                if (lane == 0) atomicOr(status, 2u);
                continue;
            }
        }
        if (lane == 0) {
            h = atomicAdd(&ctr[1], 1u);
            SweepHeader H;
            H.cone = (uint32_t)i;
            H.kc = (uint32_t)kc;
            H.base = base;
            H.px0 = px0;
            H.py0 = py0;
            for (int k = 0; k < 3; ++k) {
                H.blo[k] = blo[k];
                H.bhi[k] = bhi[k];
            }
            H.pad = 0u;
            H.mask = mask;
            for (int c = 0; c < 4; ++c) H.gmax[c] = 0.0f;  // set after the reverse sweep
            hdr[h] = H;
        }
        f4 gR = {go.x * inv * albedo[0] * M.diffuse_intensity * ct, go.y * inv * albedo[1] * M.diffuse_intensity * ct,
                 go.z * inv * albedo[2] * M.diffuse_intensity * ct, 0.0f};
        float gO = -(go.w * inv * M.ao_factor * ct);
        float amax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = kc - 1; k >= 0; --k) {
            float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (k < n) {
                f4 s;
                if (kStore) {
                    s = smp[kStore ? k : 0];
                } else {
                    float pos[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) pos[a] = cstart[a] + dir[a] * s_step[k];
                    s = level_sample<true>(data, P, pos, s_lvl[k], face, w);
                }
                out = reverse_step(gR, gO, s, s_corr[k], s_iq[k], st_pc[k], st_rw[k], st_occ[k]);
            }
            amax[0] = fmaxf(amax[0], fabsf(out.x));
            amax[1] = fmaxf(amax[1], fabsf(out.y));
            amax[2] = fmaxf(amax[2], fabsf(out.z));
            amax[3] = fmaxf(amax[3], fabsf(out.w));
            rec[base + (uint32_t)k * 64u + (uint32_t)lane] = out;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) amax[c] = wave_max(amax[c]);
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) hdr[h].gmax[c] = amax[c];
        }
    }
}

// The reverse sweep of one (wave, cone) tape slot inside the fused backward
// (kSweepTapeRegs' arithmetic): replays the recurrence from the stored
// samples, overwrites them with the sample gradients and fills the header
// fields the splat reads. Returns false when the slot holds no step.
__device__ __forceinline__ bool sweep_slot(const Pyramid& P, const March& M, float4* __restrict__ tape,
                                           const uint8_t* __restrict__ counts, uint32_t cap, size_t slot, int cone,
                                           int nsteps, const float* s_corr, const float* s_iq,
                                           bool on, const float start[3], const float normal[3],
                                           const float albedo[3], const float4& go, int lane,
                                           uint32_t* __restrict__ status, SweepHeader& H)
{
    const float* d = cone_dir(M, cone);
    const float dir[3] = {d[0], d[1], d[2]};
    const float ct = normal[0] * dir[0] + normal[1] * dir[1] + normal[2] * dir[2];
    const bool lane_on = on && ct >= 0.0f;
    int nt = lane_on ? (int)counts[slot * 64u + (size_t)lane] : 0;
    if (nt > nsteps || nt > kTapeRegs) {  // kTapeOverflow, or a march longer than the step table
        atomicOr(status, 1u);
        nt = 0;
    }
    const int kc = (int)wave_max((float)nt);
    if (kc == 0) return false;
    float4* col = tape + slot * (size_t)cap * 64u + (size_t)lane;
    // steps below the wave's count are in the slot for every lane (beyond a
    // lane's own count they hold no sample of it and only feed state that is
    // never read); the reverse sweep reads them again rather than keeping
    // them live across it
    float r_rw[kTapeRegs], r_occ[kTapeRegs], r_pc[kTapeRegs];
    f4 result = {0.0f, 0.0f, 0.0f, 0.0f};
    float occlusion = 0.0f;
#pragma unroll
    for (int k = 0; k < kTapeRegs; ++k) {
        if (k >= kc) break;
        const float4 t = col[(size_t)k * 64u];
        r_rw[k] = result.w;
        r_occ[k] = occlusion;
        r_pc[k] = replay_step(result, occlusion, {t.x, t.y, t.z, t.w}, s_corr[k], s_iq[k]);
    }
    const float inv = 1.0f / (float)M.ncones;
    f4 gR = {go.x * inv * albedo[0] * M.diffuse_intensity * ct, go.y * inv * albedo[1] * M.diffuse_intensity * ct,
             go.z * inv * albedo[2] * M.diffuse_intensity * ct, 0.0f};
    float gO = -(go.w * inv * M.ao_factor * ct);
    float amax[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = kTapeRegs - 1; k >= 0; --k) {
        if (k >= kc) continue;
        float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 t = col[(size_t)k * 64u];
        if (k < nt) out = reverse_step(gR, gO, {t.x, t.y, t.z, t.w}, s_corr[k], s_iq[k], r_pc[k], r_rw[k], r_occ[k]);
        amax[0] = fmaxf(amax[0], fabsf(out.x));
        amax[1] = fmaxf(amax[1], fabsf(out.y));
        amax[2] = fmaxf(amax[2], fabsf(out.z));
        amax[3] = fmaxf(amax[3], fabsf(out.w));
        col[(size_t)k * 64u] = out;
    }
    H.cone = (uint32_t)cone;
    H.kc = (uint32_t)kc;
    H.base = (uint32_t)(slot * (size_t)cap * 64u);  // the host keeps the tape below 2^32 float4
    H.mask = __ballot(lane_on);
    // the ten reductions side by side, so their shuffles overlap: the box of
    // the start points as maxima of (-lo, hi), then the gradient bounds
    float r[10];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        r[k] = lane_on ? -start[k] : -INFINITY;
        r[3 + k] = lane_on ? start[k] : -INFINITY;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) r[6 + c] = amax[c];
    wave_max_n<10>(r);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        H.blo[k] = -r[k];
        H.bhi[k] = r[3 + k];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) H.gmax[c] = r[6 + c];
    return true;
}

// The views of one fused backward launch (same image size): each view's
// camera, G-buffer, output gradient and tape (samples, then step counts).
constexpr int kBwdViews = 16;
struct BwdViews {
    uint32_t n;      // views in this launch
    uint32_t slots;  // tape slots per view
    float vp_inv[kBwdViews][16];
    const uint32_t* diffuse[kBwdViews];
    const uint16_t* normal[kBwdViews];
    const float* depth[kBwdViews];
    const float4* grad_out[kBwdViews];
    float4* tape[kBwdViews];
    const uint8_t* counts[kBwdViews];
};

// Splat: resident workgroups whose waves pull sweep headers from a counter
// until none is left (the LDS set-up and the coarse flush are paid once per
// workgroup, not per header). Each step's sample
// gradients are spread over their trilinear footprints into an LDS window
// per level (lane-striped copies when the window is small, so lanes sharing
// a few texels do not serialize), then flushed with one add per touched
// texel, face and channel: into the workgroup's LDS accumulators for the
// coarsest levels (written to global memory once, at the end), into global
// memory with atomics for the others.
// kFused (the backward from tapes, marches of at most kTapeRegs steps): the
// waves pull (view, tape slot) items of up to kBwdViews views instead of
// headers and run each slot's reverse sweep first (sweep_slot), so the
// sweep's arithmetic interleaves with other waves' splats, no header, start
// point or record leaves the kernel, and one launch serves all the views.
template <bool kFused>
__global__ __launch_bounds__(256, 2) void k_dense_bwd_splat(Pyramid P, March M, const SweepHeader* __restrict__ hdr,
                                                         const uint32_t* __restrict__ ctr,
                                                         float4* __restrict__ rec,
                                                         const float4* __restrict__ starts, int width, int height,
                                                         float* __restrict__ grad, uint32_t* __restrict__ next,
                                                         const BwdViews VW, uint32_t cap,
                                                         uint32_t* __restrict__ status)
{
    __shared__ double s_win[4][2 * kPoolUnits];
    __shared__ double s_coarse[kCoarseFloats];
    __shared__ float s_step[kAdjMaxSteps], s_seg[kAdjMaxSteps], s_lvl[kAdjMaxSteps];
    __shared__ float s_corr[kAdjMaxSteps], s_iq[kAdjMaxSteps];
    __shared__ int s_nsteps;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Coarse C = coarse_layout(P);
    int* win = (int*)s_win[wave];
    for (int i = lane; i < 4 * kPoolUnits; i += 64) win[i] = 0;
    for (int i = tid; i < kCoarseFloats; i += 256) s_coarse[i] = 0.0;
    if (tid == 0) {
        const int cnt = step_table(P, M, s_step, s_seg, s_lvl);
        s_nsteps = cnt;
        if (kFused && cnt > 0) step_coefs(P, M, cnt, s_step, s_seg, s_lvl, s_corr, s_iq);
    }
    __syncthreads();
    if (kFused && s_nsteps < 0) {
        if (tid == 0) atomicOr(status, 1u);
        return;
    }
    const uint32_t nh = kFused ? VW.n * VW.slots : ctr[1];
    const unsigned tiles_x = ((unsigned)width + 15u) / 16u;
    for (;;) {
        uint32_t h = 0;
        if (lane == 0) h = atomicAdd(next, 1u);
        h = __builtin_amdgcn_readfirstlane(h);  // lane 0 took it
        if (h >= nh) break;
        SweepHeader H;
        float start[3] = {0.0f, 0.0f, 0.0f};
        const float4* vrec = rec;  // the records: the item's view's tape when fused
        if (kFused) {
            // item h: view v, slot (wave of 8x8 pixels, cone), wave-major
            const uint32_t v = h / VW.slots, slot = h - v * VW.slots;
            vrec = VW.tape[v];
            March Mv = M;
#pragma unroll
            for (int k = 0; k < 16; ++k) Mv.vp_inv[k] = VW.vp_inv[v][k];
            const uint32_t wave_id = slot / (uint32_t)M.ncones, tile = wave_id >> 2, wv = wave_id & 3u;
            const int px = (int)(tile % tiles_x) * 16 + (int)(wv & 1u) * 8 + (lane & 7);
            const int py = (int)(tile / tiles_x) * 16 + (int)(wv >> 1) * 8 + (lane >> 3);
            bool on = px < width && py < height;
            float normal[3] = {0.0f, 0.0f, 1.0f}, albedo[3] = {0.0f, 0.0f, 0.0f};
            float4 go = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (on) {
                const size_t idx = (size_t)py * (size_t)width + (size_t)px;
                const float depth = VW.depth[v][idx];
                on = depth != 1.0f;
                if (on) {
                    pixel_setup(P, Mv, px, py, width, height, depth, VW.normal[v] + 4 * idx, VW.diffuse[v][idx], start,
                                normal, albedo);
                    go = VW.grad_out[v][idx];
                }
            }
            if (!sweep_slot(P, M, VW.tape[v], VW.counts[v], cap, slot, (int)(slot % (uint32_t)M.ncones), s_nsteps,
                            s_corr, s_iq, on, start, normal, albedo, go, lane, status, H))
                continue;
        } else {
            H = hdr[h];
            const int px = H.px0 + (lane & 7), py = H.py0 + (lane >> 3);
            if (px < width && py < height) {
                const float4 st = starts[(size_t)py * (size_t)width + (size_t)px];
                start[0] = st.x;
                start[1] = st.y;
                start[2] = st.z;
            }
        }
        const float* d = cone_dir(M, (int)H.cone);
        const float dir[3] = {d[0], d[1], d[2]};
        const int face[3] = {dir[0] > 0.0f ? 0 : 1, dir[1] > 0.0f ? 2 : 3, dir[2] > 0.0f ? 4 : 5};
        const float w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
        const float lead = P.vs * M.start_offset * 0.5f;
        float cstart[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) cstart[k] = start[k] + dir[k] * P.vs * M.start_offset * 0.5f;
        const float blo[3] = {H.blo[0], H.blo[1], H.blo[2]}, bhi[3] = {H.bhi[0], H.bhi[1], H.bhi[2]};
        // Level m's window lives while the sweep is on m's steps; the two
        // levels in use at a step are consecutive, so m & 1 names its slot.
        // Even levels take the wave's pool from the bottom, odd ones from the
        // top, each up to what the other live window leaves.
        int m0 = -1, m1 = -1;  // the levels held by the even and the odd slot
        Window w0, w1;
        const bool lane_in = ((H.mask >> lane) & 1ull) != 0ull;
        const int kc = (int)H.kc;
        float4 gnext = vrec[H.base + (uint32_t)(kc - 1) * 64u + (uint32_t)lane];
        for (int k = kc - 1; k >= 0; --k) {
            const float4 g4 = gnext;
            if (k > 0) gnext = vrec[H.base + (uint32_t)(k - 1) * 64u + (uint32_t)lane];
            const f4 gs = {g4.x, g4.y, g4.z, g4.w};
            const bool has = g4.x != 0.0f || g4.y != 0.0f || g4.z != 0.0f || g4.w != 0.0f;
            const float stp = s_step[k], lvl = s_lvl[k];
            float pos[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) pos[a] = cstart[a] + dir[a] * stp;
            const int lo = (int)floorf(lvl), hi = (int)ceilf(lvl);
            const float fl = lvl - floorf(lvl);
            const int nlev = hi != lo ? 2 : 1;
            const float sc_lo = hi != lo ? 1.0f - fl : 1.0f, sc_hi = fl;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (j >= nlev) break;
                const int m = j ? hi : lo;
                // steps k1..k that use level m (the level rises with k)
                int k1 = k;
                while (k1 > 0 && (int)ceilf(s_lvl[k1 - 1]) >= m) --k1;
                if (m & 1) {
                    if (m1 == m) continue;
                    if (m1 >= 0 && w1.lds) {
                        wave_lds_sync();
                        flush_window(grad, win, s_coarse, C, P, m1, w1, face, w, lane);
                        wave_lds_sync();
                    }
                    w1 = window_of(P, m, start, lane_in, blo, bhi, dir, lead + s_step[k1], lead + stp, 0,
                                   kPoolUnits - (m0 >= 0 ? w0.used : 0));
                    // the top of the pool, in elements (4 per 16-byte unit)
                    const int shift = 4 * (kPoolUnits - w1.used);
                    w1.part[0].base += shift;
                    w1.part[1].base += shift;
#pragma unroll
                    for (int c = 0; c < 4; ++c) w1.q[c] = fixed_scale(H.gmax[c], k - k1 + 1);
                    m1 = m;
                } else {
                    if (m0 == m) continue;
                    if (m0 >= 0 && w0.lds) {
                        wave_lds_sync();
                        flush_window(grad, win, s_coarse, C, P, m0, w0, face, w, lane);
                        wave_lds_sync();
                    }
                    w0 = window_of(P, m, start, lane_in, blo, bhi, dir, lead + s_step[k1], lead + stp, 0,
                                   kPoolUnits - (m1 >= 0 ? w1.used : 0));
#pragma unroll
                    for (int c = 0; c < 4; ++c) w0.q[c] = fixed_scale(H.gmax[c], k - k1 + 1);
                    m0 = m;
                }
            }
            if (has) {
                if (lo & 1) {
                    if (sc_lo != 0.0f) splat_level(grad, win, s_coarse, C, P, lo, w1, sc_lo, start, pos, gs, face, w, lane);
                    if (nlev == 2 && sc_hi != 0.0f)
                        splat_level(grad, win, s_coarse, C, P, hi, w0, sc_hi, start, pos, gs, face, w, lane);
                } else {
                    if (sc_lo != 0.0f) splat_level(grad, win, s_coarse, C, P, lo, w0, sc_lo, start, pos, gs, face, w, lane);
                    if (nlev == 2 && sc_hi != 0.0f)
                        splat_level(grad, win, s_coarse, C, P, hi, w1, sc_hi, start, pos, gs, face, w, lane);
                }
            }
        }
        wave_lds_sync();
        if (m0 >= 0 && w0.lds) flush_window(grad, win, s_coarse, C, P, m0, w0, face, w, lane);
        if (m1 >= 0 && w1.lds) flush_window(grad, win, s_coarse, C, P, m1, w1, face, w, lane);
        wave_lds_sync();
    }
    // the workgroup's coarse levels to global memory
    __syncthreads();
    for (int m = C.first; m < P.nlevels; ++m) {
        const int n = P.n[m], n3 = n * n * n;
        const unsigned long long D = (unsigned long long)n + 2;
        for (int e = tid; e < 24 * n3; e += 256) {
            const float v = (float)s_coarse[coarse_base(C, m) + e];
            if (v == 0.0f) continue;
            const int t = e % n3, ch = (e / n3) & 3, f = e / (4 * n3);
            const int x = t % n + 1, y = (t / n) % n + 1, z = t / (n * n) + 1;
            const unsigned long long texel =
                P.offset[m] + (unsigned long long)f * P.block[m] + ((unsigned long long)z * D + y) * D + x;
            unsafeAtomicAdd(grad + 4ull * texel + ch, v);
        }
    }
}

__global__ void k_bwd_reset(uint32_t* __restrict__ ctr)
{
    if (threadIdx.x < 4) ctr[threadIdx.x] = 0u;
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
    const int steps = dense::march_step_count(P, M);
    if (steps + 1 > dense::kAdjMaxSteps)
        return fail(VCT_ERR_UNSUPPORTED, "the march takes more than %d steps", dense::kAdjMaxSteps - 1);
    int coarse = 0;
    for (int m = 1; m < P.nlevels; ++m)
        if (P.n[m] <= dense::kCoarseN) coarse += 24 * P.n[m] * P.n[m] * P.n[m];
    if (coarse > dense::kCoarseFloats) return fail(VCT_ERR_UNSUPPORTED, "coarse levels exceed the tile accumulator");
    // Records of one chunk of tiles: every (wave, cone) may take the whole march.
    const uint32_t tiles = ((gb->width + 15) / 16) * ((gb->height + 15) / 16);
    // (two steps of margin over the host's count of the march)
    const size_t per_tile = (size_t)4 * (size_t)M.ncones * (size_t)(steps + 2) * 64 * sizeof(float4);
    const size_t cap = (size_t)4 << 30;
    const uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(tiles, cap / per_tile));
    VCT_TRY(ensure(g->bwd_rec, per_tile * chunk));
    VCT_TRY(ensure(g->bwd_hdr, sizeof(dense::SweepHeader) * 4 * (size_t)M.ncones * chunk));
    VCT_TRY(ensure(g->bwd_start, sizeof(float4) * (size_t)gb->width * gb->height));
    VCT_TRY(ensure(g->bwd_ctr, 16));
    hipStream_t s = (hipStream_t)stream;
    auto* ctr = (uint32_t*)g->bwd_ctr.ptr;
    const uint32_t rec_cap = (uint32_t)(per_tile * chunk / sizeof(float4));
    // the splat's resident grid: two workgroups per CU (its LDS allows two)
    int dev = 0, cus = 0;
    VCT_HIP(hipGetDevice(&dev));
    VCT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t splat_grid = (uint32_t)std::max(1, 2 * cus);
    // the sweep takes two cones per workgroup (measured on 16 views of 512 x 512:
    // 48.2 ms forward + backward against 51.4 with all 16 cones per workgroup)
    const uint32_t cone_groups = (uint32_t)((M.ncones + 1) / 2);
    for (uint32_t t0 = 0; t0 < tiles; t0 += chunk) {
        const uint32_t nt = std::min(chunk, tiles - t0);
        hipLaunchKernelGGL(dense::k_bwd_reset, dim3(1), dim3(64), 0, s, ctr);
        if (flags & VCT_ADJOINT_RECOMPUTE)
            hipLaunchKernelGGL(dense::k_dense_bwd_sweep<dense::kSweepRecompute>, dim3(nt, cone_groups), dim3(256), 0, s,
                               g->data.ptr, P, M, gb->diffuse, gb->normal, gb->depth, (int)gb->width, (int)gb->height,
                               (const float4*)grad_out, t0, rec_cap, (float4*)g->bwd_start.ptr, (float4*)g->bwd_rec.ptr,
                               (dense::SweepHeader*)g->bwd_hdr.ptr, ctr, status, nullptr, 0u);
        else
            hipLaunchKernelGGL(dense::k_dense_bwd_sweep<dense::kSweepKeep>, dim3(nt, cone_groups), dim3(256), 0, s,
                               g->data.ptr, P, M, gb->diffuse, gb->normal, gb->depth, (int)gb->width, (int)gb->height,
                               (const float4*)grad_out, t0, rec_cap, (float4*)g->bwd_start.ptr, (float4*)g->bwd_rec.ptr,
                               (dense::SweepHeader*)g->bwd_hdr.ptr, ctr, status, nullptr, 0u);
        hipLaunchKernelGGL(dense::k_dense_bwd_splat<false>, dim3(splat_grid), dim3(256), 0, s, P, M,
                           (const dense::SweepHeader*)g->bwd_hdr.ptr, (const uint32_t*)ctr, (float4*)g->bwd_rec.ptr,
                           (const float4*)g->bwd_start.ptr, (int)gb->width, (int)gb->height, grad_pyramid, ctr + 2,
                           dense::BwdViews{}, 0u, nullptr);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_dense_trace_backward_tapes(vct_dense* g, vct_stream stream, uint32_t views, const vct_gbuffer* gbs,
                                          const vct_camera_ubo* cams, const vct_dense_trace_desc* d,
                                          const float* const* grad_outs, void* const* tapes, size_t tape_bytes,
                                          float* grad_pyramid, uint32_t* status)
{
    VCT_TRY(need_f32(g));
    if (!gbs || !cams || !d || !grad_outs || !tapes || !grad_pyramid || !status)
        return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    if (views == 0) return VCT_OK;
    const uint32_t W = gbs[0].width, H = gbs[0].height;
    for (uint32_t v = 0; v < views; ++v) {
        const vct_gbuffer& gb = gbs[v];
        if (!gb.diffuse || !gb.normal || !gb.depth || !gb.width || !gb.height)
            return fail(VCT_ERR_INVALID_ARGUMENT, "view %u: G-buffer needs diffuse, normal and depth", v);
        if (gb.width != W || gb.height != H) return fail(VCT_ERR_INVALID_ARGUMENT, "views differ in image size");
        if (!grad_outs[v] || !tapes[v]) return fail(VCT_ERR_INVALID_ARGUMENT, "view %u: null gradient or tape", v);
        if ((uintptr_t)tapes[v] % 16u) return fail(VCT_ERR_INVALID_ARGUMENT, "the tapes must be 16-byte aligned");
    }
    dense::Pyramid P;
    dense::March M;
    VCT_TRY(dense::make_params(g, &cams[0], d, &P, &M));
    const dense::TapeLayout T = dense::tape_layout(P, M, W, H, sizeof(dense::SweepHeader));
    if (T.cap > (uint32_t)dense::kAdjMaxSteps)
        return fail(VCT_ERR_UNSUPPORTED, "the march takes more than %d steps", dense::kAdjMaxSteps - 2);
    if (tape_bytes < T.bytes)
        return fail(VCT_ERR_INVALID_ARGUMENT, "the tape holds %zu bytes; a view's takes %zu", tape_bytes, T.bytes);
    if ((T.sample_bytes / sizeof(float4)) >> 32) return fail(VCT_ERR_UNSUPPORTED, "tape above 2^32 samples");
    int coarse = 0;
    for (int m = 1; m < P.nlevels; ++m)
        if (P.n[m] <= dense::kCoarseN) coarse += 24 * P.n[m] * P.n[m] * P.n[m];
    if (coarse > dense::kCoarseFloats) return fail(VCT_ERR_UNSUPPORTED, "coarse levels exceed the tile accumulator");
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 0;
    VCT_HIP(hipGetDevice(&dev));
    VCT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t splat_grid = (uint32_t)std::max(1, 2 * cus);
    if (T.cap <= (uint32_t)dense::kTapeRegs) {
        // one kernel per kBwdViews views: each slot's reverse sweep, then its splat
        for (uint32_t v0 = 0; v0 < views; v0 += dense::kBwdViews) {
            dense::BwdViews VW{};
            VW.n = std::min<uint32_t>(views - v0, dense::kBwdViews);
            VW.slots = (uint32_t)T.slots;
            if ((uint64_t)VW.n * VW.slots >> 32) return fail(VCT_ERR_UNSUPPORTED, "too many tape slots per launch");
            for (uint32_t k = 0; k < VW.n; ++k) {
                const uint32_t v = v0 + k;
                for (int e = 0; e < 16; ++e) VW.vp_inv[k][e] = cams[v].view_proj_inv[e];
                VW.diffuse[k] = gbs[v].diffuse;
                VW.normal[k] = gbs[v].normal;
                VW.depth[k] = gbs[v].depth;
                VW.grad_out[k] = (const float4*)grad_outs[v];
                VW.tape[k] = (float4*)tapes[v];
                VW.counts[k] = (const uint8_t*)tapes[v] + T.sample_bytes;
            }
            auto* ctr = (uint32_t*)((uint8_t*)tapes[v0] + T.ctr_off);
            hipLaunchKernelGGL(dense::k_bwd_reset, dim3(1), dim3(64), 0, s, ctr);
            hipLaunchKernelGGL(dense::k_dense_bwd_splat<true>, dim3(splat_grid), dim3(256), 0, s, P, M, nullptr, nullptr,
                               nullptr, nullptr, (int)W, (int)H, grad_pyramid, ctr + 2, VW, T.cap, status);
        }
        VCT_HIP(hipGetLastError());
        return VCT_OK;
    }
    // longer marches: per view, the sweep kernel writes headers into the
    // tape's work area, then the splat
    const uint32_t tiles = ((W + 15) / 16) * ((H + 15) / 16);
    const uint32_t cone_groups = (uint32_t)((M.ncones + 1) / 2);
    for (uint32_t v = 0; v < views; ++v) {
        dense::March Mv = M;
        for (int e = 0; e < 16; ++e) Mv.vp_inv[e] = cams[v].view_proj_inv[e];
        auto* base = (uint8_t*)tapes[v];
        auto* hdr = (dense::SweepHeader*)(base + T.header_off);
        auto* starts = (float4*)(base + T.start_off);
        auto* ctr = (uint32_t*)(base + T.ctr_off);
        hipLaunchKernelGGL(dense::k_bwd_reset, dim3(1), dim3(64), 0, s, ctr);
        hipLaunchKernelGGL(dense::k_dense_bwd_sweep<dense::kSweepTape>, dim3(tiles, cone_groups), dim3(256), 0, s,
                           g->data.ptr, P, Mv, gbs[v].diffuse, gbs[v].normal, gbs[v].depth, (int)W, (int)H,
                           (const float4*)grad_outs[v], 0u, 0u, starts, (float4*)base, hdr, ctr, status,
                           (const uint8_t*)base + T.sample_bytes, T.cap);
        hipLaunchKernelGGL(dense::k_dense_bwd_splat<false>, dim3(splat_grid), dim3(256), 0, s, P, Mv,
                           (const dense::SweepHeader*)hdr, (const uint32_t*)ctr, (float4*)base, (const float4*)starts,
                           (int)W, (int)H, grad_pyramid, ctr + 2, dense::BwdViews{}, 0u, nullptr);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_dense_trace_backward_tape(vct_dense* g, vct_stream stream, const vct_gbuffer* gb,
                                         const vct_camera_ubo* cam, const vct_dense_trace_desc* d,
                                         const float* grad_out, void* tape, size_t tape_bytes, float* grad_pyramid,
                                         uint32_t* status)
{
    if (!gb || !cam || !grad_out || !tape) return fail(VCT_ERR_INVALID_ARGUMENT, "null argument");
    return vct_dense_trace_backward_tapes(g, stream, 1, gb, cam, d, &grad_out, &tape, tape_bytes, grad_pyramid,
                                          status);
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
