// vct_raster.hpp - the dominant-axis triangle rasterizer shared by the
// opacity voxelizer (K1), the radiance injector (K2) and the SVO fragment
// voxelizer (K9).
//
// Rule (identical in every kernel): the triangle is projected on the plane of
// its dominant axis (msaaVoxelizer.geom:27-32); the plane is cut into raster
// cells one voxel wide; each cell carries S sample points (S=1: the cell
// centre; S=4: the standard Vulkan 4x positions); a sample is covered when
// all three edge functions, evaluated in double precision, are positive, or
// zero on a top-left edge; the covered sample's position along the dominant
// axis is interpolated with its barycentrics in double precision. Work items
// are enumerated per triangle as (cell row, cell column, sample), so the
// global work index orders fragments exactly as the serial order
// (triangle, row, column, sample).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vct {

__device__ __constant__ static const float c_sx4[4] = {0.375f, 0.875f, 0.125f, 0.625f};
__device__ __constant__ static const float c_sy4[4] = {0.125f, 0.375f, 0.625f, 0.875f};

struct TriSetup {
    double Au, Av, Bu, Bv, Cu, Cv, area;
    int axis, u, v;
    int i0, i1, i2;  // caller-order vertex of the A, B, C corners
    int cu_lo, cv_lo, nu, nv;
};

__device__ __forceinline__ int dominant_axis(const float* p)
{
#pragma clang fp contract(off)
    float e1x = p[3] - p[0], e1y = p[4] - p[1], e1z = p[5] - p[2];
    float e2x = p[6] - p[0], e2y = p[7] - p[1], e2z = p[8] - p[2];
    float nx = fabsf(e1y * e2z - e1z * e2y);
    float ny = fabsf(e1z * e2x - e1x * e2z);
    float nz = fabsf(e1x * e2y - e1y * e2x);
    return (nx > ny && nx > nz) ? 0 : (ny > nz) ? 1 : 2;
}

__device__ __forceinline__ double edge_fn(double au, double av, double bu, double bv, double qu, double qv)
{
#pragma clang fp contract(off)
    return (bu - au) * (qv - av) - (bv - av) * (qu - au);
}

__device__ __forceinline__ bool edge_covered(double w, double du, double dv)
{
    return w > 0.0 || (w == 0.0 && (dv > 0.0 || (dv == 0.0 && du < 0.0)));
}

// Triangle setup against the inclusive cell box [lo, hi]. Returns false for
// degenerate triangles or an empty cell range.
__device__ __forceinline__ bool tri_setup(const float* p, float vs, const int lo[3], const int hi[3], TriSetup& s)
{
#pragma clang fp contract(off)
    s.axis = dominant_axis(p);
    s.u = (s.axis == 0) ? 1 : 0;
    s.v = (s.axis == 2) ? 1 : 2;
    const double pu0 = p[s.u], pu1 = p[3 + s.u], pu2 = p[6 + s.u];
    const double pv0 = p[s.v], pv1 = p[3 + s.v], pv2 = p[6 + s.v];
    double area = edge_fn(pu0, pv0, pu1, pv1, pu2, pv2);
    if (area == 0.0) return false;
    s.i0 = 0;
    if (area < 0.0) {
        s.i1 = 2;
        s.i2 = 1;
        area = -area;
        s.Bu = pu2; s.Bv = pv2; s.Cu = pu1; s.Cv = pv1;
    } else {
        s.i1 = 1;
        s.i2 = 2;
        s.Bu = pu1; s.Bv = pv1; s.Cu = pu2; s.Cv = pv2;
    }
    s.Au = pu0;
    s.Av = pv0;
    s.area = area;
    const float umin = fminf(p[s.u], fminf(p[3 + s.u], p[6 + s.u]));
    const float umax = fmaxf(p[s.u], fmaxf(p[3 + s.u], p[6 + s.u]));
    const float vmin = fminf(p[s.v], fminf(p[3 + s.v], p[6 + s.v]));
    const float vmax = fmaxf(p[s.v], fmaxf(p[3 + s.v], p[6 + s.v]));
    // clamp before the integer conversion so that far-away geometry cannot overflow
    const float lim = 1.0e9f;
    long cu_lo = (long)floorf(fmaxf(umin / vs, -lim)) - 1, cu_hi = (long)floorf(fminf(umax / vs, lim)) + 1;
    long cv_lo = (long)floorf(fmaxf(vmin / vs, -lim)) - 1, cv_hi = (long)floorf(fminf(vmax / vs, lim)) + 1;
    if (cu_lo < lo[s.u]) cu_lo = lo[s.u];
    if (cu_hi > hi[s.u]) cu_hi = hi[s.u];
    if (cv_lo < lo[s.v]) cv_lo = lo[s.v];
    if (cv_hi > hi[s.v]) cv_hi = hi[s.v];
    if (cu_hi < cu_lo || cv_hi < cv_lo) return false;
    s.cu_lo = (int)cu_lo;
    s.cv_lo = (int)cv_lo;
    s.nu = (int)(cu_hi - cu_lo + 1);
    s.nv = (int)(cv_hi - cv_lo + 1);
    return true;
}

// Evaluates local work item `local` = ((row * nu) + col) * S + sample.
// On coverage returns true with the fragment position and barycentrics in
// the caller's vertex order.
__device__ __forceinline__ bool tri_sample(const float* p, const TriSetup& s, float vs, int samples, uint32_t local,
                                           float pos[3], double lam[3])
{
#pragma clang fp contract(off)
    const uint32_t si = samples == 4 ? (local & 3u) : 0u;
    const uint32_t cell = samples == 4 ? (local >> 2) : local;
    const int cu = s.cu_lo + (int)(cell % (uint32_t)s.nu);
    const int cv = s.cv_lo + (int)(cell / (uint32_t)s.nu);
    const float sx = samples == 4 ? c_sx4[si] : 0.5f;
    const float sy = samples == 4 ? c_sy4[si] : 0.5f;
    const float qu = ((float)cu + sx) * vs;
    const float qv = ((float)cv + sy) * vs;
    const double w0 = edge_fn(s.Bu, s.Bv, s.Cu, s.Cv, qu, qv);
    const double w1 = edge_fn(s.Cu, s.Cv, s.Au, s.Av, qu, qv);
    const double w2 = edge_fn(s.Au, s.Av, s.Bu, s.Bv, qu, qv);
    if (!edge_covered(w0, s.Cu - s.Bu, s.Cv - s.Bv)) return false;
    if (!edge_covered(w1, s.Au - s.Cu, s.Av - s.Cv)) return false;
    if (!edge_covered(w2, s.Bu - s.Au, s.Bv - s.Av)) return false;
    double l[3];
    l[s.i0] = w0 / s.area;
    l[s.i1] = w1 / s.area;
    l[s.i2] = w2 / s.area;
    lam[0] = l[0];
    lam[1] = l[1];
    lam[2] = l[2];
    pos[s.u] = qu;
    pos[s.v] = qv;
    pos[s.axis] = (float)(l[0] * (double)p[s.axis] + l[1] * (double)p[3 + s.axis] + l[2] * (double)p[6 + s.axis]);
    return true;
}

// The triangle of work item wi: the last t in [lo, hi] with offsets[t] <= wi,
// where offsets is the exclusive scan of the triangles' work counts
// (nondecreasing) and offsets[lo] <= wi.
__device__ __forceinline__ uint32_t work_triangle(const unsigned long long* __restrict__ offsets, uint32_t lo,
                                                  uint32_t hi, unsigned long long wi)
{
    ++hi;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= wi) lo = mid; else hi = mid;
    }
    return lo;
}

// Bounds of the triangles of consecutive work items [first, last] of a block
// (n >= 1 triangles): the triangles of the two ends, searched with
// block-uniform values, so each lane then searches only between them.
struct TriRange {
    uint32_t lo, hi;
};

__device__ __forceinline__ TriRange block_triangles(const unsigned long long* __restrict__ offsets, uint32_t n,
                                                    unsigned long long first, unsigned long long last)
{
    return {work_triangle(offsets, 0u, n - 1u, first), work_triangle(offsets, 0u, n - 1u, last)};
}

}  // namespace vct
