/*
 * orc_scene.c - ray-cast G-buffer and shadow-map producers used to make
 * inputs for the trace and injection checks (the reference rasterizes them in
 * GBufferPass / ReflectiveShadowMapPass, which are outside the hot path).
 * Uniform-grid traversal (Amanatides & Woo 1987) with Moller-Trumbore
 * intersection in double precision. Test infrastructure only.
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double lo[3], hi[3], cell[3];
    int n[3];
    uint32_t* start;
    uint32_t* items;
} ugrid;

static void grid_build(ugrid* g, const orc_tris* t)
{
    for (int k = 0; k < 3; ++k) {
        g->lo[k] = DBL_MAX;
        g->hi[k] = -DBL_MAX;
    }
    for (uint32_t i = 0; i < t->count * 3u; ++i)
        for (int k = 0; k < 3; ++k) {
            double v = t->pos[3u * i + k];
            if (v < g->lo[k]) g->lo[k] = v;
            if (v > g->hi[k]) g->hi[k] = v;
        }
    int res = (int)ceil(cbrt((double)t->count) * 2.0);
    if (res < 4) res = 4;
    if (res > 128) res = 128;
    for (int k = 0; k < 3; ++k) {
        double ext = g->hi[k] - g->lo[k];
        if (ext <= 0) ext = 1e-3;
        g->lo[k] -= ext * 1e-4;
        g->hi[k] += ext * 1e-4;
        g->n[k] = res;
        g->cell[k] = (g->hi[k] - g->lo[k]) / res;
    }
    size_t nc = (size_t)res * res * res;
    uint32_t* cnt = (uint32_t*)calloc(nc + 1, sizeof(uint32_t));
    int (*rng)[6] = (int(*)[6])malloc(sizeof(int[6]) * t->count);
    for (uint32_t i = 0; i < t->count; ++i) {
        const float* p = t->pos + 9u * i;
        for (int k = 0; k < 3; ++k) {
            double mn = fmin(p[k], fmin(p[3 + k], p[6 + k])), mx = fmax(p[k], fmax(p[3 + k], p[6 + k]));
            int a = (int)floor((mn - g->lo[k]) / g->cell[k]), b = (int)floor((mx - g->lo[k]) / g->cell[k]);
            if (a < 0) a = 0;
            if (b >= res) b = res - 1;
            rng[i][2 * k] = a;
            rng[i][2 * k + 1] = b;
        }
        for (int z = rng[i][4]; z <= rng[i][5]; ++z)
            for (int y = rng[i][2]; y <= rng[i][3]; ++y)
                for (int x = rng[i][0]; x <= rng[i][1]; ++x) cnt[((size_t)z * res + y) * res + x + 1]++;
    }
    for (size_t c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
    g->start = cnt;
    g->items = (uint32_t*)malloc(sizeof(uint32_t) * (cnt[nc] + 1));
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * nc);
    memcpy(fill, cnt, sizeof(uint32_t) * nc);
    for (uint32_t i = 0; i < t->count; ++i)
        for (int z = rng[i][4]; z <= rng[i][5]; ++z)
            for (int y = rng[i][2]; y <= rng[i][3]; ++y)
                for (int x = rng[i][0]; x <= rng[i][1]; ++x) g->items[fill[((size_t)z * res + y) * res + x]++] = i;
    free(fill);
    free(rng);
}

static void grid_free(ugrid* g)
{
    free(g->start);
    free(g->items);
}

/* Moller-Trumbore; returns t or -1 and the barycentrics of v1, v2. */
static double intersect(const double o[3], const double d[3], const float* p, double* bu, double* bv)
{
    double e1[3], e2[3], pv[3], tv[3], qv[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = (double)p[3 + k] - p[k];
        e2[k] = (double)p[6 + k] - p[k];
    }
    pv[0] = d[1] * e2[2] - d[2] * e2[1];
    pv[1] = d[2] * e2[0] - d[0] * e2[2];
    pv[2] = d[0] * e2[1] - d[1] * e2[0];
    double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (fabs(det) < 1e-30) return -1.0;
    double inv = 1.0 / det;
    for (int k = 0; k < 3; ++k) tv[k] = o[k] - p[k];
    double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (u < 0.0 || u > 1.0) return -1.0;
    qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
    qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
    qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
    double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return -1.0;
    double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    *bu = u;
    *bv = v;
    return t;
}

/* Nearest hit with t in (tmin, tmax]; ties go to the lower triangle index. */
static long grid_cast(const ugrid* g, const orc_tris* t, const double o[3], const double d[3], double tmin,
                      double tmax, double* t_out, double* bu_out, double* bv_out)
{
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        if (fabs(d[k]) < 1e-300) {
            if (o[k] < g->lo[k] || o[k] > g->hi[k]) return -1;
            continue;
        }
        double a = (g->lo[k] - o[k]) / d[k], b = (g->hi[k] - o[k]) / d[k];
        if (a > b) { double tt = a; a = b; b = tt; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
    }
    if (t0 > t1) return -1;
    int c[3], step[3];
    double tnext[3], tdelta[3];
    for (int k = 0; k < 3; ++k) {
        double pk = o[k] + d[k] * t0;
        int ci = (int)floor((pk - g->lo[k]) / g->cell[k]);
        if (ci < 0) ci = 0;
        if (ci >= g->n[k]) ci = g->n[k] - 1;
        c[k] = ci;
        if (d[k] > 0) {
            step[k] = 1;
            tnext[k] = (g->lo[k] + (ci + 1) * g->cell[k] - o[k]) / d[k];
            tdelta[k] = g->cell[k] / d[k];
        } else if (d[k] < 0) {
            step[k] = -1;
            tnext[k] = (g->lo[k] + ci * g->cell[k] - o[k]) / d[k];
            tdelta[k] = -g->cell[k] / d[k];
        } else {
            step[k] = 0;
            tnext[k] = DBL_MAX;
            tdelta[k] = DBL_MAX;
        }
    }
    long best = -1;
    double bt = t1, bu = 0, bv = 0;
    for (;;) {
        size_t cell = ((size_t)c[2] * g->n[1] + c[1]) * g->n[0] + c[0];
        for (uint32_t i = g->start[cell]; i < g->start[cell + 1]; ++i) {
            uint32_t tri = g->items[i];
            double u, v;
            double th = intersect(o, d, t->pos + 9u * tri, &u, &v);
            if (th <= tmin || th > tmax) continue;
            if (th < bt || (th == bt && best >= 0 && (long)tri < best) || (th == bt && best < 0)) {
                bt = th;
                bu = u;
                bv = v;
                best = tri;
            }
        }
        int k = (tnext[0] < tnext[1]) ? (tnext[0] < tnext[2] ? 0 : 2) : (tnext[1] < tnext[2] ? 1 : 2);
        double texit = tnext[k];
        if (best >= 0 && bt <= texit) break;
        if (texit > t1) break;
        c[k] += step[k];
        if (c[k] < 0 || c[k] >= g->n[k]) break;
        tnext[k] += tdelta[k];
    }
    if (best < 0) return -1;
    *t_out = bt;
    *bu_out = bu;
    *bv_out = bv;
    return best;
}

static void mat_vec(const float m[16], const double v[4], double out[4])
{
    for (int k = 0; k < 4; ++k) out[k] = m[k] * v[0] + m[4 + k] * v[1] + m[8 + k] * v[2] + m[12 + k] * v[3];
}

static uint8_t u8(double f)
{
    double c = f > 0.0 ? (f < 1.0 ? f : 1.0) : 0.0;
    return (uint8_t)rint(c * 255.0);
}

void orc_raycast_gbuffer(const orc_tris* tris, const orc_material* mats, const float* rm,
                         const float view_proj_inv[16], const float view_proj[16], int width, int height,
                         uint32_t* diffuse, uint16_t* normal, uint32_t* specular, uint16_t* emission,
                         float* depth)
{
    ugrid g;
    grid_build(&g, tris);
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < height; ++py)
        for (int px = 0; px < width; ++px) {
            size_t idx = (size_t)py * width + px;
            double nx = ((px + 0.5) / width) * 2.0 - 1.0, ny = ((py + 0.5) / height) * 2.0 - 1.0;
            double a[4] = {nx, ny, 0.0, 1.0}, b[4] = {nx, ny, 1.0, 1.0}, pa[4], pb[4];
            mat_vec(view_proj_inv, a, pa);
            mat_vec(view_proj_inv, b, pb);
            double o[3], d[3];
            for (int k = 0; k < 3; ++k) {
                o[k] = pa[k] / pa[3];
                d[k] = pb[k] / pb[3] - o[k];
            }
            double t, u, v;
            long tri = grid_cast(&g, tris, o, d, 0.0, 1.0, &t, &u, &v);
            diffuse[idx] = 0;
            specular[idx] = 0;
            for (int k = 0; k < 4; ++k) {
                normal[4 * idx + k] = 0;
                emission[4 * idx + k] = 0;
            }
            if (tri < 0) {
                depth[idx] = 1.0f;
                continue;
            }
            double h[4] = {o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t, 1.0}, c[4];
            mat_vec(view_proj, h, c);
            depth[idx] = (float)(c[2] / c[3]);
            const float* n = tris->nrm + 9u * (size_t)tri;
            double w0 = 1.0 - u - v, nn[3];
            for (int k = 0; k < 3; ++k) nn[k] = w0 * n[k] + u * n[3 + k] + v * n[6 + k];
            double len = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            for (int k = 0; k < 3; ++k) normal[4 * idx + k] = orc_float_to_half((float)(nn[k] / len * 0.5 + 0.5));
            normal[4 * idx + 3] = orc_float_to_half(1.0f);
            uint32_t m = tris->mat[tri];
            const orc_material* mt = &mats[m];
            const float* r = rm + 5u * m; /* roughness, metallic, specular rgb */
            diffuse[idx] = (uint32_t)u8(mt->base_color[0]) | ((uint32_t)u8(mt->base_color[1]) << 8) |
                           ((uint32_t)u8(mt->base_color[2]) << 16) | ((uint32_t)u8(r[0]) << 24);
            specular[idx] = (uint32_t)u8(r[2]) | ((uint32_t)u8(r[3]) << 8) | ((uint32_t)u8(r[4]) << 16) |
                            ((uint32_t)u8(r[1]) << 24);
            for (int k = 0; k < 3; ++k) emission[4 * idx + k] = orc_float_to_half(mt->emissive[k]);
            emission[4 * idx + 3] = orc_float_to_half(1.0f);
        }
    grid_free(&g);
}

void orc_raycast_shadow(const orc_tris* tris, const float view[16], const float proj[16], int width,
                        int height, float* depth)
{
    /* light clip = proj * view; rays run from NDC z = 0 to z = 1 */
    double vp[16], inv[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += (double)proj[k * 4 + r] * view[c * 4 + k];
            vp[c * 4 + r] = s;
        }
    /* general 4x4 inverse (cofactor expansion) */
    {
        const double* m = vp;
        double o[16];
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
        double det = m[0] * o[0] + m[1] * o[4] + m[2] * o[8] + m[3] * o[12];
        for (int i = 0; i < 16; ++i) inv[i] = o[i] / det;
    }
    ugrid g;
    grid_build(&g, tris);
#pragma omp parallel for schedule(dynamic, 8)
    for (int j = 0; j < height; ++j)
        for (int i = 0; i < width; ++i) {
            double nx = ((i + 0.5) / width) * 2.0 - 1.0, ny = ((j + 0.5) / height) * 2.0 - 1.0;
            double a[4] = {nx, ny, 0.0, 1.0}, b[4] = {nx, ny, 1.0, 1.0}, pa[4], pb[4];
            for (int k = 0; k < 4; ++k) {
                pa[k] = inv[k] * a[0] + inv[4 + k] * a[1] + inv[8 + k] * a[2] + inv[12 + k] * a[3];
                pb[k] = inv[k] * b[0] + inv[4 + k] * b[1] + inv[8 + k] * b[2] + inv[12 + k] * b[3];
            }
            double o[3], d[3];
            for (int k = 0; k < 3; ++k) {
                o[k] = pa[k] / pa[3];
                d[k] = pb[k] / pb[3] - o[k];
            }
            double t, u, v;
            long tri = grid_cast(&g, tris, o, d, 0.0, 1.0, &t, &u, &v);
            depth[(size_t)j * width + i] = tri < 0 ? 1.0f : (float)t;
        }
    grid_free(&g);
}
