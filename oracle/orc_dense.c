/*
 * orc_dense.c - the "plain" dense-grid method: the reference's voxelization,
 * anisotropic 2x2x2 downsample and cone march applied to one res^3 grid with
 * a full 6-face mip pyramid down to 1^3 instead of a clipmap. This is a
 * build-defined extension of the reference kernels (the reference only has
 * clipmaps), used for the plain VXAO, batched and adjoint workloads.
 * Test infrastructure only (see vct_oracle.h).
 *
 * Storage: level 0 holds one face (voxelization writes the same value to all
 * six faces, msaaVoxelizer.frag:69-73); level m >= 1 holds six faces; every
 * face block has a one-texel zero border, (n+2)^3 texels, x fastest.
 * Texels are RGBA8 (uint32) or RGBA32F (4 floats).
 *
 * Voxel colour: the first fragment in rasterization order wins; the texel is
 * (base colour, 1).
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static size_t dense_block(int n)
{
    return (size_t)(n + 2) * (size_t)(n + 2) * (size_t)(n + 2);
}

static size_t dense_idx(int n, int face, int x, int y, int z)
{
    size_t d = (size_t)(n + 2);
    return (size_t)face * dense_block(n) + (((size_t)(z + 1) * d) + (size_t)(y + 1)) * d + (size_t)(x + 1);
}

size_t orc_dense_level_texels(int res, int level)
{
    int n = res >> level;
    return dense_block(n) * (level == 0 ? 1u : 6u);
}

static void dload(const orc_dense* g, int level, size_t i, float out[4])
{
    if (g->fmt == ORC_FMT_RGBA8) {
        orc_rgba8_to_float4(((const uint32_t*)g->levels[level])[i], out);
    } else {
        const float* t = (const float*)g->levels[level] + 4u * i;
        for (int k = 0; k < 4; ++k) out[k] = t[k];
    }
}

static void dstore(orc_dense* g, int level, size_t i, const float v[4])
{
    if (g->fmt == ORC_FMT_RGBA8) {
        ((uint32_t*)g->levels[level])[i] = orc_float_to_unorm8_rgba(v);
    } else {
        float* t = (float*)g->levels[level] + 4u * i;
        for (int k = 0; k < 4; ++k) t[k] = v[k];
    }
}

/* ---------------- voxelization ---------------- */

typedef struct {
    orc_dense* g;
    const orc_tris* tris;
    const orc_material* mats;
    float vs;
    uint32_t* winner; /* per voxel: triangle index + 1 of the first fragment (0 = empty) */
} dvox_ctx;

static void dense_frag(void* user, uint32_t tri, const float pos[3], const double lam[3])
{
    (void)lam;
    dvox_ctx* c = (dvox_ctx*)user;
    const int n = c->g->res;
    int ic[3];
    for (int k = 0; k < 3; ++k) {
        float f = floorf(pos[k] / c->vs);
        if (!(f >= 0.0f) || f >= (float)n) return;
        ic[k] = (int)f;
    }
    size_t v = ((size_t)ic[2] * n + ic[1]) * n + ic[0];
    if (c->winner[v] == 0) c->winner[v] = tri + 1u;
}

/* World-space triangles are moved into the grid frame by one float
 * subtraction of grid_min per vertex coordinate, then rasterized with the
 * voxel size extent / res. */
void orc_dense_voxelize(orc_dense* g, const orc_tris* tris, const orc_material* mats, int samples)
{
    const int n = g->res;
    const float vs = g->extent / (float)n;
    size_t nv = (size_t)n * n * n;
    dvox_ctx c = {g, tris, mats, vs, (uint32_t*)calloc(nv, 4)};
    int lo[3] = {0, 0, 0}, hi[3] = {n - 1, n - 1, n - 1};
    for (uint32_t t = 0; t < tris->count; ++t) {
        float p[9];
        for (int k = 0; k < 9; ++k) p[k] = tris->pos[9u * t + k] - g->grid_min[k % 3];
        orc_raster_triangle(p, vs, lo, hi, samples, dense_frag, &c, t);
    }
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                size_t v = ((size_t)z * n + y) * n + x;
                if (!c.winner[v]) continue;
                const float* bc = mats[tris->mat[c.winner[v] - 1u]].base_color;
                float col[4] = {bc[0], bc[1], bc[2], 1.0f};
                dstore(g, 0, dense_idx(n, 0, x, y, z), col);
            }
    free(c.winner);
}

/* ---------------- anisotropic mip pyramid ---------------- */

static const int DPAIRS[6][4][2] = {
    {{0, 1}, {2, 3}, {4, 5}, {6, 7}}, {{1, 0}, {3, 2}, {5, 4}, {7, 6}},
    {{0, 2}, {1, 3}, {4, 6}, {5, 7}}, {{2, 0}, {3, 1}, {6, 4}, {7, 5}},
    {{0, 4}, {1, 5}, {2, 6}, {3, 7}}, {{4, 0}, {5, 1}, {6, 2}, {7, 3}},
};

/* radianceDownSample.comp:95-139 applied level by level over the whole grid
 * (no clipmap transition band): per face, 4 front-to-back pairs averaged. */
void orc_dense_mip(orc_dense* g)
{
    for (int m = 1; m < g->nlevels; ++m) {
        const int n = g->res >> m, np = g->res >> (m - 1);
        for (int f = 0; f < 6; ++f)
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x) {
                        float v[8][4];
                        for (int i = 0; i < 8; ++i)
                            dload(g, m - 1,
                                  dense_idx(np, m - 1 == 0 ? 0 : f, 2 * x + (i & 1), 2 * y + ((i >> 1) & 1),
                                            2 * z + (i >> 2)),
                                  v[i]);
                        float out[4];
                        for (int ch = 0; ch < 4; ++ch) {
                            float s = 0.0f;
                            for (int p = 0; p < 4; ++p) {
                                const float* a = v[DPAIRS[f][p][0]];
                                const float* b = v[DPAIRS[f][p][1]];
                                s = s + a[ch];
                                s = s + (1.0f - a[3]) * b[ch];
                            }
                            out[ch] = s * 0.25f;
                        }
                        dstore(g, m, dense_idx(n, f, x, y, z), out);
                    }
    }
}

/* ---------------- cone trace ---------------- */

/* One trilinear fetch of level m at grid coordinate p (translated), with
 * zero outside the grid (CLAMP_TO_BORDER, black). */
static void dense_tex(const orc_dense* g, int m, int face, const double p[3], double out[4])
{
    const int n = g->res >> m;
    const double vs = (double)g->extent / (double)n;
    double u[3], fr[3];
    int i0[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = p[k] / vs - 0.5;
        double f = floor(u[k]);
        fr[k] = u[k] - f;
        if (f < -1.0 || f > (double)(n - 1)) {
            for (int c = 0; c < 4; ++c) out[c] = 0.0;
            return;
        }
        i0[k] = (int)f;
    }
    for (int c = 0; c < 4; ++c) out[c] = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                double w = (dx ? fr[0] : 1 - fr[0]) * (dy ? fr[1] : 1 - fr[1]) * (dz ? fr[2] : 1 - fr[2]);
                float t[4];
                dload(g, m, dense_idx(n, face, i0[0] + dx, i0[1] + dy, i0[2] + dz), t);
                for (int c = 0; c < 4; ++c) out[c] += w * (double)t[c];
            }
}

static void dense_sample(const orc_dense* g, int m, const double p[3], const int face[3], const double w[3],
                         double out[4])
{
    for (int c = 0; c < 4; ++c) out[c] = 0.0;
    if (m == 0) {
        double t[4];
        dense_tex(g, 0, 0, p, t);
        for (int c = 0; c < 4; ++c) out[c] = t[c] * (w[0] + w[1] + w[2]);
        return;
    }
    for (int a = 0; a < 3; ++a) {
        double t[4];
        dense_tex(g, m, face[a], p, t);
        for (int c = 0; c < 4; ++c) out[c] += t[c] * w[a];
    }
}

/* traceCone (voxelConeTracing.frag:341-392) with the dense level rule
 * level = clamp(log2(diameter / voxelSize), 0, levels - 1); p is relative to
 * the grid minimum. Computed in double precision. */
static void dense_cone(const orc_dense* g, const orc_dense_params* P, const double start_in[3], const double dir[3],
                       double out[4], uint64_t* steps)
{
    const double vs = (double)g->extent / (double)g->res;
    const double coef = 2.0 * tan((double)P->aperture * 0.5);
    double start[3];
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * vs * (double)P->start_offset * 0.5;
    double step = 0.0, diameter = vs, occlusion = 0.0, seg = vs;
    double result[4] = {0, 0, 0, 0};
    const int face[3] = {dir[0] > 0 ? 0 : 1, dir[1] > 0 ? 2 : 3, dir[2] > 0 ? 4 : 5};
    const double w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
    const double max_level = (double)(g->nlevels - 1);
    while (step < (double)P->max_dist && occlusion < 1.0) {
        double pos[3];
        for (int k = 0; k < 3; ++k) pos[k] = start[k] + dir[k] * step;
        double lvl = fmin(fmax(log2(diameter / vs), 0.0), max_level);
        int lo = (int)floor(lvl), hi = (int)ceil(lvl);
        double a[4], b[4], s[4];
        dense_sample(g, lo, pos, face, w, a);
        dense_sample(g, hi, pos, face, w, b);
        double fl = lvl - floor(lvl);
        for (int c = 0; c < 4; ++c) s[c] = a[c] * (1.0 - fl) + b[c] * fl;
        double vsl = vs * exp2(lvl);
        double corr = seg / vsl;
        double op = fmin(fmax(1.0 - pow(fmax(1.0 - s[3], 0.0), corr), 0.0), 1.0);
        double blend = fmin(fmax(1.0 - result[3], 0.0), 1.0);
        for (int c = 0; c < 3; ++c) result[c] += blend * (s[c] * corr);
        result[3] += blend * op;
        occlusion += (1.0 - occlusion) * op / (1.0 + (step + vsl) * (double)P->decay);
        double prev = step;
        step += fmax(diameter, vs) * (double)P->step_factor;
        seg = step - prev;
        diameter = step * coef;
        if (steps) ++*steps;
    }
    out[0] = result[0];
    out[1] = result[1];
    out[2] = result[2];
    out[3] = 1.0 - occlusion;
}

/* voxelConeTracing.frag:78-135 cone tables */
static const double ORC_DIRS16_F64[16][3] = {
    {0.57735, 0.57735, 0.57735},       {0.57735, -0.57735, -0.57735},     {-0.57735, 0.57735, -0.57735},
    {-0.57735, -0.57735, 0.57735},     {-0.903007, -0.182696, -0.388844}, {-0.903007, 0.182696, 0.388844},
    {0.903007, -0.182696, 0.388844},   {0.903007, 0.182696, -0.388844},   {-0.388844, -0.903007, -0.182696},
    {0.388844, -0.903007, 0.182696},   {0.388844, 0.903007, -0.182696},   {-0.388844, 0.903007, 0.182696},
    {-0.182696, -0.388844, -0.903007}, {0.182696, 0.388844, -0.903007},   {-0.182696, 0.388844, 0.903007},
    {0.182696, -0.388844, 0.903007}};
static const double ORC_DIRS32_F64[32][3] = {
    {0.898904, 0.435512, 0.0479745},   {0.898904, -0.435512, -0.0479745}, {0.898904, 0.0479745, -0.435512},
    {0.898904, -0.0479745, 0.435512},  {-0.898904, 0.435512, -0.0479745}, {-0.898904, -0.435512, 0.0479745},
    {-0.898904, 0.0479745, 0.435512},  {-0.898904, -0.0479745, -0.435512}, {0.0479745, 0.898904, 0.435512},
    {-0.0479745, 0.898904, -0.435512}, {-0.435512, 0.898904, 0.0479745},  {0.435512, 0.898904, -0.0479745},
    {-0.0479745, -0.898904, 0.435512}, {0.0479745, -0.898904, -0.435512}, {0.435512, -0.898904, 0.0479745},
    {-0.435512, -0.898904, -0.0479745}, {0.435512, 0.0479745, 0.898904},  {-0.435512, -0.0479745, 0.898904},
    {0.0479745, -0.435512, 0.898904},  {-0.0479745, 0.435512, 0.898904},  {0.435512, -0.0479745, -0.898904},
    {-0.435512, 0.0479745, -0.898904}, {0.0479745, 0.435512, -0.898904},  {-0.0479745, -0.435512, -0.898904},
    {0.57735, 0.57735, 0.57735},       {0.57735, 0.57735, -0.57735},      {0.57735, -0.57735, 0.57735},
    {0.57735, -0.57735, -0.57735},     {-0.57735, 0.57735, 0.57735},      {-0.57735, 0.57735, -0.57735},
    {-0.57735, -0.57735, 0.57735},     {-0.57735, -0.57735, -0.57735}};

const double* orc_dense_dir16(int i) { return ORC_DIRS16_F64[i]; }
const double* orc_dense_dir32(int i) { return ORC_DIRS32_F64[i]; }

/* Per pixel: indirect diffuse rgb (x albedo x intensity) and ambient
 * occlusion in alpha, as voxelConeTracing.frag:166-203 with the dense march.
 * out: 4 doubles per pixel. */
void orc_dense_trace(const orc_dense* g, const orc_gbuffer* gb, const orc_dense_params* P, int row_begin,
                     int row_end, double* out, uint64_t* samples)
{
    const int W = gb->width, H = gb->height;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int py = row_begin; py < row_end; ++py)
        for (int px = 0; px < W; ++px) {
            size_t idx = (size_t)py * W + px;
            double* o = out + 4u * idx;
            o[0] = o[1] = o[2] = 0.0;
            o[3] = 1.0;
            double depth = gb->depth[idx];
            if (depth == 1.0) continue;
            double q[4] = {((px + 0.5) / W) * 2.0 - 1.0, ((py + 0.5) / H) * 2.0 - 1.0, depth, 1.0}, wp[4];
            for (int k = 0; k < 4; ++k)
                wp[k] = (double)P->view_proj_inv[k] * q[0] + (double)P->view_proj_inv[4 + k] * q[1] +
                        (double)P->view_proj_inv[8 + k] * q[2] + (double)P->view_proj_inv[12 + k] * q[3];
            double world[3] = {wp[0] / wp[3] - g->grid_min[0], wp[1] / wp[3] - g->grid_min[1],
                               wp[2] / wp[3] - g->grid_min[2]};
            double n[3];
            for (int k = 0; k < 3; ++k) n[k] = (double)orc_half_to_float(gb->normal[4u * idx + k]) * 2.0 - 1.0;
            double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int k = 0; k < 3; ++k) n[k] /= nl;
            uint32_t dt = gb->diffuse[idx];
            double albedo[3] = {(dt & 0xff) / 255.0, ((dt >> 8) & 0xff) / 255.0, ((dt >> 16) & 0xff) / 255.0};
            const double vs = (double)g->extent / (double)g->res;
            double start[3];
            for (int k = 0; k < 3; ++k) start[k] = world[k] + n[k] * vs * (double)P->start_offset;
            double ind[4] = {0, 0, 0, 1};
            const int nc = P->cones32 ? 32 : 16;
            for (int i = 0; i < nc; ++i) {
                const double* d = P->cones32 ? ORC_DIRS32_F64[i] : ORC_DIRS16_F64[i];
                double ct = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
                if (ct < 0.0) continue;
                double c[4];
                uint64_t s = 0;
                dense_cone(g, P, start, d, c, &s);
                total += s;
                for (int k = 0; k < 4; ++k) ind[k] += c[k] * ct;
            }
            for (int k = 0; k < 4; ++k) ind[k] /= (double)nc;
            o[3] = ind[3] * (double)P->ao_factor;
            for (int k = 0; k < 3; ++k) o[k] = ind[k] * albedo[k] * (double)P->diffuse_intensity;
        }
    if (samples) *samples = total;
}
