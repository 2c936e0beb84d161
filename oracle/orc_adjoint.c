/*
 * orc_adjoint.c - FP64 reference for the adjoint of the dense method (no
 * reference counterpart; the reference has no gradients). The forward model
 * is orc_dense_mip followed by orc_dense_trace on an RGBA grid held in double
 * precision; the adjoint returns d(sum(out * grad_out)) / d(level 0) by
 * reverse-mode differentiation of the march (voxelConeTracing.frag:341-392
 * recurrence) and of the anisotropic 2x2x2 downsample
 * (radianceDownSample.comp:95-139). Test infrastructure only (see
 * vct_oracle.h). It is pinned by central finite differences of its own FP64
 * forward in tests/test_oracle_adjoint.py.
 *
 * Storage: as orc_dense.c (level 0 one face, levels >= 1 six faces, each face
 * block (n+2)^3 texels with a zero border, x fastest), 4 doubles per texel.
 * Gradients of border texels are dropped: borders are constants.
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_adj {
    int res, nlevels;
    double extent;
    double grid_min[3];
    double* lv[16];
    double* gr[16];
};

static size_t adj_block(int n) { return (size_t)(n + 2) * (size_t)(n + 2) * (size_t)(n + 2); }

static size_t adj_idx(int n, int face, int x, int y, int z)
{
    size_t d = (size_t)(n + 2);
    return (size_t)face * adj_block(n) + (((size_t)(z + 1) * d) + (size_t)(y + 1)) * d + (size_t)(x + 1);
}

orc_adj* orc_adj_create(int res, float extent, const float grid_min[3])
{
    orc_adj* a = (orc_adj*)calloc(1, sizeof(orc_adj));
    a->res = res;
    int nl = 0;
    while ((1 << nl) < res) ++nl;
    a->nlevels = nl + 1;
    a->extent = extent;
    for (int k = 0; k < 3; ++k) a->grid_min[k] = grid_min[k];
    for (int m = 0; m < a->nlevels; ++m) {
        size_t t = adj_block(res >> m) * (m == 0 ? 1u : 6u) * 4u;
        a->lv[m] = (double*)calloc(t, sizeof(double));
        a->gr[m] = (double*)calloc(t, sizeof(double));
    }
    return a;
}

void orc_adj_free(orc_adj* a)
{
    if (!a) return;
    for (int m = 0; m < a->nlevels; ++m) {
        free(a->lv[m]);
        free(a->gr[m]);
    }
    free(a);
}

/* rgba: res^3 texels x 4 doubles, x fastest. */
void orc_adj_set_level0(orc_adj* a, const double* rgba)
{
    const int n = a->res;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x)
                for (int c = 0; c < 4; ++c)
                    a->lv[0][4u * adj_idx(n, 0, x, y, z) + c] = rgba[4u * (((size_t)z * n + y) * n + x) + c];
}

/* Children of a parent texel: i = x | y << 1 | z << 2; face f pairs the
 * children that differ in the bit of axis f >> 1, front = bit clear for +
 * faces (even f), set for - faces. */
static int pair_child(int f, int p, int back)
{
    const int axis = f >> 1;
    const int c = ((p >> axis) << (axis + 1)) | (p & ((1 << axis) - 1));
    const int set = (f & 1) ? !back : back;
    return set ? c | (1 << axis) : c;
}

static const double* child_texel(const orc_adj* a, int m, int f, int x, int y, int z, int i)
{
    const int n = a->res >> (m - 1);
    return &a->lv[m - 1][4u * adj_idx(n, m - 1 == 0 ? 0 : f, 2 * x + (i & 1), 2 * y + ((i >> 1) & 1),
                                       2 * z + (i >> 2))];
}

void orc_adj_mip(orc_adj* a)
{
    for (int m = 1; m < a->nlevels; ++m) {
        const int n = a->res >> m;
        for (int f = 0; f < 6; ++f)
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x) {
                        double s[4] = {0, 0, 0, 0};
                        for (int p = 0; p < 4; ++p) {
                            const double* A = child_texel(a, m, f, x, y, z, pair_child(f, p, 0));
                            const double* B = child_texel(a, m, f, x, y, z, pair_child(f, p, 1));
                            for (int c = 0; c < 4; ++c) s[c] += A[c] + (1.0 - A[3]) * B[c];
                        }
                        double* o = &a->lv[m][4u * adj_idx(n, f, x, y, z)];
                        for (int c = 0; c < 4; ++c) o[c] = s[c] * 0.25;
                    }
    }
}

/* ---- the march, with its texel footprints (for the scatter) ---- */

typedef struct {
    int valid;
    size_t base;  /* texel index of corner (0,0,0) within face 0 of level m */
    size_t dy, dz, face_stride;
    double fr[3];
} footprint;

static footprint adj_footprint(const orc_adj* a, int m, const double p[3])
{
    footprint F;
    const int n = a->res >> m;
    const double vs = a->extent / (double)n;
    int i0[3];
    F.valid = 1;
    for (int k = 0; k < 3; ++k) {
        double u = p[k] / vs - 0.5, f = floor(u);
        F.fr[k] = u - f;
        if (f < -1.0 || f > (double)(n - 1)) F.valid = 0;
        i0[k] = F.valid ? (int)f : 0;
    }
    const size_t d = (size_t)(n + 2);
    F.base = ((size_t)(i0[2] + 1) * d + (size_t)(i0[1] + 1)) * d + (size_t)(i0[0] + 1);
    F.dy = d;
    F.dz = d * d;
    F.face_stride = adj_block(n);
    return F;
}

static double corner_w(const footprint* F, int c)
{
    return ((c & 1) ? F->fr[0] : 1 - F->fr[0]) * ((c & 2) ? F->fr[1] : 1 - F->fr[1]) *
           ((c & 4) ? F->fr[2] : 1 - F->fr[2]);
}

static size_t corner_idx(const footprint* F, int c)
{
    return F->base + (size_t)(c & 1) + ((c & 2) ? F->dy : 0) + ((c & 4) ? F->dz : 0);
}

/* Level-m sample: one face at level 0 scaled by the weight sum, else the
 * three direction faces weighted; ws[a] is the weight of face fa[a]. */
static void level_terms(int m, const int face[3], const double w[3], int* nf, int fa[3], double ws[3])
{
    if (m == 0) {
        *nf = 1;
        fa[0] = 0;
        ws[0] = w[0] + w[1] + w[2];
    } else {
        *nf = 3;
        for (int k = 0; k < 3; ++k) {
            fa[k] = face[k];
            ws[k] = w[k];
        }
    }
}

static void adj_sample_level(const orc_adj* a, int m, const double p[3], const int face[3], const double w[3],
                             double out[4])
{
    for (int c = 0; c < 4; ++c) out[c] = 0.0;
    footprint F = adj_footprint(a, m, p);
    if (!F.valid) return;
    int nf, fa[3];
    double ws[3];
    level_terms(m, face, w, &nf, fa, ws);
    for (int k = 0; k < nf; ++k)
        for (int c8 = 0; c8 < 8; ++c8) {
            const double* t = &a->lv[m][4u * (F.face_stride * (size_t)fa[k] + corner_idx(&F, c8))];
            const double wt = ws[k] * corner_w(&F, c8);
            for (int c = 0; c < 4; ++c) out[c] += wt * t[c];
        }
}

static void adj_scatter_level(orc_adj* a, int m, const double p[3], const int face[3], const double w[3],
                              double scale, const double g[4])
{
    footprint F = adj_footprint(a, m, p);
    if (!F.valid) return;
    int nf, fa[3];
    double ws[3];
    level_terms(m, face, w, &nf, fa, ws);
    for (int k = 0; k < nf; ++k)
        for (int c8 = 0; c8 < 8; ++c8) {
            double* t = &a->gr[m][4u * (F.face_stride * (size_t)fa[k] + corner_idx(&F, c8))];
            const double wt = scale * ws[k] * corner_w(&F, c8);
            for (int c = 0; c < 4; ++c) t[c] += wt * g[c];
        }
}

typedef struct {
    double coef, vs, max_level;
    int face[3];
    double w[3];
    double start[3];
    const double* dir;
} cone_geo;

static void cone_setup(const orc_adj* a, const orc_dense_params* P, const double start_in[3], const double* dir,
                       cone_geo* G)
{
    G->vs = a->extent / (double)a->res;
    G->coef = 2.0 * tan((double)P->aperture * 0.5);
    G->max_level = (double)(a->nlevels - 1);
    G->dir = dir;
    for (int k = 0; k < 3; ++k) {
        G->start[k] = start_in[k] + dir[k] * G->vs * (double)P->start_offset * 0.5;
        G->w[k] = dir[k] * dir[k];
    }
    G->face[0] = dir[0] > 0 ? 0 : 1;
    G->face[1] = dir[1] > 0 ? 2 : 3;
    G->face[2] = dir[2] > 0 ? 4 : 5;
}

static void cone_sample(const orc_adj* a, const cone_geo* G, double step, double diameter, double s[4], double* lvl_out)
{
    double pos[3];
    for (int k = 0; k < 3; ++k) pos[k] = G->start[k] + G->dir[k] * step;
    double lvl = fmin(fmax(log2(diameter / G->vs), 0.0), G->max_level);
    int lo = (int)floor(lvl), hi = (int)ceil(lvl);
    double fl = lvl - floor(lvl), A[4], B[4];
    adj_sample_level(a, lo, pos, G->face, G->w, A);
    adj_sample_level(a, hi, pos, G->face, G->w, B);
    for (int c = 0; c < 4; ++c) s[c] = A[c] * (1.0 - fl) + B[c] * fl;
    *lvl_out = lvl;
}

#define ADJ_MAX_STEPS 4096

/* One cone forward (out: rgb, 1 - occlusion) and, when gC != NULL, its
 * reverse sweep scattering d/d(pyramid) of gC . out. */
static void cone_run(orc_adj* a, const orc_dense_params* P, const double start_in[3], const double* dir,
                     double out[4], const double* gC)
{
    cone_geo G;
    cone_setup(a, P, start_in, dir, &G);
    static __thread double st_step[ADJ_MAX_STEPS], st_seg[ADJ_MAX_STEPS], st_diam[ADJ_MAX_STEPS],
        st_rw[ADJ_MAX_STEPS], st_occ[ADJ_MAX_STEPS];
    double step = 0.0, diameter = G.vs, occ = 0.0, seg = G.vs;
    double R[4] = {0, 0, 0, 0};
    int N = 0;
    while (step < (double)P->max_dist && occ < 1.0 && N < ADJ_MAX_STEPS) {
        double s[4], lvl;
        cone_sample(a, &G, step, diameter, s, &lvl);
        const double vsl = G.vs * exp2(lvl), c = seg / vsl, q = 1.0 + (step + vsl) * (double)P->decay;
        st_step[N] = step;
        st_seg[N] = seg;
        st_diam[N] = diameter;
        st_rw[N] = R[3];
        st_occ[N] = occ;
        ++N;
        const double op = 1.0 - pow(fmax(1.0 - s[3], 0.0), c);
        const double b = 1.0 - R[3];
        for (int k = 0; k < 3; ++k) R[k] += b * s[k] * c;
        R[3] += b * op;
        occ += (1.0 - occ) * op / q;
        const double prev = step;
        step += fmax(diameter, G.vs) * (double)P->step_factor;
        seg = step - prev;
        diameter = step * G.coef;
    }
    out[0] = R[0];
    out[1] = R[1];
    out[2] = R[2];
    out[3] = 1.0 - occ;
    if (!gC) return;
    double gR[4] = {gC[0], gC[1], gC[2], 0.0}, gO = -gC[3];
    for (int k = N - 1; k >= 0; --k) {
        double s[4], lvl;
        cone_sample(a, &G, st_step[k], st_diam[k], s, &lvl);
        const double vsl = G.vs * exp2(lvl), c = st_seg[k] / vsl;
        const double q = 1.0 + (st_step[k] + vsl) * (double)P->decay;
        const double base = fmax(1.0 - s[3], 0.0);
        const double op = 1.0 - pow(base, c);
        const double b = 1.0 - st_rw[k];
        const double g_op = gR[3] * b + gO * (1.0 - st_occ[k]) / q;
        double gs[4];
        for (int ch = 0; ch < 3; ++ch) gs[ch] = gR[ch] * b * c;
        gs[3] = base > 0.0 ? g_op * c * pow(base, c - 1.0) : 0.0;
        const double dot = gR[0] * s[0] + gR[1] * s[1] + gR[2] * s[2];
        gR[3] = gR[3] * (1.0 - op) - c * dot;
        gO = gO * (1.0 - op / q);
        /* scatter gs into the two levels of the sample */
        double pos[3];
        for (int i = 0; i < 3; ++i) pos[i] = G.start[i] + G.dir[i] * st_step[k];
        const int lo = (int)floor(lvl), hi = (int)ceil(lvl);
        const double fl = lvl - floor(lvl);
        adj_scatter_level(a, lo, pos, G.face, G.w, 1.0 - fl, gs);
        adj_scatter_level(a, hi, pos, G.face, G.w, fl, gs);
    }
}

/* voxelConeTracing.frag:143-203 per pixel, dense form (orc_dense_trace). */
static void pixel_setup(const orc_adj* a, const orc_gbuffer* gb, const orc_dense_params* P, int px, int py,
                        double start[3], double n[3], double albedo[3])
{
    const int W = gb->width, H = gb->height;
    const size_t idx = (size_t)py * W + px;
    const double depth = gb->depth[idx];
    double q[4] = {((px + 0.5) / W) * 2.0 - 1.0, ((py + 0.5) / H) * 2.0 - 1.0, depth, 1.0}, wp[4];
    for (int k = 0; k < 4; ++k)
        wp[k] = (double)P->view_proj_inv[k] * q[0] + (double)P->view_proj_inv[4 + k] * q[1] +
                (double)P->view_proj_inv[8 + k] * q[2] + (double)P->view_proj_inv[12 + k] * q[3];
    for (int k = 0; k < 3; ++k) n[k] = (double)orc_half_to_float(gb->normal[4u * idx + k]) * 2.0 - 1.0;
    const double nl = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const uint32_t dt = gb->diffuse[idx];
    albedo[0] = (dt & 0xff) / 255.0;
    albedo[1] = ((dt >> 8) & 0xff) / 255.0;
    albedo[2] = ((dt >> 16) & 0xff) / 255.0;
    const double vs = a->extent / (double)a->res;
    for (int k = 0; k < 3; ++k) {
        n[k] /= nl;
        start[k] = (wp[k] / wp[3] - a->grid_min[k]) + n[k] * vs * (double)P->start_offset;
    }
}

static void render_pixel(orc_adj* a, const orc_gbuffer* gb, const orc_dense_params* P, int px, int py, double o[4],
                         const double* gout)
{
    const size_t idx = (size_t)py * gb->width + px;
    o[0] = o[1] = o[2] = 0.0;
    o[3] = 1.0;
    if ((double)gb->depth[idx] == 1.0) return;
    double start[3], n[3], albedo[3];
    pixel_setup(a, gb, P, px, py, start, n, albedo);
    const int nc = P->cones32 ? 32 : 16;
    const double inv = 1.0 / (double)nc;
    double ind[4] = {0, 0, 0, 1};
    for (int i = 0; i < nc; ++i) {
        const double* d = P->cones32 ? orc_dense_dir32(i) : orc_dense_dir16(i);
        const double ct = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
        if (ct < 0.0) continue;
        double c[4], gC[4];
        if (gout) {
            for (int k = 0; k < 3; ++k) gC[k] = gout[k] * inv * albedo[k] * (double)P->diffuse_intensity * ct;
            gC[3] = gout[3] * inv * (double)P->ao_factor * ct;
        }
        cone_run(a, P, start, d, c, gout ? gC : NULL);
        for (int k = 0; k < 4; ++k) ind[k] += c[k] * ct;
    }
    o[3] = ind[3] * inv * (double)P->ao_factor;
    for (int k = 0; k < 3; ++k) o[k] = ind[k] * inv * albedo[k] * (double)P->diffuse_intensity;
}

void orc_adj_render(orc_adj* a, const orc_gbuffer* gb, const orc_dense_params* P, double* out)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int py = 0; py < gb->height; ++py)
        for (int px = 0; px < gb->width; ++px) render_pixel(a, gb, P, px, py, out + 4u * ((size_t)py * gb->width + px), NULL);
}

/* grad_level0: res^3 x 4 doubles (x fastest), d(sum(out * grad_out)) / d(level 0). */
void orc_adj_backward(orc_adj* a, const orc_gbuffer* gb, const orc_dense_params* P, const double* grad_out,
                      double* grad_level0)
{
    for (int m = 0; m < a->nlevels; ++m)
        memset(a->gr[m], 0, adj_block(a->res >> m) * (m == 0 ? 1u : 6u) * 4u * sizeof(double));
    for (int py = 0; py < gb->height; ++py)
        for (int px = 0; px < gb->width; ++px) {
            double o[4];
            render_pixel(a, gb, P, px, py, o, grad_out + 4u * ((size_t)py * gb->width + px));
        }
    /* mips, coarsest first: level m's gradient flows to level m - 1 */
    for (int m = a->nlevels - 1; m >= 1; --m) {
        const int n = a->res >> m, nc = a->res >> (m - 1);
        for (int f = 0; f < 6; ++f)
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y)
                    for (int x = 0; x < n; ++x) {
                        const double* g = &a->gr[m][4u * adj_idx(n, f, x, y, z)];
                        for (int p = 0; p < 4; ++p) {
                            const int ia = pair_child(f, p, 0), ib = pair_child(f, p, 1);
                            const int cf = m - 1 == 0 ? 0 : f;
                            const size_t ka = 4u * adj_idx(nc, cf, 2 * x + (ia & 1), 2 * y + ((ia >> 1) & 1), 2 * z + (ia >> 2));
                            const size_t kb = 4u * adj_idx(nc, cf, 2 * x + (ib & 1), 2 * y + ((ib >> 1) & 1), 2 * z + (ib >> 2));
                            const double* A = &a->lv[m - 1][ka];
                            const double* B = &a->lv[m - 1][kb];
                            double* gA = &a->gr[m - 1][ka];
                            double* gB = &a->gr[m - 1][kb];
                            double dot = 0.0;
                            for (int c = 0; c < 4; ++c) dot += g[c] * B[c];
                            for (int c = 0; c < 4; ++c) gA[c] += 0.25 * g[c];
                            gA[3] -= 0.25 * dot;
                            for (int c = 0; c < 4; ++c) gB[c] += 0.25 * (1.0 - A[3]) * g[c];
                        }
                    }
    }
    const int n = a->res;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x)
                for (int c = 0; c < 4; ++c)
                    grad_level0[4u * (((size_t)z * n + y) * n + x) + c] = a->gr[0][4u * adj_idx(n, 0, x, y, z) + c];
}
