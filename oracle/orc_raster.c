/*
 * orc_raster.c - triangle voxelization (K1), shadow visibility (PCF16) and
 * radiance injection (K2). Test infrastructure only (see vct_oracle.h).
 *
 * Coverage rule. The reference rasterizes each triangle on the viewport of
 * its dominant axis (msaaVoxelizer.geom:27-50) with MSAA and sample shading,
 * and the resulting voxel set depends on the driver (sample placement,
 * invocation rate, Voxelizer.cpp:276-280 quadrant mapping). This restatement
 * pins a driver-independent rule with the same structure: one raster cell per
 * voxel of the dominant plane, S sample points per cell (S=1: the centre;
 * S=4: the standard Vulkan 4x locations), coverage by edge functions in
 * double precision with a top-left tie rule, and the fragment position
 * interpolated with the sample's barycentrics. Every covered sample is one
 * fragment-shader invocation.
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>

static const float SX1[1] = {0.5f}, SY1[1] = {0.5f};
static const float SX4[4] = {0.375f, 0.875f, 0.125f, 0.625f};
static const float SY4[4] = {0.125f, 0.375f, 0.625f, 0.875f};

/* msaaVoxelizer.geom:27-32 getDominantAxis */
int orc_dominant_axis(const float* p)
{
    float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
    float e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
    float n[3] = {fabsf(e1[1] * e2[2] - e1[2] * e2[1]),
                  fabsf(e1[2] * e2[0] - e1[0] * e2[2]),
                  fabsf(e1[0] * e2[1] - e1[1] * e2[0])};
    return (n[0] > n[1] && n[0] > n[2]) ? 0 : (n[1] > n[2]) ? 1 : 2;
}

static double edge(double au, double av, double bu, double bv, double qu, double qv)
{
    return (bu - au) * (qv - av) - (bv - av) * (qu - au);
}

static int top_left(double du, double dv)
{
    return dv > 0.0 || (dv == 0.0 && du < 0.0);
}

static int covered(double w, double du, double dv)
{
    return w > 0.0 || (w == 0.0 && top_left(du, dv));
}

void orc_raster_triangle(const float* p, float vs, const int lo[3], const int hi[3], int samples,
                         orc_frag_fn fn, void* user, uint32_t tri)
{
    const int a = orc_dominant_axis(p);
    const int u = (a == 0) ? 1 : 0;
    const int v = (a == 2) ? 1 : 2;
    int idx[3] = {0, 1, 2};
    double pu[3], pv[3];
    for (int k = 0; k < 3; ++k) {
        pu[k] = (double)p[3 * k + u];
        pv[k] = (double)p[3 * k + v];
    }
    double area = edge(pu[0], pv[0], pu[1], pv[1], pu[2], pv[2]);
    if (area == 0.0) return;
    if (area < 0.0) {
        idx[1] = 2;
        idx[2] = 1;
        area = -area;
    }
    const double Au = pu[idx[0]], Av = pv[idx[0]];
    const double Bu = pu[idx[1]], Bv = pv[idx[1]];
    const double Cu = pu[idx[2]], Cv = pv[idx[2]];

    float umin = fminf(p[u], fminf(p[3 + u], p[6 + u]));
    float umax = fmaxf(p[u], fmaxf(p[3 + u], p[6 + u]));
    float vmin = fminf(p[v], fminf(p[3 + v], p[6 + v]));
    float vmax = fmaxf(p[v], fmaxf(p[3 + v], p[6 + v]));
    long cu_lo = (long)floorf(umin / vs) - 1, cu_hi = (long)floorf(umax / vs) + 1;
    long cv_lo = (long)floorf(vmin / vs) - 1, cv_hi = (long)floorf(vmax / vs) + 1;
    if (cu_lo < lo[u]) cu_lo = lo[u];
    if (cu_hi > hi[u]) cu_hi = hi[u];
    if (cv_lo < lo[v]) cv_lo = lo[v];
    if (cv_hi > hi[v]) cv_hi = hi[v];

    const float* sx = samples == 4 ? SX4 : SX1;
    const float* sy = samples == 4 ? SY4 : SY1;
    const int ns = samples == 4 ? 4 : 1;

    for (long cv = cv_lo; cv <= cv_hi; ++cv)
        for (long cu = cu_lo; cu <= cu_hi; ++cu)
            for (int s = 0; s < ns; ++s) {
                float qu = ((float)cu + sx[s]) * vs;
                float qv = ((float)cv + sy[s]) * vs;
                double w0 = edge(Bu, Bv, Cu, Cv, qu, qv);
                double w1 = edge(Cu, Cv, Au, Av, qu, qv);
                double w2 = edge(Au, Av, Bu, Bv, qu, qv);
                if (!covered(w0, Cu - Bu, Cv - Bv)) continue;
                if (!covered(w1, Au - Cu, Av - Cv)) continue;
                if (!covered(w2, Bu - Au, Bv - Av)) continue;
                double lam[3];
                lam[idx[0]] = w0 / area;
                lam[idx[1]] = w1 / area;
                lam[idx[2]] = w2 / area;
                float pos[3];
                pos[u] = qu;
                pos[v] = qv;
                pos[a] = (float)(lam[0] * (double)p[a] + lam[1] * (double)p[3 + a] + lam[2] * (double)p[6 + a]);
                fn(user, tri, pos, lam);
            }
}

static int voxel_cell(const orc_vox_desc* d, float p) { return (int)floorf(p / d->voxel_size); }

/* msaaVoxelizer.frag:43-55 / msaaInjectRadiance.frag:162-174 calculateImageCoords.
 * Parity (ORC_QUIRK_CLAMP_X_RANGE): clamp(worldPos, float(uRegionMinCorner + c),
 * float(uRegionMaxCorner - c)) - float(vec3) is the x component, so every axis
 * uses the x range - then fract(p / extent) * res. Intent: half-open voxel
 * cells, texel = floor(p / voxel_size) mod res with the cell clamped into the
 * region's cells; the texel of a point then does not depend on the region. */
static void image_coords(const orc_vox_desc* d, const float pos[3], uint32_t quirks, int ic[3])
{
    const int res = d->clipmap_resolution;
    if (quirks & ORC_QUIRK_CLAMP_X_RANGE) {
        float c = d->voxel_size * 0.25f;
        for (int k = 0; k < 3; ++k) {
            float lo = d->region_min[0] + c;
            float hi = d->region_max[0] - c;
            float w = fminf(fmaxf(pos[k], lo), hi);
            float q = w / d->clip_max_extent;
            float fr = q - floorf(q);
            int iv = ((int)(fr * (float)res)) % res;
            ic[k] = iv + 1;
        }
    } else {
        for (int k = 0; k < 3; ++k) {
            const int lo = d->region_cell_min[k], hi = d->region_cell_min[k] + d->region_cell_extent[k] - 1;
            int g = voxel_cell(d, pos[k]);
            g = g < lo ? lo : (g > hi ? hi : g);
            ic[k] = ((g % res) + res) % res + 1;
        }
    }
    ic[1] += (res + 2) * d->clip_level;
}

/* msaaVoxelizer.frag:59-60: the reference keeps fragments inside the closed
 * region; intent keeps the region's half-open cells. */
static int in_region(const orc_vox_desc* d, const float pos[3], uint32_t quirks)
{
    for (int k = 0; k < 3; ++k) {
        if (quirks & ORC_QUIRK_CLAMP_X_RANGE) {
            if (pos[k] < d->region_min[k] || pos[k] > d->region_max[k]) return 0;
        } else {
            const int g = voxel_cell(d, pos[k]);
            if (g < d->region_cell_min[k] || g >= d->region_cell_min[k] + d->region_cell_extent[k]) return 0;
        }
    }
    return 1;
}

static void extended_cells(const orc_vox_desc* d, int lo[3], int hi[3])
{
    for (int k = 0; k < 3; ++k) {
        lo[k] = d->region_cell_min[k] - 1;
        hi[k] = d->region_cell_min[k] + d->region_cell_extent[k];
    }
}

/* ---------------- K1 opacity ---------------- */

/* ---------------- scene textures ---------------- */

/* REPEAT addressing of texel coordinate f (an integral float): f mod n in
 * [0, n), with f first clamped to +-2^30 so the integer arithmetic stays in
 * range (only texcoords beyond 2^30 / n texture sizes are affected). */
static uint32_t repeat_texel(float f, uint32_t n)
{
    float c = fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);
    int m = (int)c % (int)n;
    if (m < 0) m += (int)n;
    return (uint32_t)m;
}

/* Vulkan LINEAR filtering of level 0 with REPEAT on both axes: the sampler
 * the reference creates for every glTF texture (GLTFScene.cpp:339,
 * Sampler.cpp:29-50: maxLod 0, so textureLod's lod clamps to level 0).
 * Weights (1-a)(1-b), a(1-b), (1-a)b, ab applied in that order. */
void orc_sample_texture(const orc_texture* t, const float uv[2], float out[4])
{
    const float x = uv[0] * (float)t->width - 0.5f;
    const float y = uv[1] * (float)t->height - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const uint32_t xs[2] = {repeat_texel(fx, t->width), (repeat_texel(fx, t->width) + 1u) % t->width};
    const uint32_t ys[2] = {repeat_texel(fy, t->height), (repeat_texel(fy, t->height) + 1u) % t->height};
    const float w[4] = {(1.0f - a) * (1.0f - b), a * (1.0f - b), (1.0f - a) * b, a * b};
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int q = 0; q < 4; ++q) {
        const uint32_t texel = t->texels[(size_t)ys[q >> 1] * t->width + xs[q & 1]];
        for (int k = 0; k < 4; ++k) {
            const float v = orc_unorm8_to_float((texel >> (8 * k)) & 0xffu);
            acc[k] = q == 0 ? w[0] * v : acc[k] + w[q] * v;
        }
    }
    for (int k = 0; k < 4; ++k) out[k] = acc[k];
}

/* fs_in.texCoord: the vertices' texcoords weighted by the fragment's
 * barycentrics, summed in double and rounded once. */
void orc_fragment_uv(const orc_tris* tris, uint32_t tri, const double lam[3], float uv[2])
{
    const float* t = tris->uv + 6u * tri;
    for (int c = 0; c < 2; ++c)
        uv[c] = (float)(lam[0] * (double)t[c] + lam[1] * (double)t[2 + c] + lam[2] * (double)t[4 + c]);
}

/* msaaVoxelizer.frag:64, msaaInjectRadiance.frag:73, voxelizer.frag:52:
 * discard where the occlusion texture's red channel is below 0.1. */
int orc_has_texture(const orc_tris* tris, int32_t idx)
{
    return tris->textures && idx >= 0 && (uint32_t)idx < tris->texture_count;
}

int orc_occlusion_discard(const orc_tris* tris, const orc_material* m, const float uv[2])
{
    if (!orc_has_texture(tris, m->occlusion_texture)) return 0;
    float s[4];
    orc_sample_texture(&tris->textures[m->occlusion_texture], uv, s);
    return s[0] < 0.1f;
}

typedef struct {
    orc_volume* vol;
    const orc_vox_desc* d;
    uint32_t quirks;
    const orc_tris* tris;
    const orc_material* mats;
} opa_ctx;

static void opacity_frag(void* user, uint32_t tri, const float pos[3], const double lam[3])
{
    opa_ctx* c = (opa_ctx*)user;
    const orc_vox_desc* d = c->d;
    if (!in_region(d, pos, c->quirks)) return;
    if (c->mats && c->tris->textures) {
        float uv[2];
        orc_fragment_uv(c->tris, tri, lam, uv);
        if (orc_occlusion_discard(c->tris, &c->mats[c->tris->mat[tri]], uv)) return;
    }
    int ic[3];
    image_coords(d, pos, c->quirks, ic);
    const int rwb = d->clipmap_resolution + 2;
    /* msaaVoxelizer.frag:69-73: imageStore(vec4(1.0)) into 6 faces */
    for (int f = 0; f < 6; ++f) {
        if (c->vol->fmt == ORC_FMT_RGBA8) {
            *orc_texel_u32(c->vol, ic[0] + rwb * f, ic[1], ic[2]) = 0xffffffffu;
        } else {
            static const float one[4] = {1.0f, 1.0f, 1.0f, 1.0f};
            orc_store(c->vol, ic[0] + rwb * f, ic[1], ic[2], one);
        }
    }
}

void orc_voxelize_opacity_textured(orc_volume* vol, const orc_tris* tris, const orc_material* mats,
                                   const orc_vox_desc* d, int samples, uint32_t quirks)
{
    opa_ctx c = {vol, d, quirks, tris, mats};
    int lo[3], hi[3];
    extended_cells(d, lo, hi);
    for (uint32_t t = 0; t < tris->count; ++t)
        orc_raster_triangle(tris->pos + 9u * t, d->voxel_size, lo, hi, samples, opacity_frag, &c, t);
}

void orc_voxelize_opacity(orc_volume* vol, const orc_tris* tris, const orc_vox_desc* d, int samples,
                          uint32_t quirks)
{
    orc_voxelize_opacity_textured(vol, tris, NULL, d, samples, quirks);
}

/* ---------------- shadow.glsl ---------------- */

static float fetch_depth(const orc_shadow* sm, int i, int j)
{
    if (i < 0 || j < 0 || i >= sm->width || j >= sm->height) return 0.0f; /* CLAMP_TO_BORDER, opaque black */
    return sm->depth[(size_t)j * (size_t)sm->width + (size_t)i];
}

/* Bilinear texture() of a 2-D map (Vulkan LINEAR filtering, u = s*W - 0.5). */
static float bilinear(const orc_shadow* sm, float s, float t, int compare, float ref)
{
    float uu = s * (float)sm->width - 0.5f;
    float vv = t * (float)sm->height - 0.5f;
    float fu = floorf(uu), fv = floorf(vv);
    float a = uu - fu, b = vv - fv;
    fu = fminf(fmaxf(fu, -4.0f), (float)sm->width + 4.0f);
    fv = fminf(fmaxf(fv, -4.0f), (float)sm->height + 4.0f);
    int i0 = (int)fu, j0 = (int)fv;
    float t00 = fetch_depth(sm, i0, j0), t10 = fetch_depth(sm, i0 + 1, j0);
    float t01 = fetch_depth(sm, i0, j0 + 1), t11 = fetch_depth(sm, i0 + 1, j0 + 1);
    if (compare) {
        /* depth-compare filtering: 1 where the receiver is not behind the
         * stored depth; texels outside the map count as lit. */
        int in00 = i0 >= 0 && j0 >= 0 && i0 < sm->width && j0 < sm->height;
        int in10 = i0 + 1 >= 0 && j0 >= 0 && i0 + 1 < sm->width && j0 < sm->height;
        int in01 = i0 >= 0 && j0 + 1 >= 0 && i0 < sm->width && j0 + 1 < sm->height;
        int in11 = i0 + 1 >= 0 && j0 + 1 >= 0 && i0 + 1 < sm->width && j0 + 1 < sm->height;
        t00 = (!in00 || ref <= t00) ? 1.0f : 0.0f;
        t10 = (!in10 || ref <= t10) ? 1.0f : 0.0f;
        t01 = (!in01 || ref <= t01) ? 1.0f : 0.0f;
        t11 = (!in11 || ref <= t11) ? 1.0f : 0.0f;
    }
    float r = ((1.0f - a) * (1.0f - b)) * t00;
    r = r + (a * (1.0f - b)) * t10;
    r = r + ((1.0f - a) * b) * t01;
    r = r + (a * b) * t11;
    return r;
}

/* shadow.glsl:28-36 calcVisibility with calcShadowCoefficientPCF16 (:14-26).
 * With the raw-depth quirk the 16 bilinear taps average the stored depth, as
 * the reference's non-compare sampler does; otherwise each tap is a depth
 * comparison against the receiver's light-space NDC depth. */
float orc_visibility(const orc_shadow* sm, const orc_light* L, const float pos[3], uint32_t quirks)
{
    const float* V = L->view;
    const float* P = L->proj;
    float ls[3];
    for (int k = 0; k < 3; ++k) ls[k] = V[0 + k] * pos[0] + V[4 + k] * pos[1] + V[8 + k] * pos[2] + V[12 + k];
    float ref = 0.0f;
    int compare = !(quirks & ORC_QUIRK_SHADOW_RAW_DEPTH);
    if (compare) {
        /* NDC depth of the receiver: (proj * view * p).z with w = 1 (ortho) */
        ref = P[2] * ls[0] + P[6] * ls[1] + P[10] * ls[2] + P[14];
        ref = ref - 0.001f;
    }
    ls[2] = ls[2] / (L->z_far - L->z_near);
    float xy[2];
    for (int k = 0; k < 2; ++k) xy[k] = P[0 + k] * ls[0] + P[4 + k] * ls[1] + P[12 + k];
    xy[0] = xy[0] * 0.5f + 0.5f;
    xy[1] = xy[1] * 0.5f + 0.5f;
    float scale_x = 1.0f / (float)sm->width;
    float scale_y = 1.0f / (float)sm->height;
    float sum = 0.0f;
    for (float y = -1.5f; y <= 1.5f; y += 1.0f)
        for (float x = -1.5f; x <= 1.5f; x += 1.0f)
            sum = sum + bilinear(sm, xy[0] + x * scale_x, xy[1] + y * scale_y, compare, ref);
    return sum * 0.0625f;
}

/* ---------------- K2 radiance injection ---------------- */

typedef struct {
    orc_volume* vol;
    float* acc; /* RGBA16F volumes: (sum.rgb, count) per atlas texel */
    const orc_tris* tris;
    const orc_material* mats;
    const orc_shadow* sm;
    const orc_light* L;
    const orc_vox_desc* d;
    uint32_t quirks;
    float light_dir[3];
    float box_lo[3], box_hi[3];
} inj_ctx;

static void normalize3(const float in[3], float out[3])
{
    float len = sqrtf(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    out[0] = in[0] / len;
    out[1] = in[1] / len;
    out[2] = in[2] / len;
}


/* imageAtomicRGBA8Avg (msaaInjectRadiance.frag:185-201) for RGBA8 volumes,
 * serialised in fragment order; for RGBA16F volumes an FP32 (sum, count)
 * accumulation in the same order. */
static void sink(inj_ctx* c, int x, int y, int z, const float color[4])
{
    if (c->vol->fmt == ORC_FMT_RGBA8) {
        float v[4] = {color[0] * 255.0f, color[1] * 255.0f, color[2] * 255.0f, color[3]};
        uint32_t* t = orc_texel_u32(c->vol, x, y, z);
        *t = orc_rgba8_avg(*t, v);
    } else {
        /* the accumulator holds the clip level's 6 face blocks only */
        const int rwb = c->vol->res + 2;
        const size_t face = (size_t)(x / rwb), xi = (size_t)(x % rwb);
        const size_t j = (size_t)(y - c->d->clip_level * rwb);
        float* a = c->acc + 4u * ((((face * rwb + (size_t)z) * rwb) + j) * rwb + xi);
        a[0] = a[0] + color[0];
        a[1] = a[1] + color[1];
        a[2] = a[2] + color[2];
        a[3] = a[3] + 1.0f;
    }
}

static void inject_frag(void* user, uint32_t tri, const float pos[3], const double lam[3])
{
    inj_ctx* c = (inj_ctx*)user;
    for (int k = 0; k < 3; ++k)
        if (pos[k] < c->box_lo[k] || pos[k] > c->box_hi[k]) return; /* depth clip of the extended region */
    const orc_material* m = &c->mats[c->tris->mat[tri]];
    float uv[2] = {0.0f, 0.0f};
    if (c->tris->textures) {
        orc_fragment_uv(c->tris, tri, lam, uv);
        if (orc_occlusion_discard(c->tris, m, uv)) return;
    }
    int ic[3];
    image_coords(c->d, pos, c->quirks, ic);
    const int rwb = c->d->clipmap_resolution + 2;
    if (m->emissive[0] > 0.0f || m->emissive[1] > 0.0f || m->emissive[2] > 0.0f) {
        /* msaaInjectRadiance.frag:76-86: emissive factor + emissive texture */
        float e[4], tex[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const int textured = orc_has_texture(c->tris, m->emissive_texture);
        if (textured) orc_sample_texture(&c->tris->textures[m->emissive_texture], uv, tex);
        for (int k = 0; k < 3; ++k) {
            const float v = textured ? m->emissive[k] + tex[k] : m->emissive[k];
            e[k] = fminf(fmaxf(v, 0.0f), 1.0f);
        }
        e[3] = 1.0f;
        for (int f = 0; f < 6; ++f) sink(c, ic[0] + rwb * f, ic[1], ic[2], e);
        return;
    }
    /* msaaInjectRadiance.frag:131-153 */
    const float* n0 = c->tris->nrm + 9u * tri;
    float ni[3];
    for (int k = 0; k < 3; ++k)
        ni[k] = (float)(lam[0] * (double)n0[k] + lam[1] * (double)n0[3 + k] + lam[2] * (double)n0[6 + k]);
    float n[3];
    normalize3(ni, n);
    float ndl = n[0] * c->light_dir[0] + n[1] * c->light_dir[1] + n[2] * c->light_dir[2];
    ndl = fminf(fmaxf(ndl, 0.001f), 1.0f);
    float vis = orc_visibility(c->sm, c->L, pos, c->quirks);
    float lc[3];
    for (int k = 0; k < 3; ++k) lc[k] = ndl * vis * c->L->color[k] * c->L->intensity;
    if (lc[0] == 0.0f && lc[1] == 0.0f && lc[2] == 0.0f) return;
    /* msaaInjectRadiance.frag:131-136: base colour factor x base colour texture */
    float col[4];
    for (int k = 0; k < 4; ++k) col[k] = m->base_color[k];
    if (orc_has_texture(c->tris, m->base_color_texture)) {
        float tex[4];
        orc_sample_texture(&c->tris->textures[m->base_color_texture], uv, tex);
        for (int k = 0; k < 4; ++k) col[k] = col[k] * tex[k];
    }
    float rad[3];
    for (int k = 0; k < 3; ++k) rad[k] = fminf(fmaxf(lc[k] * col[k] * col[3], 0.0f), 1.0f);
    int face[3] = {(-n[0] > 0.0f) ? 0 : 1, (-n[1] > 0.0f) ? 2 : 3, (-n[2] > 0.0f) ? 4 : 5};
    for (int k = 0; k < 3; ++k) {
        float w = fabsf(n[k]);
        float col[4] = {rad[0] * w, rad[1] * w, rad[2] * w, 1.0f};
        sink(c, ic[0] + rwb * face[k], ic[1], ic[2], col);
    }
}

void orc_inject_radiance(orc_volume* vol, const orc_tris* tris, const orc_material* mats,
                         const orc_shadow* sm, const orc_light* L, const orc_vox_desc* d, int samples,
                         uint32_t quirks)
{
    inj_ctx c;
    c.vol = vol;
    c.acc = NULL;
    c.tris = tris;
    c.mats = mats;
    c.sm = sm;
    c.L = L;
    c.d = d;
    c.quirks = quirks;
    float nd[3] = {-L->direction[0], -L->direction[1], -L->direction[2]};
    normalize3(nd, c.light_dir);
    int lo[3], hi[3];
    extended_cells(d, lo, hi);
    for (int k = 0; k < 3; ++k) {
        c.box_lo[k] = (float)lo[k] * d->voxel_size;
        c.box_hi[k] = (float)(hi[k] + 1) * d->voxel_size;
    }
    const size_t rwb = (size_t)vol->res + 2u;
    const size_t n = 6u * rwb * rwb * rwb; /* the level's texels, (face, z, y, x) */
    if (vol->fmt != ORC_FMT_RGBA8) c.acc = (float*)calloc(n * 4u, sizeof(float));
    for (uint32_t t = 0; t < tris->count; ++t)
        orc_raster_triangle(tris->pos + 9u * t, d->voxel_size, lo, hi, samples, inject_frag, &c, t);
    if (c.acc) {
        /* resolve: mean colour, alpha 1 marks an injected texel */
        uint16_t* tex = (uint16_t*)vol->texels;
        const size_t W = 6u * rwb, Hh = rwb * (size_t)vol->levels;
        for (size_t i = 0; i < n; ++i) {
            const float* a = c.acc + 4u * i;
            if (a[3] == 0.0f) continue;
            const size_t xi = i % rwb, j = (i / rwb) % rwb, z = (i / (rwb * rwb)) % rwb, face = i / (rwb * rwb * rwb);
            const size_t y = (size_t)d->clip_level * rwb + j;
            const size_t at = (z * Hh + y) * W + face * rwb + xi;
            tex[4u * at + 0] = orc_float_to_half(a[0] / a[3]);
            tex[4u * at + 1] = orc_float_to_half(a[1] / a[3]);
            tex[4u * at + 2] = orc_float_to_half(a[2] / a[3]);
            tex[4u * at + 3] = orc_float_to_half(1.0f);
        }
        free(c.acc);
    }
}

/* ---------------- test hook: list covered samples ---------------- */

typedef struct {
    float* out; /* 4 floats per sample: pos[3], triangle index */
    uint32_t cap;
    uint32_t n;
} list_ctx;

static void list_frag(void* user, uint32_t tri, const float pos[3], const double lam[3])
{
    (void)lam;
    list_ctx* c = (list_ctx*)user;
    if (c->n < c->cap) {
        float* o = c->out + 4u * c->n;
        o[0] = pos[0];
        o[1] = pos[1];
        o[2] = pos[2];
        o[3] = (float)tri;
    }
    c->n++;
}

/* Runs the coverage rule over `count` triangles and writes every covered
 * sample's interpolated position (and triangle index); returns the count. */
uint32_t orc_raster_list(const float* pos, uint32_t count, float voxel_size, const int lo[3], const int hi[3],
                         int samples, float* out, uint32_t cap)
{
    list_ctx c = {out, cap, 0};
    for (uint32_t t = 0; t < count; ++t) orc_raster_triangle(pos + 9u * t, voxel_size, lo, hi, samples, list_frag, &c, t);
    return c.n;
}
