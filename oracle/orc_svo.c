/*
 * orc_svo.c - sparse voxel octree path: fragment-list voxelization
 * (voxelizer.vert/geom/frag), the per-level node init/flag/alloc/modifyArg
 * loop (octreeNode*.comp driven by OctreeBuilder.cpp:213-305) and the leaf /
 * mipmap colour writes (octreeNodeLeafWrite.comp, octreeNodeMipmapWrite.comp).
 * Test infrastructure only (see vct_oracle.h).
 *
 * Serialisation: fragments take list slots in rasterization order
 * (triangle, cell row, cell column, sample); the node allocator hands out
 * child blocks to flagged nodes in increasing pool index, which is the
 * outcome of the reference's atomicAdd when invocations run in index order.
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const orc_tris* tris;
    const orc_material* mats;
    const orc_shadow* sm;
    const orc_light* L;
    const float* npos; /* normalised positions, 9 floats per triangle */
    int res;
    uint32_t quirks;
    uint32_t* frags;
    uint32_t cap;
    uint32_t counter;
    float light_dir[3];
} svo_ctx;

/* voxelizer.frag:99-100: x | y << 12 | (z & 0xff) << 24 and
 * (z >> 8) << 28 | (RGBA8 & 0x0fffffff) */
static void pack_fragment(const uint32_t vp[3], uint32_t rgba8, uint32_t out[2])
{
    out[0] = vp[0] | (vp[1] << 12u) | ((vp[2] & 0xffu) << 24u);
    out[1] = ((vp[2] >> 8u) << 28u) | (rgba8 & 0x0fffffffu);
}

static void svo_frag(void* user, uint32_t tri, const float pos[3], const double lam[3])
{
    svo_ctx* c = (svo_ctx*)user;
    /* fs_in.position = biasAndScale(p) (voxelizer.geom:27-30, 51) */
    float q[3];
    for (int k = 0; k < 3; ++k) q[k] = (pos[k] + 1.0f) * 0.5f;
    const orc_material* m = &c->mats[c->tris->mat[tri]];
    const int textured = c->tris->textures != NULL;
    float uv[2] = {0.0f, 0.0f};
    if (textured) {
        orc_fragment_uv(c->tris, tri, lam, uv);
        if (orc_occlusion_discard(c->tris, m, uv)) return; /* voxelizer.frag:52-53, before the slot */
    }
    uint32_t cur = c->counter++; /* atomicAdd before the light discard (voxelizer.frag:55) */
    if (cur >= c->cap) return;
    float rad[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (m->emissive[0] > 0.0f || m->emissive[1] > 0.0f || m->emissive[2] > 0.0f) {
        /* voxelizer.frag:60-69 */
        float tex[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const int has = orc_has_texture(c->tris, m->emissive_texture);
        if (has) orc_sample_texture(&c->tris->textures[m->emissive_texture], uv, tex);
        for (int k = 0; k < 3; ++k) rad[k] = fminf(fmaxf(has ? m->emissive[k] + tex[k] : m->emissive[k], 0.0f), 1.0f);
    } else {
        /* voxelizer.frag:72-77: textureLod(base colour) * factor; without a
         * base colour texture the colour stays (0, 0, 0, 1) under the quirk */
        float color[4] = {0.0f, 0.0f, 0.0f, 1.0f};
        const int has = orc_has_texture(c->tris, m->base_color_texture);
        if (has) {
            float tex[4];
            orc_sample_texture(&c->tris->textures[m->base_color_texture], uv, tex);
            for (int k = 0; k < 4; ++k) color[k] = tex[k] * m->base_color[k];
        } else if (!(c->quirks & ORC_QUIRK_SVO_UNTEXTURED_BLACK)) {
            for (int k = 0; k < 4; ++k) color[k] = m->base_color[k];
        }
        const float* n0 = c->tris->nrm + 9u * tri;
        float ni[3];
        for (int k = 0; k < 3; ++k)
            ni[k] = (float)(lam[0] * (double)n0[k] + lam[1] * (double)n0[3 + k] + lam[2] * (double)n0[6 + k]);
        float len = sqrtf(ni[0] * ni[0] + ni[1] * ni[1] + ni[2] * ni[2]);
        float n[3] = {ni[0] / len, ni[1] / len, ni[2] / len};
        float ndl = n[0] * c->light_dir[0] + n[1] * c->light_dir[1] + n[2] * c->light_dir[2];
        ndl = fminf(fmaxf(ndl, 0.001f), 1.0f);
        /* calcVisibility at fs_in.position, the normalised [0,1] position (voxelizer.frag:86) */
        float vis = orc_visibility(c->sm, c->L, q, c->quirks);
        float lc[3];
        for (int k = 0; k < 3; ++k) lc[k] = ndl * vis * c->L->color[k] * c->L->intensity;
        if (lc[0] == 0.0f && lc[1] == 0.0f && lc[2] == 0.0f) return; /* discard: the slot stays zero */
        for (int k = 0; k < 3; ++k) rad[k] = fminf(fmaxf(lc[k] * color[k] * color[3], 0.0f), 1.0f);
        rad[3] = 1.0f;
    }
    uint32_t vp[3];
    for (int k = 0; k < 3; ++k) {
        float f = q[k] * (float)c->res;
        uint32_t u = f > 0.0f ? (uint32_t)f : 0u;
        if (u > (uint32_t)c->res) u = (uint32_t)c->res;
        vp[k] = u;
    }
    float v255[4] = {rad[0] * 255.0f, rad[1] * 255.0f, rad[2] * 255.0f, rad[3] * 255.0f};
    pack_fragment(vp, orc_pack_rgba8(v255), c->frags + 2u * cur);
}

uint32_t orc_svo_fragments(const orc_tris* tris, const orc_material* mats, const orc_shadow* sm,
                           const orc_light* light, const float bb_min[3], const float bb_max[3], int resolution,
                           int samples, uint32_t quirks, uint32_t* frags, uint32_t cap)
{
    /* voxelizer.vert:42-47: normalise by the scene box's largest half extent */
    float ext[3], center[3];
    for (int k = 0; k < 3; ++k) {
        ext[k] = bb_max[k] - bb_min[k];
        center[k] = (bb_min[k] + bb_max[k]) * 0.5f;
    }
    float ev = fmaxf(ext[0], fmaxf(ext[1], ext[2])) * 0.5f;
    float* np = (float*)malloc(sizeof(float) * 9u * (tris->count ? tris->count : 1u));
    for (uint32_t i = 0; i < 9u * tris->count; ++i) np[i] = (tris->pos[i] - center[i % 3u]) / ev;
    orc_tris nt = *tris;
    nt.pos = np;
    svo_ctx c;
    c.tris = &nt;
    c.mats = mats;
    c.sm = sm;
    c.L = light;
    c.npos = np;
    c.res = resolution;
    c.quirks = quirks;
    c.frags = frags;
    c.cap = cap;
    c.counter = 0;
    float nd[3] = {-light->direction[0], -light->direction[1], -light->direction[2]};
    float len = sqrtf(nd[0] * nd[0] + nd[1] * nd[1] + nd[2] * nd[2]);
    for (int k = 0; k < 3; ++k) c.light_dir[k] = nd[k] / len;
    if (frags && cap) memset(frags, 0, sizeof(uint32_t) * 2u * cap);
    /* viewport R x R over NDC [-1,1]: raster cells of size 2/R in normalised space */
    float vs = 2.0f / (float)resolution;
    int lo[3] = {-resolution / 2, -resolution / 2, -resolution / 2};
    int hi[3] = {resolution / 2 - 1, resolution / 2 - 1, resolution / 2 - 1};
    for (uint32_t t = 0; t < tris->count; ++t) orc_raster_triangle(np + 9u * t, vs, lo, hi, samples, svo_frag, &c, t);
    free(np);
    return c.counter;
}

static void frag_pos(const uint32_t* f, uint32_t quirks, uint32_t p[3])
{
    p[0] = f[0] & 0xfffu;
    p[1] = (f[0] >> 12u) & 0xfffu;
    p[2] = ((f[0] >> 24u) & 0xffu) | ((f[1] >> 20u) & 0xf00u);
    if (quirks & ORC_QUIRK_SVO_HALF_COORDS)
        for (int k = 0; k < 3; ++k) p[k] >>= 1;
}

/* The shared descent of octreeNodeFlag/LeafWrite/MipmapWrite: returns idx,
 * and the resolution at which it stopped. */
static uint32_t walk(const uint32_t* pool, uint32_t p[3], uint32_t res, uint32_t stop, uint32_t* res_out)
{
    uint32_t idx = 0, cur = 0;
    do {
        res >>= 1;
        uint32_t cx = p[0] >= res, cy = p[1] >= res, cz = p[2] >= res;
        idx = cur + (cz | (cx << 1) | (cy << 2));
        cur = pool[2u * idx] & 0x7fffffffu;
        p[0] -= cx * res;
        p[1] -= cy * res;
        p[2] -= cz * res;
    } while (cur != 0u && res > stop);
    *res_out = res;
    return idx;
}

int orc_svo_build(const uint32_t* frags, uint32_t nfrag, int level, uint32_t quirks, uint32_t* pool,
                  uint32_t pool_cap, uint32_t* node_count)
{
    const uint32_t R = 1u << level;
    uint32_t alloc_begin = 0, alloc_num = 8, counter = 0;
    if (pool_cap < 8) return -1;
    for (int i = 1; i <= level; ++i) {
        /* octreeNodeInit.comp */
        if ((uint64_t)alloc_begin + alloc_num > pool_cap) return -1;
        memset(pool + 2u * alloc_begin, 0, sizeof(uint32_t) * 2u * alloc_num);
        /* octreeNodeFlag.comp */
        for (uint32_t f = 0; f < nfrag; ++f) {
            uint32_t p[3], r;
            frag_pos(frags + 2u * f, quirks, p);
            uint32_t idx = walk(pool, p, R, 1u, &r);
            pool[2u * idx] |= 0x80000000u;
        }
        if (i != level) {
            /* octreeNodeAlloc.comp, serialised in pool-index order */
            for (uint32_t n = alloc_begin; n < alloc_begin + alloc_num; ++n) {
                if (pool[2u * n] & 0x80000000u) {
                    uint32_t cv = counter++;
                    pool[2u * n] = ((cv + 1u) << 3u) | 0x80000000u;
                    pool[2u * n + 1] = 0u;
                }
            }
            /* octreeNodeModifyArg.comp */
            alloc_begin += alloc_num;
            alloc_num = (counter << 3) + 8u - alloc_begin;
        }
    }
    uint32_t total = (counter << 3) + 8u;
    /* octreeNodeLeafWrite.comp: CAS running average into the leaf, in fragment order */
    for (uint32_t f = 0; f < nfrag; ++f) {
        uint32_t p[3], r;
        frag_pos(frags + 2u * f, quirks, p);
        uint32_t idx = walk(pool, p, R, 1u, &r);
        if (r == 1u) {
            float v[4];
            orc_unpack_rgba8(frags[2u * f + 1] & 0x0fffffffu, v);
            pool[2u * idx + 1] = orc_rgba8_avg(pool[2u * idx + 1], v);
        }
    }
    /* octreeNodeMipmapWrite.comp for i = 2..level (OctreeBuilder.cpp:328-344).
     * The shader compares the descent resolution with the level index itself
     * (:50-52); without that quirk the target is the resolution 2^(i-1). */
    for (int i = 2; i <= level; ++i) {
        uint32_t target = (quirks & ORC_QUIRK_SVO_MIP_LEVEL_TEST) ? (uint32_t)i : (1u << (i - 1));
        for (uint32_t f = 0; f < nfrag; ++f) {
            uint32_t p[3], r;
            frag_pos(frags + 2u * f, quirks, p);
            uint32_t idx = walk(pool, p, R, target, &r);
            if (r == target) {
                uint32_t child = pool[2u * idx] & 0x7fffffffu;
                float sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                for (uint32_t j = 0; j < 8; ++j) {
                    float v[4];
                    orc_unpack_rgba8(pool[2u * (child + j) + 1], v);
                    for (int k = 0; k < 4; ++k) sum[k] = sum[k] + v[k] / 8.0f;
                }
                pool[2u * idx + 1] = orc_pack_rgba8(sum);
            }
        }
    }
    if (node_count) *node_count = total;
    return 0;
}

/* test hooks: the fragment word pair of voxelizer.frag:99-100 and the
 * position decode of octreeNodeFlag.comp:28-32 over arrays */
void orc_svo_pack_fragments(const uint32_t* xyz, const uint32_t* rgba8, uint32_t n, uint32_t* out)
{
    for (uint32_t i = 0; i < n; ++i) pack_fragment(xyz + 3u * i, rgba8[i], out + 2u * i);
}

void orc_svo_unpack_fragments(const uint32_t* frags, uint32_t n, uint32_t quirks, uint32_t* xyz)
{
    for (uint32_t i = 0; i < n; ++i) frag_pos(frags + 2u * i, quirks, xyz + 3u * i);
}

/* ---------------- 3^3 RGBA8 bricks (build-defined; the reference only
 * declares brick constants, EngineConfig.h:24-25) ----------------
 * Dense-grid restatement in FP64: every node of the pool gets its depth d
 * and cell c (at resolution 2^d) from a walk down the child pointers; V_d is
 * the dense grid of the depth-d node colours (zero where no node); texel
 * (i, j, k) of the node's brick is the lattice point 2c + (i, j, k) in
 * half-cell units, whose value per channel is the mean over the depth-d cells
 * touching it - per axis {c} for i = 1, {c-1, c} for i = 0, {c, c+1} for
 * i = 2, cells outside the grid counting as zero - truncated to an integer.
 * bricks[n * 27 + (k * 3 + j) * 3 + i]. */
int orc_svo_bricks(const uint32_t* pool, uint32_t node_count, int level, uint32_t* bricks)
{
    if (level < 1 || level > 10 || node_count < 8 || node_count % 8u) return -1;
    int32_t* depth = (int32_t*)calloc(node_count, sizeof(int32_t));
    uint32_t* cell = (uint32_t*)calloc((size_t)node_count * 3u, sizeof(uint32_t));
    size_t total = 0, off[12];
    for (int d = 1; d <= level; ++d) {
        off[d] = total;
        total += (size_t)1 << (3 * d);
    }
    uint32_t* grid = (uint32_t*)calloc(total, sizeof(uint32_t));
    if (!depth || !cell || !grid) {
        free(depth);
        free(cell);
        free(grid);
        return -1;
    }
    /* root children, then every child block in pool order (a block always
     * follows its parent's block) */
    for (uint32_t s = 0; s < 8; ++s) {
        depth[s] = 1;
        cell[3u * s + 0] = (s >> 1) & 1u;
        cell[3u * s + 1] = (s >> 2) & 1u;
        cell[3u * s + 2] = s & 1u;
    }
    for (uint32_t n = 0; n < node_count; ++n) {
        if (!depth[n]) continue;
        const uint32_t child = pool[2u * n] & 0x7fffffffu;
        if (!child || depth[n] >= level || child + 8u > node_count) continue;
        for (uint32_t s = 0; s < 8; ++s) {
            depth[child + s] = depth[n] + 1;
            cell[3u * (child + s) + 0] = 2u * cell[3u * n + 0] + ((s >> 1) & 1u);
            cell[3u * (child + s) + 1] = 2u * cell[3u * n + 1] + ((s >> 2) & 1u);
            cell[3u * (child + s) + 2] = 2u * cell[3u * n + 2] + (s & 1u);
        }
    }
    for (uint32_t n = 0; n < node_count; ++n) {
        const int d = depth[n];
        if (!d) continue;
        const size_t r = (size_t)1 << d;
        grid[off[d] + ((size_t)cell[3u * n + 2] * r + cell[3u * n + 1]) * r + cell[3u * n + 0]] = pool[2u * n + 1];
    }
    /* every brick reads only the grids: independent per node */
#pragma omp parallel for schedule(static)
    for (uint32_t n = 0; n < node_count; ++n) {
        const int d = depth[n];
        uint32_t* b = bricks + 27u * (size_t)n;
        if (!d) {
            memset(b, 0, 27u * sizeof(uint32_t));
            continue;
        }
        const long r = 1L << d;
        const long c[3] = {(long)cell[3u * n + 0], (long)cell[3u * n + 1], (long)cell[3u * n + 2]};
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) {
                    const int t[3] = {i, j, k};
                    long lo[3], hi[3];
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = t[a] == 2 ? c[a] : c[a] - (t[a] == 0);
                        hi[a] = t[a] == 0 ? c[a] : c[a] + (t[a] == 2);
                    }
                    double sum[4] = {0.0, 0.0, 0.0, 0.0};
                    int count = 0;
                    for (long z = lo[2]; z <= hi[2]; ++z)
                        for (long y = lo[1]; y <= hi[1]; ++y)
                            for (long x = lo[0]; x <= hi[0]; ++x) {
                                ++count;
                                if (x < 0 || y < 0 || z < 0 || x >= r || y >= r || z >= r) continue;
                                const uint32_t v = grid[off[d] + ((size_t)z * (size_t)r + (size_t)y) * (size_t)r + (size_t)x];
                                for (int ch = 0; ch < 4; ++ch) sum[ch] += (double)((v >> (8 * ch)) & 0xffu);
                            }
                    uint32_t out = 0;
                    for (int ch = 0; ch < 4; ++ch) out |= ((uint32_t)floor(sum[ch] / (double)count) & 0xffu) << (8 * ch);
                    b[(k * 3 + j) * 3 + i] = out;
                }
    }
    free(depth);
    free(cell);
    free(grid);
    return 0;
}
