/*
 * orc_trace.c - instantiates the voxelConeTracing.frag restatement
 * (orc_trace_impl.h) in FP32 and FP64. Test infrastructure only.
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>

#define R float
#define FN(name) name##_f32
#define M_FLOOR floorf
#define M_CEIL ceilf
#define M_SQRT sqrtf
#define M_EXP2 exp2f
#define M_LOG2 log2f
#define M_POW powf
#define M_TAN tanf
#define M_FABS fabsf
#define M_FMIN fminf
#define M_FMAX fmaxf
#include "orc_trace_impl.h"
#undef R
#undef FN
#undef M_FLOOR
#undef M_CEIL
#undef M_SQRT
#undef M_EXP2
#undef M_LOG2
#undef M_POW
#undef M_TAN
#undef M_FABS
#undef M_FMIN
#undef M_FMAX

#define R double
#define FN(name) name##_f64
#define M_FLOOR floor
#define M_CEIL ceil
#define M_SQRT sqrt
#define M_EXP2 exp2
#define M_LOG2 log2
#define M_POW pow
#define M_TAN tan
#define M_FABS fabs
#define M_FMIN fmin
#define M_FMAX fmax
#include "orc_trace_impl.h"

float orc_calc_min_level(const orc_trace_params* p, const float pos[3])
{
    return calc_min_level_f32(p, pos);
}

void orc_trace_cone(const orc_volume* rad, const orc_trace_params* p, const float start[3],
                    const float dir[3], float aperture, float max_dist, float start_level,
                    float step_factor, float out[4], uint32_t* steps)
{
    uint64_t n = 0;
    trace_cone_f32(rad, p, start, dir, aperture, max_dist, start_level, step_factor, out, &n);
    if (steps) *steps = (uint32_t)n;
}

void orc_brdf(const float n[3], const float v[3], const float l[3], const float diffuse_color[3],
              const float specular_color[3], float perceptual_roughness, float metallic, float out[3])
{
    (void)metallic;
    brdf_f32(n, v, l, diffuse_color, specular_color, perceptual_roughness, out);
}

void orc_trace(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
               const orc_shadow* sm, const orc_light* light, uint32_t quirks, int row_begin,
               int row_end, float* diffuse_out, float* specular_out, uint64_t* samples)
{
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int r = row_begin; r < row_end; ++r) {
        uint64_t s = 0;
        trace_rows_f32(rad, g, p, sm, light, quirks, r, r + 1, diffuse_out, specular_out, &s);
        total += s;
    }
    if (samples) *samples = total;
}

void orc_trace_rowlist(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
                       const orc_shadow* sm, const orc_light* light, uint32_t quirks, const int* rows,
                       int nrows, float* diffuse_out, float* specular_out, uint64_t* samples)
{
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int i = 0; i < nrows; ++i) {
        uint64_t s = 0;
        trace_rows_f32(rad, g, p, sm, light, quirks, rows[i], rows[i] + 1, diffuse_out, specular_out, &s);
        total += s;
    }
    if (samples) *samples = total;
}

void orc_trace_rowlist_f64(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
                           const orc_shadow* sm, const orc_light* light, uint32_t quirks, const int* rows,
                           int nrows, double* diffuse_out, double* specular_out, uint64_t* samples)
{
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int i = 0; i < nrows; ++i) {
        uint64_t s = 0;
        trace_rows_f64(rad, g, p, sm, light, quirks, rows[i], rows[i] + 1, diffuse_out, specular_out, &s);
        total += s;
    }
    if (samples) *samples = total;
}

void orc_trace_f64(const orc_volume* rad, const orc_gbuffer* g, const orc_trace_params* p,
                   const orc_shadow* sm, const orc_light* light, uint32_t quirks, int row_begin,
                   int row_end, double* diffuse_out, double* specular_out, uint64_t* samples)
{
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int r = row_begin; r < row_end; ++r) {
        uint64_t s = 0;
        trace_rows_f64(rad, g, p, sm, light, quirks, r, r + 1, diffuse_out, specular_out, &s);
        total += s;
    }
    if (samples) *samples = total;
}

/* ---------------- K16 voxelConeTracing_Octree.frag ---------------- */

/* voxelConeTracing_Octree.frag:319-346 sampleSVO. The scene box is a
 * parameter here (the shader hard-codes Sponza's, :321-322). uvec3() of a
 * negative value saturates to 0, as the hardware conversion does. */
static void sample_svo(const uint32_t* pool, const orc_trace_params* p, const float bb_min[3],
                       const float bb_max[3], const float pos[3], uint32_t level, float out[4])
{
    float ext[3], center[3];
    for (int k = 0; k < 3; ++k) {
        ext[k] = bb_max[k] - bb_min[k];
        center[k] = (bb_min[k] + bb_max[k]) * 0.5f;
    }
    float ev = fmaxf(ext[0], fmaxf(ext[1], ext[2])) * 0.5f;
    uint32_t res = (uint32_t)p->volume_dimension;
    uint32_t fp[3];
    for (int k = 0; k < 3; ++k) {
        float s = (pos[k] - center[k]) / ev;
        float f = ((s + 1.0f) * 0.5f) * (float)res;
        uint32_t u = f > 0.0f ? (f < 4294967040.0f ? (uint32_t)f : 0xffffffffu) : 0u;
        fp[k] = u > res ? res : u;
    }
    uint32_t target = 1u << level;
    uint32_t idx = 0, cur = 0;
    do {
        res >>= 1;
        uint32_t cx = fp[0] >= res, cy = fp[1] >= res, cz = fp[2] >= res;
        idx = cur + (cz | (cx << 1) | (cy << 2));
        cur = pool[2u * idx] & 0x7fffffffu;
        fp[0] -= cx * res;
        fp[1] -= cy * res;
        fp[2] -= cz * res;
    } while (cur != 0u && res > target);
    uint32_t t = pool[2u * idx + 1];
    for (int k = 0; k < 4; ++k) out[k] = (float)((t >> (8 * k)) & 0xffu) / 255.0f;
}

static void svo_trace_cone(const uint32_t* pool, const orc_trace_params* p, const float bb_min[3],
                           const float bb_max[3], const float start_in[3], const float dir[3], float aperture,
                           float max_dist, float start_level, float step_factor, float out[4], uint64_t* steps)
{
    float result[4] = {0, 0, 0, 0};
    float cone_coef = 2.0f * tanf(aperture * 0.5f);
    float cur_level = start_level;
    float voxel_size = p->voxel_size * exp2f(cur_level);
    float start[3];
    for (int k = 0; k < 3; ++k) start[k] = start_in[k] + dir[k] * voxel_size * p->trace_start_offset * 0.5f;
    float step = 0.0f;
    float diameter = fmaxf(step * cone_coef, p->voxel_size);
    float occlusion = 0.0f;
    float seg = voxel_size;
    float min_radius = p->voxel_size * p->volume_dimension * 0.5f;
    while (step < max_dist && occlusion < 1.0f) {
        float pos[3];
        for (int k = 0; k < 3; ++k) pos[k] = start[k] + dir[k] * step;
        float dc[3] = {p->volume_center[0] - pos[0], p->volume_center[1] - pos[1], p->volume_center[2] - pos[2]};
        float dist = sqrtf(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
        float min_level = ceilf(log2f(dist / min_radius));
        cur_level = log2f(diameter / p->voxel_size);
        cur_level = fminf(fmaxf(fmaxf(start_level, cur_level), min_level), (float)(p->clip_levels - 1));
        float lo[4], hi[4], s[4];
        sample_svo(pool, p, bb_min, bb_max, pos, (uint32_t)floorf(cur_level), lo);
        sample_svo(pool, p, bb_min, bb_max, pos, (uint32_t)ceilf(cur_level), hi);
        float fl = cur_level - floorf(cur_level);
        for (int k = 0; k < 4; ++k) s[k] = lo[k] * (1.0f - fl) + hi[k] * fl;
        voxel_size = p->voxel_size * exp2f(cur_level);
        float correction = seg / voxel_size;
        float opacity = fminf(fmaxf(1.0f - powf(fmaxf(1.0f - s[3], 0.0f), correction), 0.0f), 1.0f);
        float blend = fminf(fmaxf(1.0f - result[3], 0.0f), 1.0f);
        for (int k = 0; k < 3; ++k) result[k] += blend * (s[k] * correction);
        result[3] += blend * opacity;
        occlusion += (1.0f - occlusion) * opacity / (1.0f + (step + voxel_size) * p->occlusion_decay);
        float prev = step;
        step += fmaxf(diameter, p->voxel_size) * step_factor;
        seg = step - prev;
        diameter = step * cone_coef;
        if (steps) ++*steps;
    }
    out[0] = fminf(fmaxf(result[0], 0.0f), 1.0f);
    out[1] = fminf(fmaxf(result[1], 0.0f), 1.0f);
    out[2] = fminf(fmaxf(result[2], 0.0f), 1.0f);
    out[3] = fminf(fmaxf(1.0f - occlusion, 0.0f), 1.0f);
}

static float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

/* voxelConeTracing_Octree.frag:148-300 main */
void orc_svo_trace(const uint32_t* pool, const orc_gbuffer* g, const orc_trace_params* p, const float bb_min[3],
                   const float bb_max[3], const orc_shadow* sm, const orc_light* L, uint32_t quirks, int row_begin,
                   int row_end, float* dout_all, float* sout_all, uint64_t* samples)
{
    const int W = g->width, H = g->height;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int py = row_begin; py < row_end; ++py)
        for (int px = 0; px < W; ++px) {
            size_t idx = (size_t)py * (size_t)W + (size_t)px;
            float* dO = dout_all + 4u * idx;
            float* sO = sout_all + 4u * idx;
            dO[0] = dO[1] = dO[2] = 0; dO[3] = 1;
            sO[0] = sO[1] = sO[2] = 0; sO[3] = 1;
            float depth = g->depth[idx];
            if (depth == 1.0f) continue;
            float tc[2] = {((float)px + 0.5f) / (float)W, ((float)py + 0.5f) / (float)H};
            float q[4] = {tc[0] * 2 - 1, tc[1] * 2 - 1, depth, 1};
            float wp[4];
            for (int k = 0; k < 4; ++k)
                wp[k] = p->view_proj_inv[k] * q[0] + p->view_proj_inv[4 + k] * q[1] + p->view_proj_inv[8 + k] * q[2] +
                        p->view_proj_inv[12 + k] * q[3];
            float world[3] = {wp[0] / wp[3], wp[1] / wp[3], wp[2] / wp[3]};
            float view[3] = {p->eye_pos[0] - world[0], p->eye_pos[1] - world[1], p->eye_pos[2] - world[2]};
            normalize_f32(view);
            uint32_t dt = g->diffuse[idx], st = g->specular[idx];
            float diffuse_color[3], spec_color[3], normal[3], emission[3];
            for (int k = 0; k < 3; ++k) {
                diffuse_color[k] = (float)((dt >> (8 * k)) & 0xffu) / 255.0f;
                spec_color[k] = (float)((st >> (8 * k)) & 0xffu) / 255.0f;
                normal[k] = orc_half_to_float(g->normal[4u * idx + k]) * 2.0f - 1.0f;
                emission[k] = orc_half_to_float(g->emission[4u * idx + k]);
            }
            float rough = (float)((dt >> 24) & 0xffu) / 255.0f;
            float metallic = (float)((st >> 24) & 0xffu) / 255.0f;
            normalize_f32(normal);
            int has_emission = emission[0] > 0 || emission[1] > 0 || emission[2] > 0;
            float min_level = calc_min_level_f32(p, world);
            float vs = p->voxel_size * exp2f(min_level);
            float start[3];
            for (int k = 0; k < 3; ++k) start[k] = world[k] + normal[k] * vs * p->trace_start_offset;
            float ind[4] = {0, 0, 0, 1};
            float valid = 0.0f;
            const int ncones = p->enable_32_cones ? 32 : 16;
            const float aperture = p->enable_32_cones ? 0.628319f : 0.872665f;
            const float sf = fmaxf(0.2f, p->min_trace_step_factor);
            for (int i = 0; i < ncones; ++i) {
                const float* d = p->enable_32_cones ? DIRS32_f32[i] : DIRS16_f32[i];
                float ct = normal[0] * d[0] + normal[1] * d[1] + normal[2] * d[2];
                if (ct < 0.0f) continue;
                float c[4];
                svo_trace_cone(pool, p, bb_min, bb_max, start, d, aperture, 30.0f, min_level, sf, c, &total);
                for (int k = 0; k < 4; ++k) ind[k] += c[k] * ct;
                valid += ct;
            }
            for (int k = 0; k < 4; ++k) ind[k] /= valid;
            ind[3] *= p->ambient_occlusion_factor;
            for (int k = 0; k < 3; ++k) ind[k] *= diffuse_color[k] * p->indirect_diffuse_intensity;
            for (int k = 0; k < 4; ++k) ind[k] = clamp01(ind[k]);
            float spec[3] = {0, 0, 0};
            if ((spec_color[0] > 1e-6f || spec_color[1] > 1e-6f || spec_color[2] > 1e-6f) && metallic > 1e-6f) {
                float I[3] = {-view[0], -view[1], -view[2]};
                float dni = normal[0] * I[0] + normal[1] * I[1] + normal[2] * I[2];
                float dir[3] = {I[0] - 2 * dni * normal[0], I[1] - 2 * dni * normal[1], I[2] - 2 * dni * normal[2]};
                normalize_f32(dir);
                float c[4];
                svo_trace_cone(pool, p, bb_min, bb_max, start, dir, fmaxf(rough, 0.05f), 30.0f, min_level,
                               fmaxf(0.2f, p->voxel_size), c, &total);
                for (int k = 0; k < 3; ++k) spec[k] = c[k] * spec_color[k] * p->indirect_specular_intensity;
            }
            float direct[3] = {0, 0, 0};
            if (has_emission) {
                for (int k = 0; k < 3; ++k) direct[k] = emission[k];
            } else {
                float l[3] = {-L->direction[0], -L->direction[1], -L->direction[2]};
                normalize_f32(l);
                float b[3];
                brdf_f32(normal, view, l, diffuse_color, spec_color, rough, b);
                float vis = visibility_f32(sm, L, world, quirks);
                for (int k = 0; k < 3; ++k) direct[k] = b[k] * vis;
            }
            for (int k = 0; k < 3; ++k) direct[k] = clamp01(direct[k]);
            switch (p->rendering_mode) {
            case 0: for (int k = 0; k < 3; ++k) dO[k] = diffuse_color[k]; break;
            case 1: for (int k = 0; k < 3; ++k) dO[k] = spec_color[k]; break;
            case 2: for (int k = 0; k < 3; ++k) dO[k] = normal[k] * 0.5f + 0.5f; break;
            case 3: {
                static const float colors[7][4] = {{1, 0, 0, 1}, {0, 1, 0, 1}, {0, 0, 1, 1}, {1, 1, 0, 1},
                                                   {0, 1, 1, 1}, {1, 0, 1, 1}, {1, 1, 1, 1}};
                float out4[4] = {0, 0, 0, 0};
                if (min_level < 6.0f) {
                    int lower = (int)floorf(min_level);
                    if (lower < 0) lower = 0;
                    float fr = min_level - floorf(min_level);
                    for (int k = 0; k < 4; ++k) out4[k] = colors[lower][k] * (1.0f - fr) + colors[lower + 1][k] * fr;
                }
                for (int k = 0; k < 4; ++k) dO[k] = out4[k] * 0.5f;
                break;
            }
            case 4: for (int k = 0; k < 3; ++k) dO[k] = direct[k] * ind[3]; break;
            case 5: for (int k = 0; k < 3; ++k) dO[k] = ind[k]; break;
            case 6: for (int k = 0; k < 3; ++k) sO[k] = spec[k]; break;
            case 7: for (int k = 0; k < 3; ++k) dO[k] = ind[3]; break;
            case 8:
                for (int k = 0; k < 3; ++k) {
                    dO[k] = direct[k] * ind[3] + ind[k];
                    sO[k] = spec[k];
                }
                break;
            default: break;
            }
        }
    if (samples) *samples = total;
}

/* test hook: one sampleSVO of voxelConeTracing_Octree.frag:319-346 */
void orc_sample_svo(const uint32_t* pool, const orc_trace_params* p, const float bb_min[3], const float bb_max[3],
                    const float pos[3], uint32_t level, float out[4])
{
    sample_svo(pool, p, bb_min, bb_max, pos, level, out);
}

/* test hook: one sampleClipmap of voxelConeTracing.frag:313-327 */
void orc_sample_clipmap(const orc_volume* vol, const orc_trace_params* p, const float pos[3], int level,
                        const float dir[3], float out[4])
{
    float face_off[3] = {(dir[0] > 0 ? 0.0f : 1.0f) / 6.0f, (dir[1] > 0 ? 2.0f : 3.0f) / 6.0f,
                         (dir[2] > 0 ? 4.0f : 5.0f) / 6.0f};
    float w[3] = {dir[0] * dir[0], dir[1] * dir[1], dir[2] * dir[2]};
    sample_clipmap_f32(vol, p, pos, level, face_off, w, out);
}
