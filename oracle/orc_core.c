/*
 * orc_core.c - numeric helpers and the atlas maintenance kernels
 * (clear, copy-alpha, border wrap, anisotropic downsample) restated from the
 * reference GLSL. Test infrastructure only (see vct_oracle.h).
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <string.h>

/* ---------------- numeric helpers ---------------- */

/* atomic.glsl:4-12 convRGBA8ToVec4 */
void orc_unpack_rgba8(uint32_t v, float out[4])
{
    out[0] = (float)(v & 0x000000ffu);
    out[1] = (float)((v & 0x0000ff00u) >> 8u);
    out[2] = (float)((v & 0x00ff0000u) >> 16u);
    out[3] = (float)((v & 0xff000000u) >> 24u);
}

/* atomic.glsl:14-22 convVec4ToRGBA8 (uint() truncates toward zero) */
uint32_t orc_pack_rgba8(const float v[4])
{
    return (((uint32_t)v[3] & 0xffu) << 24u) | (((uint32_t)v[2] & 0xffu) << 16u) |
           (((uint32_t)v[1] & 0xffu) << 8u) | ((uint32_t)v[0] & 0xffu);
}

/* Vulkan float -> UNORM8 store: clamp to [0,1] (NaN -> 0), scale by 255,
 * round to nearest (ties to even). */
uint8_t orc_unorm8(float f)
{
    float c = f > 0.0f ? (f < 1.0f ? f : 1.0f) : 0.0f;
    return (uint8_t)rintf(c * 255.0f);
}

float orc_unorm8_to_float(uint32_t b)
{
    return (float)b / 255.0f;
}

uint32_t orc_float_to_unorm8_rgba(const float v[4])
{
    return (uint32_t)orc_unorm8(v[0]) | ((uint32_t)orc_unorm8(v[1]) << 8) |
           ((uint32_t)orc_unorm8(v[2]) << 16) | ((uint32_t)orc_unorm8(v[3]) << 24);
}

void orc_rgba8_to_float4(uint32_t t, float out[4])
{
    out[0] = orc_unorm8_to_float(t & 0xffu);
    out[1] = orc_unorm8_to_float((t >> 8) & 0xffu);
    out[2] = orc_unorm8_to_float((t >> 16) & 0xffu);
    out[3] = orc_unorm8_to_float((t >> 24) & 0xffu);
}

/* IEEE 754 binary32 -> binary16, round to nearest even. */
uint16_t orc_float_to_half(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xffu;
    uint32_t mant = x & 0x7fffffu;
    if (exp == 0xffu) return (uint16_t)(sign | 0x7c00u | (mant ? 0x200u : 0u));
    int e = (int)exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        int shift = 14 - e;
        uint32_t hm = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (hm & 1u))) hm++;
        return (uint16_t)(sign | hm);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)h;
}

float orc_half_to_float(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t exp = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t mant = (uint32_t)h & 0x3ffu;
    uint32_t out;
    if (exp == 0) {
        if (mant == 0) {
            out = sign;
        } else {
            int e = -1;
            do { e++; mant <<= 1; } while (!(mant & 0x400u));
            out = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        out = sign | 0x7f800000u | (mant << 13);
    } else {
        out = sign | ((exp + 127 - 15) << 23) | (mant << 13);
    }
    float f;
    memcpy(&f, &out, 4);
    return f;
}

/* One uncontended pass of the imageAtomicCompSwap running-average loop
 * (msaaInjectRadiance.frag:185-201, octreeNodeLeafWrite.comp:73-89):
 * the first CAS against 0 stores pack(value); otherwise the stored average is
 * denormalised by its count, the value added and renormalised. */
uint32_t orc_rgba8_avg(uint32_t stored, const float value[4])
{
    if (stored == 0u) return orc_pack_rgba8(value);
    float r[4];
    orc_unpack_rgba8(stored, r);
    r[0] = r[0] * r[3];
    r[1] = r[1] * r[3];
    r[2] = r[2] * r[3];
    float c[4] = {r[0] + value[0], r[1] + value[1], r[2] + value[2], r[3] + value[3]};
    c[0] = c[0] / c[3];
    c[1] = c[1] / c[3];
    c[2] = c[2] / c[3];
    return orc_pack_rgba8(c);
}

/* ---------------- atlas access ---------------- */

size_t orc_volume_texels(int res, int levels)
{
    size_t d = (size_t)(res + 2);
    return d * 6u * d * (size_t)levels * d;
}

static size_t atlas_index(const orc_volume* v, int x, int y, int z)
{
    size_t w = (size_t)(v->res + 2) * 6u;
    size_t h = (size_t)(v->res + 2) * (size_t)v->levels;
    return ((size_t)z * h + (size_t)y) * w + (size_t)x;
}

void orc_load(const orc_volume* v, int x, int y, int z, float out[4])
{
    size_t i = atlas_index(v, x, y, z);
    if (v->fmt == ORC_FMT_RGBA8) {
        orc_rgba8_to_float4(((const uint32_t*)v->texels)[i], out);
    } else {
        const uint16_t* t = (const uint16_t*)v->texels + 4u * i;
        for (int k = 0; k < 4; ++k) out[k] = orc_half_to_float(t[k]);
    }
}

void orc_store(orc_volume* v, int x, int y, int z, const float val[4])
{
    size_t i = atlas_index(v, x, y, z);
    if (v->fmt == ORC_FMT_RGBA8) {
        ((uint32_t*)v->texels)[i] = orc_float_to_unorm8_rgba(val);
    } else {
        uint16_t* t = (uint16_t*)v->texels + 4u * i;
        for (int k = 0; k < 4; ++k) t[k] = orc_float_to_half(val[k]);
    }
}

uint32_t* orc_texel_u32(orc_volume* v, int x, int y, int z)
{
    return (uint32_t*)v->texels + atlas_index(v, x, y, z);
}

/* ---------------- K3 clipmapCleaning.comp:17-31 ---------------- */

void orc_clear_region(orc_volume* vol, const int min_corner[3], const int extent[3],
                      int clip_level, int face_count)
{
    const int res = vol->res;
    const int rwb = res + 2;
    static const float zero[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int gz = 0; gz < extent[2]; ++gz)
        for (int gy = 0; gy < extent[1]; ++gy)
            for (int gx = 0; gx < extent[0]; ++gx) {
                /* toroidal (non-negative) modulo; res is a power of two */
                int px = ((gx + min_corner[0]) & (res - 1)) + 1;
                int py = ((gy + min_corner[1]) & (res - 1)) + 1 + clip_level * rwb;
                int pz = ((gz + min_corner[2]) & (res - 1)) + 1;
                for (int f = 0; f < face_count; ++f) orc_store(vol, px + rwb * f, py, pz, zero);
            }
}

/* ---------------- K6 copyAlphaImage.comp:16-30 ---------------- */

void orc_copy_alpha(orc_volume* radiance, const orc_volume* opacity, int clip_level, int face_count)
{
    const int res = radiance->res;
    const int rwb = res + 2;
    /* texels are independent */
#pragma omp parallel for collapse(2) schedule(static)
    for (int gz = 0; gz < res; ++gz)
        for (int gy = 0; gy < res; ++gy)
            for (int gx = 0; gx < res; ++gx) {
                int px = gx + 1, py = gy + 1 + clip_level * rwb, pz = gz + 1;
                for (int f = 0; f < face_count; ++f) {
                    float dst[4], src[4];
                    orc_load(radiance, px + rwb * f, py, pz, dst);
                    orc_load(opacity, px + rwb * f, py, pz, src);
                    dst[3] = src[3];
                    orc_store(radiance, px + rwb * f, py, pz, dst);
                }
            }
}

/* ---------------- K7 borderWrapping.comp:14-37 ---------------- */
/* The reference dispatches ((res + border) >> 3) groups of 8 per axis
 * (BorderWrapper.cpp:139-140), which reaches invocation ids 0..res-1 only;
 * skip_far reproduces that, otherwise every id in [0, res + border) runs. */
void orc_wrap_border(orc_volume* vol, int border_width_total, int face_count, int region_count,
                     int skip_far)
{
    const int res = vol->res;
    const int rwb = res + border_width_total;
    const int n = skip_far ? 8 * ((res + border_width_total) >> 3) : rwb;
    for (int gz = 0; gz < n; ++gz)
        for (int gy = 0; gy < n; ++gy)
            for (int gx = 0; gx < n; ++gx) {
                if (gx >= rwb || gy >= rwb || gz >= rwb) continue;
                int interior = gx < res + 1 && gy < res + 1 && gz < res + 1 && gx > 0 && gy > 0 && gz > 0;
                if (interior) continue;
                int rx = ((gx + res - 1) & (res - 1)) + 1;
                int ry = ((gy + res - 1) & (res - 1)) + 1;
                int rz = ((gz + res - 1) & (res - 1)) + 1;
                int wx = gx, wy = gy, wz = gz;
                for (int i = 0; i < region_count; ++i) {
                    for (int j = 0; j < face_count; ++j) {
                        if (vol->fmt == ORC_FMT_RGBA8) {
                            *orc_texel_u32(vol, wx + rwb * j, wy, wz) = *orc_texel_u32(vol, rx + rwb * j, ry, rz);
                        } else {
                            float t[4];
                            orc_load(vol, rx + rwb * j, ry, rz, t);
                            orc_store(vol, wx + rwb * j, wy, wz, t);
                        }
                    }
                    ry += rwb;
                    wy += rwb;
                }
            }
}

/* ---------------- K4/K5 opacityDownSample.comp / radianceDownSample.comp ---------------- */

static const int OFFSETS[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

/* opacityDownSample.comp:89-92 */
static int image_coord(int p, int res)
{
    int q = p / res; /* truncates toward zero, as GLSL int division */
    if (q < 0) q = -q;
    return (p + res * (q + 1)) & (res - 1);
}

/* Front/back texel pairs per face: +x, -x, +y, -y, +z, -z
 * (opacityDownSample.comp:94-138). */
static const int PAIRS[6][4][2] = {
    {{0, 1}, {2, 3}, {4, 5}, {6, 7}},
    {{1, 0}, {3, 2}, {5, 4}, {7, 6}},
    {{0, 2}, {1, 3}, {4, 6}, {5, 7}},
    {{2, 0}, {3, 1}, {6, 4}, {7, 5}},
    {{0, 4}, {1, 5}, {2, 6}, {3, 7}},
    {{4, 0}, {5, 1}, {6, 2}, {7, 3}},
};

static float lerp_factor(const int cur[3], const int prev_min[3], int half_res, int region_size)
{
    float center[3], dist[3];
    for (int k = 0; k < 3; ++k) {
        center[k] = (float)(prev_min[k] >> 1) + (float)(half_res >> 1);
        dist[k] = fabsf((float)cur[k] + 0.5f - center[k]) - 0.5f;
    }
    float lerp = 0.0f;
    float inv = 1.0f / ((float)region_size + 1.0f);
    float thr = (float)((half_res >> 1) - region_size);
    if (dist[0] >= thr || dist[1] >= thr || dist[2] >= thr) {
        float m = fmaxf(dist[0], fmaxf(dist[1], dist[2]));
        lerp = m - (float)((half_res >> 1) - region_size) + 1.0f;
        lerp = lerp * inv;
    }
    return lerp;
}

static float mixf(float x, float y, float a)
{
    return x * (1.0f - a) + y * a;
}

static void downsample_common(orc_volume* vol, const int prev_min[3], int clip_level,
                              int region_size, int radiance)
{
    const int res = vol->res;
    const int half = res >> 1;
    const int rwb = res + 2;
    /* every destination texel reads level clip_level - 1 and writes only
     * itself, so the texels are independent */
#pragma omp parallel for collapse(2) schedule(static)
    for (int gz = 0; gz < half; ++gz)
        for (int gy = 0; gy < half; ++gy)
            for (int gx = 0; gx < half; ++gx) {
                int g[3] = {gx, gy, gz};
                int cur[3], wp[3], ps[3];
                for (int k = 0; k < 3; ++k) {
                    cur[k] = (prev_min[k] >> 1) + g[k];
                    wp[k] = image_coord(cur[k], res) + 1;
                    ps[k] = image_coord(cur[k] << 1, res);
                }
                wp[1] += rwb * clip_level;
                int pos[8][3];
                for (int i = 0; i < 8; ++i) {
                    for (int k = 0; k < 3; ++k) pos[i][k] = ps[k] + OFFSETS[i][k] + 1;
                    pos[i][1] += rwb * (clip_level - 1);
                }
                float lerp = lerp_factor(cur, prev_min, half, region_size);
                for (int f = 0; f < 6; ++f) {
                    float v[8][4];
                    for (int i = 0; i < 8; ++i) orc_load(vol, pos[i][0] + rwb * f, pos[i][1], pos[i][2], v[i]);
                    float self[4];
                    orc_load(vol, wp[0] + rwb * f, wp[1], wp[2], self);
                    if (!radiance) {
                        float s = 0.0f;
                        for (int p = 0; p < 4; ++p) {
                            const float* a = v[PAIRS[f][p][0]];
                            const float* b = v[PAIRS[f][p][1]];
                            s = s + a[3];
                            s = s + (1.0f - a[3]) * b[3];
                        }
                        float down = s * 0.25f;
                        self[3] = mixf(down, self[0], lerp);
                        self[1] = self[3];
                        orc_store(vol, wp[0] + rwb * f, wp[1], wp[2], self);
                    } else {
                        float out[4];
                        for (int c = 0; c < 4; ++c) {
                            float s = 0.0f;
                            for (int p = 0; p < 4; ++p) {
                                const float* a = v[PAIRS[f][p][0]];
                                const float* b = v[PAIRS[f][p][1]];
                                s = s + a[c];
                                s = s + (1.0f - a[3]) * b[c];
                            }
                            out[c] = mixf(s * 0.25f, self[c], lerp);
                        }
                        orc_store(vol, wp[0] + rwb * f, wp[1], wp[2], out);
                    }
                }
            }
}

void orc_downsample_opacity(orc_volume* vol, const int prev_min_corner[3], int clip_level, int region_size)
{
    downsample_common(vol, prev_min_corner, clip_level, region_size, 0);
}

/* radianceDownSample.comp:28-139. The shader declares lerpFactor inside the
 * transition branch and uses it outside (:58-71), which does not compile; the
 * opacity shader's form (0 outside the band, :57-64) is restated here. */
void orc_downsample_radiance(orc_volume* vol, const int prev_min_corner[3], int clip_level, int region_size)
{
    downsample_common(vol, prev_min_corner, clip_level, region_size, 1);
}
