/*
 * orc_filter.c - the step after the cone trace: SpecularFilterPass
 * (specularFilter.frag:25-53) = diffuse + filtered specular, then the
 * optional Uncharted2 tonemap (tonemapping.glsl:9-26). Filters of
 * filter.glsl: the 32-direction x 8-ring "gaussian" blur (:9-24) and the
 * 15x15 bilateral filter (:26-63). Textures are sampled as the reference's
 * "VCTSampler" does: bilinear, clamp to edge. FP32, GLSL operation order.
 * Test infrastructure only (see vct_oracle.h).
 */
#include "vct_oracle.h"
#include "orc_internal.h"

#include <math.h>

static void tex_clamp_linear(const float* img, int W, int H, float s, float t, float out[4])
{
    float u = s * (float)W - 0.5f, v = t * (float)H - 0.5f;
    float fu = floorf(u), fv = floorf(v);
    float a = u - fu, b = v - fv;
    int i0 = (int)fu, j0 = (int)fv;
    for (int c = 0; c < 4; ++c) out[c] = 0.0f;
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            int i = i0 + dx, j = j0 + dy;
            i = i < 0 ? 0 : (i > W - 1 ? W - 1 : i);
            j = j < 0 ? 0 : (j > H - 1 ? H - 1 : j);
            const float w = (dx ? a : 1.0f - a) * (dy ? b : 1.0f - b);
            const float* p = img + 4u * ((size_t)j * W + i);
            for (int c = 0; c < 4; ++c) out[c] += w * p[c];
        }
}

/* filter.glsl:9-24; the direction loop runs on a float accumulator, exactly
 * as the shader does (the iteration count follows from float rounding). */
static void gaussian_blur(const float* img, int W, int H, float s, float t, float blur, float out[4])
{
    const float two_pi = 6.28318530718f;
    const float rx = blur / (float)W, ry = blur / (float)H;
    float c[4];
    tex_clamp_linear(img, W, H, s, t, c);
    for (float d = 0.0f; d < two_pi; d += two_pi / 32.0f)
        for (float i = 1.0f / 8.0f; i <= 1.0f; i += 1.0f / 8.0f) {
            float q[4];
            tex_clamp_linear(img, W, H, s + cosf(d) * rx * i, t + sinf(d) * ry * i, q);
            for (int k = 0; k < 4; ++k) c[k] += q[k];
        }
    for (int k = 0; k < 4; ++k) out[k] = c[k] / (8.0f * 32.0f - 15.0f);
}

static const float KERNEL15[15] = {0.031225216f, 0.033322271f, 0.035206333f, 0.036826804f, 0.038138565f,
                                   0.039104044f, 0.039695028f, 0.039894000f, 0.039695028f, 0.039104044f,
                                   0.038138565f, 0.036826804f, 0.035206333f, 0.033322271f, 0.031225216f};

static float normpdf(float x, float sigma) { return 0.39894f * expf(-0.5f * x * x / (sigma * sigma)) / sigma; }

/* filter.glsl:26-63 */
static void bilateral(const float* img, int W, int H, float s, float t, float out[3])
{
    const float bsigma = 10.0f;
    float center[4];
    tex_clamp_linear(img, W, H, s, t, center);
    float Z = 0.0f, acc[3] = {0, 0, 0};
    const float bZ = 1.0f / normpdf(0.0f, bsigma);
    const float ix = 1.0f / (float)W, iy = 1.0f / (float)H;
    for (int i = -7; i <= 7; ++i)
        for (int j = -7; j <= 7; ++j) {
            float q[4];
            tex_clamp_linear(img, W, H, s + (float)i * ix, t + (float)j * iy, q);
            const float dv[3] = {q[0] - center[0], q[1] - center[1], q[2] - center[2]};
            const float dd = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            const float f = (0.39894f * expf(-0.5f * dd / (bsigma * bsigma)) / bsigma) * bZ * KERNEL15[7 + j] *
                            KERNEL15[7 + i];
            Z += f;
            for (int k = 0; k < 3; ++k) acc[k] += f * q[k];
        }
    for (int k = 0; k < 3; ++k) out[k] = acc[k] / Z;
}

static float uncharted2(float x)
{
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

void orc_specular_filter(const float* diffuse, const float* specular, int W, int H, float gamma, float exposure,
                         int tonemap_enable, int filter_method, float* out)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            const float s = ((float)px + 0.5f) / (float)W, t = ((float)py + 0.5f) / (float)H;
            float col[4], sc[4] = {0.0f, 0.0f, 0.0f, 1.0f};
            tex_clamp_linear(diffuse, W, H, s, t, col);
            if (filter_method == 1) {
                gaussian_blur(specular, W, H, s, t, 0.01f, sc);
            } else {
                float b[3];
                bilateral(specular, W, H, s, t, b);
                for (int k = 0; k < 3; ++k) sc[k] += b[k];
            }
            for (int k = 0; k < 3; ++k) col[k] += sc[k];
            float* o = out + 4u * ((size_t)py * W + px);
            if (tonemap_enable == 1) {
                const float white = 1.0f / uncharted2(11.2f);
                for (int k = 0; k < 3; ++k) o[k] = powf(uncharted2(col[k] * exposure) * white, 1.0f / gamma);
                o[3] = col[3];
            } else {
                for (int k = 0; k < 4; ++k) o[k] = col[k];
            }
        }
}
