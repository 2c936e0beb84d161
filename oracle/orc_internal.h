/* orc_internal.h - helpers shared by the oracle's translation units. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <stdint.h>
#include "vct_oracle.h"

uint8_t orc_unorm8(float f);
float orc_unorm8_to_float(uint32_t b);
int orc_has_texture(const orc_tris* tris, int32_t idx);
int orc_occlusion_discard(const orc_tris* tris, const orc_material* m, const float uv[2]);
void orc_rgba8_to_float4(uint32_t t, float out[4]);
void orc_load(const orc_volume* v, int x, int y, int z, float out[4]);
void orc_store(orc_volume* v, int x, int y, int z, const float val[4]);
uint32_t* orc_texel_u32(orc_volume* v, int x, int y, int z);

/* Triangle rasterizer shared by K1, K2 and the SVO voxelizer. For every
 * sample inside the triangle it calls fn with the interpolated world
 * position and the barycentrics (in the caller's vertex order). */
typedef void (*orc_frag_fn)(void* user, uint32_t tri, const float pos[3], const double lam[3]);
void orc_raster_triangle(const float* p, float voxel_size, const int lo[3], const int hi[3],
                         int samples, orc_frag_fn fn, void* user, uint32_t tri);

/* voxelConeTracing.frag:78-135 cone directions in double (orc_dense.c). */
const double* orc_dense_dir16(int i);
const double* orc_dense_dir32(int i);

#endif
