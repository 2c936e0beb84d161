// vct_dense.hpp - the dense grid object shared by the dense kernels.
#pragma once

#include "vct_internal.hpp"

struct vct_dense {
    vct_dense_desc desc{};
    int res = 0;
    int nlevels = 0;
    size_t texel_bytes = 4;
    size_t offset[16]{};  // texel offset of each level
    size_t block[16]{};   // texels per face block of each level
    size_t texels = 0;
    vct::DeviceBuffer data;
    vct::DeviceBuffer winner;
    vct::DeviceBuffer bwd_rec, bwd_hdr, bwd_start, bwd_ctr;  // the trace adjoint's sweep records
    vct_ctx* scratch = nullptr;  // vertex stage / raster scratch (no clipmap volumes)
};
