// gs_planes.h - host-side fills of the device scene's plane layout (used by
// the uploads in gs_api.cpp; implemented in gs_loaders.cpp, host code only).
#pragma once

#include <stdint.h>

#include "../../include/gs_api.h"

namespace gs {

// Entries [0, count) of the source into the 15 float4 planes of `planes`,
// plane stride `stride` entries (entry i of plane p at (p * stride + i) * 4).
int fill_planes_gaussians(const GsGaussian* g, uint32_t count, uint32_t stride, float* planes);
int fill_planes_ply(const float* rows, uint32_t count, const GsPlyLayout* layout, uint32_t stride, float* planes);
// SOG entries [first, first + count) into entries [0, count) of `planes`.
int fill_planes_sog(const GsSogData* sog, uint32_t first, uint32_t count, uint32_t stride, float* planes);
// Entries [first, first + count) of a plane array with `src_stride` entries a
// plane into entries [0, count) of `planes`.
int fill_planes_copy(const float* src, uint32_t src_stride, uint32_t first, uint32_t count, uint32_t stride,
                     float* planes);

}  // namespace gs
