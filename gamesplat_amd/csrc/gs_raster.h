// gs_raster.h - the coverage rule of the reference's splat quads.
//
// splat_vs (shaders/splat.glsl:40-67) draws two triangles spanning
// centre +- radius; the rasteriser covers pixel (i, j) when the pixel centre
// (i + 0.5, j + 0.5) lies inside.  Here the corners go through the vertex
// shader's clip-space round trip (ndc = 2 pos / viewport - 1, window =
// (ndc + 1) viewport / 2), are snapped to 1/256 pixel, and the left/bottom
// edges are inclusive: pixel i is covered iff x0 <= 256 i + 128 < x1 in 24.8
// fixed point.  Shared by the cull (pixel rectangle for binning) and the
// compositors (per-pixel test).
#pragma once

#include <stdint.h>

#include "gs_math.h"

namespace gs {

__device__ __forceinline__ int32_t snap256(float w) {
    float s = w * 256.0f;
    if (s > 1073741824.0f) return 1073741824;
    if (s < -1073741824.0f) return -1073741824;
    return (int32_t)__builtin_rintf(s);
}

__device__ __forceinline__ int32_t ceil_div256(int32_t a) { return -((-a) >> 8); }

// Quad edges x0, x1, y0, y1 in 24.8 fixed point window coordinates.
__device__ __forceinline__ int4 quad_fixed(float cx, float cy, float rx, float ry, float vw, float vh) {
    float x0 = cx + -1.0f * rx, x1 = cx + 1.0f * rx;
    float y0 = cy + -1.0f * ry, y1 = cy + 1.0f * ry;
    float hw = 0.5f * vw, hh = 0.5f * vh;
    int4 q;
    q.x = snap256((2.0f * x0 / vw - 1.0f + 1.0f) * hw);
    q.y = snap256((2.0f * x1 / vw - 1.0f + 1.0f) * hw);
    q.z = snap256((2.0f * y0 / vh - 1.0f + 1.0f) * hh);
    q.w = snap256((2.0f * y1 / vh - 1.0f + 1.0f) * hh);
    return q;
}

// Covered pixels [i0, i1) x [j0, j1) inside the image, packed
// (i0 | i1 << 16, j0 | j1 << 16); all zero when no fragment can survive:
// the fragment test power >= log((1/255) / max(o, 1e-8)) fails everywhere
// when that log is positive, since power <= 0.
__device__ __forceinline__ uint2 pixel_rect(float cx, float cy, float rx, float ry, float opacity, int width,
                                            int height) {
    if ((1.0f / 255.0f) / maxf(opacity, 1e-8f) > 1.0f) return make_uint2(0u, 0u);
    int4 q = quad_fixed(cx, cy, rx, ry, (float)width, (float)height);
    int i0 = max(ceil_div256(q.x - 128), 0), i1 = min(ceil_div256(q.y - 128), width);
    int j0 = max(ceil_div256(q.z - 128), 0), j1 = min(ceil_div256(q.w - 128), height);
    if (i0 >= i1 || j0 >= j1) return make_uint2(0u, 0u);
    return make_uint2((uint32_t)i0 | ((uint32_t)i1 << 16), (uint32_t)j0 | ((uint32_t)j1 << 16));
}

}  // namespace gs
