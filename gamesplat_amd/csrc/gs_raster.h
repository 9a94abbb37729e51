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

// splat_vs emits gl_Position = (ndc, ndc_z, 1) (shaders/splat.glsl:63-67) and
// no pipeline enables depth clamping (src/renderer.cpp:277-320), so the
// rasteriser clips a quad (its four corners share ndc_z) unless
// -w <= z <= w: -1 <= ndc_z <= 1 in GL clip space, 0 <= ndc_z <= 1 when
// clip_z_01 selects the [0, 1] depth range.  With the camera's far plane at
// 100 (src/camera.cpp:12) this drops every splat deeper than 100.
__device__ __forceinline__ bool depth_clipped(float ndc_z, float clip_z_01) {
    const float lo = clip_z_01 > 0.5f ? 0.0f : -1.0f;
    return !(ndc_z <= 1.0f && ndc_z >= lo);
}

// Quad edges x0, x1, y0, y1 in 24.8 fixed point window coordinates; the
// corners are centre + corner * radius + jitter, in that order (TAA jitter
// of StochasticSplats samples, zero otherwise: adding 0 changes no value).
__device__ __forceinline__ int4 quad_fixed(float cx, float cy, float rx, float ry, float vw, float vh,
                                           float jx = 0.0f, float jy = 0.0f) {
    float x0 = (cx + -1.0f * rx) + jx, x1 = (cx + 1.0f * rx) + jx;
    float y0 = (cy + -1.0f * ry) + jy, y1 = (cy + 1.0f * ry) + jy;
    float hw = 0.5f * vw, hh = 0.5f * vh;
    int4 q;
    q.x = snap256((2.0f * x0 / vw - 1.0f + 1.0f) * hw);
    q.y = snap256((2.0f * x1 / vw - 1.0f + 1.0f) * hw);
    q.z = snap256((2.0f * y0 / vh - 1.0f + 1.0f) * hh);
    q.w = snap256((2.0f * y1 / vh - 1.0f + 1.0f) * hh);
    return q;
}

// Pixel bounds [i0, i1) x [j0, j1) (not cut to the image) outside which the
// fragment test cannot pass; false when the bound does not apply.
//
// The fragment test keeps power = -q(d)/2 >= cutoff = log((1/255) / o), i.e.
// q(d) <= c = -2 cutoff, with q(d) = d^T K d and K the conic.  When the
// determinant clamp of the conic is not hit, K = S^-1 for the covariance
// S = [[sa, sb], [sb, sc]], and {q <= c} is an ellipse whose axis-aligned box
// has half extents sqrt(c sa) and sqrt(c sc).  With rho the correlation,
// 1 - rho^2 = det / (sa sc); the float evaluation of power differs from the
// exact form by at most ~10 eps q / (1 - |rho|)^2 <= 40 eps q / (1 - rho^2)^2,
// so the box is grown by that relative margin (plus 0.3% and an absolute
// 1e-5 on c, and 1e-3 px) and is only used for 1 - rho^2 >= 0.01.  Every
// pixel that can pass the fragment test stays inside, so composites with and
// without the cut are identical; it only shortens the tile lists.
// The grown bound cg on q(d) of every pixel that can pass the fragment test
// (see ellipse_pixel_bounds); false when the bound does not apply.
__device__ __forceinline__ bool ellipse_cutoff_bound(float sa, float sb, float sc, float opacity, float* cg) {
    // reciprocals to 1 ulp: their error is far inside the 0.3% and 2x slack
    const float det = sa * sc - sb * sb;
    const float r2 = det * __builtin_amdgcn_rcpf(sa * sc);
    if (!(det >= 1e-6f && r2 >= 0.01f)) return false;
    const float c = maxf(-2.0f * __logf((1.0f / 255.0f) / maxf(opacity, 1e-8f)), 0.0f);
    *cg = c * (1.003f + 4.8e-6f * __builtin_amdgcn_rcpf(r2 * r2)) + 1e-5f;
    return true;
}

// Pixel bounds of the ellipse {q <= cg} around (cx, cy).
__device__ __forceinline__ int4 ellipse_box(float cx, float cy, float sa, float sc, float cg) {
    const float ex = __builtin_sqrtf(cg * sa) + 1e-3f;
    const float ey = __builtin_sqrtf(cg * sc) + 1e-3f;
    // pixel i can pass only if |i + 0.5 - cx| <= ex
    int4 e;
    e.x = (int)__builtin_ceilf(clampf(cx - ex - 0.5f, -2.0f, 65536.0f));
    e.y = (int)__builtin_floorf(clampf(cx + ex - 0.5f, -2.0f, 65536.0f)) + 1;
    e.z = (int)__builtin_ceilf(clampf(cy - ey - 0.5f, -2.0f, 65536.0f));
    e.w = (int)__builtin_floorf(clampf(cy + ey - 0.5f, -2.0f, 65536.0f)) + 1;
    return e;
}

__device__ __forceinline__ bool ellipse_pixel_bounds(float cx, float cy, float sa, float sb, float sc, float opacity,
                                                     int4* e) {
    float cg;
    if (!ellipse_cutoff_bound(sa, sb, sc, opacity, &cg)) return false;
    *e = ellipse_box(cx, cy, sa, sc, cg);
    return true;
}

// Half widths of the same ellipse across the diagonals: every d with
// d^T S^-1 d <= cg has |dx + dy| <= wp and |dx - dy| <= wm (Cauchy-Schwarz:
// |n.d| <= sqrt(cg n^T S n) for n = (1, +-1)).  n^T S n = sa + sc +- 2 sb is
// formed in float to within 2 eps (sa + sc + 2 |sb|), which is added; the
// square root is grown by 1e-6 relative, and 4e-3 px covers the rounding of
// the pixel-centre offsets the caller compares with.
__device__ __forceinline__ float2 ellipse_diagonal_halfwidths(float sa, float sb, float sc, float cg) {
    const float s = sa + sc, t = 2.0f * sb, slack = 4.8e-7f * (s + __builtin_fabsf(t));
    const float wp = __builtin_sqrtf(fmaxf(cg * ((s + t) + slack), 0.0f)) * 1.000001f + 4e-3f;
    const float wm = __builtin_sqrtf(fmaxf(cg * ((s - t) + slack), 0.0f)) * 1.000001f + 4e-3f;
    return make_float2(wp, wm);
}

// Pixels [i0, i1) x [j0, j1) (not cut to the image) whose centre the quad
// covers.
__device__ __forceinline__ int4 quad_pixel_bounds(float cx, float cy, float rx, float ry, int width, int height,
                                                  float jx = 0.0f, float jy = 0.0f) {
    const int4 q = quad_fixed(cx, cy, rx, ry, (float)width, (float)height, jx, jy);
    return make_int4(ceil_div256(q.x - 128), ceil_div256(q.y - 128), ceil_div256(q.z - 128), ceil_div256(q.w - 128));
}

// Pixels [i0, i1) x [j0, j1) inside the image whose centre the quad covers,
// optionally cut down to the ellipse bounds above.  A jitter (jx, jy) moves
// the quad and the fragment centre (cx + jx, cy + jy).  pad grows the
// rectangle by that many pixels on each side before it is cut to the image:
// with pad 1 it holds the rectangle of every jitter up to half a pixel (each
// bound moves by less than one pixel, rounding included).
__device__ __forceinline__ int4 splat_pixel_rect(float cx, float cy, float sa, float sb, float sc, float rx,
                                                 float ry, float opacity, int width, int height, bool tight,
                                                 float jx = 0.0f, float jy = 0.0f, int pad = 0) {
    int4 q = quad_pixel_bounds(cx, cy, rx, ry, width, height, jx, jy);
    int4 e;
    if (tight && ellipse_pixel_bounds(cx + jx, cy + jy, sa, sb, sc, opacity, &e)) {
        q.x = max(q.x, e.x);
        q.y = min(q.y, e.y);
        q.z = max(q.z, e.z);
        q.w = min(q.w, e.w);
    }
    return make_int4(max(q.x - pad, 0), min(q.y + pad, width), max(q.z - pad, 0), min(q.w + pad, height));
}

// The rectangle above packed (i0 | i1 << 16, j0 | j1 << 16); all zero when
// no fragment can survive: the fragment test power >= log((1/255) / max(o,
// 1e-8)) fails everywhere when that log is positive, since power <= 0.
__device__ __forceinline__ uint2 pixel_rect(float cx, float cy, float sa, float sb, float sc, float rx, float ry,
                                            float opacity, int width, int height, bool tight, float jx = 0.0f,
                                            float jy = 0.0f, int pad = 0) {
    if ((1.0f / 255.0f) / maxf(opacity, 1e-8f) > 1.0f) return make_uint2(0u, 0u);
    int4 r = splat_pixel_rect(cx, cy, sa, sb, sc, rx, ry, opacity, width, height, tight, jx, jy, pad);
    if (r.x >= r.y || r.z >= r.w) return make_uint2(0u, 0u);
    return make_uint2((uint32_t)r.x | ((uint32_t)r.y << 16), (uint32_t)r.z | ((uint32_t)r.w << 16));
}

// A pixel rectangle from pixel_rect as its tile rectangle at tile size
// 1 << shift, x0 | x1 << 8 | y0 << 16 | y1 << 24 (0 when empty; the view must
// span at most 255 tiles a side).
__device__ __forceinline__ uint32_t pack_tile_rect(uint2 px, int shift) {
    const uint32_t x0 = (px.x & 0xFFFFu) >> shift, x1 = (((px.x >> 16) - 1u) >> shift) + 1u;
    const uint32_t y0 = (px.y & 0xFFFFu) >> shift, y1 = (((px.y >> 16) - 1u) >> shift) + 1u;
    return px.x ? x0 | (x1 << 8) | (y0 << 16) | (y1 << 24) : 0u;
}

}  // namespace gs
