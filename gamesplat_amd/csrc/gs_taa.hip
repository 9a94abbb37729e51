// gs_taa.hip - temporal accumulation of StochasticSplats samples.
//
// Replaces taa_accum_fs (shaders/accum.glsl:47-179), the fullscreen pass of
// src/renderer.cpp:2425-2478.  One thread per pixel: the pixel's world
// position comes from the current depth and the inverse view, its
// reprojection into the previous view picks the history (bilinear over the
// four neighbouring texels, clamped to the image), and a valid history is
// blended with weight n = min(history alpha, 128).  The arithmetic is the
// shader's in source order (GLSL mix, column-major mat4 * vec4 summed left
// to right) so results match a CPU evaluation of the same expressions.
//
// Traffic per pixel: 16 B colour + 4 B depth in, 2 x 4 texels of 16 B
// history (neighbouring threads share them through the caches, so about
// 32 B from HBM), 32 B out: ~84 B, an HBM-bound elementwise pass.
#include "gs_internal.h"
#include "gs_math.h"

namespace gs {

namespace {

// GLSL mat4 * vec4 for a column-major matrix.
__device__ __forceinline__ float4 mat4_mul(const float* m, float x, float y, float z, float w) {
    float4 o;
    o.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
    o.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
    o.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
    o.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
    return o;
}

__device__ __forceinline__ float4 texel(const float4* img, int w, int h, int x, int y) {
    x = min(max(x, 0), w - 1);
    y = min(max(y, 0), h - 1);
    return img[(size_t)y * w + x];
}

__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
    return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}

// sample_history_*_bilinear (shaders/accum.glsl:120-152).
__device__ __forceinline__ float4 bilinear(const float4* img, int w, int h, float u, float v) {
    const float px = u * (float)w - 0.5f, py = v * (float)h - 0.5f;
    const float flx = __builtin_floorf(px), fly = __builtin_floorf(py);
    const int x0 = (int)flx, y0 = (int)fly;
    const float fx = px - flx, fy = py - fly;
    const float4 c00 = texel(img, w, h, x0, y0), c10 = texel(img, w, h, x0 + 1, y0);
    const float4 c01 = texel(img, w, h, x0, y0 + 1), c11 = texel(img, w, h, x0 + 1, y0 + 1);
    return mix4(mix4(c00, c10, fx), mix4(c01, c11, fx), fy);
}

__global__ void __launch_bounds__(256) taa_kernel(TaaLaunch a) {
    const int i = (int)(blockIdx.x * 16u + (threadIdx.x & 15u));
    const int j = (int)(blockIdx.y * 16u + (threadIdx.x >> 4));
    const int W = a.width, H = a.height;
    if (i >= W || j >= H) return;
    const GsTaaUniforms& u = a.u;
    const float* vo = u.viewport_orthographic;
    const float* fc = u.focal_clip;
    const float* pfc = u.prev_focal_clip;
    const size_t p = (size_t)j * W + i;
    const float uvx = ((float)i + 0.5f) / (float)W, uvy = ((float)j + 0.5f) / (float)H;
    const float4 cc = a.color[p];
    const float depth = a.depth[p];

    // reconstruct_view_pos
    const float z_ndc = depth * 2.0f - 1.0f;
    const float view_z = -u.proj_z[1] / (z_ndc + u.proj_z[0]);
    const float pxx = uvx * vo[0];
    float pxy = uvy * vo[1];
    const float unflip = (1.0f - fc[2]) * 0.5f;
    pxy = mixf(pxy, vo[1] - pxy, unflip);
    const float cx = pxx - vo[0] * 0.5f, cy = pxy - vo[1] * 0.5f;
    const float denom = mixf(fc[0] / view_z, -fc[1], vo[2]);
    const float4 world = mat4_mul(u.inv_view, cx / denom, cy / denom, view_z, 1.0f);

    // project_prev_uv
    const float4 pv = mat4_mul(u.prev_view, world.x, world.y, world.z, 1.0f);
    float pu = -1.0f, pvv = -1.0f;
    if (!(pv.z >= -0.001f)) {
        const float t = vo[3];
        const float ppx = pfc[0] * pv.x / pv.z + vo[0] * 0.5f;
        const float ppy = vo[1] * 0.5f + pfc[0] * pv.y / pv.z;
        const float opx = -pfc[1] * pv.x + vo[0] * 0.5f;
        const float opy = vo[1] * 0.5f - pfc[1] * pv.y;
        const float qx = mixf(ppx, opx, t);
        float qy = mixf(ppy, opy, t);
        const float y_flip = (1.0f - pfc[2]) * 0.5f;
        qy = mixf(qy, vo[1] - qy, y_flip);
        pu = qx / vo[0];
        pvv = qy / vo[1];
    }
    bool valid = u.history_valid != 0.0f && pu >= 0.0f && pu <= 1.0f && pvv >= 0.0f && pvv <= 1.0f;
    float4 hc = make_float4(0.0f, 0.0f, 0.0f, 0.0f), hx = hc;
    if (valid) {
        hc = bilinear(a.hist_color, W, H, pu, pvv);
        hx = bilinear(a.hist_xyz, W, H, pu, pvv);
        const float dx = world.x - hx.x, dy = world.y - hx.y, dz = world.z - hx.z;
        valid = hc.w > 0.0f && (u.view_changed == 0.0f || __builtin_sqrtf((dx * dx + dy * dy) + dz * dz) < 100.0f);
    }
    if (valid) {
        const float n = minf(hc.w, 128.0f);
        const float d = n + 1.0f;
        a.out_color[p] = make_float4((hc.x * n + cc.x) / d, (hc.y * n + cc.y) / d, (hc.z * n + cc.z) / d,
                                     minf(n + 1.0f, 128.0f));
        a.out_xyz[p] = make_float4((hx.x * n + world.x) / d, (hx.y * n + world.y) / d, (hx.z * n + world.z) / d,
                                   1.0f);
    } else {
        a.out_color[p] = make_float4(cc.x, cc.y, cc.z, 1.0f);
        a.out_xyz[p] = make_float4(world.x, world.y, world.z, 1.0f);
    }
}

}  // namespace

hipError_t launch_taa(const TaaLaunch& a, hipStream_t s) {
    if (a.width <= 0 || a.height <= 0) return hipSuccess;
    const dim3 grid((unsigned)(a.width + 15) / 16u, (unsigned)(a.height + 15) / 16u);
    hipLaunchKernelGGL(taa_kernel, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gs
