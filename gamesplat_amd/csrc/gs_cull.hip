// gs_cull.hip - per-Gaussian cull, EWA projection and SH colour.
//
// Replaces cull_reset_cs and cull_cs (shaders/cull.glsl:166-474).  One
// thread per Gaussian, 256-thread blocks, up to 64 views per launch.  For
// each view the cull writes the projected splat, a dense depth key
// (0xFFFFFFFF when culled, as the reset leaves it) and a per-block visible
// count; after a scan of those counts compact_kernel moves the visible
// (key, id) pairs to one list in index order per view, the order a stable
// sort over the reference's sentinel-padded arrays keeps.
//
// Float policy: the library is built with -ffp-contract=off and every
// expression below keeps the operand order of the GLSL source, so each
// value is the same IEEE binary32 result the CPU restatement computes.
#include <hip/hip_fp16.h>

#include "gs_internal.h"
#include "gs_math.h"
#include "gs_raster.h"

namespace gs {

namespace {

struct M3 { float m[3][3]; };  // m[column][row], as GLSL

__device__ __forceinline__ M3 mul3(const M3& a, const M3& b) {
    M3 o;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            o.m[c][r] = (a.m[0][r] * b.m[c][0] + a.m[1][r] * b.m[c][1]) + a.m[2][r] * b.m[c][2];
    return o;
}

__device__ __forceinline__ M3 transpose3(const M3& a) {
    M3 o;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) o.m[c][r] = a.m[r][c];
    return o;
}

__device__ __forceinline__ float hash11(float p) { return fractf(sinf_(p * 127.1f) * 43758.5453123f); }

constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f;
constexpr float kShC2_1 = -1.0925484305920792f;
constexpr float kShC2_2 = 0.31539156525252005f;
constexpr float kShC2_3 = -1.0925484305920792f;
constexpr float kShC2_4 = 0.5462742152960396f;
constexpr float kShC3_0 = -0.5900435899266435f;
constexpr float kShC3_1 = 2.890611442640554f;
constexpr float kShC3_2 = -0.4570457994644658f;
constexpr float kShC3_3 = 0.3731763325901154f;
constexpr float kShC3_4 = -0.4570457994644658f;
constexpr float kShC3_5 = 1.445305721320277f;
constexpr float kShC3_6 = -0.5900435899266435f;

// SH result of one channel before the +0.5 offset (cull.glsl:354-408); sh
// holds the rest coefficients as RGB triples (coefficient k, channel c at
// 3 k + c).
__device__ __forceinline__ float sh_result(const float* dc, const float* sh, int ch, int eff, float dx, float dy,
                                           float dz) {
    float res = kShC0 * dc[ch];
    if (eff >= 1) {
        res += -kShC1 * dy * sh[0 * 3 + ch];
        res += kShC1 * dz * sh[1 * 3 + ch];
        res += -kShC1 * dx * sh[2 * 3 + ch];
        if (eff >= 2) {
            float xx = dx * dx, yy = dy * dy, zz = dz * dz;
            float xy = dx * dy, yz = dy * dz, xz = dz * dx;
            res += kShC2_0 * xy * sh[3 * 3 + ch];
            res += kShC2_1 * yz * sh[4 * 3 + ch];
            res += kShC2_2 * (2.0f * zz - xx - yy) * sh[5 * 3 + ch];
            res += kShC2_3 * xz * sh[6 * 3 + ch];
            res += kShC2_4 * (xx - yy) * sh[7 * 3 + ch];
            if (eff >= 3) {
                res += kShC3_0 * dy * (3.0f * xx - yy) * sh[8 * 3 + ch];
                res += kShC3_1 * xy * dz * sh[9 * 3 + ch];
                res += kShC3_2 * dy * (4.0f * zz - xx - yy) * sh[10 * 3 + ch];
                res += kShC3_3 * dz * (2.0f * zz - 3.0f * xx - 3.0f * yy) * sh[11 * 3 + ch];
                res += kShC3_4 * dx * (4.0f * zz - xx - yy) * sh[12 * 3 + ch];
                res += kShC3_5 * dz * (xx - yy) * sh[13 * 3 + ch];
                res += kShC3_6 * dx * (xx - 3.0f * yy) * sh[14 * 3 + ch];
            }
        }
    }
    return res;
}

// Channel-major GSAnim coefficient index of rest value r = 3 k + c:
// 3 + 15 c + k (cull.glsl:141); DC channel c is coefficient c.
__host__ __device__ constexpr int rest_coef(int r) { return 3 + (r % 3) * 15 + r / 3; }

// Signed byte j of a 48-byte delta row held in three uint4.
__device__ __forceinline__ int delta_byte(const uint4* q, int j) {
    const uint4 v = q[j / 16];
    const int w = (j % 16) / 4;
    const uint32_t word = w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w));
    return (int)(int8_t)(uint8_t)((word >> ((j % 4) * 8)) & 255u);
}

// Half coefficient j of a 96-byte layer row held in six uint4.
__device__ __forceinline__ float layer_half(const uint4* h, int j) {
    const uint4 v = h[j / 8];
    const int w = (j % 8) / 2;
    const uint32_t word = w == 0 ? v.x : (w == 1 ? v.y : (w == 2 ? v.z : v.w));
    return __half2float(__ushort_as_half((unsigned short)((word >> ((j % 2) * 16)) & 65535u)));
}

// animation_coefficient + the state lerp (cull.glsl:125-132,150-157).
__device__ __forceinline__ float anim_coef(const CullLaunch& a, float base, const uint4* qa, const uint4* qb, int j) {
    float va = a.anim_state_a != 0 ? base + (float)delta_byte(qa, j) * a.anim_scales[j] : base;
    if (a.anim_weight > 0.0f) {
        const float vb = a.anim_state_b != 0 ? base + (float)delta_byte(qb, j) * a.anim_scales[j] : base;
        va = (1.0f - a.anim_weight) * va + a.anim_weight * vb;
    }
    return va;
}

// One thread per Gaussian, projected into every view of the launch.  The
// Gaussian's planes are read once: P0 (and P1, P2 under the effect) for the
// visibility tests of all views, then P1-P3 and the SH planes the deepest
// visible view needs.  The 3D covariance does not depend on the view and is
// formed once; projection, SH colour, record, pixel rectangle and dense key
// are per view, with the operations of the single-view cull.  ANIM selects
// the GSAnim appearance branch: 1 applies the state deltas to every
// coefficient once, 2 sums the clamped SH of each coefficient layer per view.
template <int ANIM>
__global__ void __launch_bounds__(256) cull_kernel(CullLaunch a) {
    __shared__ uint32_t s_cnt[kMaxViews][4];
    __shared__ uint32_t s_diag[8];
    __shared__ uint32_t s_klo[kMaxViews][4], s_khi[kMaxViews][4];  // visible key range per view and wave

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid < 8) s_diag[tid] = 0u;
    __syncthreads();
    const uint32_t block = blockIdx.x;
    const uint32_t idx = block * 256u + tid;
    const uint32_t n = a.n;
    const uint32_t views = a.views;

    uint64_t vis = 0ull;
    float position[3] = {0.0f, 0.0f, 0.0f};
    float opacity = 0.0f;
    float effect_wave = 0.0f;
    float4 p1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), p2 = p1, p3 = p1;
    typedef float f32x4v_t __attribute__((ext_vector_type(4)));
    f32x4v_t shp[11];
    float scale[3] = {0.0f, 0.0f, 0.0f};
    bool loaded = false;
    // Without an SH LOD distance (or layer appearance) every visible
    // Gaussian needs the planes of the cull's SH degree: they are loaded
    // together with P0, before the visibility test, so a Gaussian costs one
    // memory round trip (the few culled ones read planes they do not use).
    const bool prefetch = ANIM != 2 && !(a.sh_lod_distance > 0.0f);
    const int prefetch_planes = a.sh_degree <= 0 ? 0 : (a.sh_degree == 1 ? 2 : (a.sh_degree == 2 ? 6 : 11));
    if (idx < n) {
        float4 p0 = a.planes[idx];
        if (prefetch) {
            p1 = a.planes[(size_t)1 * n + idx];
            p2 = a.planes[(size_t)2 * n + idx];
            p3 = a.planes[(size_t)3 * n + idx];
#pragma unroll
            for (int j = 0; j < 11; ++j)
                if (j < prefetch_planes) shp[j] = reinterpret_cast<const f32x4v_t*>(a.planes)[(size_t)(4 + j) * n + idx];
        }
        position[0] = p0.x; position[1] = p0.y; position[2] = p0.z;
        opacity = p0.w;

        if (a.effect_active) {
            // shockwave (cull.glsl:183-207)
            const GsSplatEffectParams& fx = a.effect;
            if (!prefetch) {
                p1 = a.planes[(size_t)1 * n + idx];
                p2 = a.planes[(size_t)2 * n + idx];
            }
            loaded = true;
            scale[0] = p2.x; scale[1] = p2.y; scale[2] = p2.z;
            float duration = maxf(fx.params[1], 0.001f);
            float progress = clampf(fx.params[0] / duration, 0.0f, 1.0f);
            float scene_radius = maxf(fx.center_radius[3], 0.001f);
            float to_splat[3], jitter[3];
            for (int k = 0; k < 3; ++k) to_splat[k] = position[k] - fx.center_radius[k];
            float jscale = scene_radius * 0.015f;
            float fidx = (float)idx;
            for (int k = 0; k < 3; ++k) jitter[k] = (hash11(fidx + (float)(k + 1)) * 2.0f - 1.0f) * jscale;
            float dist = __builtin_sqrtf((to_splat[0] * to_splat[0] + to_splat[1] * to_splat[1]) + to_splat[2] * to_splat[2]);
            float wave_radius = progress * scene_radius * 1.15f;
            float band_width = scene_radius * 0.085f;
            float band = (dist - wave_radius) / band_width;
            float wave = expf_(-band * band);
            float attack = smoothstepf(0.0f, 0.08f, progress);
            float release = 1.0f - smoothstepf(0.82f, 1.0f, progress);
            effect_wave = wave * attack * release;
            float d[3];
            d[0] = (to_splat[0] + jitter[0]) + 0.0f;
            d[1] = (to_splat[1] + jitter[1]) + scene_radius * 0.01f;
            d[2] = (to_splat[2] + jitter[2]) + 0.0f;
            float inv = 1.0f / __builtin_sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
            for (int k = 0; k < 3; ++k) d[k] = d[k] * inv;
            float signed_ripple = sinf_(progress * 18.849556f + dist * 8.0f / scene_radius);
            float push = effect_wave * fx.params[2] * scene_radius;
            for (int k = 0; k < 3; ++k) position[k] = position[k] + d[k] * push;
            float ripple = effect_wave * signed_ripple * 0.012f * scene_radius;
            for (int k = 0; k < 3; ++k) position[k] = position[k] + d[k] * ripple;
            float sgrow = 1.0f + effect_wave * 0.75f;
            for (int k = 0; k < 3; ++k) scale[k] = scale[k] * sgrow;
            opacity = opacity * (1.0f - effect_wave * 0.20f);
        }

        // visibility in each view (cull.glsl:209-220)
        for (uint32_t v = 0; v < views; ++v) {
            const GsCameraUniforms& cam = a.cams[v];
            const float* V = cam.view;
            const float* P = cam.proj;
            float pv[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
                pv[r] = ((V[0 * 4 + r] * position[0] + V[1 * 4 + r] * position[1]) + V[2 * 4 + r] * position[2]) + V[3 * 4 + r] * 1.0f;
            bool visible = pv[2] <= -0.2f;
            if (visible) {
                float inv_z = -1.0f / pv[2];
                float o = cam.orthographic;
                float ndc_x = (P[0] * pv[0]) * inv_z * (1.0f - o) + P[0] * pv[0] * o;
                float ndc_y = (P[5] * pv[1]) * inv_z * (1.0f - o) + P[5] * pv[1] * o;
                visible = __builtin_fabsf(ndc_x) <= 1.3f && __builtin_fabsf(ndc_y) <= 1.3f;
            }
            if (visible) vis |= 1ull << v;
        }
    }

    // view-independent part: the remaining planes and the 3D covariance
    // (cull.glsl:234-254); SH planes only up to the degree the deepest
    // visible view evaluates (2 planes for degree 1, 6 for 2, 11 for 3)
    float dc[3] = {0.0f, 0.0f, 0.0f};
    M3 cov3d;
    float sh[46];
    if (vis) {
        if (!loaded) {
            if (!prefetch) {
                p1 = a.planes[(size_t)1 * n + idx];
                p2 = a.planes[(size_t)2 * n + idx];
            }
            scale[0] = p2.x; scale[1] = p2.y; scale[2] = p2.z;
        }
        if (!prefetch) p3 = a.planes[(size_t)3 * n + idx];
        float qw = p1.x, qx = p1.y, qy = p1.z, qz = p1.w;
        dc[0] = p2.w; dc[1] = p3.x; dc[2] = p3.y;
        M3 R;
        R.m[0][0] = 1.0f - 2.0f * (qy * qy + qz * qz);
        R.m[0][1] = 2.0f * (qx * qy + qw * qz);
        R.m[0][2] = 2.0f * (qx * qz - qw * qy);
        R.m[1][0] = 2.0f * (qx * qy - qw * qz);
        R.m[1][1] = 1.0f - 2.0f * (qx * qx + qz * qz);
        R.m[1][2] = 2.0f * (qy * qz + qw * qx);
        R.m[2][0] = 2.0f * (qx * qz + qw * qy);
        R.m[2][1] = 2.0f * (qy * qz - qw * qx);
        R.m[2][2] = 1.0f - 2.0f * (qx * qx + qy * qy);
        M3 S;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) S.m[c][r] = 0.0f;
        S.m[0][0] = scale[0]; S.m[1][1] = scale[1]; S.m[2][2] = scale[2];
        M3 M = mul3(R, S);
        cov3d = mul3(M, transpose3(M));

        int need = 0;
        for (uint32_t v = 0; v < views; ++v) {
            if (!((vis >> v) & 1ull)) continue;
            int eff = a.sh_degree;
            if (a.sh_lod_distance > 0.0f) {
                const float* V = a.cams[v].view;
                float z = ((V[2] * position[0] + V[6] * position[1]) + V[10] * position[2]) + V[14] * 1.0f;
                if ((-z) > a.sh_lod_distance) eff = 0;
            }
            need = eff > need ? eff : need;
        }
        const int planes_needed = ANIM == 2 ? 0 : (need <= 0 ? 0 : (need == 1 ? 2 : (need == 2 ? 6 : 11)));
        sh[0] = p3.z; sh[1] = p3.w;
#pragma unroll
        for (int j = 0; j < 11; ++j) {
            float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            // with prefetch, planes_needed == prefetch_planes
            if (j < planes_needed)
                t = prefetch ? make_float4(shp[j].x, shp[j].y, shp[j].z, shp[j].w) : a.planes[(size_t)(4 + j) * n + idx];
            sh[2 + 4 * j] = t.x;
            sh[3 + 4 * j] = t.y;
            sh[4 + 4 * j] = t.z;
            sh[5 + 4 * j] = t.w;
        }
        if constexpr (ANIM == 1) {
            // the 48 deltas of this row in state_a and state_b
            uint4 qa[3], qb[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) qa[k] = qb[k] = make_uint4(0u, 0u, 0u, 0u);
            if (a.anim_state_a != 0) {
                const uint4* r = a.anim_data + ((size_t)(a.anim_state_a - 1) * n + idx) * 3u;
#pragma unroll
                for (int k = 0; k < 3; ++k) qa[k] = r[k];
            }
            if (a.anim_weight > 0.0f && a.anim_state_b != 0) {
                const uint4* r = a.anim_data + ((size_t)(a.anim_state_b - 1) * n + idx) * 3u;
#pragma unroll
                for (int k = 0; k < 3; ++k) qb[k] = r[k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) dc[c] = anim_coef(a, dc[c], qa, qb, c);
            const int loaded = 2 + 4 * planes_needed;
#pragma unroll
            for (int r = 0; r < 45; ++r)
                if (r < loaded) sh[r] = anim_coef(a, sh[r], qa, qb, rest_coef(r));
        }
    }

    for (uint32_t v = 0; v < views; ++v) {
        const bool visible = (vis >> v) & 1ull;
        uint32_t key = 0xFFFFFFFFu;
        const size_t slot = (size_t)v * n + idx;
        if (visible) {
            const GsCameraUniforms& cam = a.cams[v];
            const float* V = cam.view;
            const float* P = cam.proj;
            float pv[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
                pv[r] = ((V[0 * 4 + r] * position[0] + V[1 * 4 + r] * position[1]) + V[2 * 4 + r] * position[2]) + V[3 * 4 + r] * 1.0f;
            key = 0xFFFFFFFFu - fbits(maxf(-pv[2], 0.0f));

            // EWA projection (cull.glsl:256-277)
            float z = pv[2], x = pv[0], y = pv[1];
            float fx_p = cam.persp_focal, fy_p = cam.persp_focal;
            float fx_o = cam.ortho_focal, fy_o = cam.ortho_focal;
            float ortho = cam.orthographic;
            float J00 = mixf(fx_p / z, -fx_o, ortho);
            float J11 = mixf(fy_p / z, -fy_o, ortho);
            float J02 = mixf(-fx_p * x / (z * z), 0.0f, ortho);
            float J12 = mixf(-fy_p * y / (z * z), 0.0f, ortho);
            M3 W;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) W.m[c][r] = V[c * 4 + r];
            M3 Wt = transpose3(W);
            M3 WcW = mul3(mul3(W, cov3d), Wt);
            float ca = J00 * J00 * WcW.m[0][0] + 2.0f * J00 * J02 * WcW.m[0][2] + J02 * J02 * WcW.m[2][2];
            float cb = J00 * J11 * WcW.m[0][1] + J00 * J12 * WcW.m[0][2] + J02 * J11 * WcW.m[1][2] + J02 * J12 * WcW.m[2][2];
            float cc = J11 * J11 * WcW.m[1][1] + 2.0f * J11 * J12 * WcW.m[1][2] + J12 * J12 * WcW.m[2][2];
            ca += 0.3f;
            cc += 0.3f;
            float det = ca * cc - cb * cb;
            if (det < 1e-6f) det = 1e-6f;

            // raster centre (cull.glsl:279-291)
            float pcx = fx_p * x / z + cam.viewport[0] * 0.5f;
            float pcy = cam.viewport[1] * 0.5f + fy_p * y / z;
            float ocx = -fx_o * x + cam.viewport[0] * 0.5f;
            float ocy = cam.viewport[1] * 0.5f - fy_o * y;
            float cx = mixf(pcx, ocx, ortho);
            float cy = mixf(pcy, ocy, ortho);
            float y_flip = (1.0f - cam.clip_y_sign) * 0.5f;
            float rcy = mixf(cy, cam.viewport[1] - cy, y_flip);
            float radius_x = __builtin_ceilf(3.0f * __builtin_sqrtf(ca));
            float radius_y = __builtin_ceilf(3.0f * __builtin_sqrtf(cc));

            // diagnostics (cull.glsl:296-318), single-view launches only
            if (a.collect_stats) {
                float quad_area = maxf(0.0f, 4.0f * radius_x * radius_y);
                uint32_t qpx = f2u_sat(minf(quad_area, 4294967295.0f));
                float gps_points = 6.28318530718f * __builtin_sqrtf(maxf(det, 0.0f)) * dilog(clampf(opacity, 0.0f, 1.0f));
                uint32_t gpc = f2u_sat(minf(__builtin_ceilf(gps_points), 4294967295.0f));
                atomicAdd(&s_diag[0], (qpx + 1023u) >> 10u);
                atomicMax(&s_diag[1], qpx);
                if (qpx > 1024u) atomicAdd(&s_diag[2], 1u);
                if (qpx > 16384u) atomicAdd(&s_diag[3], 1u);
                atomicAdd(&s_diag[4], (gpc + 1023u) >> 10u);
                atomicMax(&s_diag[5], gpc);
                if (gpc > 1024u) atomicAdd(&s_diag[6], 1u);
                if (gpc > 16384u) atomicAdd(&s_diag[7], 1u);
            }

            // ndc depth (cull.glsl:320-321)
            float ndc_z = (P[2 * 4 + 2] * z + P[3 * 4 + 2]) / (-z);
            ndc_z = mixf(ndc_z, ndc_z * 0.5f + 0.5f, cam.clip_z_01);

            // view direction (cull.glsl:323-324)
            float cpos[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
                cpos[r] = -((Wt.m[0][r] * V[12] + Wt.m[1][r] * V[13]) + Wt.m[2][r] * V[14]);
            float dv[3] = {position[0] - cpos[0], position[1] - cpos[1], position[2] - cpos[2]};
            float dinv = 1.0f / __builtin_sqrtf((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]);
            float dx = dv[0] * dinv, dy = dv[1] * dinv, dz = dv[2] * dinv;

            // SH (cull.glsl:348-416)
            int eff = a.sh_degree;
            if (a.sh_lod_distance > 0.0f && (-z) > a.sh_lod_distance) eff = 0;
            float color[3];
            if constexpr (ANIM == 2) {
                // coefficient layers: sum of max(SH, 0) x RGB weight, no offset
                color[0] = color[1] = color[2] = 0.0f;
                for (int l = 0; l < a.anim_layers; ++l) {
                    const uint4* row = a.anim_data + ((size_t)l * n + idx) * 6u;
                    uint4 h[6];
#pragma unroll
                    for (int k = 0; k < 6; ++k) h[k] = row[k];
                    float ldc[3], lsh[45];
#pragma unroll
                    for (int c = 0; c < 3; ++c) ldc[c] = layer_half(h, c);
#pragma unroll
                    for (int r = 0; r < 45; ++r) lsh[r] = layer_half(h, rest_coef(r));
#pragma unroll
                    for (int ch = 0; ch < 3; ++ch) {
                        const float res = sh_result(ldc, lsh, ch, eff, dx, dy, dz);
                        color[ch] = color[ch] + maxf(res, 0.0f) * a.layer_weights[l * 3 + ch];
                    }
                }
            } else {
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) {
                    float res = sh_result(dc, sh, ch, eff, dx, dy, dz);
                    res += 0.5f;
                    color[ch] = maxf(res, 0.0f) * 1.0f;
                }
            }
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                float ecol = a.effect_active ? a.effect.color[ch] : 0.0f;
                float ea = a.effect_active ? a.effect.color[3] : 0.0f;
                color[ch] += ecol * (effect_wave * ea);
            }

            float4* rec = a.records + slot * kRecordVec4;
            rec[0] = make_float4(cx, rcy, opacity, ndc_z);
            rec[1] = make_float4(ca, cb * cam.clip_y_sign, cc, radius_x);
            rec[2] = make_float4(color[0], color[1], color[2], radius_y);
            // the rasteriser drops a quad outside the clip volume's depth
            // range (far plane): it covers no pixel, so it gets no tile
            const bool drawn = !depth_clipped(ndc_z, cam.clip_z_01);
            if (a.rect_px)
                a.rect_px[slot] = drawn ? pixel_rect(cx, rcy, ca, cb * cam.clip_y_sign, cc, radius_x, radius_y,
                                                     opacity, (int)cam.viewport[0], (int)cam.viewport[1],
                                                     a.tight_rects != 0)
                                        : make_uint2(0u, 0u);
            if (a.packed_rects)
                a.packed_rects[slot] = drawn ? pack_tile_rect(
                    pixel_rect(cx, rcy, ca, cb * cam.clip_y_sign, cc, radius_x, radius_y, opacity,
                               (int)cam.viewport[0], (int)cam.viewport[1], a.tight_rects != 0, 0.0f, 0.0f,
                               a.packed_pad),
                    a.packed_shift) : 0u;
        }
        // dense key (culled = 0xFFFFFFFF, as cull_reset_cs leaves it) and the
        // block's visible count; compaction runs after the counts are scanned
        if (idx < n) a.dense_keys[slot] = key;
        const uint64_t bal = __ballot(visible);
        if (lane == 0) s_cnt[v][wave] = (uint32_t)__popcll(bal);
        if (a.block_range) {
            uint32_t lo = visible ? key : 0xFFFFFFFFu, hi = visible ? key : 0u;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                lo = min(lo, (uint32_t)__shfl_xor((int)lo, d, 64));
                hi = max(hi, (uint32_t)__shfl_xor((int)hi, d, 64));
            }
            if (lane == 0) {
                s_klo[v][wave] = lo;
                s_khi[v][wave] = hi;
            }
        }
    }
    __syncthreads();
    for (uint32_t v = tid; v < views; v += 256u) {
        a.block_counts[(size_t)v * gridDim.x + block] = s_cnt[v][0] + s_cnt[v][1] + s_cnt[v][2] + s_cnt[v][3];
        if (a.block_range)
            a.block_range[(size_t)v * gridDim.x + block] =
                make_uint2(min(min(s_klo[v][0], s_klo[v][1]), min(s_klo[v][2], s_klo[v][3])),
                           max(max(s_khi[v][0], s_khi[v][1]), max(s_khi[v][2], s_khi[v][3])));
    }
    if (a.collect_stats && tid < 8) a.block_diag[(size_t)block * 8 + tid] = s_diag[tid];
}

// Order-preserving compaction: block (b, v) moves the visible entries of its
// 256 Gaussians in view v to offsets[v * blocks + b] + rank.  The value is
// v * n + index, which addresses the view's records and rectangles, and the
// list holds view 0's visible Gaussians in index order, then view 1's, ...
// With vis_rects, the splat's pixel rectangle goes along as its tile
// rectangle at tile size 1 << tile_shift, packed into 8-bit fields (PACKED:
// the cull's packed rectangle is moved as it is).
template <bool PACKED>
__global__ void __launch_bounds__(256) compact_kernel(const uint32_t* __restrict__ dense_keys, uint32_t n,
                                                      const uint32_t* __restrict__ offsets, uint32_t* vis_keys,
                                                      uint32_t* vis_ids, const uint2* __restrict__ rect_px,
                                                      int tile_shift, uint32_t* vis_rects,
                                                      const uint32_t* __restrict__ packed) {
    __shared__ uint32_t s_scan[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t v = blockIdx.y;
    const uint32_t idx = blockIdx.x * 256u + tid;
    const size_t slot = (size_t)v * n + idx;
    const uint32_t key = idx < n ? dense_keys[slot] : 0xFFFFFFFFu;
    // the rectangle is loaded with the key (one memory round trip), and used
    // only for visible splats
    uint2 px = make_uint2(0u, 0u);
    uint32_t pre = 0u;
    if (vis_rects && idx < n) {
        if constexpr (PACKED) pre = packed[slot];
        else px = rect_px[slot];
    }
    const uint32_t block_base = offsets[(size_t)v * gridDim.x + blockIdx.x];
    const bool visible = key != 0xFFFFFFFFu;
    const uint64_t bal = __ballot(visible);
    if (lane == 0) s_scan[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w)
        if ((uint32_t)w < wave) before += s_scan[w];
    if (visible) {
        const uint32_t pos = block_base + before +
                             (uint32_t)__popcll(bal & lanemask_lt());
        if (vis_keys) vis_keys[pos] = key;
        vis_ids[pos] = (uint32_t)slot;
        if (vis_rects) vis_rects[pos] = PACKED ? pre : pack_tile_rect(px, tile_shift);
    }
}

// Box of the Gaussian centres (plane P0): block b reduces the centres
// b, b + grid, ... into out[8 b ..]: min x, y, z, max x, y, z, and 1 when a
// coordinate is not finite (else 0).
__global__ void __launch_bounds__(256) centre_bounds_kernel(const float4* __restrict__ p0, uint32_t n, float* out) {
    __shared__ float s_v[7][4];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const float inf = __builtin_inff();
    float v[7] = {inf, inf, inf, -inf, -inf, -inf, 0.0f};
    for (uint32_t i = blockIdx.x * 256u + tid; i < n; i += gridDim.x * 256u) {
        const float4 q = p0[i];
        if (!(__builtin_isfinite(q.x) && __builtin_isfinite(q.y) && __builtin_isfinite(q.z))) v[6] = 1.0f;
        v[0] = fminf(v[0], q.x); v[1] = fminf(v[1], q.y); v[2] = fminf(v[2], q.z);
        v[3] = fmaxf(v[3], q.x); v[4] = fmaxf(v[4], q.y); v[5] = fmaxf(v[5], q.z);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = fminf(v[k], __shfl_xor(v[k], d, 64));
#pragma unroll
        for (int k = 3; k < 7; ++k) v[k] = fmaxf(v[k], __shfl_xor(v[k], d, 64));
    }
    if (lane == 0u)
        for (int k = 0; k < 7; ++k) s_v[k][wave] = v[k];
    __syncthreads();
    if (tid < 8u) {
        float r = 0.0f;
        if (tid < 3u) r = fminf(fminf(s_v[tid][0], s_v[tid][1]), fminf(s_v[tid][2], s_v[tid][3]));
        else if (tid < 7u) r = fmaxf(fmaxf(s_v[tid][0], s_v[tid][1]), fmaxf(s_v[tid][2], s_v[tid][3]));
        out[blockIdx.x * 8u + tid] = r;
    }
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

}  // namespace

hipError_t launch_cull(const CullLaunch& a, hipStream_t s) {
    if (a.n == 0 || a.views == 0) return hipSuccess;
    if (a.views > (uint32_t)kMaxViews) return hipErrorInvalidValue;
    uint32_t blocks = (a.n + 255u) / 256u;
    if (a.anim == 1)
        hipLaunchKernelGGL(cull_kernel<1>, dim3(blocks), dim3(256), 0, s, a);
    else if (a.anim == 2)
        hipLaunchKernelGGL(cull_kernel<2>, dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(cull_kernel<0>, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_compact(const uint32_t* dense_keys, uint32_t n, uint32_t views, const uint32_t* offsets,
                          uint32_t* vis_keys, uint32_t* vis_ids, const uint2* rect_px, int pack_tile,
                          uint32_t* vis_rects, hipStream_t s, const uint32_t* packed) {
    if (n == 0 || views == 0) return hipSuccess;
    int shift = 0;
    if (pack_tile) {
        if ((!rect_px && !packed) || !vis_rects || (pack_tile != 8 && pack_tile != 16 && pack_tile != 32))
            return hipErrorInvalidValue;
        while ((1 << shift) < pack_tile) ++shift;
    } else {
        vis_rects = nullptr;
        packed = nullptr;
    }
    if (packed)
        hipLaunchKernelGGL(compact_kernel<true>, dim3((n + 255u) / 256u, views), dim3(256), 0, s, dense_keys, n,
                           offsets, vis_keys, vis_ids, rect_px, shift, vis_rects, packed);
    else
        hipLaunchKernelGGL(compact_kernel<false>, dim3((n + 255u) / 256u, views), dim3(256), 0, s, dense_keys, n,
                           offsets, vis_keys, vis_ids, rect_px, shift, vis_rects, packed);
    return hipGetLastError();
}

hipError_t launch_centre_bounds(const float4* p0, uint32_t n, float* partials, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(centre_bounds_kernel, dim3(kCentreBoundBlocks), dim3(256), 0, s, p0, n, partials);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, count);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t value, size_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, value, count);
    return hipGetLastError();
}

}  // namespace gs
