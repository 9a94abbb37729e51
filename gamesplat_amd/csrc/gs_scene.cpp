// gs_scene.cpp - seeded synthetic scenes and reference-view cameras (host).
//
// The generator draws the raw 3DGS PLY fields and activates them exactly as
// the reference's PLY loader does (src/gaussian_loader_ply.cpp:143-203):
// exp on scale, sigmoid on opacity, quaternion normalised when its length
// exceeds 1e-8, channel-major f_rest reordered into RGB triples.
//
// Random numbers: splitmix64, 53-bit uniforms, Box-Muller normals in double
// rounded to float.  Cluster centres come from the stream seeded by `seed`;
// Gaussian i draws from its own stream seeded by splitmix64 of
// seed + 0x632BE59BD9B4E019 * (i + 1), so generation is parallel and the
// first k Gaussians of a larger scene equal the k Gaussians of a smaller one
// with the same seed and cluster count.  Draw order per Gaussian: cluster
// index, position offset xyz, log-scale xyz, quaternion wxyz, opacity logit,
// DC rgb, f_rest_0..44 (PLY channel-major order).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/gs_api.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    bool has_spare = false;
    double spare = 0.0;
    double normal() {
        if (has_spare) { has_spare = false; return spare; }
        double u1 = 1.0 - uniform();  // (0, 1]
        double u2 = uniform();
        double r = sqrt(-2.0 * log(u1));
        double t = 6.283185307179586 * u2;
        spare = r * sin(t);
        has_spare = true;
        return r * cos(t);
    }
};

uint64_t mix64(uint64_t x) {
    SplitMix m{x};
    return m.next();
}

void generate_range(uint64_t seed, uint32_t begin, uint32_t end, uint32_t clusters, const float* centres,
                    float opacity_mean, GsGaussian* out) {
    for (uint32_t i = begin; i < end; ++i) {
        SplitMix r{mix64(seed + 0x632BE59BD9B4E019ull * ((uint64_t)i + 1ull))};
        uint32_t k = (uint32_t)(r.uniform() * clusters);
        if (k >= clusters) k = clusters - 1;
        float raw_pos[3], raw_scale[3], raw_rot[4], raw_op, raw_dc[3], raw_rest[45];
        for (int a = 0; a < 3; ++a) raw_pos[a] = (float)((double)centres[k * 3 + a] + 1.5 * r.normal());
        for (int a = 0; a < 3; ++a) raw_scale[a] = (float)(-4.5 + 0.8 * r.normal());
        for (int a = 0; a < 4; ++a) raw_rot[a] = (float)r.normal();
        raw_op = (float)((double)opacity_mean + 2.0 * r.normal());
        for (int a = 0; a < 3; ++a) raw_dc[a] = (float)(0.5 * r.normal());
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 15; ++j) {
                int l = j < 3 ? 1 : (j < 8 ? 2 : 3);
                raw_rest[c * 15 + j] = (float)(0.1 / (l + 1) * r.normal());
            }
        // activation as the PLY loader applies it
        GsGaussian& g = out[i];
        memcpy(g.position, raw_pos, sizeof(raw_pos));
        g.scale[0] = expf(raw_scale[0]);
        g.scale[1] = expf(raw_scale[1]);
        g.scale[2] = expf(raw_scale[2]);
        float rw = raw_rot[0], rx = raw_rot[1], ry = raw_rot[2], rz = raw_rot[3];
        float len = sqrtf(rw * rw + rx * rx + ry * ry + rz * rz);
        if (len > 1e-8f) { rw /= len; rx /= len; ry /= len; rz /= len; }
        g.rotation[0] = rw; g.rotation[1] = rx; g.rotation[2] = ry; g.rotation[3] = rz;
        g.opacity = 1.0f / (1.0f + expf(-raw_op));
        memcpy(g.color, raw_dc, sizeof(raw_dc));
        for (int j = 0; j < 15; ++j) {
            g.sh_rest[j * 3 + 0] = raw_rest[j];
            g.sh_rest[j * 3 + 1] = raw_rest[15 + j];
            g.sh_rest[j * 3 + 2] = raw_rest[30 + j];
        }
    }
}

bool normalize3(float v[3]) {
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len <= 1e-8f) return false;
    v[0] /= len; v[1] /= len; v[2] /= len;
    return true;
}

void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace

extern "C" int gs_generate_scene(uint64_t seed, uint32_t count, uint32_t clusters, float opacity_logit_mean,
                                 GsGaussian* out) {
    if (!out || clusters == 0) return GS_ERR_INVALID;
    if (count == 0) return GS_OK;
    std::vector<float> centres(clusters * 3);
    SplitMix r{seed};
    for (auto& c : centres) c = (float)(-8.0 + 16.0 * r.uniform());
    unsigned hw = std::thread::hardware_concurrency();
    unsigned threads = std::max(1u, std::min(hw ? hw : 1u, 16u));
    if (count < 65536) threads = 1;
    std::vector<std::thread> pool;
    uint32_t per = (count + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        uint32_t b = t * per, e = std::min(count, b + per);
        if (b >= e) break;
        pool.emplace_back(generate_range, seed, b, e, clusters, centres.data(), opacity_logit_mean, out);
    }
    for (auto& th : pool) th.join();
    return GS_OK;
}

// baked_camera_uniforms (src/main.cpp:165-193) with the projection of
// camera_get_proj_matrix (src/camera.cpp:217-245) for camera_init defaults
// (near 0.1, far 100, ortho size 1, perspective).  The fields the baked view
// does not set keep the fly camera's GL values (src/main.cpp:1268-1281):
// ortho_focal = height / 2, clip_y_sign = 1, clip_z_01 = 0.
extern "C" int gs_camera_from_pose(const float p[16], float focal_over_height, float width, float height,
                                   GsCameraUniforms* cam) {
    if (!p || !cam || !(focal_over_height > 0.0f) || !(width >= 1.0f) || !(height >= 1.0f)) return GS_ERR_INVALID;
    memset(cam, 0, sizeof(*cam));
    cam->viewport[0] = width;
    cam->viewport[1] = height;
    cam->ortho_focal = height / (2.0f * 1.0f);
    cam->clip_y_sign = 1.0f;
    cam->clip_z_01 = 0.0f;

    float back[3] = {p[2], p[6], p[10]}, up[3] = {p[1], p[5], p[9]}, right[3];
    normalize3(back);
    cross3(up, back, right);
    normalize3(right);
    cross3(back, right, up);
    for (int k = 0; k < 3; ++k) {
        cam->view[k * 4] = -right[k];
        cam->view[k * 4 + 1] = -up[k];
        cam->view[k * 4 + 2] = back[k];
        cam->view[12] += right[k] * p[k * 4 + 3];
        cam->view[13] += up[k] * p[k * 4 + 3];
        cam->view[14] -= back[k] * p[k * 4 + 3];
    }
    cam->view[15] = 1.0f;

    float fov_y = 2.0f * atanf(0.5f / focal_over_height);
    float aspect = cam->viewport[0] / cam->viewport[1];
    float n = 0.1f, fa = 100.0f, t = 0.0f;
    float persp[16] = {}, ortho[16] = {};
    float f = 1.0f / tanf(fov_y * 0.5f);
    persp[0] = -f / aspect;
    persp[5] = -f;
    persp[10] = (n + fa) / (n - fa);
    persp[11] = -1.0f;
    persp[14] = (2.0f * n * fa) / (n - fa);
    float half_h = 1.0f, half_w = half_h * aspect;
    ortho[0] = -1.0f / half_w;
    ortho[5] = -1.0f / half_h;
    ortho[10] = 2.0f / (n - fa);
    ortho[14] = (n + fa) / (n - fa);
    ortho[15] = 1.0f;
    for (int i = 0; i < 16; ++i) cam->proj[i] = persp[i] * (1.0f - t) + ortho[i] * t;
    cam->orthographic = 0.0f;
    cam->persp_focal = focal_over_height * cam->viewport[1];
    return GS_OK;
}

extern "C" int gs_look_at_origin_pose(const float position[3], float m[16]) {
    if (!position || !m) return GS_ERR_INVALID;
    float back[3] = {position[0], position[1], position[2]};
    if (!normalize3(back)) return GS_ERR_INVALID;
    float world_up[3] = {0.0f, 0.0f, 1.0f};
    float right[3];
    cross3(world_up, back, right);
    if (!normalize3(right)) {
        float alt[3] = {0.0f, 1.0f, 0.0f};
        cross3(alt, back, right);
        normalize3(right);
    }
    float up[3];
    cross3(back, right, up);
    const float cols[4][3] = {{right[0], right[1], right[2]}, {up[0], up[1], up[2]},
                              {back[0], back[1], back[2]}, {position[0], position[1], position[2]}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) m[r * 4 + c] = cols[c][r];
    m[12] = m[13] = m[14] = 0.0f;
    m[15] = 1.0f;
    return GS_OK;
}
