// gs_loaders.cpp - per-Gaussian decode of the scene formats (host).
//
// The container work (PLY header, ZIP archive, lossless WebP images, JSON
// metadata) is left to the caller; these functions do the arithmetic the
// reference loaders apply to each Gaussian, with the same float operations
// and libm functions, so the same input gives the same Gaussians.
//   gs_decode_ply   src/gaussian_loader_ply.cpp:143-203
//   gs_decode_sog   src/gaussian_loader_sog.cpp:403-510
// Each also writes straight into the device scene's plane layout, for the
// pipelined uploads (gs_upload_ply, gs_upload_sog) and gs_decode_ply_planes.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/gs_api.h"
#include "gs_planes.h"

namespace {

// Runs fn(begin, end) over [0, count) on the host's cores; returns the first
// non-zero status.
template <typename Fn>
int parallel_rows(uint32_t count, Fn fn) {
    const uint32_t workers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t per = (count + workers - 1) / workers;
    std::vector<int> status(workers, GS_OK);
    std::vector<std::thread> pool;
    for (uint32_t w = 0; w < workers; ++w) {
        const uint32_t b = w * per, e = std::min(count, b + per);
        if (b >= e) break;
        pool.emplace_back([&, w, b, e] { status[w] = fn(b, e); });
    }
    for (auto& t : pool) t.join();
    for (int s : status)
        if (s != GS_OK) return s;
    return GS_OK;
}

float sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }

float lerp(float a, float b, float t) { return a + (b - a) * t; }

// Writes one Gaussian into the plane layout of the device scene: 15 float4
// planes with `stride` entries each, plane p of entry i at (p * stride + i)
// * 4 floats: P0 position + opacity, P1 rotation w x y z, P2 scale + DC r,
// P3 DC g b + rest 0..1, P4+j rest 2+4j..5+4j, P14 rest 42..44 and 0.
void put_planes(const GsGaussian& g, float* planes, size_t stride, size_t i) {
    auto plane = [&](int p) { return planes + ((size_t)p * stride + i) * 4; };
    auto put = [](float* d, float a, float b, float c, float e) {
        d[0] = a;
        d[1] = b;
        d[2] = c;
        d[3] = e;
    };
    const float* r = g.sh_rest;
    put(plane(0), g.position[0], g.position[1], g.position[2], g.opacity);
    put(plane(1), g.rotation[0], g.rotation[1], g.rotation[2], g.rotation[3]);
    put(plane(2), g.scale[0], g.scale[1], g.scale[2], g.color[0]);
    put(plane(3), g.color[1], g.color[2], r[0], r[1]);
    for (int j = 0; j < 10; ++j) put(plane(4 + j), r[2 + 4 * j], r[3 + 4 * j], r[4 + 4 * j], r[5 + 4 * j]);
    put(plane(14), r[42], r[43], r[44], 0.0f);
}

int check_ply_layout(const GsPlyLayout* l) {
    if (!l) return GS_ERR_INVALID;
    if (l->x < 0 || l->y < 0 || l->z < 0) return GS_ERR_INVALID;
    const int32_t* offs[] = {&l->x, &l->y, &l->z, l->scale, l->scale + 1, l->scale + 2, l->rot, l->rot + 1,
                             l->rot + 2, l->rot + 3, &l->opacity, l->dc, l->dc + 1, l->dc + 2};
    for (const int32_t* o : offs)
        if (*o >= (int32_t)l->stride) return GS_ERR_INVALID;
    for (int k = 0; k < 45; ++k)
        if (l->rest[k] >= (int32_t)l->stride) return GS_ERR_INVALID;
    // scale, rotation and DC are used as whole groups: all present or all absent
    auto whole = [](const int32_t* o, int n) {
        bool any = false, all = true;
        for (int k = 0; k < n; ++k) {
            any |= o[k] >= 0;
            all &= o[k] >= 0;
        }
        return all || !any;
    };
    if (!whole(l->scale, 3) || !whole(l->rot, 4) || !whole(l->dc, 3)) return GS_ERR_INVALID;
    return GS_OK;
}

// One PLY vertex row activated as the reference's PLY loader does
// (src/gaussian_loader_ply.cpp:143-203).
void decode_ply_row(const float* v, const GsPlyLayout& l, GsGaussian& g) {
    memset(&g, 0, sizeof(g));
    g.position[0] = v[l.x];
    g.position[1] = v[l.y];
    g.position[2] = v[l.z];
    if (l.scale[0] >= 0) {
        for (int a = 0; a < 3; ++a) g.scale[a] = expf(v[l.scale[a]]);
    } else {
        g.scale[0] = g.scale[1] = g.scale[2] = 0.01f;
    }
    if (l.rot[0] >= 0) {
        float q[4];
        for (int k = 0; k < 4; ++k) q[k] = v[l.rot[k]];
        const float len = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        if (len > 1e-8f)
            for (int k = 0; k < 4; ++k) q[k] /= len;
        for (int k = 0; k < 4; ++k) g.rotation[k] = q[k];
    } else {
        g.rotation[0] = 1.0f;
    }
    g.opacity = l.opacity >= 0 ? sigmoid(v[l.opacity]) : 1.0f;
    for (int c = 0; c < 3; ++c) g.color[c] = l.dc[0] >= 0 ? v[l.dc[c]] : 0.0f;
    // f_rest is channel-major (all red coefficients, then green, then blue);
    // the Gaussian keeps RGB triples per coefficient
    for (int k = 0; k < 15; ++k)
        for (int c = 0; c < 3; ++c) {
            const int32_t o = l.rest[c * 15 + k];
            g.sh_rest[k * 3 + c] = o >= 0 ? v[o] : 0.0f;
        }
}

int check_sog(const GsSogData* s, int* coefs_out) {
    if (!s) return GS_ERR_INVALID;
    if (s->count && (!s->means_l || !s->means_u || !s->scales || !s->quats || !s->sh0)) return GS_ERR_INVALID;
    const int coefs = s->shN_bands == 1 ? 3 : (s->shN_bands == 2 ? 8 : (s->shN_bands == 3 ? 15 : 0));
    if (s->shN_bands) {
        if (!coefs || !s->shN_centroids || !s->shN_labels) return GS_ERR_INVALID;
        if (s->shN_count <= 0 || s->shN_count > 65536) return GS_ERR_INVALID;
        if (s->centroids_width < 64 * coefs || s->centroids_height < (s->shN_count + 63) / 64)
            return GS_ERR_INVALID;
    }
    *coefs_out = coefs;
    return GS_OK;
}

// Entry i of a SOG v2 scene (src/gaussian_loader_sog.cpp:403-510): each
// image holds one RGBA8 texel per Gaussian.
//   position: a 16-bit value per axis (high byte in means_u, low in
//     means_l) placed between the bounds, then mapped back from the log
//     domain: x = sign(n) (e^|n| - 1);
//   scale: e^(codebook value); DC colour: its codebook value;
//   rotation: three components stored at 8 bits in [-1/sqrt2, 1/sqrt2], the
//     fourth (the largest, slot qt[3] - 252) restored from the unit norm;
//   opacity: the sh0 alpha byte / 255;
//   higher bands: a 16-bit palette label; palette entry e stores its
//     coefficients in the centroid image at column (e % 64) * coefs + k,
//     row e / 64, one codebook index per channel.
int decode_sog_entry(const GsSogData& s, uint32_t i, int coefs, GsGaussian& g) {
    const float root2 = sqrtf(2.0f);
    memset(&g, 0, sizeof(g));
    const size_t t = (size_t)i * 4;  // the texel of entry i
    for (int a = 0; a < 3; ++a) {
        const uint32_t word = ((uint32_t)s.means_u[t + a] << 8) | (uint32_t)s.means_l[t + a];
        const float n = lerp(s.means_mins[a], s.means_maxs[a], (float)word / 65535.0f);
        g.position[a] = n < 0.0f ? -(expf(-n) - 1.0f) : (expf(n) - 1.0f);
        g.scale[a] = expf(s.scales_codebook[s.scales[t + a]]);
        g.color[a] = s.sh0_codebook[s.sh0[t + a]];
    }
    const int largest = (int)s.quats[t + 3] - 252;
    if (largest < 0 || largest > 3) return GS_ERR_INVALID;
    float q3[3];
    for (int k = 0; k < 3; ++k) q3[k] = ((float)s.quats[t + k] / 255.0f - 0.5f) * 2.0f / root2;
    const float rest_norm = q3[0] * q3[0] + q3[1] * q3[1] + q3[2] * q3[2];
    const float restored = sqrtf(fmaxf(0.0f, 1.0f - rest_norm));
    for (int slot = 0, k = 0; slot < 4; ++slot) g.rotation[slot] = slot == largest ? restored : q3[k++];
    g.opacity = (float)s.sh0[t + 3] / 255.0f;
    if (coefs) {
        const uint32_t entry = (uint32_t)s.shN_labels[t] | ((uint32_t)s.shN_labels[t + 1] << 8);
        if (entry >= (uint32_t)s.shN_count) return GS_ERR_INVALID;
        const size_t row = entry / 64u, col0 = (size_t)(entry % 64u) * (size_t)coefs;
        for (int k = 0; k < coefs; ++k) {
            const uint8_t* texel = s.shN_centroids + (row * (size_t)s.centroids_width + col0 + (size_t)k) * 4;
            for (int c = 0; c < 3; ++c) g.sh_rest[k * 3 + c] = s.shN_codebook[texel[c]];
        }
    }
    return GS_OK;
}

}  // namespace

extern "C" int gs_decode_ply(const float* rows, uint32_t count, const GsPlyLayout* l, GsGaussian* out) {
    if (!out || (count && !rows)) return GS_ERR_INVALID;
    const int st = check_ply_layout(l);
    if (st != GS_OK) return st;
    return parallel_rows(count, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) decode_ply_row(rows + (size_t)i * l->stride, *l, out[i]);
        return (int)GS_OK;
    });
}

extern "C" int gs_decode_sog(const GsSogData* s, GsGaussian* out) {
    if (!s || !out) return GS_ERR_INVALID;
    int coefs = 0;
    const int st = check_sog(s, &coefs);
    if (st != GS_OK) return st;
    return parallel_rows(s->count, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const int r = decode_sog_entry(*s, i, coefs, out[i]);
            if (r != GS_OK) return r;
        }
        return (int)GS_OK;
    });
}

extern "C" int gs_planes_from_gaussians(const GsGaussian* g, uint32_t count, float* planes) {
    if (!planes || (count && !g)) return GS_ERR_INVALID;
    return gs::fill_planes_gaussians(g, count, count, planes);
}

extern "C" int gs_decode_ply_planes(const float* rows, uint32_t count, const GsPlyLayout* layout, float* planes) {
    if (!planes || (count && !rows)) return GS_ERR_INVALID;
    return gs::fill_planes_ply(rows, count, layout, count, planes);
}

namespace gs {

int fill_planes_gaussians(const GsGaussian* g, uint32_t count, uint32_t stride, float* planes) {
    return parallel_rows(count, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) put_planes(g[i], planes, stride, i);
        return (int)GS_OK;
    });
}

int fill_planes_copy(const float* src, uint32_t src_stride, uint32_t first, uint32_t count, uint32_t stride,
                     float* planes) {
    return parallel_rows(count, [&](uint32_t b, uint32_t e) {
        for (int p = 0; p < 15; ++p)
            memcpy(planes + ((size_t)p * stride + b) * 4, src + ((size_t)p * src_stride + first + b) * 4,
                   (size_t)(e - b) * 16);
        return (int)GS_OK;
    });
}

int fill_planes_ply(const float* rows, uint32_t count, const GsPlyLayout* l, uint32_t stride, float* planes) {
    const int st = check_ply_layout(l);
    if (st != GS_OK) return st;
    return parallel_rows(count, [&](uint32_t b, uint32_t e) {
        GsGaussian g;
        for (uint32_t i = b; i < e; ++i) {
            decode_ply_row(rows + (size_t)i * l->stride, *l, g);
            put_planes(g, planes, stride, i);
        }
        return (int)GS_OK;
    });
}

int fill_planes_sog(const GsSogData* s, uint32_t first, uint32_t count, uint32_t stride, float* planes) {
    int coefs = 0;
    const int st = check_sog(s, &coefs);
    if (st != GS_OK) return st;
    return parallel_rows(count, [&](uint32_t b, uint32_t e) {
        GsGaussian g;
        for (uint32_t i = b; i < e; ++i) {
            const int r = decode_sog_entry(*s, first + i, coefs, g);
            if (r != GS_OK) return r;
            put_planes(g, planes, stride, i);
        }
        return (int)GS_OK;
    });
}

}  // namespace gs
