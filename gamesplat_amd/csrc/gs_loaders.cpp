// gs_loaders.cpp - per-Gaussian decode of the scene formats (host).
//
// The container work (PLY header, ZIP archive, lossless WebP images, JSON
// metadata) is left to the caller; these functions do the arithmetic the
// reference loaders apply to each Gaussian, with the same float operations
// and libm functions, so the same input gives the same Gaussians.
//   gs_decode_ply   src/gaussian_loader_ply.cpp:143-203
//   gs_decode_sog   src/gaussian_loader_sog.cpp:403-510
#include <math.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "../../include/gs_api.h"

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

}  // namespace

extern "C" int gs_decode_ply(const float* rows, uint32_t count, const GsPlyLayout* l, GsGaussian* out) {
    if (!l || !out || (count && !rows)) return GS_ERR_INVALID;
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
    return parallel_rows(count, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const float* v = rows + (size_t)i * l->stride;
            GsGaussian& g = out[i];
            memset(&g, 0, sizeof(g));
            g.position[0] = v[l->x];
            g.position[1] = v[l->y];
            g.position[2] = v[l->z];
            if (l->scale[0] >= 0) {
                for (int a = 0; a < 3; ++a) g.scale[a] = expf(v[l->scale[a]]);
            } else {
                g.scale[0] = g.scale[1] = g.scale[2] = 0.01f;
            }
            if (l->rot[0] >= 0) {
                float rw = v[l->rot[0]], rx = v[l->rot[1]], ry = v[l->rot[2]], rz = v[l->rot[3]];
                const float len = sqrtf(rw * rw + rx * rx + ry * ry + rz * rz);
                if (len > 1e-8f) {
                    rw /= len;
                    rx /= len;
                    ry /= len;
                    rz /= len;
                }
                g.rotation[0] = rw;
                g.rotation[1] = rx;
                g.rotation[2] = ry;
                g.rotation[3] = rz;
            } else {
                g.rotation[0] = 1.0f;
            }
            g.opacity = l->opacity >= 0 ? sigmoid(v[l->opacity]) : 1.0f;
            for (int c = 0; c < 3; ++c) g.color[c] = l->dc[0] >= 0 ? v[l->dc[c]] : 0.0f;
            // channel-major f_rest (R k=0..14, G, B) to RGB triples per coefficient
            for (int k = 0; k < 15; ++k)
                for (int c = 0; c < 3; ++c) {
                    const int32_t o = l->rest[c * 15 + k];
                    g.sh_rest[k * 3 + c] = o >= 0 ? v[o] : 0.0f;
                }
        }
        return (int)GS_OK;
    });
}

extern "C" int gs_decode_sog(const GsSogData* s, GsGaussian* out) {
    if (!s || !out) return GS_ERR_INVALID;
    if (!s->count) return GS_OK;
    if (!s->means_l || !s->means_u || !s->scales || !s->quats || !s->sh0) return GS_ERR_INVALID;
    const int coefs = s->shN_bands == 1 ? 3 : (s->shN_bands == 2 ? 8 : (s->shN_bands == 3 ? 15 : 0));
    if (s->shN_bands) {
        if (!coefs || !s->shN_centroids || !s->shN_labels) return GS_ERR_INVALID;
        if (s->shN_count <= 0 || s->shN_count > 65536) return GS_ERR_INVALID;
        if (s->centroids_width < 64 * coefs || s->centroids_height < (s->shN_count + 63) / 64)
            return GS_ERR_INVALID;
    }
    const float sqrt2 = sqrtf(2.0f);
    return parallel_rows(s->count, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const uint8_t* ml = s->means_l + (size_t)i * 4;
            const uint8_t* mu = s->means_u + (size_t)i * 4;
            const uint8_t* sc = s->scales + (size_t)i * 4;
            const uint8_t* qt = s->quats + (size_t)i * 4;
            const uint8_t* s0 = s->sh0 + (size_t)i * 4;
            GsGaussian& g = out[i];
            memset(&g, 0, sizeof(g));
            for (int a = 0; a < 3; ++a) {
                const uint32_t q = ((uint32_t)mu[a] << 8) | (uint32_t)ml[a];
                const float n = lerp(s->means_mins[a], s->means_maxs[a], (float)q / 65535.0f);
                g.position[a] = n < 0.0f ? -(expf(-n) - 1.0f) : (expf(n) - 1.0f);
                g.scale[a] = expf(s->scales_codebook[sc[a]]);
                g.color[a] = s->sh0_codebook[s0[a]];
            }
            // smallest three: three stored components in [-1/sqrt2, 1/sqrt2]
            float stored[3];
            for (int k = 0; k < 3; ++k) stored[k] = ((float)qt[k] / 255.0f - 0.5f) * 2.0f / sqrt2;
            const int mode = (int)qt[3] - 252;
            if (mode < 0 || mode > 3) return (int)GS_ERR_INVALID;
            const float omitted =
                sqrtf(fmaxf(0.0f, 1.0f - (stored[0] * stored[0] + stored[1] * stored[1] + stored[2] * stored[2])));
            // the omitted component takes slot `mode`, the stored ones fill the rest in order
            for (int m = 0, k = 0; m < 4; ++m) g.rotation[m] = m == mode ? omitted : stored[k++];
            g.opacity = (float)s0[3] / 255.0f;
            if (coefs) {
                const uint8_t* lp = s->shN_labels + (size_t)i * 4;
                const uint32_t label = (uint32_t)lp[0] | ((uint32_t)lp[1] << 8);
                if (label >= (uint32_t)s->shN_count) return (int)GS_ERR_INVALID;
                for (int cf = 0; cf < coefs; ++cf) {
                    const int u = (int)(label % 64) * coefs + cf, v = (int)(label / 64);
                    const uint8_t* px = s->shN_centroids + ((size_t)v * s->centroids_width + (size_t)u) * 4;
                    for (int ch = 0; ch < 3; ++ch) g.sh_rest[cf * 3 + ch] = s->shN_codebook[px[ch]];
                }
            }
        }
        return (int)GS_OK;
    });
}
