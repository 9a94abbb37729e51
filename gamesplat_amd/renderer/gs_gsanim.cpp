// gs_gsanim.cpp - GSAnim playback sampling for the standalone build.
//
// Inside the reference tree these come from the asset loader
// (src/gaussian_loader_gsanim.cpp:495-531); this file supplies the same
// semantics when the renderer shim is built on its own:
//   - the playback time wraps into [0, duration) (negative times too);
//   - the active key is the last one at or before that time (the first
//     key while none is);
//   - a linear key blends towards the next key (the first one after the
//     last), weight = elapsed fraction of the interval, which ends at the
//     next key's time or at duration for the last key;
//   - invalid input (no keys, non-finite times, duration <= 0) gives state
//     0 and, for layers, no weights.
#ifndef GS_RENDERER_IN_TREE

#include <math.h>

#include "gs_renderer.h"

namespace {

// Playback time wrapped into [0, duration), or -1 when it cannot be.
double wrap_time(double seconds, double duration) {
    if (!isfinite(seconds) || !isfinite(duration) || duration <= 0.0) return -1.0;
    double u = fmod(seconds, duration);
    if (u < 0.0) u += duration;
    if (u >= duration) u = 0.0;
    return u;
}

// Index of the last of `count` ascending key times at or before u (0 when
// every key is later): the largest i with time(i) <= u.
template <typename TimeAt>
size_t active_key(size_t count, double u, TimeAt time_at) {
    size_t lo = 0, hi = count;  // keys [0, lo) are <= u, keys [hi, count) are > u
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (time_at(mid) <= u)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo == 0 ? 0 : lo - 1;
}

}  // namespace

GSAnimSample gsanim_sample(const GSAnim& a, double seconds) {
    const double u = wrap_time(seconds, a.duration);
    if (u < 0.0 || a.keys.empty()) return {0, 0, 0.0f};
    const size_t n = a.keys.size();
    const size_t i = active_key(n, u, [&](size_t k) { return a.keys[k].time_seconds; });
    const int state = (int)a.keys[i].state;
    const bool linear = a.version == 2 ? a.keys[i].linear : a.linear;
    if (!linear || n == 1) return {state, state, 0.0f};
    const size_t j = (i + 1) % n;
    const double end = j ? a.keys[j].time_seconds : a.duration;
    const double t0 = a.keys[i].time_seconds;
    return {state, (int)a.keys[j].state, (float)((u - t0) / (end - t0))};
}

std::vector<std::array<float, 3>> gsanim_layer_weights(const GSAnim& a, double seconds) {
    std::vector<std::array<float, 3>> out;
    const double u = a.version == 3 ? wrap_time(seconds, a.duration) : -1.0;
    if (u < 0.0) return out;
    out.reserve(a.layers.size());
    for (const GSAnimLayer& layer : a.layers) {
        const size_t n = layer.keys.size();
        if (n == 0) return {};
        const size_t i = active_key(n, u, [&](size_t k) { return layer.keys[k].time_seconds; });
        const size_t j = (i + 1) % n;
        const double end = j ? layer.keys[j].time_seconds : a.duration;
        const double t0 = layer.keys[i].time_seconds;
        const double w = layer.keys[i].linear && n > 1 ? (u - t0) / (end - t0) : 0.0;
        std::array<float, 3> rgb;
        for (int c = 0; c < 3; ++c)
            rgb[c] = (float)((1.0 - w) * layer.keys[i].weight[c] + w * layer.keys[j].weight[c]);
        out.push_back(rgb);
    }
    return out;
}

#endif  // GS_RENDERER_USE_REFERENCE_TYPES
