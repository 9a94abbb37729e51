// gs_renderer.cpp - renderer.h splat path over the gs_* C ABI (headless).
//
// Pass order and state rules follow renderer_draw_frame
// (src/renderer.cpp:2139-2289): cull every frame; sorted alpha sorts and
// composites into the RGBA16F appearance image; GPS resets accumulation on
// any CameraUniforms byte change, increments the frame seed before each
// sample, reads the work statistics at most twice per second and keeps a
// running mean when accumulation is enabled.
#include "gs_renderer.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace {

// What the splat path keeps per Renderer beyond the reference's fields: the
// gs context, the status of the last frame and the time of the last GPS
// work-statistics readback.  A side table keyed by the Renderer's address,
// so the reference's struct stays as it is.
struct SplatState {
    gs_ctx* ctx = nullptr;
    int last_status = GS_OK;
    double last_stats_time = 0.0;
    // learned denoiser on the GPS path, judged against sorted alpha
    gs_denoiser* den = nullptr;
    float min_psnr_db = 40.0f;
    float max_abs_error = 1.0f;
    uint32_t judge_every = 1;
    bool fallback = false;      // the last judged frame missed a bar: show sorted alpha
    bool have_denoised = false; // the display holds this frame's denoised image or its fallback
    uint32_t frames = 0, judged = 0, fallbacks = 0;
    double last_psnr_db = 0.0;
    float last_max_abs = 0.0f;
};
std::mutex g_state_mutex;
std::unordered_map<const Renderer*, SplatState> g_states;

SplatState& state(const Renderer* r) {
    std::lock_guard<std::mutex> lock(g_state_mutex);
    return g_states[r];  // element references stay valid across rehashing
}

void drop_state(const Renderer* r) {
    std::lock_guard<std::mutex> lock(g_state_mutex);
    g_states.erase(r);
}

void log_error(const char* what, int status) {
    fprintf(stderr, "[renderer] %s failed (%d): %s\n", what, status, gs_last_error());
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

float half_to_float(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1fu, man = h & 0x3ffu;
    uint32_t bits;
    if (exp == 0) {
        if (man == 0) {
            bits = sign;
        } else {
            float v = (float)man * (1.0f / 16777216.0f);
            return sign ? -v : v;
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (man << 13);
    } else {
        bits = sign | ((exp + 112u) << 23) | (man << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

}  // namespace

gs_ctx* renderer_gs_context(const Renderer* r) { return state(r).ctx; }

int renderer_last_status(const Renderer* r) { return state(r).last_status; }

bool renderer_init(Renderer* r) {
    memset(r, 0, sizeof(*r));
    // defaults of renderer_init (src/renderer.cpp:588-598)
    r->splat_render_mode = SplatRenderMode::GaussianPointSplatting;
    r->stochastic_samples_per_frame = 1;
    r->stochastic_taa_current_samples = 4;
    r->gps_supersample_factor = 2;
    r->gps_max_work_items = 16u * 1024u * 1024u;
    r->stochastic_accumulation_enabled = true;
    r->stochastic_taa_enabled = true;
    r->sh_degree = 3;
    r->sh_lod_distance = 0.0f;
    SplatState& ss = state(r);
    if (ss.den) gs_denoiser_destroy(ss.den);  // a re-initialised Renderer at the same address
    if (ss.ctx) gs_destroy(ss.ctx);
    ss = SplatState{};
    int st = gs_create(&ss.ctx, 0);
    if (st != GS_OK) {
        log_error("gs_create", st);
        ss.ctx = nullptr;
        return false;
    }
    r->gps_capabilities = {true, true, true, true, true};
    r->gps_gpu.max_work_items = r->gps_max_work_items;
    r->gps_gpu.supersample_factor = r->gps_supersample_factor;
    return true;
}

bool renderer_upload_gaussians(Renderer* r, const GaussianScene* scene) {
    gs_ctx* ctx = state(r).ctx;
    if (!ctx || !scene) return false;
    gs_set_memory_budget(ctx, r->gpu_budget_bytes);
    int st = gs_set_gps_budget(ctx, r->gps_max_work_items ? r->gps_max_work_items : 1u);
    if (st == GS_OK)
        st = gs_upload_gaussians(ctx, reinterpret_cast<const GsGaussian*>(scene->gaussians), scene->gaussian_count);
    if (st != GS_OK) {
        log_error("upload", st);
        return false;
    }
    r->gaussian_count = scene->gaussian_count;
    r->sort_capacity = scene->gaussian_count;
    renderer_reset_stochastic_accumulation(r);
    if (scene->animation) {
        // appearance tracks next to the scene (src/renderer.cpp:751-893)
        const GSAnim& an = *scene->animation;
        const size_t row_coefs = (size_t)scene->gaussian_count * 48u;
        GsAnimation ga;
        memset(&ga, 0, sizeof(ga));
        ga.rows = scene->gaussian_count;
        if (an.version == 3) {
            ga.layer_count = (uint32_t)an.layers.size();
            ga.layer_data = an.layer_data.data();
            ga.lut_size = an.lut_size;
            ga.display_lut = an.lut_size ? an.display_lut.data() : nullptr;
            ga.log2_min = an.log2_min;
            ga.log2_max = an.log2_max;
            if (an.layer_data.size() != ga.layer_count * row_coefs ||
                an.display_lut.size() != (size_t)an.lut_size * an.lut_size * an.lut_size * 3u) {
                fprintf(stderr, "[renderer] GSAnim layer or LUT sizes do not match the scene\n");
                return false;
            }
        } else {
            ga.state_count = an.state_count;
            ga.deltas = an.deltas.data();
            memcpy(ga.scales, an.scales, sizeof(ga.scales));
            if (an.deltas.size() != an.state_count * row_coefs) {
                fprintf(stderr, "[renderer] GSAnim delta size does not match the scene\n");
                return false;
            }
        }
        if (ga.state_count || ga.layer_count) {
            st = gs_upload_animation(ctx, &ga);
            if (st != GS_OK) {
                log_error("animation upload", st);
                return false;
            }
        }
        r->animation_state_a = r->animation_state_b = 0;
        r->animation_weight = 0.0f;
        r->animation_time = 0.0;
    }
    return true;
}

bool renderer_set_gps_work_budget(Renderer* r, uint32_t max_work_items) {
    r->gps_work_budget_resize_failed = false;
    if (max_work_items == r->gps_gpu.max_work_items) return true;
    gs_ctx* ctx = state(r).ctx;
    int st = ctx ? gs_set_gps_budget(ctx, max_work_items) : GS_ERR_STATE;
    if (st != GS_OK) {
        log_error("GPS work budget resize", st);
        r->gps_work_budget_resize_failed = true;
        return false;
    }
    r->gps_gpu.max_work_items = max_work_items;
    r->gps_max_work_items = max_work_items;
    r->gps_gpu.work_stats_valid = false;
    r->gps_gpu.work_stats_elapsed = 0.0f;
    renderer_reset_stochastic_accumulation(r);
    return true;
}

void renderer_reset_stochastic_accumulation(Renderer* r) {
    r->stochastic_sample_count = 0;
    r->stochastic_accum_write_index = 0;
    r->stochastic_prev_cam_valid = false;
    r->stochastic_prev_effect_valid = false;
    r->stochastic_taa_prev_view_proj_valid = false;
}

static int denoise_frame(Renderer* r, SplatState& ss, gs_ctx* ctx, const CameraUniforms* cam);

static uint32_t clamp_samples(uint32_t n) { return n < 1u ? 1u : (n > 16u ? 16u : n); }

// StochasticSplats (src/renderer.cpp:2290-2480): per frame, render_samples
// samples with consecutive frame seeds; with TAA each is Halton-jittered,
// they are averaged, and the average is reprojected into the history; without
// TAA they fold into the running mean, which a camera change resets.
static int draw_stochastic(Renderer* r, gs_ctx* ctx, const CameraUniforms* cam, const SplatEffectParams* splat_effect,
                           float bg) {
    const GsCameraUniforms* gcam = reinterpret_cast<const GsCameraUniforms*>(cam);
    // new target sizes start a new accumulation (renderer_ensure_stochastic_targets)
    if (r->stochastic_width != (uint32_t)cam->viewport[0] || r->stochastic_height != (uint32_t)cam->viewport[1])
        renderer_reset_stochastic_accumulation(r);
    SplatEffectParams no_effect;
    memset(&no_effect, 0, sizeof(no_effect));
    const SplatEffectParams* effect = splat_effect ? splat_effect : &no_effect;
    const CameraUniforms prev_cam = r->stochastic_prev_cam;
    bool prev_cam_valid = r->stochastic_prev_cam_valid;
    const bool cam_changed = r->stochastic_prev_cam_valid && memcmp(&r->stochastic_prev_cam, cam, sizeof(*cam)) != 0;
    const bool effect_changed =
        r->stochastic_prev_effect_valid && memcmp(&r->stochastic_prev_effect, effect, sizeof(*effect)) != 0;
    if ((!r->stochastic_taa_enabled && cam_changed) || effect_changed) {
        renderer_reset_stochastic_accumulation(r);
        prev_cam_valid = false;
    }
    r->stochastic_prev_cam = *cam;
    r->stochastic_prev_effect = *effect;
    r->stochastic_prev_cam_valid = true;
    r->stochastic_prev_effect_valid = true;

    const bool use_taa = r->stochastic_taa_enabled;
    const uint32_t samples = clamp_samples(use_taa ? r->stochastic_taa_current_samples
                                                   : r->stochastic_samples_per_frame);
    GsTaaUniforms taa;
    if (use_taa) {
        const bool has_history = r->stochastic_taa_prev_view_proj_valid && prev_cam_valid;
        const GsCameraUniforms* prev = has_history ? reinterpret_cast<const GsCameraUniforms*>(&prev_cam) : nullptr;
        if (gs_taa_uniforms(gcam, prev, cam_changed ? 1 : 0, &taa) != GS_OK) {
            // a view without an inverse: the uniforms fall back to no history
            renderer_reset_stochastic_accumulation(r);
        }
    }
    int st = GS_OK;
    for (uint32_t i = 0; i < samples && st == GS_OK; ++i) {
        r->stochastic_frame_seed++;
        float jitter[2] = {0.0f, 0.0f};
        if (use_taa) gs_taa_jitter(r->stochastic_frame_seed, jitter);
        st = gs_stochastic_sample(ctx, r->stochastic_frame_seed, bg, jitter);
        if (st != GS_OK) break;
        if (use_taa && samples > 1u) {
            st = gs_frame_average(ctx, i);
        } else if (!use_taa && (r->stochastic_accumulation_enabled || samples > 1u)) {
            const uint32_t next = r->stochastic_accumulation_enabled ? r->stochastic_sample_count + 1u : i + 1u;
            st = gs_accumulate(ctx, next);
            if (st == GS_OK && r->stochastic_accumulation_enabled) {
                r->stochastic_sample_count = next;
                r->stochastic_accum_write_index = (r->stochastic_accum_write_index + 1u) & 1u;
            }
        }
    }
    if (st == GS_OK && use_taa) {
        st = gs_taa_accumulate(ctx, &taa, samples > 1u ? 1 : 0);
        if (st == GS_OK) {
            r->stochastic_taa_prev_view_proj_valid = true;
            r->stochastic_sample_count++;
            r->stochastic_accum_write_index = (r->stochastic_accum_write_index + 1u) & 1u;
        }
    }
    if (st != GS_OK) return st;
    r->stochastic_width = (uint32_t)cam->viewport[0];
    r->stochastic_height = (uint32_t)cam->viewport[1];
    return GS_OK;
}

void renderer_draw_frame(Renderer* r, GaussianScene* scene, const CameraUniforms* cam, const OverlayParams*,
                         const NodeRenderParams*, const SplatEffectParams* splat_effect, float, const CameraUniforms*) {
    SplatState& ss = state(r);
    gs_ctx* ctx = ss.ctx;
    ss.last_status = GS_OK;
    if (!ctx || !scene || !cam) {
        ss.last_status = GS_ERR_STATE;
        return;
    }
    const float bg = r->black_background || scene->animation ? 0.0f : 0.1f;
    if (scene->animation) {
        // playback state (src/renderer.cpp:2148-2161) and AnimationUBO
        // (src/renderer.cpp:1823-1835)
        const GSAnim& an = *scene->animation;
        GSAnimSample sample = gsanim_sample(an, an.playback_time);
        if (an.selected_state >= 0 && (uint32_t)an.selected_state < an.state_count)
            sample = {an.selected_state, an.selected_state, 0.0f};
        if (sample.state_a != r->animation_state_a || sample.state_b != r->animation_state_b ||
            sample.weight != r->animation_weight || (an.version == 3 && an.playback_time != r->animation_time))
            renderer_reset_stochastic_accumulation(r);
        r->animation_state_a = sample.state_a;
        r->animation_state_b = sample.state_b;
        r->animation_weight = sample.weight;
        r->animation_time = an.playback_time;
        std::vector<std::array<float, 3>> weights;
        uint32_t layers = 0;
        if (an.version == 3) {
            weights = gsanim_layer_weights(an, an.playback_time);
            layers = (uint32_t)weights.size();
            if (r->first_layer_only && layers > 1u) layers = 1u;
        }
        const int st = gs_set_animation_frame(ctx, an.version == 3 ? 0 : sample.state_a,
                                              an.version == 3 ? 0 : sample.state_b,
                                              an.version == 3 ? 0.0f : sample.weight, layers,
                                              layers ? weights[0].data() : nullptr);
        if (st != GS_OK) {
            log_error("animation frame", st);
            ss.last_status = st;
            return;
        }
    }
    const bool gps = r->splat_render_mode == SplatRenderMode::GaussianPointSplatting;
    const bool stochastic = r->splat_render_mode == SplatRenderMode::StochasticSplats;
    if (scene->gaussian_count == 0 || r->gaussian_count == 0) {
        scene->visible_count = 0;
        return;
    }
    const GsCameraUniforms* gcam = reinterpret_cast<const GsCameraUniforms*>(cam);
    GsCullOptions opts = {r->sh_degree, r->sh_lod_distance, r->splat_diagnostics_enabled ? 1 : 0};
    // animated scenes draw without the shockwave (src/renderer.cpp:2168)
    const SplatEffectParams* cull_effect = scene->animation ? nullptr : splat_effect;
    int st = gs_cull(ctx, gcam, reinterpret_cast<const GsSplatEffectParams*>(cull_effect), &opts);
    if (st != GS_OK) {
        log_error("cull", st);
        ss.last_status = st;
        return;
    }
    r->splat_diagnostics.valid = false;
    if (r->splat_diagnostics_enabled) {
        GsStats s;
        if (gs_stats(ctx, &s) == GS_OK && s.diagnostics_valid) {
            memcpy(&r->splat_diagnostics, s.diagnostics, sizeof(s.diagnostics));
            r->splat_diagnostics.valid = true;
        }
    }
    // no indirect draw count in the reference either (src/renderer.cpp:2177,2187)
    scene->visible_count = scene->gaussian_count;

    if (!gps && !stochastic) {
        st = gs_sort(ctx);
        if (st == GS_OK) st = gs_composite_alpha(ctx, 0);
        if (st != GS_OK) {
            log_error("sorted-alpha frame", st);
            ss.last_status = st;
            return;
        }
        r->appearance_width = (int)cam->viewport[0];
        r->appearance_height = (int)cam->viewport[1];
        return;
    }

    if (stochastic) {
        st = draw_stochastic(r, ctx, cam, splat_effect, bg);
        if (st != GS_OK) {
            log_error("stochastic splats", st);
            ss.last_status = st;
        }
        return;
    }

    if (r->stochastic_prev_cam_valid && memcmp(&r->stochastic_prev_cam, cam, sizeof(*cam)) != 0)
        renderer_reset_stochastic_accumulation(r);
    memcpy(&r->stochastic_prev_cam, cam, sizeof(*cam));
    r->stochastic_prev_cam_valid = true;
    r->stochastic_frame_seed++;

    const uint32_t sf = r->gps_supersample_factor < 1u ? 1u : (r->gps_supersample_factor > 4u ? 4u : r->gps_supersample_factor);
    const int linear_color = scene->animation && scene->animation->version == 3 ? 1 : 0;
    st = gs_gps_sample(ctx, r->stochastic_frame_seed, sf, linear_color, bg);
    if (st != GS_OK) {
        log_error("GPS sample", st);
        ss.last_status = st;
        return;
    }
    r->gps_gpu.width = (uint32_t)cam->viewport[0];
    r->gps_gpu.height = (uint32_t)cam->viewport[1];
    r->gps_gpu.supersample_factor = sf;
    r->stochastic_width = r->gps_gpu.width;
    r->stochastic_height = r->gps_gpu.height;
    double t = now_s();
    if (!r->gps_gpu.work_stats_valid || t - ss.last_stats_time >= 0.5) {
        GsStats s;
        if (gs_stats(ctx, &s) == GS_OK) {
            r->gps_gpu.requested_work_items = s.gps_requested_work_items;
            r->gps_gpu.work_budget_exhausted = s.gps_work_budget_exhausted != 0;
            r->gps_gpu.work_stats_valid = true;
            r->gps_gpu.work_stats_elapsed = 0.0f;
            ss.last_stats_time = t;
        }
    }
    if (!r->stochastic_accumulation_enabled) {
        r->stochastic_sample_count = 0;
        return;
    }
    uint32_t next = r->stochastic_sample_count + 1u;
    st = gs_accumulate(ctx, next);
    if (st != GS_OK) {
        log_error("accumulate", st);
        ss.last_status = st;
        return;
    }
    r->stochastic_sample_count = next;
    r->stochastic_accum_write_index = (r->stochastic_accum_write_index + 1u) & 1u;
    if (ss.den) {
        st = denoise_frame(r, ss, ctx, cam);
        if (st != GS_OK) {
            log_error("denoised frame", st);
            ss.last_status = st;
        }
    }
}

// The learned denoiser's frame: denoise the running mean; on a judge frame
// (every judge_every-th), and on every frame while the last judgement fell
// back, render the sorted-alpha image of the same cull; a judge frame
// measures the denoised image against it, and a PSNR below min_psnr_db or
// a max abs error above max_abs_error makes sorted alpha the display until
// a later judge frame passes.
static int denoise_frame(Renderer* r, SplatState& ss, gs_ctx* ctx, const CameraUniforms* cam) {
    ss.have_denoised = false;
    int st = gs_denoise(ss.den);
    if (st != GS_OK) return st;
    const bool judge = ss.judge_every && ss.frames % ss.judge_every == 0;
    ss.frames++;
    if (judge || ss.fallback) {
        st = gs_sort(ctx);
        if (st == GS_OK) st = gs_composite_alpha(ctx, 0);
        if (st != GS_OK) return st;
        r->appearance_width = (int)cam->viewport[0];
        r->appearance_height = (int)cam->viewport[1];
        if (judge) {
            st = gs_denoiser_judge(ss.den, &ss.last_psnr_db, &ss.last_max_abs);
            if (st != GS_OK) return st;
            ss.judged++;
            ss.fallback = !(ss.last_psnr_db >= (double)ss.min_psnr_db) || !(ss.last_max_abs <= ss.max_abs_error);
            if (ss.fallback) ss.fallbacks++;
        }
    }
    ss.have_denoised = true;
    return GS_OK;
}

bool renderer_enable_denoiser(Renderer* r, const float* params, size_t count, float min_psnr_db, float max_abs_error,
                              uint32_t judge_every) {
    SplatState& ss = state(r);
    if (!ss.ctx || !params) return false;
    if (!ss.den) {
        const int st = gs_denoiser_create(ss.ctx, 1, &ss.den);
        if (st != GS_OK) {
            log_error("denoiser create", st);
            ss.den = nullptr;
            return false;
        }
    }
    const int st = gs_denoiser_set_params(ss.den, params, count);
    if (st != GS_OK) {
        log_error("denoiser weights", st);
        return false;
    }
    ss.min_psnr_db = min_psnr_db;
    ss.max_abs_error = max_abs_error;
    ss.judge_every = judge_every;
    ss.fallback = ss.have_denoised = false;
    ss.frames = ss.judged = ss.fallbacks = 0;
    return true;
}

void renderer_disable_denoiser(Renderer* r) {
    SplatState& ss = state(r);
    if (ss.den) gs_denoiser_destroy(ss.den);
    ss.den = nullptr;
    ss.have_denoised = ss.fallback = false;
}

bool renderer_denoise_report(const Renderer* r, RendererDenoiseReport* out) {
    if (!out) return false;
    const SplatState& ss = state(r);
    out->enabled = ss.den != nullptr;
    out->showing_denoised = ss.have_denoised && !ss.fallback;
    out->frames = ss.frames;
    out->judged = ss.judged;
    out->fallbacks = ss.fallbacks;
    out->last_psnr_db = ss.last_psnr_db;
    out->last_max_abs = ss.last_max_abs;
    return true;
}

// The image the reference harness captures (tests/gsanim_render.cpp:148-156).
static int linear_image(const Renderer* r) {
    if (r->splat_render_mode == SplatRenderMode::AlphaBlendSorted) return GS_IMAGE_ALPHA_RGBA16F;
    const bool stochastic = r->splat_render_mode == SplatRenderMode::StochasticSplats;
    if (r->stochastic_accumulation_enabled || (stochastic && r->stochastic_taa_enabled))
        return stochastic && r->stochastic_taa_enabled ? GS_IMAGE_TAA_RGBA32F : GS_IMAGE_ACCUM_RGBA32F;
    return GS_IMAGE_GPS_SAMPLE_RGBA32F;  // the GPS resolved image or the StochasticSplats sample
}

bool renderer_read_linear(Renderer* r, float* out, size_t floats) {
    const SplatState& ss = state(r);
    gs_ctx* ctx = ss.ctx;
    if (!ctx || !out) return false;
    int which = linear_image(r);
    if (ss.den && ss.have_denoised && r->splat_render_mode == SplatRenderMode::GaussianPointSplatting) {
        if (ss.fallback) {
            which = GS_IMAGE_ALPHA_RGBA16F;  // the judge's verdict: sorted alpha
        } else {
            const size_t px = (size_t)r->stochastic_width * r->stochastic_height;
            if (floats != px * 4) return false;
            std::vector<uint16_t> h(px * 4);
            const int st = gs_denoiser_read_image(ss.den, h.data(), h.size() * 2);
            if (st != GS_OK) {
                log_error("read denoised image", st);
                return false;
            }
            for (size_t i = 0; i < h.size(); ++i) out[i] = half_to_float(h[i]);
            return true;
        }
    }
    if (which == GS_IMAGE_ALPHA_RGBA16F) {
        size_t px = (size_t)r->appearance_width * r->appearance_height;
        if (floats != px * 4) return false;
        std::vector<uint16_t> h(px * 4);
        int st = gs_read_image(ctx, GS_IMAGE_ALPHA_RGBA16F, h.data(), h.size() * 2);
        if (st != GS_OK) {
            log_error("read appearance image", st);
            return false;
        }
        for (size_t i = 0; i < h.size(); ++i) out[i] = half_to_float(h[i]);
        return true;
    }
    size_t px = (size_t)r->stochastic_width * r->stochastic_height;
    if (floats != px * 4) return false;
    int st = gs_read_image(ctx, which, out, floats * 4);
    if (st != GS_OK) {
        log_error("read stochastic image", st);
        return false;
    }
    return true;
}

bool renderer_read_display(Renderer* r, const GaussianScene* scene, float* out, size_t floats) {
    const GSAnim* an = scene ? scene->animation : nullptr;
    if (!an || an->version != 3 || !an->lut_size) return renderer_read_linear(r, out, floats);
    gs_ctx* ctx = state(r).ctx;
    if (!ctx || !out) return false;
    const int source = linear_image(r);
    const size_t px = source == GS_IMAGE_ALPHA_RGBA16F ? (size_t)r->appearance_width * r->appearance_height
                                                       : (size_t)r->stochastic_width * r->stochastic_height;
    if (floats != px * 4) return false;
    int st = gs_display_layers(ctx, source, r->exposure_ev);
    if (st == GS_OK) st = gs_read_image(ctx, GS_IMAGE_DISPLAY_RGBA32F, out, floats * 4);
    if (st != GS_OK) {
        log_error("layer display", st);
        return false;
    }
    return true;
}

void renderer_destroy(Renderer* r) {
    SplatState& ss = state(r);
    if (ss.den) gs_denoiser_destroy(ss.den);
    if (ss.ctx) gs_destroy(ss.ctx);
    drop_state(r);
}
