// gs_headless.cpp - headless frame driver and capture harness.
//
// Reproduces the call sequence of app_init / app_frame (src/main.cpp:276-573,
// 1060-1977) for the splat path and the measurement protocol of the
// reference harness (tests/gsanim_render.cpp:73-188): key=value flags,
// warm-up then timed interval, per-frame GPU time, a RESULT JSON line and a
// raw float32 RGBA linear capture (top row first).
//
// Flags (defaults in brackets):
//   ply=<path>                 3DGS binary PLY; otherwise a synthetic scene:
//   gaussians=[100000] clusters=[16] seed=[0x5EED0001] opacity_mean=[0]
//   width=[1280] height=[720] render_mode=[alpha|gps|stochastic] stochastic_spp=[1]
//   view_seed=[seed] radius=[15] orbit_deg=[0]   camera on a seeded direction
//   gps_ss=[1] gps_budget_m=[16] gps_accumulation=[1] sh_degree=[3]
//   gpu_budget_mb=[0] diagnostics=[0]
//   warmup=[3] frames=[20]     untimed then timed frame counts
//   taa=[off|on|0|1] taa_samples=[4]   StochasticSplats with TAA
//   linear_capture=<path>      raw f32 RGBA of the last frame
//   display_capture=<path>     raw f32 RGBA of what the frame displays (LUT for layered GSAnim)
//   capture=<path>             binary PPM (P6) of the displayed frame, 8-bit, top row first
//   csv=<path>                 per timed frame: wall_ms, cpu_submit_ms, gpu_elapsed_ms,
//                              animation_time, state, selected_state, playing
//   fixed_time=<s>             GSAnim playback pinned at s (not playing)
//   exercise_changes=1         GSAnim: frames cycle selected_state, the last frame plays
//   exercise_modes=1           the first three frames render alpha, stochastic, GPS, then render_mode
// GSAnim appearance from raw files (the loader's arrays):
//   anim_deltas=<path> anim_scales=<path> anim_states=N       int8 deltas, 48 f32 scales
//   anim_layers=<path> anim_layer_count=L                      half coefficients
//   anim_keys=t:state:linear,...  anim_version=[2] anim_linear=[0] anim_duration=[1]
//   anim_layer_keys=t:r:g:b:linear,...;... (one list per layer)
//   anim_lut=<path> anim_lut_size=S anim_log2_min anim_log2_max
//   anim_time=[0] anim_time_step=[0]  playback time of frame f: time + f * step
//   first_layer_only=[0] exposure_ev=[0]
// Learned denoiser on the GPS path (renderer_enable_denoiser):
//   denoiser=<weights path> (float32, gs_denoiser_get_params layout)
//   denoise_min_psnr=[40] denoise_max_abs=[1] denoise_judge_every=[1]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "gs_renderer.h"

namespace {

std::map<std::string, std::string> g_args;

const char* arg(const char* k, const char* dflt) {
    auto it = g_args.find(k);
    return it == g_args.end() ? dflt : it->second.c_str();
}
double argf(const char* k, double dflt) {
    auto it = g_args.find(k);
    return it == g_args.end() ? dflt : strtod(it->second.c_str(), nullptr);
}
long long argi(const char* k, long long dflt) {
    auto it = g_args.find(k);
    return it == g_args.end() ? dflt : strtoll(it->second.c_str(), nullptr, 0);
}

// Binary little-endian 3DGS PLY with float properties; the rows are
// activated by gs_decode_ply as src/gaussian_loader_ply.cpp:143-203 does.
bool load_ply(const char* path, std::vector<Gaussian>& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char line[512];
    uint32_t count = 0;
    std::vector<std::string> props;
    bool binary_le = false;
    while (fgets(line, sizeof(line), f)) {
        if (!strncmp(line, "format binary_little_endian", 27)) binary_le = true;
        if (!strncmp(line, "element vertex ", 15)) count = (uint32_t)strtoul(line + 15, nullptr, 10);
        if (!strncmp(line, "property float ", 15)) {
            std::string n(line + 15);
            while (!n.empty() && (n.back() == '\n' || n.back() == '\r' || n.back() == ' ')) n.pop_back();
            props.push_back(n);
        } else if (!strncmp(line, "property ", 9)) {
            fprintf(stderr, "unsupported non-float property: %s", line);
            fclose(f);
            return false;
        }
        if (!strncmp(line, "end_header", 10)) break;
    }
    if (!binary_le || !count || props.empty()) {
        fclose(f);
        return false;
    }
    auto find = [&](const std::string& n) {
        for (size_t i = 0; i < props.size(); ++i)
            if (props[i] == n) return (int)i;
        return -1;
    };
    int ox = find("x"), oy = find("y"), oz = find("z");
    if (ox < 0 || oy < 0 || oz < 0) {
        fclose(f);
        return false;
    }
    GsPlyLayout lay;
    memset(&lay, 0, sizeof(lay));
    lay.stride = (uint32_t)props.size();
    lay.x = ox;
    lay.y = oy;
    lay.z = oz;
    for (int a = 0; a < 3; ++a) lay.scale[a] = find("scale_" + std::to_string(a));
    for (int a = 0; a < 4; ++a) lay.rot[a] = find("rot_" + std::to_string(a));
    lay.opacity = find("opacity");
    for (int a = 0; a < 3; ++a) lay.dc[a] = find("f_dc_" + std::to_string(a));
    for (int k = 0; k < 45; ++k) lay.rest[k] = find("f_rest_" + std::to_string(k));
    const size_t stride = props.size();
    out.resize(count);
    std::vector<float> rows(stride * 65536);
    for (uint32_t i = 0; i < count; i += 65536) {
        const uint32_t b = std::min(65536u, count - i);
        if (fread(rows.data(), stride * 4, b, f) != b ||
            gs_decode_ply(rows.data(), b, &lay, reinterpret_cast<GsGaussian*>(out.data() + i)) != GS_OK) {
            fclose(f);
            return false;
        }
    }
    fclose(f);
    return true;
}

template <typename T>
bool read_raw(const char* path, std::vector<T>& out, size_t count) {
    out.resize(count);
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    const bool ok = fread(out.data(), sizeof(T), count, f) == count;
    fclose(f);
    return ok;
}

// The GSAnim asset from the anim_* flags; false when a file is missing.
bool load_animation(GSAnim& an, uint32_t rows) {
    an.duration = argf("anim_duration", 1.0);
    an.version = (int)argi("anim_version", 2);
    an.linear = argi("anim_linear", 0) != 0;
    an.gaussian_count = rows;
    for (const char* p = arg("anim_keys", ""); *p;) {
        GSAnimKey k{};
        int linear = 0, used = 0;
        if (sscanf(p, "%lf:%u:%d%n", &k.time_seconds, &k.state, &linear, &used) != 3) return false;
        k.linear = linear != 0;
        an.keys.push_back(k);
        p += used;
        if (*p == ',') ++p;
    }
    if (g_args.count("anim_layers")) {
        an.version = 3;
        const uint32_t layers = (uint32_t)argi("anim_layer_count", 1);
        if (!read_raw(arg("anim_layers", ""), an.layer_data, (size_t)layers * rows * 48)) return false;
        an.layers.resize(layers);
        const char* p = arg("anim_layer_keys", "");
        for (uint32_t l = 0; l < layers && *p; ++l) {
            while (*p && *p != ';') {
                GSAnimLayerKey k;
                int linear = 0, used = 0;
                if (sscanf(p, "%lf:%f:%f:%f:%d%n", &k.time_seconds, &k.weight[0], &k.weight[1], &k.weight[2],
                           &linear, &used) != 5)
                    return false;
                k.linear = linear != 0;
                an.layers[l].keys.push_back(k);
                p += used;
                if (*p == ',') ++p;
            }
            if (*p == ';') ++p;
        }
        an.lut_size = (uint32_t)argi("anim_lut_size", 0);
        if (an.lut_size &&
            !read_raw(arg("anim_lut", ""), an.display_lut, (size_t)an.lut_size * an.lut_size * an.lut_size * 3))
            return false;
        an.log2_min = (float)argf("anim_log2_min", -8.0);
        an.log2_max = (float)argf("anim_log2_max", 4.0);
    } else {
        an.state_count = (uint32_t)argi("anim_states", 1);
        if (!read_raw(arg("anim_deltas", ""), an.deltas, (size_t)an.state_count * rows * 48)) return false;
        std::vector<float> scales;
        if (!read_raw(arg("anim_scales", ""), scales, 48)) return false;
        memcpy(an.scales, scales.data(), sizeof(an.scales));
    }
    return true;
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    return v[(size_t)((v.size() - 1) * p)];
}

}  // namespace

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        const char* eq = strchr(argv[i], '=');
        if (!eq) {
            fprintf(stderr, "flags are key=value, got %s\n", argv[i]);
            return 2;
        }
        g_args[std::string(argv[i], eq - argv[i])] = eq + 1;
    }
    auto t_start = std::chrono::steady_clock::now();
    const int width = (int)argi("width", 1280), height = (int)argi("height", 720);

    std::vector<Gaussian> gaussians;
    if (g_args.count("ply")) {
        if (!load_ply(arg("ply", ""), gaussians)) {
            fprintf(stderr, "could not load %s\n", arg("ply", ""));
            return 1;
        }
    } else {
        uint32_t n = (uint32_t)argi("gaussians", 100000);
        gaussians.resize(n);
        if (gs_generate_scene((uint64_t)argi("seed", 0x5EED0001), n, (uint32_t)argi("clusters", 16),
                              (float)argf("opacity_mean", 0.0), reinterpret_cast<GsGaussian*>(gaussians.data())) != GS_OK)
            return 1;
    }
    GaussianScene scene{};
    scene.gaussians = gaussians.data();
    scene.gaussian_count = (uint32_t)gaussians.size();
    GSAnim animation;
    if (g_args.count("anim_deltas") || g_args.count("anim_layers")) {
        if (!load_animation(animation, scene.gaussian_count)) {
            fprintf(stderr, "could not load the GSAnim flags\n");
            return 1;
        }
        scene.animation = &animation;
    }

    Renderer r;
    if (!renderer_init(&r)) return 1;
    const std::string mode = arg("render_mode", "alpha");
    r.splat_render_mode = mode == "gps"          ? SplatRenderMode::GaussianPointSplatting
                          : mode == "stochastic" ? SplatRenderMode::StochasticSplats
                                                 : SplatRenderMode::AlphaBlendSorted;
    const std::string taa = arg("taa", "0");
    r.stochastic_taa_enabled = taa == "on" || taa == "1";
    r.stochastic_taa_current_samples = (uint32_t)argi("taa_samples", 4);
    r.stochastic_samples_per_frame = (uint32_t)argi("stochastic_spp", 1);
    r.black_background = true;
    r.sh_degree = (int)argi("sh_degree", 3);
    r.gpu_budget_bytes = (size_t)argi("gpu_budget_mb", 0) * 1048576u;
    r.gps_supersample_factor = (uint32_t)argi("gps_ss", 1);
    r.gps_max_work_items = (uint32_t)(argf("gps_budget_m", 16.0) * 1048576.0);
    r.stochastic_accumulation_enabled = argi("gps_accumulation", 1) != 0;
    r.splat_diagnostics_enabled = argi("diagnostics", 0) != 0;
    r.first_layer_only = argi("first_layer_only", 0) != 0;
    r.exposure_ev = (float)argf("exposure_ev", 0.0);
    size_t free0 = 0, total0 = 0;
    hipMemGetInfo(&free0, &total0);
    if (!renderer_upload_gaussians(&r, &scene)) return 1;
    const char* weights_path = arg("denoiser", "");
    if (weights_path[0]) {
        std::vector<float> w(gs_denoiser_param_count());
        FILE* wf = fopen(weights_path, "rb");
        if (!wf || fread(w.data(), 4, w.size(), wf) != w.size()) {
            fprintf(stderr, "could not read %zu denoiser parameters from %s\n", w.size(), weights_path);
            return 1;
        }
        fclose(wf);
        if (!renderer_enable_denoiser(&r, w.data(), w.size(), (float)argf("denoise_min_psnr", 40.0),
                                      (float)argf("denoise_max_abs", 1.0), (uint32_t)argi("denoise_judge_every", 1)))
            return 1;
    }
    double startup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();

    // seeded camera direction (splitmix64 -> normal via Box-Muller), radius
    uint64_t s = (uint64_t)argi("view_seed", argi("seed", 0x5EED0001));
    auto next = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    double d[3];
    for (int k = 0; k < 3; k += 2) {
        double u1 = 1.0 - next(), u2 = next();
        double rr = sqrt(-2.0 * log(u1));
        d[k] = rr * cos(6.283185307179586 * u2);
        if (k + 1 < 3) d[k + 1] = rr * sin(6.283185307179586 * u2);
    }
    double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double radius = argf("radius", 15.0), orbit = argf("orbit_deg", 0.0);

    const int warmup = (int)argi("warmup", 3), frames = (int)argi("frames", 20);
    const SplatRenderMode final_mode = r.splat_render_mode;
    hipStream_t stream = (hipStream_t)gs_stream(renderer_gs_context(&r));
    std::vector<hipEvent_t> ev(2 * frames);
    for (auto& e : ev) hipEventCreate(&e);
    struct FrameRecord {
        double wall_ms, cpu_ms, animation_time;
        int state, selected_state, playing;
    };
    std::vector<FrameRecord> rec;
    CameraUniforms cam;
    auto previous = std::chrono::steady_clock::now();
    for (int f = 0; f < warmup + frames; ++f) {
        const auto begin = std::chrono::steady_clock::now();
        double a = orbit * M_PI / 180.0 * f, c = cos(a), sn = sin(a);
        float pos[3] = {(float)((c * d[0] - sn * d[1]) / dn * radius), (float)((sn * d[0] + c * d[1]) / dn * radius),
                        (float)(d[2] / dn * radius)};
        float pose[16];
        gs_look_at_origin_pose(pos, pose);
        gs_camera_from_pose(pose, 0.8660254f, (float)width, (float)height, reinterpret_cast<GsCameraUniforms*>(&cam));
        if (scene.animation) {
            if (g_args.count("fixed_time")) {
                animation.playing = false;
                animation.playback_time = argf("fixed_time", 0.0);
            } else {
                animation.playback_time = argf("anim_time", 0.0) + argf("anim_time_step", 0.0) * f;
            }
            if (argi("exercise_changes", 0) && animation.state_count) {
                const bool last = f == warmup + frames - 1;
                animation.selected_state = last ? -1 : f % (int)animation.state_count;
            }
        }
        if (argi("exercise_modes", 0)) {
            // the harness's mode exercise: alpha, stochastic, GPS, then the requested mode
            const SplatRenderMode m = f < 3 ? SplatRenderMode(f) : final_mode;
            if (r.splat_render_mode != m) {
                r.splat_render_mode = m;
                renderer_reset_stochastic_accumulation(&r);
            }
        }
        const bool timed = f >= warmup;
        if (timed) hipEventRecord(ev[2 * (f - warmup)], stream);
        renderer_draw_frame(&r, &scene, &cam, nullptr, nullptr);
        if (timed) hipEventRecord(ev[2 * (f - warmup) + 1], stream);
        if (renderer_last_status(&r) != GS_OK) return 1;
        const auto submitted = std::chrono::steady_clock::now();
        if (timed)
            rec.push_back({std::chrono::duration<double, std::milli>(begin - previous).count(),
                           std::chrono::duration<double, std::milli>(submitted - begin).count(),
                           scene.animation ? animation.playback_time : 0.0, r.animation_state_a,
                           scene.animation ? animation.selected_state : -1,
                           scene.animation && animation.playing ? 1 : 0});
        previous = begin;
    }
    gs_synchronize(renderer_gs_context(&r));
    std::vector<double> wall, cpu, gpu;
    for (int f = 0; f < frames; ++f) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, ev[2 * f], ev[2 * f + 1]);
        gpu.push_back(ms);
        wall.push_back(rec[f].wall_ms);
        cpu.push_back(rec[f].cpu_ms);
    }
    for (auto& e : ev) hipEventDestroy(e);
    const char* csv_path = arg("csv", "");
    if (csv_path[0]) {
        FILE* f = fopen(csv_path, "w");
        if (!f) return 1;
        fprintf(f, "wall_ms,cpu_submit_ms,gpu_elapsed_ms,animation_time,state,selected_state,playing\n");
        for (int k = 0; k < frames; ++k)
            fprintf(f, "%.6f,%.6f,%.6f,%.9f,%d,%d,%d\n", rec[k].wall_ms, rec[k].cpu_ms, gpu[k], rec[k].animation_time,
                    rec[k].state, rec[k].selected_state, rec[k].playing);
        fclose(f);
    }
    const char* ppm = arg("capture", "");
    if (ppm[0]) {
        // displayed values, 8-bit UNORM (round to nearest), as glReadPixels of the swapchain
        std::vector<float> px((size_t)width * height * 4);
        if (!renderer_read_display(&r, &scene, px.data(), px.size())) return 1;
        std::vector<unsigned char> rgb((size_t)width * height * 3);
        for (size_t i = 0; i < (size_t)width * height; ++i)
            for (int k = 0; k < 3; ++k) {
                float v = px[i * 4 + k];
                v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
                rgb[i * 3 + k] = (unsigned char)lrintf(v * 255.0f);
            }
        FILE* f = fopen(ppm, "wb");
        if (!f) return 1;
        fprintf(f, "P6\n%d %d\n255\n", width, height);
        if (fwrite(rgb.data(), 1, rgb.size(), f) != rgb.size()) return 1;
        fclose(f);
    }

    const char* cap = arg("linear_capture", "");
    if (cap[0]) {
        std::vector<float> px((size_t)width * height * 4);
        if (!renderer_read_linear(&r, px.data(), px.size())) return 1;
        FILE* f = fopen(cap, "wb");
        if (!f || fwrite(px.data(), 4, px.size(), f) != px.size()) return 1;
        fclose(f);
    }
    const char* dcap = arg("display_capture", "");
    if (dcap[0]) {
        std::vector<float> px((size_t)width * height * 4);
        if (!renderer_read_display(&r, &scene, px.data(), px.size())) return 1;
        FILE* f = fopen(dcap, "wb");
        if (!f || fwrite(px.data(), 4, px.size(), f) != px.size()) return 1;
        fclose(f);
    }
    GsStats st;
    gs_stats(renderer_gs_context(&r), &st);
    GsFrameInfo fi;
    gs_frame_info(renderer_gs_context(&r), &fi);
    size_t free1 = 0, total1 = 0;
    hipMemGetInfo(&free1, &total1);
    RendererDenoiseReport dr;
    renderer_denoise_report(&r, &dr);
    if (dr.enabled)
        printf("DENOISE {\"frames\":%u,\"judged\":%u,\"fallbacks\":%u,\"showing_denoised\":%d,"
               "\"last_psnr_db\":%.6f,\"last_max_abs\":%.9g}\n",
               dr.frames, dr.judged, dr.fallbacks, dr.showing_denoised ? 1 : 0, dr.last_psnr_db,
               (double)dr.last_max_abs);
    printf("RESULT {\"frames\":%zu,\"wall_median_ms\":%.6f,\"wall_p95_ms\":%.6f,\"cpu_median_ms\":%.6f,"
           "\"cpu_p95_ms\":%.6f,\"gpu_median_ms\":%.6f,\"gpu_p95_ms\":%.6f,\"startup_ms\":%.3f,"
           "\"vram_free_delta_mib\":%.3f,\"scene_buffers_mib\":%.3f,\"gaussians\":%u,\"visible\":%u,"
           "\"tile_entries\":%u,\"gps_requested\":%u,\"gps_exhausted\":%u}\n",
           wall.size(), percentile(wall, .5), percentile(wall, .95), percentile(cpu, .5), percentile(cpu, .95),
           percentile(gpu, .5), percentile(gpu, .95), startup_ms, ((double)free0 - (double)free1) / 1048576.0,
           (double)fi.held_bytes / 1048576.0, st.gaussian_count, st.visible_count, st.tile_entries,
           st.gps_requested_work_items, st.gps_work_budget_exhausted);
    renderer_destroy(&r);
    return 0;
}
