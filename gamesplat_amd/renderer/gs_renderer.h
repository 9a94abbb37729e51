// gs_renderer.h - the splat path of gamesplat's renderer.h (src/renderer.h:249-266)
// implemented over the gs_* C ABI.
//
// Two builds of one source (gs_renderer.cpp):
//   - in the reference tree (-DGS_RENDERER_IN_TREE, with the reference's
//     src/ and vendor/ on the include path): every type and every renderer_*
//     declaration comes from the reference's own renderer.h, gaussian.h,
//     camera.h and gsanim.h, so src/main.cpp and the harness
//     (tests/gsanim_render.cpp) compile against their own Renderer; the gs
//     context of each Renderer lives in a side table, not in the struct;
//   - standalone (default, used by gs_headless): the same types are declared
//     here with the reference's names and layouts, keeping the Renderer
//     fields src/main.cpp and the harness read or write
//     (src/main.cpp:341-470,1344-1470, tests/gsanim_render.cpp:46,97-113,
//     148-156,186); sokol handles are left out.
// Mesh, overlay and ImGui passes are not part of this build in either mode.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <array>
#include <string>
#include <vector>

#include "../../include/gs_api.h"

#ifdef GS_RENDERER_IN_TREE
#include "gsanim.h"
#include "renderer.h"
#else
#define GAUSSIAN_SH_REST_FLOATS 45
struct Gaussian {
    float position[3];
    float scale[3];
    float rotation[4];  // w, x, y, z
    float color[3];
    float opacity;
    float sh_rest[GAUSSIAN_SH_REST_FLOATS];
};
// GSAnim as the loader fills it (src/gsanim.h:9-48): the renderer reads the
// coefficient data, the keys and the playback state.
struct GSAnimKey {
    double time_seconds;
    uint32_t state;
    bool linear = false;  // outgoing interval interpolates (version 2)
    bool has_interpolation = false;
};
struct GSAnimLayerKey {
    double time_seconds = 0;
    std::array<float, 3> weight{};
    bool linear = false;
};
struct GSAnimLayer {
    std::string name;
    std::vector<GSAnimLayerKey> keys;
};
struct GSAnim {
    int version = 1;
    std::vector<GSAnimLayer> layers;
    std::vector<uint16_t> layer_data;   // half bits, layer/row/channel-major, 48 per row
    std::vector<uint16_t> display_lut;  // RGB uint16, red fastest
    uint32_t lut_size = 0;
    float log2_min = 0, log2_max = 0;
    std::vector<int8_t> deltas;         // state/row/channel-major, state 0 excluded
    float scales[48] = {};
    uint32_t state_count = 0;
    uint32_t gaussian_count = 0;
    double duration = 0;
    std::vector<GSAnimKey> keys;
    bool linear = false;
    double playback_time = 0;
    bool playing = true;
    int selected_state = -1;
};
struct GSAnimSample {
    int state_a, state_b;
    float weight;
};
// The states and blend weight at `seconds` (src/gsanim.h:60-62), and the
// RGB weight of every layer of a version-3 asset (src/gsanim.h:64-65).
GSAnimSample gsanim_sample(const GSAnim& animation, double seconds);
std::vector<std::array<float, 3>> gsanim_layer_weights(const GSAnim& animation, double seconds);

struct GaussianScene {
    GSAnim* animation;
    Gaussian* gaussians;
    uint32_t gaussian_count;
    uint32_t* visible_indices;
    float* visible_depths;
    uint32_t visible_count;
    uint32_t* sorted_indices;
    uint32_t* scratch_indices;
    uint32_t* scratch_keys;
    uint32_t* scratch_keys2;
};
struct CameraUniforms {
    float view[16];
    float proj[16];
    float viewport[2];
    float orthographic;
    float persp_focal;
    float ortho_focal;
    float clip_y_sign;
    float clip_z_01;
    float pad[1];
};
struct SplatEffectParams {
    float center_radius[4];
    float params[4];
    float color[4];
};
// Passed through renderer_draw_frame and not read by the splat path
// (src/renderer.h:7-26): the overlay's sokol texture handles are left out.
struct OverlayParams {
    float camera_ray_basis[16];
    float camera_tan_half_fov[2];
    float ref_rotation[16];
    float alpha;
};
struct NodeRenderParams {
    const float* positions;
    uint32_t count;
    float half_size;
};

struct SplatDiagnostics {
    uint32_t total_quad_kpix;
    uint32_t max_quad_px;
    uint32_t splats_over_1k_px;
    uint32_t splats_over_16k_px;
    uint32_t total_gps_kpoints;
    uint32_t max_gps_points;
    uint32_t gps_splats_over_1k_points;
    uint32_t gps_splats_over_16k_points;
    bool valid;
};

struct GpsCapabilities {
    bool compute;
    bool storage_images;
    bool shader_int64;
    bool atomic_min_64;
    bool supported;
};

enum class SplatRenderMode {
    AlphaBlendSorted = 0,
    StochasticSplats = 1,
    GaussianPointSplatting = 2,
};

struct GaussianPointSplatGpu {
    uint32_t width;
    uint32_t height;
    uint32_t supersample_factor;
    uint32_t max_work_items;
    uint32_t requested_work_items;
    float work_stats_elapsed;
    bool work_stats_valid;
    bool work_budget_exhausted;
};

struct Renderer {
    SplatRenderMode splat_render_mode;
    int appearance_width, appearance_height;
    float exposure_ev;
    bool first_layer_only;
    double animation_time;
    int animation_state_a;
    int animation_state_b;
    float animation_weight;
    bool black_background;
    size_t gpu_budget_bytes;
    bool gps_work_budget_resize_failed;
    uint32_t sort_capacity;
    uint32_t gaussian_count;
    bool splat_diagnostics_enabled;
    SplatDiagnostics splat_diagnostics;
    GpsCapabilities gps_capabilities;
    int sh_degree;
    float sh_lod_distance;
    GaussianPointSplatGpu gps_gpu;
    uint32_t stochastic_width;
    uint32_t stochastic_height;
    uint32_t stochastic_sample_count;
    uint32_t stochastic_samples_per_frame;
    uint32_t stochastic_taa_current_samples;
    uint32_t gps_supersample_factor;
    uint32_t gps_max_work_items;
    uint32_t stochastic_frame_seed;
    uint32_t stochastic_accum_write_index;
    CameraUniforms stochastic_prev_cam;
    SplatEffectParams stochastic_prev_effect;
    bool stochastic_prev_cam_valid;
    bool stochastic_prev_effect_valid;
    bool stochastic_taa_prev_view_proj_valid;
    bool stochastic_accumulation_enabled;
    bool stochastic_taa_enabled;
};

bool renderer_init(Renderer* r);
bool renderer_upload_gaussians(Renderer* r, const GaussianScene* scene);
bool renderer_set_gps_work_budget(Renderer* r, uint32_t max_work_items);
void renderer_reset_stochastic_accumulation(Renderer* r);
void renderer_draw_frame(Renderer* r, GaussianScene* scene, const CameraUniforms* cam,
                         const OverlayParams* overlay, const NodeRenderParams* nodes,
                         const SplatEffectParams* splat_effect = NULL, float wireframe_occlusion = 1.0f,
                         const CameraUniforms* map_cam = NULL);
void renderer_destroy(Renderer* r);
#endif  // GS_RENDERER_IN_TREE

static_assert(sizeof(Gaussian) == sizeof(GsGaussian), "Gaussian layout");
static_assert(sizeof(CameraUniforms) == sizeof(GsCameraUniforms), "CameraUniforms layout");
static_assert(sizeof(SplatEffectParams) == sizeof(GsSplatEffectParams), "SplatEffectParams layout");
static_assert(sizeof(SplatDiagnostics) >= 8 * sizeof(uint32_t), "SplatDiagnostics layout");

// The gs context renderer_init created for r (NULL before, or after
// renderer_destroy), and the gs status of its last renderer_draw_frame
// (GS_OK when it succeeded).
gs_ctx* renderer_gs_context(const Renderer* r);
int renderer_last_status(const Renderer* r);

// The learned denoiser on the GPS path (no counterpart in the reference).
// With accumulation on, every GPS frame's running mean goes through the
// U-Net (params: gs_denoiser_get_params layout).  Every judge_every-th frame
// (1 = every frame) the sorted-alpha image of the same cull is rendered and
// judges the denoised one: a PSNR below min_psnr_db or a max abs error above
// max_abs_error makes sorted alpha the displayed image (renderer_read_linear)
// until a later judge frame passes; while it does, sorted alpha is rendered
// every frame.
struct RendererDenoiseReport {
    bool enabled;
    bool showing_denoised;   // the last frame displays the denoised image
    uint32_t frames;         // frames denoised
    uint32_t judged;         // frames judged against sorted alpha
    uint32_t fallbacks;      // judged frames that fell back to sorted alpha
    double last_psnr_db;     // the last judged frame
    float last_max_abs;
};
bool renderer_enable_denoiser(Renderer* r, const float* params, size_t count, float min_psnr_db, float max_abs_error,
                              uint32_t judge_every);
void renderer_disable_denoiser(Renderer* r);
bool renderer_denoise_report(const Renderer* r, RendererDenoiseReport* out);

// Headless readback of what the reference harness captures with
// linear_capture (tests/gsanim_render.cpp:144-165), chosen by its rule: the
// appearance image in sorted-alpha mode (GS_IMAGE_ALPHA_RGBA16F as floats);
// with accumulation on (or StochasticSplats with TAA) the accumulation image
// (GS_IMAGE_TAA_RGBA32F for TAA, else GS_IMAGE_ACCUM_RGBA32F); otherwise the
// GPS resolved image or the StochasticSplats sample image (both
// GS_IMAGE_GPS_SAMPLE_RGBA32F).  Raw float32 RGBA, top row first; out holds
// width*height*4 floats.
bool renderer_read_linear(Renderer* r, float* out, size_t floats);

// What renderer_display_appearance shows (src/renderer.cpp:2120-2137): for a
// GSAnim version-3 scene the linear image through the asset's LUT at
// exposure_ev, otherwise the linear image itself.  Same size and order as
// renderer_read_linear.
bool renderer_read_display(Renderer* r, const GaussianScene* scene, float* out, size_t floats);
