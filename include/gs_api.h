/*
 * gs_api.h - C ABI of the MI355X (gfx950) Gaussian-splatting frame pipeline.
 *
 * The library replaces the GL compute/raster work that gamesplat's renderer
 * issues from src/renderer.cpp through sokol_gfx:
 *   cull + project + SH        (shaders/cull.glsl:166-474, src/renderer.cpp:1778-1871)
 *   radix depth sort           (shaders/sort.glsl:4-164,   src/renderer.cpp:1873-1962)
 *   sorted-alpha composite     (shaders/splat.glsl:6-105,  src/renderer.cpp:1687-1720,2516-2541)
 *   GPS clear/expand/splat/resolve (shaders/gps.glsl:6-386, src/renderer.cpp:1964-2118)
 *   running-mean accumulation  (shaders/accum.glsl:19-45,  src/renderer.cpp:2242-2268)
 *
 * Conventions
 *   - Every entry point returns GS_OK (0) or a negative GS_ERR_* code and
 *     records a message readable through gs_last_error() (thread-local).
 *   - Host pointers are borrowed for the duration of the call only.
 *   - One context owns one HIP stream and all device memory it uses; calls
 *     on one context are ordered on that stream.  Calls that return data to
 *     the host synchronise the stream; launches do not.
 *   - Images returned to the host are top row first, like the reference
 *     harness' linear_capture (tests/gsanim_render.cpp:144-165).
 *   - No HIP or torch type appears in this header.
 */
#ifndef GS_API_H
#define GS_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_API_VERSION 1

/* ---- status codes ------------------------------------------------------ */
enum {
    GS_OK = 0,
    GS_ERR_INVALID = -1,  /* bad argument or unsupported option            */
    GS_ERR_OOM = -2,      /* device or host allocation failed              */
    GS_ERR_HIP = -3,      /* a HIP runtime call failed                      */
    GS_ERR_STATE = -4,    /* call out of order (e.g. sort before cull)     */
    GS_ERR_LIMIT = -5     /* request exceeds the memory budget or a limit  */
};

/* ---- data contracts (byte-compatible with the reference types) --------- */

/* Same layout as `Gaussian` (src/gaussian.h:9-16): activated scale and
 * opacity, unit quaternion (w,x,y,z), raw DC and 45 SH-rest floats stored as
 * RGB triples per coefficient. 59 floats = 236 bytes. */
typedef struct GsGaussian {
    float position[3];
    float scale[3];
    float rotation[4];
    float color[3];
    float opacity;
    float sh_rest[45];
} GsGaussian;

/* Same layout as `CameraUniforms` (src/camera.h:23-38): column-major view and
 * projection, scene-target viewport in pixels. */
typedef struct GsCameraUniforms {
    float view[16];
    float proj[16];
    float viewport[2];
    float orthographic;
    float persp_focal;
    float ortho_focal;
    float clip_y_sign;
    float clip_z_01;
    float pad[1];
} GsCameraUniforms;

/* Same layout as `SplatEffectParams` (src/renderer.h:52-56). */
typedef struct GsSplatEffectParams {
    float center_radius[4];
    float params[4];
    float color[4];
} GsSplatEffectParams;

/* Per-frame cull options: the CullUBO fields the renderer fills from its own
 * state (src/renderer.cpp:1847-1849). */
typedef struct GsCullOptions {
    int32_t sh_degree;        /* 0..3                                         */
    float sh_lod_distance;    /* <= 0 disables the distance LOD               */
    int32_t collect_stats;    /* non-zero: fill the 8 splat diagnostics       */
} GsCullOptions;

/* Counters, read back with gs_stats (src/renderer.h:28-38,104-111). */
typedef struct GsStats {
    uint32_t gaussian_count;
    uint32_t visible_count;
    uint32_t tile_entries;            /* duplicated (tile, splat) pairs of the last composite */
    uint32_t gps_requested_work_items;
    uint32_t gps_work_budget_exhausted;
    uint32_t gps_dropped_work_items;  /* max(0, requested - budget)           */
    uint32_t gps_max_work_items;
    uint32_t diagnostics[8];          /* SplatDiagnostics order               */
    uint32_t diagnostics_valid;
} GsStats;

/* Images that gs_read_image can return. */
enum {
    GS_IMAGE_ALPHA_RGBA16F = 0,   /* sorted-alpha composite, 8 B/pixel (half RGBA) */
    GS_IMAGE_GPS_SAMPLE_RGBA32F = 1, /* GPS resolved sample, premultiplied rgb + coverage */
    GS_IMAGE_GPS_DEPTH_R32F = 2,  /* GPS resolved depth                         */
    GS_IMAGE_ACCUM_RGBA32F = 3,   /* running mean of GPS samples               */
    GS_IMAGE_FRAME_AVG_RGBA32F = 4, /* mean of this frame's StochasticSplats samples */
    GS_IMAGE_TAA_RGBA32F = 5,     /* TAA history: rgb, alpha = history length  */
    GS_IMAGE_TAA_XYZ_RGBA32F = 6, /* TAA history: world position, 1            */
    GS_IMAGE_DISPLAY_RGBA32F = 7  /* GSAnim 0.3 LUT display of a linear image  */
};

/* Buffers that gs_read_buffer can return (element counts in elements). */
enum {
    GS_BUFFER_SORTED_KEYS = 0,    /* u32 x gaussian_count: sorted depth keys, culled = 0xFFFFFFFF */
    GS_BUFFER_SORTED_IDS = 1,     /* u32 x gaussian_count: sorted Gaussian ids, culled = 0xFFFFFFFF */
    GS_BUFFER_VISIBLE_IDS = 2,    /* u32 x visible_count: visible ids in index order */
    GS_BUFFER_VISIBLE_KEYS = 3,   /* u32 x visible_count: their depth keys          */
    GS_BUFFER_PROJECTED = 4,      /* 16 floats x gaussian_count, the reference's
                                     CullProjectedSplatData layout (shaders/cull.glsl:12-17);
                                     culled rows as the reference writes them    */
    GS_BUFFER_GPS_PACKED = 5      /* u64 x (ss*W)*(ss*H): depth key << 32 | payload */
};

/* Pipeline stages timed by gs_stage_times when profiling is on.  CULL is the
 * whole cull stage (cull kernel, visible-count scan, compaction);
 * CULL_KERNEL the cull kernel alone, inside it. */
enum {
    GS_STAGE_CULL = 0,
    GS_STAGE_SORT = 1,
    GS_STAGE_BIN = 2,
    GS_STAGE_COMPOSITE = 3,
    GS_STAGE_GPS_EXPAND = 4,
    GS_STAGE_GPS_SPLAT = 5,
    GS_STAGE_GPS_RESOLVE = 6,
    GS_STAGE_CULL_KERNEL = 7,
    GS_STAGE_COUNT = 8
};

typedef struct gs_ctx gs_ctx;

/* ---- lifetime ----------------------------------------------------------- */

/* Replaces renderer_init's GPU-side setup (src/renderer.cpp:588-684).
 * device: HIP device ordinal. */
int gs_create(gs_ctx** out, int device);
/* Replaces renderer_destroy (src/renderer.h:266). NULL is accepted. */
void gs_destroy(gs_ctx* ctx);
/* Message of the last failing call on this thread ("" when none). */
const char* gs_last_error(void);
/* Human-readable device description (name, CUs, HBM, LDS). */
int gs_device_info(gs_ctx* ctx, char* out, size_t out_bytes);

/* HBM rates this device reaches with the library's access patterns, in
 * GB/s (1e9 bytes), each the best of five event-timed launches over buffers
 * of `bytes` (at least 64 MiB, allocated for the call and freed): a 16-byte
 * load per thread, the same over 15 planes read together (the cull's
 * prefetch), a 16-byte store per thread, and a load and store per thread
 * (bytes read plus bytes written).  No counterpart in the reference: the
 * ceiling against which the HBM-bound kernels here are measured. */
typedef struct GsHbmRates {
    double read_gbs;
    double planes_read_gbs;
    double write_gbs;
    double copy_gbs;
} GsHbmRates;
int gs_hbm_probe(gs_ctx* ctx, size_t bytes, GsHbmRates* out);

/* Cap on device memory this context may hold (0 = no cap).  Mirrors
 * gpu_budget_bytes / renderer_gpu_buffer_limits (src/renderer.cpp:709-734). */
int gs_set_memory_budget(gs_ctx* ctx, size_t bytes);

/* ---- upload -------------------------------------------------------------- */

/* Replaces renderer_upload_gaussians (src/renderer.cpp:736-1074), whose
 * 64 MiB chunked copies of the packed AoS records it follows: the host
 * repacks each chunk of 262,144 Gaussians into the device's attribute planes
 * (GS_SCENE_PLANES below) in one of two page-locked buffers, on its cores,
 * while the copies of the previous chunk run from the other.  Refuses
 * (GS_ERR_LIMIT) a scene that exceeds the memory budget and keeps the
 * previous scene in that case. */
int gs_upload_gaussians(gs_ctx* ctx, const GsGaussian* gaussians, uint32_t count);

/* The scene's layout in HBM: GS_SCENE_PLANES planes of `count` float4 each,
 * plane p of Gaussian i at floats (p * count + i) * 4:
 *   P0 position x y z, opacity      P1 rotation w x y z
 *   P2 scale x y z, DC r            P3 DC g b, rest[0], rest[1]
 *   P4+j rest[2+4j .. 5+4j] (j = 0..9)      P14 rest[42..44], 0
 * (rest as GsGaussian.sh_rest: RGB triples per coefficient). */
#define GS_SCENE_PLANES 15
/* Uploads a scene already in that layout.  From page-locked memory (e.g.
 * gs_alloc_pinned) it is one asynchronous copy; from pageable memory it goes
 * through the upload's page-locked chunk buffers.  Same budget rules as
 * gs_upload_gaussians.  No counterpart in the reference (loaders that
 * decode straight into the planes). */
int gs_upload_planes(gs_ctx* ctx, const float* planes, uint32_t count);
/* Page-locked host memory for gs_upload_planes / gs_decode_ply_planes
 * (hipHostMalloc); gs_free_pinned(NULL) is accepted. */
int gs_alloc_pinned(size_t bytes, void** out);
void gs_free_pinned(void* p);

/* GSAnim appearance tracks of the uploaded scene (GSAnim, src/gsanim.h:27-48),
 * as renderer_upload_gaussians packs them (src/renderer.cpp:751-758,848-893):
 * either int8 coefficient deltas per state (asset versions 1 and 2) or IEEE
 * half coefficients per layer (version 3), 48 per Gaussian row in
 * channel-major SH order (DC r, g, b, then rest coefficient k of channel c at
 * 3 + 15 c + k).  rows must equal the uploaded Gaussian count. */
typedef struct GsAnimation {
    uint32_t rows;
    uint32_t state_count;         /* delta states, state 0 (the scene itself) excluded */
    const int8_t* deltas;         /* state_count * rows * 48, state/row/channel-major */
    float scales[48];             /* per-coefficient delta scales */
    uint32_t layer_count;         /* version 3: coefficient layers (then no deltas) */
    const uint16_t* layer_data;   /* layer_count * rows * 48 half bits, layer/row/channel-major */
    uint32_t lut_size;            /* version 3 display LUT edge (0 = none) */
    const uint16_t* display_lut;  /* lut_size^3 RGB uint16 UNORM triples, red fastest */
    float log2_min, log2_max;     /* LUT input range in log2 radiance */
} GsAnimation;

/* Uploads (or with NULL drops) the scene's animation.  Deltas or layers, not
 * both; at most 256 states or layers (the loader's limit).  A new
 * gs_upload_gaussians drops the animation.  Counts against the memory
 * budget like the scene. */
int gs_upload_animation(gs_ctx* ctx, const GsAnimation* animation);

/* AnimationUBO of the next culls (shaders/cull.glsl:78-85, filled by
 * renderer_cull_gaussians_gpu, src/renderer.cpp:1823-1835, from
 * gsanim_sample / gsanim_layer_weights): with deltas, every SH coefficient
 * (DC included) is base + q * scale in state_a, lerped by weight towards
 * state_b (state 0 = the base itself); with layers, the colour is the sum
 * over the first layer_count layers of max(SH(layer coefficients), 0) times
 * the layer's RGB weight (layer_weights: 3 per layer), with no +0.5 offset.
 * layer_count 0 with layered data leaves the scene's own coefficients. */
int gs_set_animation_frame(gs_ctx* ctx, int32_t state_a, int32_t state_b, float weight, uint32_t layer_count,
                           const float* layer_weights);

/* Replaces renderer_set_gps_work_budget (src/renderer.cpp:1076-1123):
 * max_work_items in [1, 2^28]; refuses 0 or an over-budget value and keeps
 * the previous budget. */
int gs_set_gps_budget(gs_ctx* ctx, uint32_t max_work_items);

/* ---- per-frame stages ---------------------------------------------------- */

/* Replaces cull_reset_cs + cull_cs (shaders/cull.glsl:166-474) as dispatched
 * by renderer_cull_gaussians_gpu (src/renderer.cpp:1778-1871).  Produces the
 * projected splats, the order-preserving list of visible ids with their
 * far-to-near depth keys, and (optionally) the 8 diagnostics.  effect may be
 * NULL. */
int gs_cull(gs_ctx* ctx, const GsCameraUniforms* cam,
            const GsSplatEffectParams* effect, const GsCullOptions* opts);

/* Replaces the 4-pass radix sort (shaders/sort.glsl:4-164,
 * src/renderer.cpp:1873-1962): stable ascending sort of the visible (key, id)
 * pairs; culled entries sort last as in the reference. */
int gs_sort(gs_ctx* ctx);

/* No counterpart in the reference: gs_sort for callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Stable ascending sort of `count`
 * (key, id) pairs the caller holds in device memory, with the kernels of
 * gs_sort (five 7-bit passes over the full 32-bit keys; the order equals
 * gs_sort's for the same pairs: ascending key, ties by input position).
 * keys_in / ids_in are only read; keys_out / ids_out receive the sorted
 * pairs and also serve as a ping-pong buffer during the call.  All four are
 * device pointers, 4-byte aligned, not overlapping.  Runs on the context's
 * stream and returns without waiting.  The rest of the sort's memory
 * (ping-pong pair, digit table, column totals, count) is a static device
 * array of about 34 MiB per GPU, held until the process ends: the call makes
 * no hipMalloc, hipFree or other allocation.  count above
 * GS_SORT_STATIC_MAX_PAIRS is refused with GS_ERR_LIMIT.  Every call on one
 * GPU shares that array: calls that overlap in time on one GPU race. */
#define GS_SORT_STATIC_MAX_PAIRS (1u << 22)
int gs_sort_pairs_static(gs_ctx* ctx, const uint32_t* keys_in, const uint32_t* ids_in, uint32_t* keys_out,
                         uint32_t* ids_out, uint32_t count);

/* Replaces the sorted instanced splat draw into the RGBA16F target
 * (shaders/splat.glsl:6-105, src/renderer.cpp:292-299,1687-1720,2516-2541):
 * tile binning of the sorted list and front-to-back compositing, black
 * (0,0,0,0) clear.  tile_size: 8, 16 or 32 (0 = tuned default). */
int gs_composite_alpha(gs_ctx* ctx, int tile_size);

/* Replaces renderer_draw_gps_sample (src/renderer.cpp:1964-2118): clear,
 * expand under the work budget, point splat with a 64-bit depth|payload
 * atomicMin, resolve to RGBA32F + R32F.  frame_seed as stochastic_frame_seed,
 * supersample 1..4, linear_color as GSAnim v3 assets, background as the
 * resolve clear value. */
int gs_gps_sample(gs_ctx* ctx, uint32_t frame_seed, uint32_t supersample,
                  int linear_color, float background);

/* Replaces the StochasticSplats sample pass (splat_stochastic_fs,
 * shaders/splat.glsl:107-146; pass src/renderer.cpp:2331-2352): every
 * covered fragment that passes the alpha test is kept when hash13(pixel
 * centre, id + frame_seed) < alpha, the kept fragment with the smallest
 * 24-bit depth wins (ties: lowest index, the reference's draw order), and
 * the sample image (GS_IMAGE_GPS_SAMPLE_RGBA32F) receives its colour (half
 * precision, alpha 1) or (background, background, background, 0); the depth
 * image receives its D24 depth / (2^24 - 1) or 1.  jitter_px (NULL = none)
 * is draw_world's splat_jitter_px: it moves every quad and fragment centre
 * by that many pixels (shaders/splat.glsl:61,72; finite, gs_taa_jitter
 * gives |jitter| <= 0.5).  Needs gs_cull (single
 * view); sorts if gs_sort has not run.  gs_accumulate averages these
 * samples as it does GPS samples. */
int gs_stochastic_sample(gs_ctx* ctx, uint32_t frame_seed, float background, const float* jitter_px);

/* TAA jitter of a StochasticSplats sample (src/renderer.cpp:2335-2340):
 * index = frame_seed % 1024 + 1, jitter = (halton(index, 2) - 0.5,
 * halton(index, 3) - 0.5) with halton of src/renderer.cpp:246-255.  Host
 * only. */
void gs_taa_jitter(uint32_t frame_seed, float jitter_px[2]);

/* Mean of the StochasticSplats samples of one TAA frame (the avg pass,
 * src/renderer.cpp:2362-2385): frame_avg = h + (s - h) / (sample_index + 1)
 * with s the current sample image; sample_index 0 copies the sample. */
int gs_frame_average(gs_ctx* ctx, uint32_t sample_index);

/* TaaAccumUBO (shaders/accum.glsl:64-75), 208 bytes; matrices column-major. */
typedef struct GsTaaUniforms {
    float inv_view[16];
    float prev_view[16];
    float viewport_orthographic[4]; /* viewport xy, orthographic now, orthographic before */
    float focal_clip[4];            /* persp focal, ortho focal, clip_y_sign, clip_z_01 */
    float prev_focal_clip[4];       /* previous persp focal, ortho focal, clip_y_sign, unused */
    float proj_z[4];                /* proj[2][2], proj[3][2], unused, unused */
    float history_valid;
    float view_changed;
    float pad[2];
} GsTaaUniforms;

/* The TAA uniforms renderer_draw_frame builds (src/renderer.cpp:2320-2327,
 * 2449-2466): inv_view = inverse(cam->view); the previous camera's fields
 * when prev_cam is non-NULL (history valid), the current camera's
 * otherwise; view_changed as given.  A view matrix that cannot be inverted
 * gives inv_view = view and no history, as the reference's fallback, and
 * GS_ERR_INVALID so the caller can reset its accumulation.  Host only. */
int gs_taa_uniforms(const GsCameraUniforms* cam, const GsCameraUniforms* prev_cam, int view_changed,
                    GsTaaUniforms* out);

/* Replaces taa_accum_fs (shaders/accum.glsl:47-179; pass
 * src/renderer.cpp:2425-2478): per pixel, the current colour (the frame
 * average when use_frame_average, else the last sample) and the last
 * sample's depth give a world position; its reprojection into the previous
 * view samples the colour and position history bilinearly; a valid history
 * (history_valid, inside the previous view, alpha > 0 and, when the view
 * changed, within 100 units) is blended with weight n = min(alpha, 128).
 * The history is double buffered: this call reads one buffer and writes the
 * other, which GS_IMAGE_TAA_RGBA32F / GS_IMAGE_TAA_XYZ_RGBA32F then read. */
int gs_taa_accumulate(gs_ctx* ctx, const GsTaaUniforms* uniforms, int use_frame_average);

/* Replaces layers_display_fs (shaders/accum.glsl:197-240; bound by
 * renderer_display_appearance, src/renderer.cpp:2120-2137): the linear
 * image `source` (GS_IMAGE_ALPHA_RGBA16F, GS_IMAGE_GPS_SAMPLE_RGBA32F,
 * GS_IMAGE_ACCUM_RGBA32F or GS_IMAGE_TAA_RGBA32F) times 2^exposure_ev
 * through the animation's log2-domain RGB LUT (exact trilinear lookup)
 * into GS_IMAGE_DISPLAY_RGBA32F (alpha 1).  Needs a LUT from
 * gs_upload_animation. */
int gs_display_layers(gs_ctx* ctx, int source, float exposure_ev);

/* Replaces accum_fs (shaders/accum.glsl:19-45): accum = h + (s - h) / n with
 * n = sample_count (1 = first sample, history ignored). */
int gs_accumulate(gs_ctx* ctx, uint32_t sample_count);

/* Whole sorted-alpha frame: gs_cull + gs_sort + gs_composite_alpha
 * (renderer_draw_frame's alpha branch, src/renderer.cpp:2139-2289). */
int gs_render_alpha(gs_ctx* ctx, const GsCameraUniforms* cam,
                    const GsSplatEffectParams* effect, const GsCullOptions* opts,
                    int tile_size);

/* Batched multi-view sorted-alpha frame; no counterpart in the reference,
 * which draws one camera per frame.  `count` cameras (1..64, any mix of
 * viewport sizes) over the resident scene: one cull launch reads each
 * Gaussian once for all views, one depth sort orders every view's visible
 * list (a stable sort restricted to one view keeps that view's order), one
 * tile sort and one composite launch cover every (view, tile).  Each view's
 * image equals what gs_render_alpha gives for that camera.  collect_stats
 * needs count == 1.  Read the images with gs_read_view_image. */
int gs_render_views(gs_ctx* ctx, const GsCameraUniforms* cams, uint32_t count,
                    const GsSplatEffectParams* effect, const GsCullOptions* opts,
                    int tile_size);

/* RGBA16F image of view `view` of the last gs_render_views (or
 * gs_render_alpha: view 0), top row first; bytes = width * height * 8. */
int gs_read_view_image(gs_ctx* ctx, uint32_t view, void* host, size_t bytes);

/* Configuration of the last frame (no counterpart in the reference). */
typedef struct GsFrameInfo {
    uint32_t views;
    uint32_t total_tiles;
    int32_t tile_size;
    int32_t radix_bits;
    int32_t blocks_per_cu;
    int32_t graph;
    uint32_t graph_captures;   /* hipGraph captures so far   */
    uint32_t graph_replays;    /* replays of a captured frame */
    uint32_t overflow_reruns;  /* frames re-binned after a tile-entry overflow */
    uint32_t pad;
    uint64_t held_bytes;       /* device memory the context holds */
} GsFrameInfo;
int gs_frame_info(gs_ctx* ctx, GsFrameInfo* out);

/* Autotuner (no counterpart in the reference): renders the given views with
 * every tile size {8, 16, 32} x radix digit width {4, 8} x composite blocks
 * per CU {0 (one block per tile), 1, 2, 4, 8}, `frames` timed frames each
 * after one warm-up frame, keeps the fastest configuration as the context's
 * options and reports it next to the default (16, 8, 0). */
typedef struct GsTuneResult {
    int32_t tile_size;
    int32_t radix_bits;
    int32_t blocks_per_cu;
    int32_t configs;           /* configurations timed */
    double best_ms;            /* per frame, fastest configuration */
    double default_ms;         /* per frame, tile 16 / 8-bit digits / one block per tile */
} GsTuneResult;
int gs_autotune(gs_ctx* ctx, const GsCameraUniforms* cams, uint32_t count,
                const GsCullOptions* opts, int frames, GsTuneResult* out);

/* ---- learned denoiser (no counterpart in the reference) ------------------
 * A 4-level U-Net (32/48/64/96 channels, 3x3 convolutions + ReLU, 2x2
 * average pooling, nearest upsampling, skip concatenation, 1x1 head added to
 * the input mean) that maps the GPS running mean (RGBA) and GPS depth of a
 * render context to its sorted-alpha image.  bf16 weights and activations on
 * the matrix cores with fp32 accumulation; fp32 parameters and Adam state.
 * Every call is ordered on the render context's stream.
 * Parameter layout (float32): per 3x3 convolution in the order e0a e0b e1a
 * e1b e2a e2b ba bb d2a d2b d1a d1b d0a d0b, W[cout][tap][cin] (tap =
 * (dy + 1) * 3 + dx + 1) then b[cout]; then the head W[4][32], b[4]. */
typedef struct gs_denoiser gs_denoiser;
uint32_t gs_denoiser_param_count(void);
/* He-normal initial weights from `seed`. */
int gs_denoiser_create(gs_ctx* ctx, uint64_t seed, gs_denoiser** out);
void gs_denoiser_destroy(gs_denoiser* d);
int gs_denoiser_get_params(gs_denoiser* d, float* host, size_t count);
/* Also clears the Adam state. */
int gs_denoiser_set_params(gs_denoiser* d, const float* host, size_t count);
/* One forward + backward pass over a host batch: input (n, h, w, 8) float
 * (channels: mean RGBA, depth, 0, 0, 0; rounded to bf16), target (n, h, w,
 * 4); h, w multiples of 8.  Returns the mean squared error, optionally the
 * output (n, h, w, 4) and the gradient of every parameter (count floats). */
int gs_denoiser_forward_backward(gs_denoiser* d, const float* input, const float* target, uint32_t n,
                                 uint32_t h, uint32_t w, float* loss, float* output, float* grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on the last gradients. */
int gs_denoiser_adam_step(gs_denoiser* d, float lr);
/* Adds `crops` random 256x256 crops of the context's current frame (GPS
 * running mean + depth as input, sorted-alpha RGBA16F as target, same camera
 * and size) to the training pool (1024 crops, oldest replaced first). */
int gs_denoiser_capture(gs_denoiser* d, uint32_t crops, uint64_t seed);
int gs_denoiser_pool_size(gs_denoiser* d, uint32_t* count);
/* `steps` Adam steps on random batches of pooled crops; losses (optional,
 * steps floats) receive each step's mean squared error. */
int gs_denoiser_train(gs_denoiser* d, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* losses);
/* Denoises the context's current GPS accumulation into an RGBA16F image. */
int gs_denoise(gs_denoiser* d);
/* The judge of a denoised frame: PSNR (peak 1, mean squared error over
 * every RGBA value) and the largest absolute per-channel difference of the
 * last gs_denoise image against the context's sorted-alpha image of the same
 * cull and size (gs_sort + gs_composite_alpha after the GPS sample).  A
 * caller falls back to the sorted-alpha image when these miss its bars.
 * Synchronises the stream. */
int gs_denoiser_judge(gs_denoiser* d, double* psnr_db, float* max_abs);
/* The last denoised image, top row first; bytes = width * height * 8. */
int gs_denoiser_read_image(gs_denoiser* d, void* host, size_t bytes);

/* Options (gs_set_option). */
enum {
    /* 1: the GPS resolve leaves the depth|payload cells in place so they can
     * be read with GS_BUFFER_GPS_PACKED; the next sample clears them first.
     * 0 (default): the resolve clears the cells it reads. */
    GS_OPT_GPS_KEEP_CELLS = 0,
    /* Default tile size of gs_composite_alpha when 0 is passed (8, 16, 32). */
    GS_OPT_TILE_SIZE = 1,
    /* 16x16 tiles: 1 (default) = per-wave 8x8 quadrant kernel that walks
     * the splats whose rectangle and alpha >= 1/255 ellipse meet the
     * quadrant; 2 = the same kernel with the rectangle test only; 0 = one
     * pixel per thread over every splat of the tile.  Images are identical
     * for 1 and 2. */
    GS_OPT_QUADRANT_CULL = 2,
    /* 1 (default): each splat's pixel rectangle (binning and compositing) is
     * cut to the box of the ellipse where its fragment test can pass, with a
     * margin that keeps every passing pixel inside; images are identical to
     * 0 = the full quad rectangle, with fewer (tile, splat) pairs.  Takes
     * effect from the next gs_cull. */
    GS_OPT_TIGHT_RECTS = 3,
    /* Radix digit width of the depth and tile sorts: 8 (default) or 4. */
    GS_OPT_RADIX_BITS = 4,
    /* Composite launch: 0 (default) = one block per tile; k > 0 = k
     * persistent blocks per CU pulling tiles from per-XCD queues. */
    GS_OPT_BLOCKS_PER_CU = 5,
    /* 1: gs_render_alpha / gs_render_views capture their frame once into a
     * hipGraph and replay it while every launch argument is unchanged
     * (cameras and view sizes live in device memory, so moving cameras
     * replay; new buffer sizes, options or effect parameters re-capture).
     * Ignored while profiling is on.  0 (default): direct launches. */
    GS_OPT_GRAPH = 6,
    /* 1: sorted-alpha composites use the reference's blend arithmetic
     * (ONE / ONE_MINUS_SRC_ALPHA into RGBA16F, src/renderer.cpp:292-299):
     * each tile list is walked back to front, the fragment shader's
     * operations run in source order (shaders/splat.glsl:84-104, with the
     * deterministic log and exp), and every channel is rounded to half
     * after every blend, as the target stores it; on 16x16 tiles whatever
     * tile size is asked.  0 (default): front to back in fp32 with early
     * termination, rounded to half once. */
    GS_OPT_REFERENCE_BLEND = 7,
    /* 1: the radix sorts and the bin address their arrays with 64-bit
     * element indices even where 32-bit byte offsets suffice (the variants
     * arrays of 2^30 or more elements take); results are identical.  For
     * testing those variants at small sizes.  0 (default): off. */
    GS_OPT_WIDE_INDEX = 8,
    /* 1 (default): frames with at most 65536 tiles (over all views) bin
     * their (tile, splat) pairs with 16-bit tile keys; 0: 32-bit keys.
     * Results are identical. */
    GS_OPT_TILE_KEY16 = 9
};
int gs_set_option(gs_ctx* ctx, int option, int value);

/* ---- readback ------------------------------------------------------------ */

int gs_read_image(gs_ctx* ctx, int which, void* host, size_t bytes);
int gs_read_buffer(gs_ctx* ctx, int which, void* host, size_t bytes);
int gs_stats(gs_ctx* ctx, GsStats* out);
int gs_synchronize(gs_ctx* ctx);

/* ---- measurement --------------------------------------------------------- */

/* Enable per-stage HIP-event timing (the counterpart of the reference's
 * PROFILE_GPU zones, src/renderer.cpp:1890-1937,1984-2088). */
int gs_set_profiling(gs_ctx* ctx, int enabled);
/* Sum of stage times (ms) and launch counts since the last reset; ms and
 * counts have GS_STAGE_COUNT entries.  Synchronises the stream. */
int gs_stage_times(gs_ctx* ctx, double* ms, uint32_t* counts);
int gs_reset_stage_times(gs_ctx* ctx);
/* Opaque hipStream_t of the context, for event timing by the caller. */
void* gs_stream(gs_ctx* ctx);

/* ---- scene decoding (host only) ------------------------------------------ */

/* Float offsets of the 3DGS properties within one binary PLY vertex row
 * (-1 = absent); rest[k] is f_rest_k (all red coefficients, then green, then
 * blue). */
typedef struct GsPlyLayout {
    uint32_t stride;  /* floats per row */
    int32_t x, y, z;
    int32_t scale[3];
    int32_t rot[4];
    int32_t opacity;
    int32_t dc[3];
    int32_t rest[45];
} GsPlyLayout;

/* Activation of PLY rows into Gaussians as the reference's PLY loader does
 * (src/gaussian_loader_ply.cpp:143-203): exp of the log scales (0.01 when
 * absent), normalised quaternion (w, x, y, z; identity when absent), sigmoid
 * opacity (1 when absent), raw DC, rest coefficients repacked from
 * channel-major to RGB triples. */
int gs_decode_ply(const float* rows, uint32_t count, const GsPlyLayout* layout, GsGaussian* out);

/* A SOG v2 scene with its WebP images already decoded to RGBA8 (lossless:
 * any decoder gives the same pixels) and meta.json parsed. */
typedef struct GsSogData {
    uint32_t count;
    float means_mins[3], means_maxs[3];
    float scales_codebook[256];
    float sh0_codebook[256];
    const uint8_t* means_l;   /* each RGBA8 with at least count pixels, row-major */
    const uint8_t* means_u;
    const uint8_t* scales;
    const uint8_t* quats;
    const uint8_t* sh0;
    int32_t shN_bands;        /* 0 = no higher bands, else 1..3 */
    int32_t shN_count;        /* palette entries, 1..65536 */
    float shN_codebook[256];
    const uint8_t* shN_centroids;  /* RGBA8, 64 * coefficients wide */
    int32_t centroids_width, centroids_height;
    const uint8_t* shN_labels;     /* RGBA8, at least count pixels */
} GsSogData;

/* Per-Gaussian decode of the reference's SOG loader
 * (src/gaussian_loader_sog.cpp:403-510): 16-bit means lerped between the
 * bounds and un-logged (sign(n) (e^|n| - 1)); scales exp(codebook); DC from
 * its codebook; smallest-three quaternion (the mode byte 252..255 names the
 * omitted component; stored x, y, z, w in rotation[0..3] as the reference
 * keeps it); opacity = alpha / 255; palette SH by 16-bit label.  An invalid
 * quaternion mode or label is GS_ERR_INVALID. */
int gs_decode_sog(const GsSogData* sog, GsGaussian* out);

/* The decoders straight into the scene's plane layout (GS_SCENE_PLANES,
 * plane stride count), with the same arithmetic: a loader's fast path
 * (src/gaussian_loader_ply.cpp:116-203 reads 4096-vertex batches, then
 * renderer_upload_gaussians repacks and copies them).  Host only. */
int gs_planes_from_gaussians(const GsGaussian* gaussians, uint32_t count, float* planes);
int gs_decode_ply_planes(const float* rows, uint32_t count, const GsPlyLayout* layout, float* planes);
/* Decode and upload in one pipeline: the host decodes chunk k of the rows
 * (or SOG texels) into page-locked plane buffers while the copies of chunk
 * k - 1 run; the result equals gs_decode_ply / gs_decode_sog followed by
 * gs_upload_gaussians.  Same budget rules as gs_upload_gaussians. */
int gs_upload_ply(gs_ctx* ctx, const float* rows, uint32_t count, const GsPlyLayout* layout);
int gs_upload_sog(gs_ctx* ctx, const GsSogData* sog);

/* ---- synthetic scenes and cameras (host only) --------------------------- */

/* Seeded synthetic scene as the reference's PLY loader would activate it
 * (src/gaussian_loader_ply.cpp:143-203).  Positions from `clusters` isotropic
 * normal clusters (std 1.5, centres uniform in [-8,8]^3); log-scale
 * N(-4.5, 0.8); quaternion N(0,1)^4 normalised; opacity logit
 * N(opacity_mean, 2); DC N(0, 0.5); rest coefficient of band l N(0, 0.1/(l+1)). */
int gs_generate_scene(uint64_t seed, uint32_t count, uint32_t clusters,
                      float opacity_logit_mean, GsGaussian* out);

/* Camera uniforms for a baked reference view (src/main.cpp:165-193): a
 * row-major camera_to_world pose, focal/height, target size. */
int gs_camera_from_pose(const float camera_to_world[16], float focal_over_height,
                        float width, float height, GsCameraUniforms* out);
/* Pose looking at the origin from `position` with world up +Z (Blender
 * convention of the baked views); row-major camera_to_world. */
int gs_look_at_origin_pose(const float position[3], float camera_to_world[16]);

#ifdef __cplusplus
}
#endif

#endif /* GS_API_H */
