// gs_api.cpp - C ABI of libgamesplat_hip: context, device memory, stage
// sequencing and readback.  The per-stage kernels live in gs_*.hip.
//
// Stage order of a frame mirrors renderer_draw_frame
// (src/renderer.cpp:2139-2289): cull -> sort -> composite for sorted alpha,
// cull -> GPS sample -> accumulate for the stochastic path.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gs_internal.h"
#include "gs_planes.h"

using namespace gs;

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int ok() {
    g_error.clear();
    return GS_OK;
}

#define GS_HIP(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(GS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                  \
    } while (0)

#define GS_CHECK(expr)             \
    do {                           \
        int r_ = (expr);           \
        if (r_ != GS_OK) return r_; \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

constexpr uint32_t kMaxGaussians = (1u << 30) - 1u;
constexpr uint32_t kMaxWorkItems = 1u << 28;
constexpr uint32_t kUploadChunk = 262144u;  // Gaussians per staging copy (about 59 MiB)

struct TimedLaunch {
    int stage;
    hipEvent_t start, stop;
};

// Per-frame constants of a (batched) frame, uploaded through a pinned ring.
constexpr uint32_t kMaxAnimLayers = 256;  // the GSAnim loader's state / layer limit
struct FrameConst {
    GsCameraUniforms cams[kMaxViews];
    ViewInfo views[kMaxViews];
    float layer_weights[kMaxAnimLayers * 3];
};
constexpr int kRing = 8;

// Every launch of one sorted-alpha frame, by value.  Two frames with the same
// bytes here launch the same kernels with the same arguments, which is what
// lets a captured hipGraph be replayed.
struct FramePlan {
    CullLaunch cull;
    uint32_t cull_blocks;    // per view
    SortLaunch depth;
    BinLaunch bin;
    SortLaunch tiles;
    CompositeLaunch comp;
    int pack_tile;           // tile size the compact packs rectangles for (0: none)
    int pack_pad;            // 1: it moves the cull's padded rectangles instead
};

}  // namespace

struct gs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    size_t budget = 0;
    size_t held = 0;

    uint32_t n = 0;
    DevBuf vrects[4];  // packed tile rectangles moved by the depth sort
    DevBuf block_range;  // [view][block]: visible key range
    DevBuf packed_rects;  // [view][n]: the cull's packed tile rectangles (grown by a pixel for StochasticSplats)
    // cull_counts, cull_offsets: the cull's per-block visible counts and
    // their offsets, kept for the compaction (block_counts and block_offsets
    // are shared by the bin and GPS scans)
    DevBuf box_partials;  // the centre-box reduction's per-block boxes
    DevBuf planes, records, keys[4], ids[4], dense_keys, cull_counts, cull_offsets, block_counts, block_offsets, block_diag, rects, rect_px,
        table,
        counters;

    DevBuf dup_keys[3], dup_vals[3], ranges, image;
    uint32_t dup_cap = 0, last_dup = 0, overflow_reruns = 0;
    uint32_t* dup_host = nullptr;     // pinned copy of the requested pair count
    hipEvent_t dup_event = nullptr;
    bool dup_pending = false;
    int composite_tile = 16;
    int composite_kind = 0;       // the pending frame: 0 sorted alpha, 1 stochastic splats
    uint32_t stoch_seed = 0;
    float stoch_bg = 0.0f;
    float stoch_jitter[2] = {0.0f, 0.0f};
    // GSAnim appearance data (anim_kind 1 deltas, 2 layers) and frame state
    DevBuf anim_data, lut, display;
    int anim_kind = 0;
    uint32_t anim_sets = 0;  // states (excluding state 0) or layers uploaded
    float anim_scales[48] = {};
    uint32_t lut_size = 0;
    float lut_log2_min = 0.0f, lut_log2_max = 0.0f;
    int32_t anim_state_a = 0, anim_state_b = 0;
    float anim_weight = 0.0f;
    uint32_t anim_layer_count = 0;
    std::vector<float> anim_layer_weights;
    int disp_w = 0, disp_h = 0;
    bool have_display = false;
    // Tile lists of StochasticSplats samples depend on the cull and sort
    // only, so the samples of one cull share them.  bin_epoch advances with
    // every launch that may rewrite the tile buffers; the lists are reused
    // while it equals stoch_bins_epoch.  stoch_bins_mode: 0 the cull's
    // rectangles (unjittered samples), 1 grown by one pixel (any jitter up
    // to half a pixel), 2 moved by stoch_bins_jitter exactly.
    uint64_t bin_epoch = 1, stoch_bins_epoch = 0;
    int stoch_bins_mode = 0;
    float stoch_bins_jitter[2] = {0.0f, 0.0f};
    int img_w = 0, img_h = 0;

    // page-locked chunk buffers of the pipelined uploads
    float* up_pin[2] = {nullptr, nullptr};
    hipEvent_t up_ev[2] = {nullptr, nullptr};
    bool up_used[2] = {false, false};

    uint32_t gps_budget = 16u * 1024u * 1024u;  // renderer_init default (src/renderer.cpp:595)
    DevBuf work, prepared, item_count, item_base, cells, sample, depth, accum;
    DevBuf gps_hint;  // per GPS count block: prepare in the count pass (see gps_count_kernel)
    int gps_w = 0, gps_h = 0, gps_ss = 1, res_w = 0, res_h = 0, acc_w = 0, acc_h = 0;
    // StochasticSplats + TAA: frame average and the double-buffered history
    DevBuf frame_avg, taa_color[2], taa_xyz[2];
    int avg_w = 0, avg_h = 0, taa_w = 0, taa_h = 0;
    int taa_last = 0;        // history buffer the last TAA pass wrote
    bool have_avg = false, have_taa = false;
    bool cells_dirty = true;
    int keep_cells = 0;
    int default_tile = 16;
    int quadrant_cull = 1;
    int reference_blend = 0;  // the composite's blend arithmetic (GS_OPT_REFERENCE_BLEND)
    int wide_index = 0;       // GS_OPT_WIDE_INDEX
    int tile_key16 = 1;       // GS_OPT_TILE_KEY16
    bool rect_px_valid = false;  // the last cull wrote pixel rectangles (else packed tile rectangles only)
    int tight_rects = 1;
    int cull_tight = 1;      // tight_rects as the last cull used it

    GsCameraUniforms cam{};  // view 0 of the current frame
    uint32_t views = 1;      // views of the current frame
    uint32_t view_cap = 0;   // views the [view][n] arrays are sized for
    std::vector<ViewInfo> vinfo;
    std::vector<GsCameraUniforms> host_cams;
    uint32_t total_tiles = 0, total_pixels = 0;
    int vinfo_tile = 0;
    DevBuf frame_dev;        // FrameConst on the device
    FrameConst* ring = nullptr;  // pinned, kRing slots
    hipEvent_t ring_event[kRing] = {};
    bool ring_used[kRing] = {};
    int ring_next = 0;
    int radix_bits = 8;
    int blocks_per_cu = 0;
    int use_graph = 0;
    FramePlan plan{};
    bool plan_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint8_t> graph_key;
    uint32_t graph_captures = 0, graph_replays = 0;
    bool have_cull = false, have_sort = false, have_composite = false, have_gps = false, have_accum = false;
    // the last cull's visible list (keys[0], ids[0], vrects[0]) is built:
    // GPS and the visible-list readbacks build it on demand; the depth sort
    // otherwise compacts in its first pass.  0: none, 1: ids only (GPS),
    // 2: keys, ids and tile rectangles
    int compacted = 0;
    // box of the scene's Gaussian centres (valid: every centre finite), which
    // bounds the depth keys' range and so the depth-sort passes launched
    bool box_valid = false;
    double box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};
    bool stats_collected = false;
    int sort_slot = 0;
    int rect_tile = 0;       // tile size of the packed rectangles the last cull wrote (0: none)
    int rect_pad = 0;        // 1: they are grown by a pixel (StochasticSplats lists)

    bool profiling = false;
    std::vector<TimedLaunch> timed;
    std::vector<hipEvent_t> free_events;
    double stage_ms[GS_STAGE_COUNT] = {};
    uint32_t stage_count[GS_STAGE_COUNT] = {};

    DevCounters* dc() { return static_cast<DevCounters*>(counters.p); }
};

namespace {

int set_device(gs_ctx* c) {
    GS_HIP(hipSetDevice(c->device));
    return GS_OK;
}

void release(gs_ctx* c, DevBuf& b) {
    if (b.p) {
        (void)hipFree(b.p);
        c->held -= b.bytes;
    }
    b.p = nullptr;
    b.bytes = 0;
}

// Grow `b` to at least `bytes` (slack: extra fraction for buffers whose size
// varies per frame).  Existing contents are not preserved.
int ensure(gs_ctx* c, DevBuf& b, size_t bytes, const char* what, double slack = 0.0) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return GS_OK;
    size_t want = bytes + (size_t)((double)bytes * slack);
    want = (want + 255) & ~(size_t)255;
    if (c->budget && c->held - b.bytes + want > c->budget)
        return fail(GS_ERR_LIMIT, "%s needs %.1f MiB; memory budget %.1f MiB with %.1f MiB held", what,
                    want / 1048576.0, c->budget / 1048576.0, c->held / 1048576.0);
    release(c, b);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(GS_ERR_OOM, "hipMalloc(%zu) for %s failed: %s", want, what, hipGetErrorString(e));
    }
    b.p = p;
    b.bytes = want;
    c->held += want;
    return GS_OK;
}

template <typename T>
T* ptr(DevBuf& b) {
    return static_cast<T*>(b.p);
}

int ensure_table(gs_ctx* c, uint32_t elems) {
    return ensure(c, c->table, (size_t)sort_table_words(elems) * 4, "radix tile table");
}

// Size the [view][n] arrays (records, dense keys, pixel rectangles, the
// visible list and its sort buffers) for `views` views of the scene.
int ensure_views(gs_ctx* c, uint32_t views) {
    if (views <= c->view_cap) return GS_OK;
    const size_t cap = (size_t)views * c->n;
    if (cap > 0xFFFFFFFFull) return fail(GS_ERR_LIMIT, "%u views x %u Gaussians exceed 2^32 splats", views, c->n);
    GS_CHECK(ensure(c, c->records, cap * kRecordVec4 * 16, "projected splats"));
    for (int k = 0; k < 4; ++k) {
        GS_CHECK(ensure(c, c->keys[k], cap * 4, "sort keys"));
        GS_CHECK(ensure(c, c->ids[k], cap * 4, "sort ids"));
        GS_CHECK(ensure(c, c->vrects[k], cap * 4, "sorted tile rectangles"));
    }
    const size_t blocks = (size_t)views * ((c->n + 255u) / 256u);
    GS_CHECK(ensure(c, c->dense_keys, cap * 4, "dense keys"));
    GS_CHECK(ensure(c, c->block_counts, blocks * 4, "block counts"));
    GS_CHECK(ensure(c, c->block_range, blocks * 8, "block key ranges"));
    GS_CHECK(ensure(c, c->block_offsets, blocks * 4, "block offsets"));
    GS_CHECK(ensure(c, c->cull_counts, blocks * 4, "cull block counts"));
    GS_CHECK(ensure(c, c->cull_offsets, blocks * 4, "cull block offsets"));
    GS_CHECK(ensure(c, c->rects, cap * 8, "bin rectangles"));
    GS_CHECK(ensure(c, c->rect_px, cap * 8, "pixel rectangles"));
    GS_CHECK(ensure(c, c->packed_rects, cap * 4, "packed tile rectangles"));
    GS_CHECK(ensure_table(c, (uint32_t)std::max<size_t>(cap, c->dup_cap)));
    c->view_cap = views;
    return GS_OK;
}

// Next pinned ring slot, once the copy that last read it has completed.
int ring_slot(gs_ctx* c, FrameConst** slot, hipEvent_t* ev) {
    const int k = c->ring_next;
    c->ring_next = (k + 1) % kRing;
    if (c->ring_used[k]) GS_HIP(hipEventSynchronize(c->ring_event[k]));
    c->ring_used[k] = true;
    *slot = c->ring + k;
    *ev = c->ring_event[k];
    return GS_OK;
}

FrameConst* frame_dev(gs_ctx* c) { return static_cast<FrameConst*>(c->frame_dev.p); }

int upload_cameras(gs_ctx* c, const GsCameraUniforms* cams, uint32_t count) {
    FrameConst* slot;
    hipEvent_t ev;
    GS_CHECK(ring_slot(c, &slot, &ev));
    memcpy(slot->cams, cams, sizeof(GsCameraUniforms) * count);
    c->host_cams.assign(cams, cams + count);
    GS_HIP(hipMemcpyAsync(frame_dev(c)->cams, slot->cams, sizeof(GsCameraUniforms) * count, hipMemcpyHostToDevice,
                          c->stream));
    if (c->anim_kind == 2 && c->anim_layer_count) {
        const size_t wb = sizeof(float) * 3 * c->anim_layer_count;
        memcpy(slot->layer_weights, c->anim_layer_weights.data(), wb);
        GS_HIP(hipMemcpyAsync(frame_dev(c)->layer_weights, slot->layer_weights, wb, hipMemcpyHostToDevice, c->stream));
    }
    GS_HIP(hipEventRecord(ev, c->stream));
    c->cam = cams[0];
    c->views = count;
    return GS_OK;
}

// Tile grid and image placement of every view for `tile`, uploaded when it
// changes.
int upload_views(gs_ctx* c, const GsCameraUniforms* cams, int tile) {
    std::vector<ViewInfo> vinfo(c->views);
    uint64_t tiles = 0, pixels = 0;
    for (uint32_t v = 0; v < c->views; ++v) {
        ViewInfo& vi = vinfo[v];
        vi.width = (int)cams[v].viewport[0];
        vi.height = (int)cams[v].viewport[1];
        vi.tiles_x = (vi.width + tile - 1) / tile;
        vi.tiles_y = (vi.height + tile - 1) / tile;
        vi.tile_base = (uint32_t)tiles;
        vi.pixel_base = (uint32_t)pixels;
        tiles += (uint64_t)vi.tiles_x * vi.tiles_y;
        pixels += (uint64_t)vi.width * vi.height;
    }
    if (tiles > 0x7FFFFFFFull || pixels > 0x7FFFFFFFull)
        return fail(GS_ERR_LIMIT, "views need %llu tiles / %llu pixels, over 2^31", (unsigned long long)tiles,
                    (unsigned long long)pixels);
    if (c->vinfo_tile == tile && c->vinfo.size() == vinfo.size() &&
        memcmp(c->vinfo.data(), vinfo.data(), sizeof(ViewInfo) * vinfo.size()) == 0)
        return GS_OK;
    c->vinfo.swap(vinfo);
    FrameConst* slot;
    hipEvent_t ev;
    GS_CHECK(ring_slot(c, &slot, &ev));
    memcpy(slot->views, c->vinfo.data(), sizeof(ViewInfo) * c->views);
    GS_HIP(hipMemcpyAsync(frame_dev(c)->views, slot->views, sizeof(ViewInfo) * c->views, hipMemcpyHostToDevice,
                          c->stream));
    GS_HIP(hipEventRecord(ev, c->stream));
    c->total_tiles = (uint32_t)tiles;
    c->total_pixels = (uint32_t)pixels;
    c->vinfo_tile = tile;
    return GS_OK;
}

int timed_begin(gs_ctx* c, int stage, TimedLaunch* t) {
    if (!c->profiling) return GS_OK;
    t->stage = stage;
    for (hipEvent_t* ev : {&t->start, &t->stop}) {
        if (!c->free_events.empty()) {
            *ev = c->free_events.back();
            c->free_events.pop_back();
        } else {
            GS_HIP(hipEventCreate(ev));
        }
    }
    GS_HIP(hipEventRecord(t->start, c->stream));
    return GS_OK;
}

int timed_end(gs_ctx* c, TimedLaunch* t) {
    if (!c->profiling) return GS_OK;
    GS_HIP(hipEventRecord(t->stop, c->stream));
    c->timed.push_back(*t);
    return GS_OK;
}

int check_camera(const GsCameraUniforms* cam) {
    if (!cam) return fail(GS_ERR_INVALID, "camera uniforms are NULL");
    for (int k = 0; k < 2; ++k) {
        float v = cam->viewport[k];
        if (!(v >= 1.0f && v <= 32768.0f) || v != (float)(int)v)
            return fail(GS_ERR_INVALID, "viewport must be integral pixel sizes in [1, 32768], got %g x %g",
                        cam->viewport[0], cam->viewport[1]);
    }
    if (cam->clip_y_sign != 1.0f && cam->clip_y_sign != -1.0f)
        return fail(GS_ERR_INVALID, "clip_y_sign must be +1 (GL) or -1 (WebGPU)");
    return GS_OK;
}

// GL framebuffer rows start at the bottom; readback returns the top row first.
bool gl_rows(const gs_ctx* c) { return c->cam.clip_y_sign > 0.0f; }

int read_u32(gs_ctx* c, const uint32_t* dev, uint32_t* host) {
    GS_HIP(hipMemcpyAsync(host, dev, 4, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    return GS_OK;
}

int copy_image(gs_ctx* c, const void* dev, void* host, int w, int h, size_t px_bytes, size_t bytes, bool flip) {
    size_t need = (size_t)w * h * px_bytes;
    if (bytes != need) return fail(GS_ERR_INVALID, "image needs %zu bytes, got %zu", need, bytes);
    std::vector<uint8_t> tmp(need);
    GS_HIP(hipMemcpyAsync(tmp.data(), dev, need, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    size_t row = (size_t)w * px_bytes;
    for (int y = 0; y < h; ++y) {
        int src = flip ? h - 1 - y : y;
        memcpy(static_cast<uint8_t*>(host) + (size_t)y * row, tmp.data() + (size_t)src * row, row);
    }
    return GS_OK;
}

}  // namespace

extern "C" {

const char* gs_last_error(void) { return g_error.c_str(); }

int gs_create(gs_ctx** out, int device) {
    if (!out) return fail(GS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    GS_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(GS_ERR_INVALID, "device %d out of range (%d devices)", device, count);
    gs_ctx* c = new gs_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&c->props, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(GS_ERR_HIP, "device setup failed: %s", hipGetErrorString(e));
    }
    // kernel attributes are set here, never while a frame is being captured;
    // the static sort scratch's address on this device is looked up once
    e = init_kernel_attributes();
    // a failed lookup leaves gs_sort_pairs_static refusing with the reason;
    // every other entry point works without it
    if (e == hipSuccess && static_scratch_init(device) != hipSuccess) (void)hipGetLastError();
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(GS_ERR_HIP, "kernel attribute setup failed: %s", hipGetErrorString(e));
    }
    int r = ensure(c, c->counters, sizeof(DevCounters), "counters");
    if (r == GS_OK) r = ensure(c, c->frame_dev, sizeof(FrameConst), "frame constants");
    if (r == GS_OK) {
        e = hipMemsetAsync(c->counters.p, 0, sizeof(DevCounters), c->stream);
        if (e == hipSuccess) e = hipHostMalloc((void**)&c->dup_host, 4, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&c->dup_event, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&c->ring, sizeof(FrameConst) * kRing, hipHostMallocDefault);
        for (int k = 0; k < kRing && e == hipSuccess; ++k)
            e = hipEventCreateWithFlags(&c->ring_event[k], hipEventDisableTiming);
        if (e != hipSuccess) r = fail(GS_ERR_HIP, "context init failed: %s", hipGetErrorString(e));
    }
    if (r != GS_OK) {
        gs_destroy(c);
        return r;
    }
    *out = c;
    return ok();
}

void gs_destroy(gs_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& t : c->timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    DevBuf* bufs[] = {&c->planes, &c->records, &c->keys[0], &c->keys[1], &c->keys[2], &c->keys[3], &c->ids[0],
                      &c->ids[1], &c->ids[2], &c->ids[3], &c->vrects[3], &c->block_range, &c->dense_keys, &c->cull_counts, &c->cull_offsets, &c->box_partials, &c->block_counts, &c->block_offsets, &c->block_diag,
                      &c->rects, &c->rect_px, &c->table, &c->counters, &c->dup_keys[0], &c->dup_keys[1], &c->dup_keys[2],
                      &c->dup_vals[0], &c->dup_vals[1], &c->dup_vals[2], &c->ranges, &c->image, &c->work,
                      &c->prepared, &c->item_count, &c->item_base, &c->cells, &c->sample, &c->depth, &c->accum,
                      &c->frame_avg, &c->taa_color[0], &c->taa_color[1], &c->taa_xyz[0], &c->taa_xyz[1],
                      &c->anim_data, &c->lut, &c->display, &c->vrects[0], &c->vrects[1], &c->vrects[2],
                      &c->packed_rects, &c->gps_hint};
    for (DevBuf* b : bufs) release(c, *b);
    release(c, c->frame_dev);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (int k = 0; k < kRing; ++k)
        if (c->ring_event[k]) (void)hipEventDestroy(c->ring_event[k]);
    if (c->ring) (void)hipHostFree(c->ring);
    if (c->dup_event) (void)hipEventDestroy(c->dup_event);
    if (c->dup_host) (void)hipHostFree(c->dup_host);
    for (int b = 0; b < 2; ++b) {
        if (c->up_ev[b]) (void)hipEventDestroy(c->up_ev[b]);
        if (c->up_pin[b]) (void)hipHostFree(c->up_pin[b]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int gs_hbm_probe(gs_ctx* c, size_t bytes, GsHbmRates* out) {
    if (!c || !out) return fail(GS_ERR_INVALID, "bad arguments");
    if (bytes < ((size_t)64 << 20)) return fail(GS_ERR_INVALID, "probe buffers need at least 64 MiB");
    GS_CHECK(set_device(c));
    GS_HIP(hipStreamSynchronize(c->stream));  // nothing else runs while the probes are timed
    const size_t n = bytes / 16;
    void *a = nullptr, *b = nullptr, *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best[4] = {1e30, 1e30, 1e30, 1e30};
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMalloc(&sink, 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * 16, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b, 0, n * 16, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int rep = 0; rep < 5 && e == hipSuccess; ++rep) {
        for (int w = 0; w < 4 && e == hipSuccess; ++w) {
            e = hipEventRecord(e0, c->stream);
            if (e == hipSuccess) e = launch_probe(w, static_cast<const float4*>(a), static_cast<float4*>(b), n,
                                                  static_cast<float*>(sink), c->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess && ms > 0.0f) best[w] = std::min(best[w], (double)ms);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (sink) (void)hipFree(sink);
    if (e != hipSuccess) return fail(GS_ERR_HIP, "HBM probe failed: %s", hipGetErrorString(e));
    const double planes_bytes = (double)(n / 15) * 15 * 16;
    out->read_gbs = (double)n * 16 / best[0] / 1e6;
    out->planes_read_gbs = planes_bytes / best[1] / 1e6;
    out->write_gbs = (double)n * 16 / best[2] / 1e6;
    out->copy_gbs = 2.0 * n * 16 / best[3] / 1e6;
    return ok();
}

int gs_device_info(gs_ctx* c, char* out, size_t out_bytes) {
    if (!c || !out || !out_bytes) return fail(GS_ERR_INVALID, "bad arguments");
    snprintf(out, out_bytes, "%s (%s): %d CUs, %.1f GiB HBM, %zu KiB LDS/block, clock %d MHz, mem clock %d MHz, bus %d bit",
             c->props.name, c->props.gcnArchName, c->props.multiProcessorCount,
             c->props.totalGlobalMem / 1073741824.0, c->props.sharedMemPerBlock / 1024, c->props.clockRate / 1000,
             c->props.memoryClockRate / 1000, c->props.memoryBusWidth);
    return ok();
}

int gs_set_memory_budget(gs_ctx* c, size_t bytes) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    c->budget = bytes;
    return ok();
}

int gs_set_option(gs_ctx* c, int option, int value) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    switch (option) {
        case GS_OPT_GPS_KEEP_CELLS:
            c->keep_cells = value ? 1 : 0;
            return ok();
        case GS_OPT_TILE_SIZE:
            if (value != 8 && value != 16 && value != 32) return fail(GS_ERR_INVALID, "tile size must be 8, 16 or 32");
            c->default_tile = value;
            return ok();
        case GS_OPT_QUADRANT_CULL:
            if (value < 0 || value > 2) return fail(GS_ERR_INVALID, "quadrant cull must be 0, 1 or 2");
            c->quadrant_cull = value;
            return ok();
        case GS_OPT_TIGHT_RECTS:
            c->tight_rects = value ? 1 : 0;
            return ok();
        case GS_OPT_RADIX_BITS:
            if (value != 4 && value != 8) return fail(GS_ERR_INVALID, "radix digit width must be 4 or 8");
            c->radix_bits = value;
            return ok();
        case GS_OPT_BLOCKS_PER_CU:
            if (value < 0 || value > 64) return fail(GS_ERR_INVALID, "blocks per CU must be in [0, 64]");
            c->blocks_per_cu = value;
            return ok();
        case GS_OPT_GRAPH:
            c->use_graph = value ? 1 : 0;
            return ok();
        case GS_OPT_REFERENCE_BLEND:
            c->reference_blend = value ? 1 : 0;
            return ok();
        case GS_OPT_WIDE_INDEX:
            c->wide_index = value ? 1 : 0;
            return ok();
        case GS_OPT_TILE_KEY16:
            c->tile_key16 = value ? 1 : 0;
            return ok();
        default:
            return fail(GS_ERR_INVALID, "unknown option %d", option);
    }
}

}  // extern "C"

namespace {

// Device planes for a scene of `count` Gaussians, after the memory-budget
// checks of renderer_upload_gaussians (src/renderer.cpp:709-734).
int reserve_planes(gs_ctx* c, uint32_t count, DevBuf* planes) {
    if (count == 0 || count > kMaxGaussians)
        return fail(GS_ERR_LIMIT, "gaussian count %u outside [1, %u]", count, kMaxGaussians);
    // Scene-resident bytes per Gaussian: planes 240 + record 48 + keys/ids 24
    // + sorted tile rectangles 12 + dense key 4 + rectangles 16 + padded tile
    // rectangle 4 + GPS prepared 48 + item count/base 8, plus the GPS work list.
    size_t required = (size_t)count * (240 + 48 + 24 + 12 + 4 + 16 + 4 + 48 + 8) + (size_t)c->gps_budget * 4;
    size_t held_other = c->held - c->planes.bytes - c->records.bytes - c->keys[0].bytes - c->keys[1].bytes -
                        c->keys[2].bytes - c->keys[3].bytes - c->ids[0].bytes - c->ids[1].bytes -
                        c->ids[2].bytes - c->ids[3].bytes - c->vrects[0].bytes - c->vrects[1].bytes -
                        c->vrects[2].bytes - c->vrects[3].bytes - c->packed_rects.bytes - c->block_range.bytes -
                        c->dense_keys.bytes - c->rects.bytes - c->rect_px.bytes - c->prepared.bytes - c->item_count.bytes -
                        c->item_base.bytes - c->work.bytes;
    if (c->budget && held_other + required > c->budget)
        return fail(GS_ERR_LIMIT, "scene of %u Gaussians needs %.1f MiB; memory budget is %.1f MiB", count,
                    required / 1048576.0, c->budget / 1048576.0);
    const size_t bytes = (size_t)count * kPlanes * 16;
    if (c->budget && c->held + bytes > c->budget) {
        // the previous scene cannot stay resident next to the new one
        release(c, c->planes);
        c->n = 0;
        c->have_cull = c->have_sort = c->have_composite = c->have_gps = false;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(GS_ERR_OOM, "hipMalloc(%zu) for Gaussian planes failed: %s", bytes, hipGetErrorString(e));
    }
    planes->p = p;
    planes->bytes = bytes;
    c->held += bytes;
    return GS_OK;
}

// The box of the uploaded centres (plane P0), reduced on the device and
// finished on the host.
int measure_centre_box(gs_ctx* c) {
    c->box_valid = false;
    if (!c->n) return GS_OK;
    GS_CHECK(ensure(c, c->box_partials, (size_t)kCentreBoundBlocks * 8 * 4, "centre box"));
    GS_HIP(launch_centre_bounds(ptr<float4>(c->planes), c->n, ptr<float>(c->box_partials), c->stream));
    std::vector<float> part((size_t)kCentreBoundBlocks * 8);
    GS_HIP(hipMemcpyAsync(part.data(), c->box_partials.p, part.size() * 4, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool finite = true;
    for (uint32_t b = 0; b < kCentreBoundBlocks; ++b) {
        const float* q = part.data() + (size_t)b * 8;
        if (q[6] != 0.0f) finite = false;
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], (double)q[k]);
            hi[k] = std::max(hi[k], (double)q[3 + k]);
        }
    }
    for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(lo[k]) && std::isfinite(hi[k]);
    if (finite) {
        for (int k = 0; k < 3; ++k) {
            c->box_lo[k] = lo[k];
            c->box_hi[k] = hi[k];
        }
        c->box_valid = true;
    }
    return GS_OK;
}

// Significant bits of the depth keys' range over the current views: the
// keys are 0xFFFFFFFF - bits(depth) for depths in [0.2, d], d the deepest
// corner of the centre box in any view (with a margin for the cull's float
// evaluation), so the range is below bits(d) - bits(0.2).  32 when the box
// is unknown or the effect moves the centres.
int depth_key_bits(const gs_ctx* c) {
    if (!c->box_valid || c->plan.cull.effect_active || c->host_cams.empty()) return 32;
    double deepest = 0.0;
    for (const GsCameraUniforms& cam : c->host_cams) {
        const float* V = cam.view;
        double z = V[14], mag = std::fabs((double)V[14]);
        for (int k = 0; k < 3; ++k) {
            const double m = V[2 + 4 * k];
            z += std::min(m * c->box_lo[k], m * c->box_hi[k]);
            mag += std::fabs(m) * std::max(std::fabs(c->box_lo[k]), std::fabs(c->box_hi[k]));
        }
        // depth = -z; the cull's float sum lies within a few ulps of mag
        deepest = std::max(deepest, -z + 1e-5 * mag + 1e-4);
    }
    if (!(deepest < 1e30)) return 32;
    if (deepest < 0.2) return 1;
    const float df = (float)deepest, nf = 0.2f;
    uint32_t bd, bn;
    memcpy(&bd, &df, 4);
    memcpy(&bn, &nf, 4);
    const uint32_t range = bd + 1u - bn;  // float rounding of deepest may round down by one
    int bits = 1;
    while (bits < 32 && (range >> bits)) ++bits;
    return bits;
}

// The new planes become the scene: per-view buffers are sized for it, and
// every frame state and the animation of the previous scene are dropped.
int install_planes(gs_ctx* c, DevBuf& planes, uint32_t count) {
    release(c, c->planes);
    c->planes = planes;
    planes = DevBuf{};
    c->n = count;
    c->view_cap = 0;
    GS_CHECK(ensure_views(c, 1));
    GS_CHECK(measure_centre_box(c));
    c->have_cull = c->have_sort = c->have_composite = c->have_gps = false;
    release(c, c->anim_data);
    release(c, c->lut);
    c->anim_kind = 0;
    c->anim_sets = 0;
    c->lut_size = 0;
    c->anim_state_a = c->anim_state_b = 0;
    c->anim_weight = 0.0f;
    c->anim_layer_count = 0;
    return GS_OK;
}

// Two page-locked chunk buffers in the plane layout (kUploadChunk entries,
// plane stride = the chunk's entry count), kept for later uploads.
int ensure_upload_staging(gs_ctx* c) {
    if (c->up_pin[0]) return GS_OK;
    for (int b = 0; b < 2; ++b) {
        GS_HIP(hipHostMalloc((void**)&c->up_pin[b], (size_t)kUploadChunk * kPlanes * 16, hipHostMallocDefault));
        GS_HIP(hipEventCreateWithFlags(&c->up_ev[b], hipEventDisableTiming));
        c->up_used[b] = false;
    }
    return GS_OK;
}

// Pipelined upload: the host fills chunk k of the planes into one pinned
// buffer (produce(first, n, planes) with plane stride n, on the host's cores)
// while the copies of chunk k - 1 run from the other; each chunk is 15
// asynchronous copies, one per plane, straight into the device planes.
template <typename Produce>
int upload_pipelined(gs_ctx* c, uint32_t count, Produce produce) {
    GS_CHECK(set_device(c));
    DevBuf planes;
    GS_CHECK(reserve_planes(c, count, &planes));
    int r = ensure_upload_staging(c);
    for (uint32_t first = 0, k = 0; first < count && r == GS_OK; first += kUploadChunk, ++k) {
        const uint32_t n = std::min(kUploadChunk, count - first);
        const int b = (int)(k & 1u);
        if (c->up_used[b] && hipEventSynchronize(c->up_ev[b]) != hipSuccess) {
            r = fail(GS_ERR_HIP, "upload staging wait failed");
            break;
        }
        r = produce(first, n, c->up_pin[b]);
        if (r != GS_OK) {
            r = fail(r, "scene decode failed at Gaussian %u", first);
            break;
        }
        hipError_t e = hipSuccess;
        for (int p = 0; p < kPlanes && e == hipSuccess; ++p)
            e = hipMemcpyAsync(static_cast<float4*>(planes.p) + (size_t)p * count + first, c->up_pin[b] + (size_t)p * n * 4,
                               (size_t)n * 16, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->up_ev[b], c->stream);
        if (e != hipSuccess) r = fail(GS_ERR_HIP, "upload copy failed: %s", hipGetErrorString(e));
        c->up_used[b] = true;
    }
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (r == GS_OK && e != hipSuccess) r = fail(GS_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
    if (r != GS_OK) {
        release(c, planes);
        return r;
    }
    GS_CHECK(install_planes(c, planes, count));
    return ok();
}

}  // namespace

extern "C" {

int gs_upload_gaussians(gs_ctx* c, const GsGaussian* g, uint32_t count) {
    if (!c || (!g && count)) return fail(GS_ERR_INVALID, "bad arguments");
    return upload_pipelined(c, count, [g](uint32_t first, uint32_t n, float* buf) {
        return fill_planes_gaussians(g + first, n, n, buf);
    });
}

int gs_upload_ply(gs_ctx* c, const float* rows, uint32_t count, const GsPlyLayout* layout) {
    if (!c || !layout || (!rows && count)) return fail(GS_ERR_INVALID, "bad arguments");
    return upload_pipelined(c, count, [rows, layout](uint32_t first, uint32_t n, float* buf) {
        return fill_planes_ply(rows + (size_t)first * layout->stride, n, layout, n, buf);
    });
}

int gs_upload_sog(gs_ctx* c, const GsSogData* sog) {
    if (!c || !sog) return fail(GS_ERR_INVALID, "bad arguments");
    return upload_pipelined(c, sog->count, [sog](uint32_t first, uint32_t n, float* buf) {
        return fill_planes_sog(sog, first, n, n, buf);
    });
}

int gs_upload_planes(gs_ctx* c, const float* host_planes, uint32_t count) {
    if (!c || (!host_planes && count)) return fail(GS_ERR_INVALID, "bad arguments");
    hipPointerAttribute_t attr{};
    const bool pinned = hipPointerGetAttributes(&attr, host_planes) == hipSuccess && attr.type == hipMemoryTypeHost;
    (void)hipGetLastError();
    if (!pinned) {
        // pageable memory: through the pinned chunk buffers
        return upload_pipelined(c, count, [host_planes, count](uint32_t first, uint32_t n, float* buf) {
            return fill_planes_copy(host_planes, count, first, n, n, buf);
        });
    }
    // page-locked: one asynchronous copy of the whole plane array
    GS_CHECK(set_device(c));
    DevBuf planes;
    GS_CHECK(reserve_planes(c, count, &planes));
    hipError_t e = hipMemcpyAsync(planes.p, host_planes, planes.bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        release(c, planes);
        return fail(GS_ERR_HIP, "upload copy failed: %s", hipGetErrorString(e));
    }
    GS_CHECK(install_planes(c, planes, count));
    return ok();
}

int gs_alloc_pinned(size_t bytes, void** out) {
    if (!out || !bytes) return fail(GS_ERR_INVALID, "bad arguments");
    *out = nullptr;
    const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(GS_ERR_OOM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return ok();
}

void gs_free_pinned(void* p) {
    if (p) (void)hipHostFree(p);
}

int gs_upload_animation(gs_ctx* c, const GsAnimation* an) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    GS_CHECK(set_device(c));
    if (!an) {
        release(c, c->anim_data);
        release(c, c->lut);
        c->anim_kind = 0;
        c->anim_sets = 0;
        c->lut_size = 0;
        c->anim_state_a = c->anim_state_b = 0;
        c->anim_weight = 0.0f;
        c->anim_layer_count = 0;
        return ok();
    }
    if (!c->n) return fail(GS_ERR_STATE, "upload the Gaussians before their animation");
    if (an->rows != c->n) return fail(GS_ERR_INVALID, "animation has %u rows, the scene %u Gaussians", an->rows, c->n);
    const bool deltas = an->state_count > 0, layers = an->layer_count > 0;
    if (deltas == layers) return fail(GS_ERR_INVALID, "an animation holds either delta states or layers");
    const uint32_t sets = deltas ? an->state_count : an->layer_count;
    if (sets > kMaxAnimLayers) return fail(GS_ERR_LIMIT, "%u animation states or layers exceed %u", sets, kMaxAnimLayers);
    if (deltas ? !an->deltas : !an->layer_data) return fail(GS_ERR_INVALID, "animation data is NULL");
    for (int j = 0; j < 48; ++j)
        if (!std::isfinite(an->scales[j])) return fail(GS_ERR_INVALID, "delta scale %d is not finite", j);
    if (an->lut_size == 1 || an->lut_size > 256 || (an->lut_size && !an->display_lut))
        return fail(GS_ERR_INVALID, "display LUT size %u outside [2, 256]", an->lut_size);
    if (an->lut_size && !(std::isfinite(an->log2_min) && std::isfinite(an->log2_max) && an->log2_max > an->log2_min))
        return fail(GS_ERR_INVALID, "display LUT range [%g, %g] is empty", (double)an->log2_min, (double)an->log2_max);
    const size_t bytes = (size_t)sets * c->n * 48 * (deltas ? 1 : 2);
    const size_t lut_bytes = (size_t)an->lut_size * an->lut_size * an->lut_size * 3 * 2;
    release(c, c->anim_data);
    release(c, c->lut);
    c->anim_kind = 0;
    GS_CHECK(ensure(c, c->anim_data, bytes, "animation coefficients"));
    GS_HIP(hipMemcpy(c->anim_data.p, deltas ? (const void*)an->deltas : (const void*)an->layer_data, bytes,
                     hipMemcpyHostToDevice));
    if (an->lut_size) {
        GS_CHECK(ensure(c, c->lut, lut_bytes, "display LUT"));
        GS_HIP(hipMemcpy(c->lut.p, an->display_lut, lut_bytes, hipMemcpyHostToDevice));
    }
    c->anim_kind = deltas ? 1 : 2;
    c->anim_sets = sets;
    memcpy(c->anim_scales, an->scales, sizeof(c->anim_scales));
    c->lut_size = an->lut_size;
    c->lut_log2_min = an->log2_min;
    c->lut_log2_max = an->log2_max;
    c->anim_state_a = c->anim_state_b = 0;
    c->anim_weight = 0.0f;
    c->anim_layer_count = 0;
    c->have_cull = c->have_sort = c->have_composite = false;
    return ok();
}

int gs_set_animation_frame(gs_ctx* c, int32_t state_a, int32_t state_b, float weight, uint32_t layer_count,
                           const float* layer_weights) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!std::isfinite(weight)) return fail(GS_ERR_INVALID, "animation weight is not finite");
    const int32_t states = c->anim_kind == 1 ? (int32_t)c->anim_sets : 0;
    if (state_a < 0 || state_a > states || state_b < 0 || state_b > states)
        return fail(GS_ERR_INVALID, "animation states %d, %d outside [0, %d]", state_a, state_b, states);
    const uint32_t layers = c->anim_kind == 2 ? c->anim_sets : 0;
    if (layer_count > layers) return fail(GS_ERR_INVALID, "%u layers requested, %u uploaded", layer_count, layers);
    if (layer_count && !layer_weights) return fail(GS_ERR_INVALID, "layer weights are NULL");
    for (uint32_t k = 0; k < layer_count * 3; ++k)
        if (!std::isfinite(layer_weights[k])) return fail(GS_ERR_INVALID, "layer weight %u is not finite", k);
    c->anim_state_a = state_a;
    c->anim_state_b = state_b;
    c->anim_weight = weight;
    c->anim_layer_count = layer_count;
    c->anim_layer_weights.assign(layer_weights, layer_weights + layer_count * 3);
    return ok();
}

int gs_set_gps_budget(gs_ctx* c, uint32_t max_work_items) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (max_work_items == 0 || max_work_items > kMaxWorkItems)
        return fail(GS_ERR_LIMIT, "GPS work budget %u outside [1, %u]; keeping %u", max_work_items, kMaxWorkItems,
                    c->gps_budget);
    if (max_work_items == c->gps_budget) return ok();
    size_t bytes = (size_t)max_work_items * 4;
    if (c->budget && c->held - c->work.bytes + bytes > c->budget)
        return fail(GS_ERR_LIMIT, "GPS work buffer of %.1f MiB exceeds the memory budget; keeping %u work items",
                    bytes / 1048576.0, c->gps_budget);
    if (c->work.p) {
        // keep the old buffer until the new one exists (src/renderer.cpp:1083-1117)
        GS_CHECK(set_device(c));
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(GS_ERR_OOM, "GPS work buffer allocation failed; keeping %u work items", c->gps_budget);
        }
        release(c, c->work);
        c->work.p = p;
        c->work.bytes = bytes;
        c->held += bytes;
    }
    c->gps_budget = max_work_items;
    return ok();
}

}  // extern "C"

namespace {

// (tile id, splat id) buffers are sized by capacity so a frame needs no host
// round trip; they grow once a frame's requested count is known to exceed it.
int grow_tile_buffers(gs_ctx* c, uint32_t entries) {
    if (entries > 0x7FFFFFFFu) return fail(GS_ERR_LIMIT, "%u tile entries exceed the 2^31 limit", entries);
    if (entries <= c->dup_cap) return GS_OK;
    c->bin_epoch++;
    for (int k = 0; k < 3; ++k) {
        GS_CHECK(ensure(c, c->dup_keys[k], (size_t)entries * 4, "tile keys"));
        GS_CHECK(ensure(c, c->dup_vals[k], (size_t)entries * 4, "tile values"));
    }
    GS_CHECK(ensure_table(c, (uint32_t)std::max<size_t>(entries, (size_t)c->view_cap * c->n)));
    c->dup_cap = entries;
    return GS_OK;
}

int check_cull_options(const GsCullOptions* opts, uint32_t views) {
    int sh_degree = opts ? opts->sh_degree : 3;
    if (sh_degree < 0 || sh_degree > 3) return fail(GS_ERR_INVALID, "sh_degree %d outside [0, 3]", sh_degree);
    if (opts && opts->collect_stats && views > 1)
        return fail(GS_ERR_INVALID, "collect_stats needs a single view");
    return GS_OK;
}

CullLaunch make_cull(gs_ctx* c, const GsSplatEffectParams* effect, const GsCullOptions* opts) {
    CullLaunch a{};
    a.planes = ptr<float4>(c->planes);
    a.n = c->n;
    a.cams = frame_dev(c)->cams;
    a.views = c->views;
    if (effect) a.effect = *effect;
    a.effect_active = effect && effect->params[3] > 0.5f;
    a.sh_degree = opts ? opts->sh_degree : 3;
    a.sh_lod_distance = opts ? opts->sh_lod_distance : 0.0f;
    a.collect_stats = opts ? opts->collect_stats : 0;
    a.records = ptr<float4>(c->records);
    a.dense_keys = ptr<uint32_t>(c->dense_keys);
    a.block_counts = ptr<uint32_t>(c->cull_counts);
    a.block_range = ptr<uint2>(c->block_range);
    a.block_diag = a.collect_stats ? ptr<uint32_t>(c->block_diag) : nullptr;
    a.rect_px = ptr<uint2>(c->rect_px);
    a.tight_rects = c->tight_rects;
    // GSAnim: the delta branch only when a state differs from the base, the
    // layer branch when layers are summed
    if (c->anim_kind == 1 && (c->anim_state_a != 0 || (c->anim_weight > 0.0f && c->anim_state_b != 0))) {
        a.anim = 1;
        a.anim_state_a = c->anim_state_a;
        a.anim_state_b = c->anim_state_b;
        a.anim_weight = c->anim_weight;
        memcpy(a.anim_scales, c->anim_scales, sizeof(a.anim_scales));
    } else if (c->anim_kind == 2 && c->anim_layer_count) {
        a.anim = 2;
        a.anim_layers = (int)c->anim_layer_count;
        a.layer_weights = frame_dev(c)->layer_weights;
    }
    if (a.anim) a.anim_data = ptr<uint4>(c->anim_data);
    return a;
}

// The cull's rectangle outputs for a plan: with a packing tile size the tile
// rectangles packed at that size (grown by a pixel with pack_pad), which the
// depth sort and the compaction carry, and no pixel rectangles (a bin at
// another tile size then recomputes them from the records); else the pixel
// rectangles, which the bin gathers.
void set_cull_rects(gs_ctx* c, FramePlan* p) {
    if (p->pack_tile) {
        p->cull.rect_px = nullptr;
        p->cull.packed_rects = ptr<uint32_t>(c->packed_rects);
        p->cull.packed_shift = 0;
        while ((1 << p->cull.packed_shift) < p->pack_tile) ++p->cull.packed_shift;
        p->cull.packed_pad = p->pack_pad;
    } else {
        p->cull.rect_px = ptr<uint2>(c->rect_px);
        p->cull.packed_rects = nullptr;
    }
    c->rect_px_valid = p->cull.rect_px != nullptr;
}

// Tile size the compact packs each visible splat's tile rectangle for (the
// depth sort carries it to the binning, which then reads it in order instead
// of gathering the pixel rectangle): `tile` when every view spans at most 255
// tiles a side, else 0.
int rect_pack_tile(const GsCameraUniforms* cams, uint32_t count, int tile) {
    for (uint32_t v = 0; v < count; ++v) {
        const long w = (long)cams[v].viewport[0], h = (long)cams[v].viewport[1];
        if ((w + tile - 1) / tile > 255 || (h + tile - 1) / tile > 255) return 0;
    }
    return tile;
}

SortLaunch make_depth_sort(gs_ctx* c) {
    SortLaunch s{};
    for (int k = 0; k < 3; ++k) {
        s.keys[k] = ptr<uint32_t>(c->keys[k]);
        s.vals[k] = ptr<uint32_t>(c->ids[k]);
        s.vals2[k] = c->rect_tile ? ptr<uint32_t>(c->vrects[k]) : nullptr;
    }
    // the keys' range (reduced by the cull's scan) sets the passes that run;
    // the result is in slot 3
    s.keys[3] = ptr<uint32_t>(c->keys[3]);
    s.vals[3] = ptr<uint32_t>(c->ids[3]);
    s.vals2[3] = c->rect_tile ? ptr<uint32_t>(c->vrects[3]) : nullptr;
    s.key_range = &c->dc()->cull.key_min;
    s.count = &c->dc()->cull.visible_count;
    s.capacity = c->views * c->n;
    s.scratch = c->dc()->depth_sort.scratch;
    s.table = ptr<uint32_t>(c->table);
    // 32-bit depth keys: 4-bit digits take 8 passes; with the 8-bit setting
    // the keys go in 9-bit passes, at most four, as many as the key range
    // needs (three for a range under 2^27: depths within a factor of 2^16)
    s.bits = c->radix_bits == 8 ? 9 : c->radix_bits;
    // launched: what the key range the scene's box allows can need (the
    // adaptive passes may still stop earlier)
    s.passes = (depth_key_bits(c) + s.bits - 1) / s.bits;
    if (s.passes < 1) s.passes = 1;
    s.wide_index = c->wide_index;
    // nothing on the device reads the sorted keys: the readback gathers them
    // from the cull's dense keys by the sorted ids
    s.skip_last_keys = 1;
    // without a visible list the first pass reads the cull's dense keys and
    // drops the culled slots (the list it keeps is the one the compaction
    // would build, in the same order)
    if (c->compacted != 2) {
        const CullLaunch& cl = c->plan.cull;
        s.dense_keys = cl.dense_keys;
        s.dense_count = c->views * c->n;
        // the cull packed the tile rectangles itself whenever they are carried
        if (c->rect_tile) s.dense_rect_packed = cl.packed_rects;
    }
    return s;
}

// Bin, tile-sort and composite launches for the current views and tile size.
void make_composite(gs_ctx* c, int tile, FramePlan* p) {
    const uint32_t ntiles = c->total_tiles;
    int bits_needed = 1;
    while (bits_needed < 31 && ((ntiles - 1u) >> bits_needed)) ++bits_needed;
    // digits of up to 9 bits with the 8-bit setting (as the depth sort), 4
    // with the 4-bit one
    const int max_bits = c->radix_bits == 8 ? 9 : c->radix_bits;
    const int passes = (bits_needed + max_bits - 1) / max_bits;
    // the pass count of that width, with the digits balanced over the passes
    // (13 tile bits: two 7-bit passes, not 9 + 4; 18 bits of 64 views: two
    // 9-bit passes): fewer digits per pass mean fewer ballots and longer runs
    // of equal digits to write
    int tile_bits = (bits_needed + passes - 1) / passes;
    tile_bits = tile_bits <= 4 ? 4 : (tile_bits <= 6 ? 6 : (tile_bits <= 7 ? 7 : (tile_bits <= 8 ? 8 : 9)));
    if (tile_bits > max_bits) tile_bits = max_bits;

    BinLaunch& b = p->bin;
    b = BinLaunch{};
    b.sorted_ids = ptr<uint32_t>(c->ids[c->sort_slot]);
    b.count = &c->dc()->cull.visible_count;
    b.capacity = c->views * c->n;
    b.n = c->n;
    b.tile = tile;
    b.views = frame_dev(c)->views;
    b.single_tiles_x = c->views == 1 ? ((uint32_t)c->cam.viewport[0] + (uint32_t)tile - 1u) / (uint32_t)tile : 0u;
    b.counters = c->dc();
    b.rect_px = ptr<uint2>(c->rect_px);
    b.sorted_rects = c->rect_tile == tile && !c->rect_pad ? ptr<uint32_t>(c->vrects[c->sort_slot]) : nullptr;
    if (!b.sorted_rects && !c->rect_px_valid) {
        // the cull packed its rectangles for another tile size and wrote no
        // pixel rectangles: they come from the records (unmoved, unpadded)
        b.jittered = 1;
        b.records = ptr<float4>(c->records);
        b.tight = c->cull_tight;
        b.clip_z_01 = c->cam.clip_z_01;
    }
    b.rects = ptr<uint2>(c->rects);
    b.block_counts = ptr<uint32_t>(c->block_counts);
    b.block_offsets = ptr<uint32_t>(c->block_offsets);
    b.clear = ptr<uint32_t>(c->ranges);
    b.clear_words = ntiles * 2u;
    b.dup_keys = ptr<uint32_t>(c->dup_keys[0]);
    b.dup_vals = ptr<uint32_t>(c->dup_vals[0]);
    b.dup_capacity = c->dup_cap;
    b.wide_index = c->wide_index;
    // tile ids below 2^16 travel as 16-bit keys (a third less traffic in
    // the emit, the tile sort and the ranges)
    b.key16 = ntiles <= 65536u && c->tile_key16 ? 1 : 0;

    SortLaunch& s = p->tiles;
    s = SortLaunch{};
    for (int k = 0; k < 3; ++k) {
        s.keys[k] = ptr<uint32_t>(c->dup_keys[k]);
        s.vals[k] = ptr<uint32_t>(c->dup_vals[k]);
    }
    s.count = &c->dc()->bin.dup_count;
    s.capacity = c->dup_cap;
    s.scratch = c->dc()->tile_sort.scratch;
    s.table = ptr<uint32_t>(c->table);
    s.passes = passes;
    s.bits = tile_bits;
    s.wide_index = c->wide_index;
    s.key16 = b.key16;

    // the sorted pairs land in buffer 2 for an even pass count, else 1
    const int slot = passes % 2 == 0 ? 2 : 1;
    CompositeLaunch& cl = p->comp;
    cl = CompositeLaunch{};
    cl.records = ptr<float4>(c->records);
    cl.tile_vals = s.vals[slot];
    cl.ranges = ptr<uint2>(c->ranges);
    cl.tile = tile;
    cl.views = frame_dev(c)->views;
    cl.nviews = c->views;
    cl.total_tiles = ntiles;
    cl.image = ptr<uint16_t>(c->image);
    cl.quadrant_cull = c->quadrant_cull;
    cl.tight_rects = c->tight_rects;
    cl.reference_blend = tile == 16 ? c->reference_blend : 0;
    cl.persistent_blocks = c->blocks_per_cu > 0 ? (uint32_t)(c->blocks_per_cu * c->props.multiProcessorCount) : 0u;
    cl.queue = c->dc()->composite_queue;
}

int stage_cull(gs_ctx* c, const FramePlan& p) {
    c->bin_epoch++;
    TimedLaunch t{}, tk{};
    GS_CHECK(timed_begin(c, GS_STAGE_CULL, &t));
    GS_CHECK(timed_begin(c, GS_STAGE_CULL_KERNEL, &tk));
    GS_HIP(launch_cull(p.cull, c->stream));
    GS_CHECK(timed_end(c, &tk));
    // the visible count, key range and diagnostics; the block offsets only
    // when the visible list is compacted
    GS_HIP(launch_reduce_counts(p.cull.block_counts, p.cull_blocks * p.cull.views, &c->dc()->cull.visible_count,
                                p.cull.block_diag, c->dc()->cull.diag, c->stream, p.cull.block_range,
                                &c->dc()->cull.key_min));
    GS_CHECK(timed_end(c, &t));
    c->compacted = 0;
    return GS_OK;
}

// The last cull's visible list in index order (keys[0], ids[0], vrects[0]),
// for the consumers that read it unsorted; `ids_only` (GPS) writes ids[0]
// alone, a third of the bytes.
int ensure_compacted(gs_ctx* c, bool ids_only = false) {
    if (c->compacted == 2 || (ids_only && c->compacted == 1)) return GS_OK;
    const FramePlan& p = c->plan;
    uint32_t* visible = &c->dc()->cull.visible_count;
    GS_HIP(launch_scan_counts(p.cull.block_counts, p.cull_blocks * p.cull.views, ptr<uint32_t>(c->cull_offsets),
                              visible, nullptr, nullptr, c->stream));
    if (ids_only)
        GS_HIP(launch_compact(p.cull.dense_keys, p.cull.n, p.cull.views, ptr<uint32_t>(c->cull_offsets), nullptr,
                              ptr<uint32_t>(c->ids[0]), nullptr, 0, nullptr, c->stream, nullptr));
    else
        GS_HIP(launch_compact(p.cull.dense_keys, p.cull.n, p.cull.views, ptr<uint32_t>(c->cull_offsets),
                              ptr<uint32_t>(c->keys[0]), ptr<uint32_t>(c->ids[0]), p.cull.rect_px, p.pack_tile,
                              ptr<uint32_t>(c->vrects[0]), c->stream, p.pack_tile ? p.cull.packed_rects : nullptr));
    c->compacted = ids_only ? 1 : 2;
    return GS_OK;
}

int stage_sort(gs_ctx* c, const FramePlan& p, int* slot) {
    c->bin_epoch++;
    TimedLaunch t{};
    GS_CHECK(timed_begin(c, GS_STAGE_SORT, &t));
    GS_HIP(launch_radix_sort(p.depth, slot, c->stream));
    GS_CHECK(timed_end(c, &t));
    return GS_OK;
}

// (tile, splat) pairs, their stable sort by tile and the tile ranges.
int stage_bin(gs_ctx* c, const FramePlan& p) {
    c->bin_epoch++;
    // the count pass of the bin clears the tile ranges
    if (p.comp.persistent_blocks) GS_HIP(launch_fill_u32(p.comp.queue, 0u, 8, c->stream));
    TimedLaunch t{};
    GS_CHECK(timed_begin(c, GS_STAGE_BIN, &t));
    GS_HIP(launch_bin(p.bin, c->stream));
    int slot = 0;
    GS_HIP(launch_radix_sort(p.tiles, &slot, c->stream));
    GS_HIP(launch_tile_ranges(p.tiles.keys[slot], p.tiles.key16, p.tiles.count, p.tiles.capacity,
                              const_cast<uint2*>(p.comp.ranges), c->stream));
    GS_CHECK(timed_end(c, &t));
    return GS_OK;
}

int stage_composite(gs_ctx* c, const FramePlan& p) {
    GS_CHECK(stage_bin(c, p));
    TimedLaunch t2{};
    GS_CHECK(timed_begin(c, GS_STAGE_COMPOSITE, &t2));
    GS_HIP(launch_composite(p.comp, c->stream));
    GS_CHECK(timed_end(c, &t2));
    return GS_OK;
}

// Buffers a composite at `tile` over the current views needs.
int prepare_composite(gs_ctx* c, int tile) {
    GS_CHECK(ensure(c, c->image, (size_t)c->total_pixels * 8, "RGBA16F images"));
    GS_CHECK(ensure(c, c->ranges, (size_t)c->total_tiles * 8, "tile ranges"));
    if (!c->dup_cap) {
        // three pairs per splat for the first view and one and a half for
        // every other one to start with (a 6M-splat 1080p frame takes 1.9
        // per splat, a batch of 64 mixed views of 3M splats 1.4 per splat
        // and view), formed in 64 bits and capped at the pair limit; an
        // overflow grows them
        const size_t want = std::min<size_t>(
            std::max<size_t>((size_t)3 * c->n + (size_t)3 * (c->views - 1) * c->n / 2, 65536), 0x7FFFFFFF);
        GS_CHECK(grow_tile_buffers(c, (uint32_t)want));
    }
    (void)tile;
    return GS_OK;
}

// The requested pair count comes back asynchronously: before any composite
// result is observed (image readback, stats, synchronize) the count is
// checked, and a frame that overflowed is re-binned with grown buffers.
int after_composite(gs_ctx* c, int tile) {
    GS_HIP(hipMemcpyAsync(c->dup_host, &c->dc()->bin.dup_total, 4, hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipEventRecord(c->dup_event, c->stream));
    c->dup_pending = true;
    c->composite_tile = tile;
    c->img_w = (int)c->cam.viewport[0];
    c->img_h = (int)c->cam.viewport[1];
    return GS_OK;
}

// StochasticSplats sample into the sample / depth images, binning first
// unless the tile lists of this cull can be reused (*binned tells which).
int enqueue_stochastic(gs_ctx* c, bool* binned) {
    const int W = (int)c->cam.viewport[0], H = (int)c->cam.viewport[1];
    const float jx = c->stoch_jitter[0], jy = c->stoch_jitter[1];
    const int mode = (jx == 0.0f && jy == 0.0f) ? 0 : (std::fabs(jx) <= 0.5f && std::fabs(jy) <= 0.5f ? 1 : 2);
    const bool lists_ok = c->stoch_bins_epoch == c->bin_epoch && !c->dup_pending && c->last_dup <= c->dup_cap;
    const bool reuse = lists_ok && (c->stoch_bins_mode == mode || (c->stoch_bins_mode == 1 && mode == 0) ||
                                    (mode == 2 && c->stoch_bins_jitter[0] == jx && c->stoch_bins_jitter[1] == jy));
    *binned = !reuse;
    if (!reuse) {
        GS_CHECK(upload_views(c, c->host_cams.data(), 16));
        GS_CHECK(prepare_composite(c, 16));
        make_composite(c, 16, &c->plan);
        // rectangles grown by a pixel came with the depth sort: their lists
        // serve samples with and without jitter up to half a pixel
        const bool padded = c->rect_pad && c->rect_tile == 16 && mode <= 1;
        if (padded) {
            c->plan.bin.sorted_rects = ptr<uint32_t>(c->vrects[c->sort_slot]);
            c->plan.bin.jittered = 0;
        } else if (mode) {
            // the tile rectangles come from the records: grown by a pixel,
            // which holds every jitter up to half a pixel, or moved exactly
            BinLaunch& b = c->plan.bin;
            b.jittered = 1;
            b.sorted_rects = nullptr;
            b.records = ptr<float4>(c->records);
            b.tight = c->cull_tight;
            b.pad = mode == 1 ? 1 : 0;
            b.jitter_x = mode == 2 ? jx : 0.0f;
            b.jitter_y = mode == 2 ? jy : 0.0f;
            b.clip_z_01 = c->cam.clip_z_01;
        }
        GS_CHECK(stage_bin(c, c->plan));
        c->stoch_bins_epoch = c->bin_epoch;
        c->stoch_bins_mode = padded ? 1 : mode;
        c->stoch_bins_jitter[0] = jx;
        c->stoch_bins_jitter[1] = jy;
    }
    GS_CHECK(ensure(c, c->sample, (size_t)W * H * 16, "sample image"));
    GS_CHECK(ensure(c, c->depth, (size_t)W * H * 4, "sample depth"));
    StochasticLaunch sl{};
    sl.tiles = c->plan.comp;
    sl.frame_seed = (float)c->stoch_seed;
    sl.background = c->stoch_bg;
    sl.clip_z_01 = c->cam.clip_z_01;
    sl.jitter_x = jx;
    sl.jitter_y = jy;
    sl.sample = ptr<float4>(c->sample);
    sl.depth = ptr<float>(c->depth);
    TimedLaunch t{};
    GS_CHECK(timed_begin(c, GS_STAGE_COMPOSITE, &t));
    GS_HIP(launch_stochastic(sl, c->stream));
    GS_CHECK(timed_end(c, &t));
    c->res_w = W;
    c->res_h = H;
    return GS_OK;
}

int enqueue_composite(gs_ctx* c, int tile) {
    if (c->composite_kind == 1) {
        bool binned = false;
        GS_CHECK(enqueue_stochastic(c, &binned));
        return binned ? after_composite(c, 16) : GS_OK;
    }
    GS_CHECK(upload_views(c, c->host_cams.data(), tile));
    GS_CHECK(prepare_composite(c, tile));
    make_composite(c, tile, &c->plan);
    GS_CHECK(stage_composite(c, c->plan));
    return after_composite(c, tile);
}

// Waits for the pending pair count (or only polls when wait is false); grows
// the buffers on overflow and, when the current composite overflowed and
// rerun is set, composites again until every pair fits.
int settle_composite(gs_ctx* c, bool wait, bool rerun) {
    while (c->dup_pending) {
        if (wait) {
            GS_HIP(hipEventSynchronize(c->dup_event));
        } else {
            hipError_t q = hipEventQuery(c->dup_event);
            if (q == hipErrorNotReady) return GS_OK;
            GS_HIP(q);
        }
        c->dup_pending = false;
        const uint32_t total = *c->dup_host;
        c->last_dup = total;
        if (total <= c->dup_cap) return GS_OK;
        GS_CHECK(grow_tile_buffers(c, total + total / 4u));
        if (!(rerun && (c->have_composite || c->composite_kind == 1))) return GS_OK;
        c->overflow_reruns++;
        GS_CHECK(enqueue_composite(c, c->composite_tile));
    }
    return GS_OK;
}

// Replays the captured frame when the plan is byte-identical to the one it
// was captured from, otherwise captures the plan's launches first.
int launch_graph(gs_ctx* c, const FramePlan& p) {
    c->bin_epoch++;
    const uint8_t* b = reinterpret_cast<const uint8_t*>(&p);
    if (!c->graph_exec || c->graph_key.size() != sizeof(FramePlan) ||
        memcmp(c->graph_key.data(), b, sizeof(FramePlan)) != 0) {
        // the last replay of the old graph may still run on the stream: the
        // executable graph is destroyed only once it has finished
        if (c->graph_exec) {
            GS_HIP(hipStreamSynchronize(c->stream));
            GS_HIP(hipGraphExecDestroy(c->graph_exec));
        }
        if (c->graph) GS_HIP(hipGraphDestroy(c->graph));
        c->graph_exec = nullptr;
        c->graph = nullptr;
        GS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int slot = 0;
        int r = stage_cull(c, p);
        if (r == GS_OK) r = stage_sort(c, p, &slot);
        if (r == GS_OK) r = stage_composite(c, p);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (r != GS_OK) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        GS_HIP(e);
        c->graph = g;
        GS_HIP(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
        c->graph_key.assign(b, b + sizeof(FramePlan));
        c->graph_captures++;
    } else {
        c->graph_replays++;
    }
    GS_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    return GS_OK;
}

}  // namespace

extern "C" {

int gs_cull(gs_ctx* c, const GsCameraUniforms* cam, const GsSplatEffectParams* effect, const GsCullOptions* opts) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->n) return fail(GS_ERR_STATE, "no Gaussians uploaded");
    GS_CHECK(check_camera(cam));
    GS_CHECK(check_cull_options(opts, 1));
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, false, false));
    const uint32_t blocks = (c->n + 255u) / 256u;
    const int collect = opts ? opts->collect_stats : 0;
    if (collect) GS_CHECK(ensure(c, c->block_diag, (size_t)blocks * 32, "block diagnostics"));
    GS_CHECK(upload_cameras(c, cam, 1));
    c->plan.cull = make_cull(c, effect, opts);
    c->plan.cull_blocks = blocks;
    // after a StochasticSplats frame the next bin is most likely another one:
    // the depth sort then carries the rectangles grown by a pixel (16-pixel
    // tiles), which serve its jittered samples; other bins gather instead
    const int pad_tile = c->composite_kind == 1 ? rect_pack_tile(cam, 1, 16) : 0;
    c->plan.pack_pad = pad_tile ? 1 : 0;
    c->plan.pack_tile = pad_tile ? pad_tile : rect_pack_tile(cam, 1, c->default_tile);
    set_cull_rects(c, &c->plan);
    c->rect_tile = c->plan.pack_tile;
    c->rect_pad = c->plan.pack_pad;
    GS_CHECK(stage_cull(c, c->plan));
    c->have_cull = true;
    c->have_sort = c->have_composite = false;
    c->stats_collected = collect != 0;
    return ok();
}

int gs_sort(gs_ctx* c) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->have_cull) return fail(GS_ERR_STATE, "gs_sort needs a gs_cull first");
    GS_CHECK(set_device(c));
    c->plan.depth = make_depth_sort(c);
    int slot = 0;
    GS_CHECK(stage_sort(c, c->plan, &slot));
    c->sort_slot = slot;
    c->have_sort = true;
    c->have_composite = false;
    return ok();
}

int gs_composite_alpha(gs_ctx* c, int tile) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->have_sort) return fail(GS_ERR_STATE, "gs_composite_alpha needs gs_cull and gs_sort first");
    if (tile == 0) tile = c->default_tile;
    if (tile != 8 && tile != 16 && tile != 32) return fail(GS_ERR_INVALID, "tile size must be 8, 16 or 32");
    if (c->reference_blend) tile = 16;  // the reference-arithmetic compositor's tile
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, false, false));
    c->have_composite = false;
    c->composite_kind = 0;
    GS_CHECK(enqueue_composite(c, tile));
    c->have_composite = true;
    return ok();
}

int gs_stochastic_sample(gs_ctx* c, uint32_t frame_seed, float background, const float* jitter_px) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (jitter_px && !(std::isfinite(jitter_px[0]) && std::isfinite(jitter_px[1])))
        return fail(GS_ERR_INVALID, "jitter must be finite");
    if (!c->have_cull) return fail(GS_ERR_STATE, "gs_stochastic_sample needs a gs_cull first");
    if (c->views != 1) return fail(GS_ERR_STATE, "gs_stochastic_sample needs a single-view cull");
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, false, false));
    if (!c->have_sort) {
        c->plan.depth = make_depth_sort(c);
        int slot = 0;
        GS_CHECK(stage_sort(c, c->plan, &slot));
        c->sort_slot = slot;
        c->have_sort = true;
    }
    c->have_composite = false;
    c->composite_kind = 1;
    c->stoch_seed = frame_seed;
    c->stoch_bg = background;
    c->stoch_jitter[0] = jitter_px ? jitter_px[0] : 0.0f;
    c->stoch_jitter[1] = jitter_px ? jitter_px[1] : 0.0f;
    GS_CHECK(enqueue_composite(c, 16));
    c->have_gps = true;  // the sample and depth images hold this frame
    return ok();
}

int gs_render_views(gs_ctx* c, const GsCameraUniforms* cams, uint32_t count, const GsSplatEffectParams* effect,
                    const GsCullOptions* opts, int tile) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->n) return fail(GS_ERR_STATE, "no Gaussians uploaded");
    if (!cams || count == 0 || count > (uint32_t)kMaxViews)
        return fail(GS_ERR_INVALID, "view count %u outside [1, %d]", count, kMaxViews);
    for (uint32_t v = 0; v < count; ++v) GS_CHECK(check_camera(cams + v));
    GS_CHECK(check_cull_options(opts, count));
    if (tile == 0) tile = c->default_tile;
    if (tile != 8 && tile != 16 && tile != 32) return fail(GS_ERR_INVALID, "tile size must be 8, 16 or 32");
    if (c->reference_blend) tile = 16;  // the reference-arithmetic compositor's tile
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, false, false));
    const uint32_t blocks = (c->n + 255u) / 256u;
    const int collect = opts ? opts->collect_stats : 0;
    if (collect) GS_CHECK(ensure(c, c->block_diag, (size_t)blocks * 32, "block diagnostics"));
    GS_CHECK(ensure_views(c, count));
    GS_CHECK(upload_cameras(c, cams, count));
    GS_CHECK(upload_views(c, cams, tile));
    GS_CHECK(prepare_composite(c, tile));

    FramePlan& p = c->plan;
    p.cull = make_cull(c, effect, opts);
    p.cull_blocks = blocks;
    p.pack_tile = rect_pack_tile(cams, count, tile);
    p.pack_pad = 0;
    set_cull_rects(c, &p);
    c->rect_tile = p.pack_tile;
    c->rect_pad = 0;
    c->compacted = 0;  // this frame's cull builds no visible list
    p.depth = make_depth_sort(c);
    // the depth sort's last needed pass writes the sorted list to buffer 3
    c->sort_slot = 3;
    make_composite(c, tile, &p);
    if (c->use_graph && !c->profiling) {
        GS_CHECK(launch_graph(c, p));
    } else {
        int slot = 0;
        GS_CHECK(stage_cull(c, p));
        GS_CHECK(stage_sort(c, p, &slot));
        GS_CHECK(stage_composite(c, p));
    }
    c->composite_kind = 0;
    GS_CHECK(after_composite(c, tile));
    c->have_cull = c->have_sort = c->have_composite = true;
    c->stats_collected = collect != 0;
    return ok();
}

int gs_render_alpha(gs_ctx* c, const GsCameraUniforms* cam, const GsSplatEffectParams* effect,
                    const GsCullOptions* opts, int tile) {
    GS_CHECK(check_camera(cam));
    return gs_render_views(c, cam, 1, effect, opts, tile);
}

int gs_read_view_image(gs_ctx* c, uint32_t view, void* host, size_t bytes) {
    if (!c || !host) return fail(GS_ERR_INVALID, "bad arguments");
    if (!c->have_composite) return fail(GS_ERR_STATE, "no composite yet");
    if (view >= c->views) return fail(GS_ERR_INVALID, "view %u of %u", view, c->views);
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, true, true));
    const ViewInfo& vi = c->vinfo[view];
    GS_CHECK(copy_image(c, ptr<uint16_t>(c->image) + (size_t)vi.pixel_base * 4, host, vi.width, vi.height, 8, bytes,
                        c->host_cams[view].clip_y_sign > 0.0f));
    return ok();
}

int gs_autotune(gs_ctx* c, const GsCameraUniforms* cams, uint32_t count, const GsCullOptions* opts, int frames,
                GsTuneResult* out) {
    if (!c || !cams || !out) return fail(GS_ERR_INVALID, "bad arguments");
    if (frames < 1 || frames > 1000) return fail(GS_ERR_INVALID, "frames %d outside [1, 1000]", frames);
    GS_CHECK(set_device(c));
    const int save_tile = c->default_tile, save_bits = c->radix_bits, save_bpc = c->blocks_per_cu;
    const bool save_prof = c->profiling;
    c->profiling = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GS_HIP(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(GS_ERR_HIP, "hipEventCreate failed: %s", hipGetErrorString(he));
    }
    const int tiles[3] = {16, 8, 32}, bits[2] = {8, 4}, bpcs[5] = {0, 1, 2, 4, 8};
    GsTuneResult best{};
    best.best_ms = 1e30;
    uint32_t best_pairs = 0;
    int r = GS_OK, configs = 0;
    for (int ti = 0; ti < 3 && r == GS_OK; ++ti)
        for (int bi = 0; bi < 2 && r == GS_OK; ++bi)
            for (int pi = 0; pi < 5 && r == GS_OK; ++pi) {
                c->radix_bits = bits[bi];
                c->blocks_per_cu = bpcs[pi];
                // warm-up frame (also settles the tile-entry capacity)
                r = gs_render_views(c, cams, count, nullptr, opts, tiles[ti]);
                if (r == GS_OK) r = settle_composite(c, true, true);
                if (r != GS_OK) break;
                if (hipEventRecord(e0, c->stream) != hipSuccess) { r = fail(GS_ERR_HIP, "event record failed"); break; }
                for (int f = 0; f < frames && r == GS_OK; ++f) r = gs_render_views(c, cams, count, nullptr, opts, tiles[ti]);
                if (r != GS_OK) break;
                float ms = 0.0f;
                if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    r = fail(GS_ERR_HIP, "event timing failed");
                    break;
                }
                r = settle_composite(c, true, true);
                const double per = (double)ms / frames;
                ++configs;
                if (ti == 0 && bi == 0 && pi == 0) best.default_ms = per;
                if (per < best.best_ms) {
                    best.best_ms = per;
                    best.tile_size = tiles[ti];
                    best.radix_bits = bits[bi];
                    best.blocks_per_cu = bpcs[pi];
                    best_pairs = c->last_dup;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->profiling = save_prof;
    if (r != GS_OK) {
        c->default_tile = save_tile;
        c->radix_bits = save_bits;
        c->blocks_per_cu = save_bpc;
        return r;
    }
    best.configs = configs;
    c->default_tile = best.tile_size;
    c->radix_bits = best.radix_bits;
    c->blocks_per_cu = best.blocks_per_cu;
    // the tile-pair buffers grew for the configuration with the most pairs
    // (8-pixel tiles): back to the chosen one's count with a quarter to spare
    const size_t keep = std::max<size_t>((size_t)best_pairs + best_pairs / 4u, 65536);
    if (keep < c->dup_cap && !c->dup_pending) {
        GS_HIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < 3; ++k) {
            release(c, c->dup_keys[k]);
            release(c, c->dup_vals[k]);
        }
        c->dup_cap = 0;
        GS_CHECK(grow_tile_buffers(c, (uint32_t)keep));
    }
    *out = best;
    return ok();
}

int gs_frame_info(gs_ctx* c, GsFrameInfo* out) {
    if (!c || !out) return fail(GS_ERR_INVALID, "bad arguments");
    memset(out, 0, sizeof(*out));
    out->views = c->have_cull ? c->views : 0;
    out->total_tiles = c->have_composite ? c->total_tiles : 0;
    out->tile_size = c->have_composite ? c->composite_tile : 0;
    out->radix_bits = c->radix_bits;
    out->blocks_per_cu = c->blocks_per_cu;
    out->graph = c->use_graph;
    out->graph_captures = c->graph_captures;
    out->graph_replays = c->graph_replays;
    out->overflow_reruns = c->overflow_reruns;
    out->held_bytes = c->held;
    return ok();
}

int gs_gps_sample(gs_ctx* c, uint32_t frame_seed, uint32_t supersample, int linear_color, float background) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->have_cull) return fail(GS_ERR_STATE, "gs_gps_sample needs a gs_cull first");
    if (c->views != 1) return fail(GS_ERR_STATE, "gs_gps_sample needs a single-view cull");
    GS_CHECK(set_device(c));
    // a pending stochastic-splat frame writes the same images: settle it first
    if (c->composite_kind == 1) GS_CHECK(settle_composite(c, true, true));
    uint32_t ss = supersample < 1u ? 1u : (supersample > 4u ? 4u : supersample);
    const int W = (int)c->cam.viewport[0], H = (int)c->cam.viewport[1];
    const int gw = W * (int)ss, gh = H * (int)ss;
    if (gw != c->gps_w || gh != c->gps_h) c->cells_dirty = true;
    GS_CHECK(ensure(c, c->cells, (size_t)gw * gh * 8, "GPS cells"));
    GS_CHECK(ensure(c, c->work, (size_t)c->gps_budget * 4, "GPS work list"));
    GS_CHECK(ensure(c, c->prepared, (size_t)c->n * 48, "GPS prepared splats"));
    GS_CHECK(ensure(c, c->item_count, (size_t)c->n * 4, "GPS item counts"));
    GS_CHECK(ensure(c, c->item_base, (size_t)c->n * 4, "GPS item offsets"));
    const uint32_t count_blocks = (c->n + 255u) / 256u;
    const void* old_hint = c->gps_hint.p;
    GS_CHECK(ensure(c, c->gps_hint, (size_t)count_blocks * 4, "GPS prepare hints"));
    if (c->gps_hint.p != old_hint)  // new blocks start by preparing in the count pass
        GS_HIP(launch_fill_u32(ptr<uint32_t>(c->gps_hint), 1u, c->gps_hint.bytes / 4, c->stream));
    GS_CHECK(ensure(c, c->sample, (size_t)W * H * 16, "GPS sample"));
    GS_CHECK(ensure(c, c->depth, (size_t)W * H * 4, "GPS depth"));
    if (c->cells_dirty) {
        GS_HIP(launch_fill_u64(ptr<uint64_t>(c->cells), 0xFFFFFFFF00000000ull, (size_t)gw * gh, c->stream));
        c->cells_dirty = false;
    }
    c->gps_w = gw;
    c->gps_h = gh;
    c->gps_ss = (int)ss;
    c->res_w = W;
    c->res_h = H;

    // the expand stage includes the compaction of the visible ids
    TimedLaunch t{};
    GS_CHECK(timed_begin(c, GS_STAGE_GPS_EXPAND, &t));
    GS_CHECK(ensure_compacted(c, true));
    GpsLaunch g{};
    g.records = ptr<float4>(c->records);
    g.vis_ids = ptr<uint32_t>(c->ids[0]);
    g.vis_count = &c->dc()->cull.visible_count;
    g.capacity = c->n;
    g.max_work_items = c->gps_budget;
    g.ss = (float)ss;
    g.clip_z_01 = c->cam.clip_z_01;
    g.linear_color = linear_color;
    g.prepared = ptr<uint4>(c->prepared);
    g.work = ptr<uint32_t>(c->work);
    g.item_count = ptr<uint32_t>(c->item_count);
    g.item_base = ptr<uint32_t>(c->item_base);
    g.block_counts = ptr<uint32_t>(c->block_counts);
    g.block_offsets = ptr<uint32_t>(c->block_offsets);
    g.prepare_in_count = ptr<uint32_t>(c->gps_hint);
    g.counters = c->dc();
    g.cells = ptr<uint64_t>(c->cells);
    g.gps_width = gw;
    g.gps_height = gh;
    g.frame_seed = frame_seed;
    g.width = W;
    g.height = H;
    g.background = background;
    g.sample = ptr<float4>(c->sample);
    g.depth = ptr<float>(c->depth);

    GS_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(&c->dc()->gps), 0u, sizeof(GpsCounters) / 4, c->stream));
    GS_HIP(launch_gps_expand(g, c->stream));
    GS_CHECK(timed_end(c, &t));
    TimedLaunch t2{};
    GS_CHECK(timed_begin(c, GS_STAGE_GPS_SPLAT, &t2));
    GS_HIP(launch_gps_splat(g, c->stream));
    GS_CHECK(timed_end(c, &t2));
    TimedLaunch t3{};
    GS_CHECK(timed_begin(c, GS_STAGE_GPS_RESOLVE, &t3));
    g.clear_cells = c->keep_cells ? 0 : 1;
    GS_HIP(launch_gps_resolve(g, c->stream));
    if (c->keep_cells) c->cells_dirty = true;
    GS_CHECK(timed_end(c, &t3));
    c->have_gps = true;
    return ok();
}

int gs_accumulate(gs_ctx* c, uint32_t sample_count) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->have_gps) return fail(GS_ERR_STATE, "gs_accumulate needs a gs_gps_sample or gs_stochastic_sample first");
    GS_CHECK(set_device(c));
    // a stochastic sample whose tile lists overflowed is redone before it is averaged
    if (c->composite_kind == 1) GS_CHECK(settle_composite(c, true, true));
    if (c->acc_w != c->res_w || c->acc_h != c->res_h) {
        if (sample_count > 1)
            return fail(GS_ERR_STATE, "accumulation history is %dx%d, sample is %dx%d; reset with sample_count 1",
                        c->acc_w, c->acc_h, c->res_w, c->res_h);
    }
    GS_CHECK(ensure(c, c->accum, (size_t)c->res_w * c->res_h * 16, "accumulation image"));
    GS_HIP(launch_accumulate(ptr<float4>(c->sample), ptr<float4>(c->accum), (uint32_t)(c->res_w * c->res_h),
                             (float)sample_count, c->stream));
    c->acc_w = c->res_w;
    c->acc_h = c->res_h;
    c->have_accum = true;
    return ok();
}

void gs_taa_jitter(uint32_t frame_seed, float jitter_px[2]) {
    if (!jitter_px) return;
    const uint32_t index = frame_seed % 1024u + 1u;
    const uint32_t bases[2] = {2u, 3u};
    for (int k = 0; k < 2; ++k) {
        // radical inverse of index in the base, in float as the reference
        float f = 1.0f, r = 0.0f;
        for (uint32_t i = index; i > 0; i /= bases[k]) {
            f /= (float)bases[k];
            r += f * (float)(i % bases[k]);
        }
        jitter_px[k] = r - 0.5f;
    }
}

int gs_frame_average(gs_ctx* c, uint32_t sample_index) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->have_gps) return fail(GS_ERR_STATE, "gs_frame_average needs a gs_stochastic_sample first");
    GS_CHECK(set_device(c));
    if (c->composite_kind == 1) GS_CHECK(settle_composite(c, true, true));
    if (sample_index > 0 && (c->avg_w != c->res_w || c->avg_h != c->res_h || !c->have_avg))
        return fail(GS_ERR_STATE, "frame average is %dx%d, sample is %dx%d; start with sample_index 0", c->avg_w,
                    c->avg_h, c->res_w, c->res_h);
    GS_CHECK(ensure(c, c->frame_avg, (size_t)c->res_w * c->res_h * 16, "frame average image"));
    GS_HIP(launch_accumulate(ptr<float4>(c->sample), ptr<float4>(c->frame_avg), (uint32_t)(c->res_w * c->res_h),
                             (float)sample_index + 1.0f, c->stream));
    c->avg_w = c->res_w;
    c->avg_h = c->res_h;
    c->have_avg = true;
    return ok();
}

namespace {

// Inverse of a column-major 4x4 matrix: Gauss-Jordan elimination with
// partial pivoting in double precision, rounded to float.  False when
// |det| <= 1e-8, the reference's singularity threshold.
bool invert4(const float* m, float* out) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            a[r][k] = m[k * 4 + r];
            a[r][4 + k] = r == k ? 1.0 : 0.0;
        }
    double det = 1.0;
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return false;
        if (piv != col) {
            for (int k = 0; k < 8; ++k) std::swap(a[piv][k], a[col][k]);
            det = -det;
        }
        det *= a[col][col];
        const double inv = 1.0 / a[col][col];
        for (int k = 0; k < 8; ++k) a[col][k] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == col || a[r][col] == 0.0) continue;
            const double f = a[r][col];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[col][k];
        }
    }
    if (std::fabs(det) <= 1e-8) return false;
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) out[k * 4 + r] = (float)a[r][4 + k];
    return true;
}

}  // namespace

int gs_taa_uniforms(const GsCameraUniforms* cam, const GsCameraUniforms* prev, int view_changed,
                    GsTaaUniforms* out) {
    if (!cam || !out) return fail(GS_ERR_INVALID, "bad arguments");
    *out = GsTaaUniforms{};
    const bool invertible = invert4(cam->view, out->inv_view);
    if (!invertible) {
        std::memcpy(out->inv_view, cam->view, sizeof(out->inv_view));
        prev = nullptr;
    }
    const GsCameraUniforms* p = prev ? prev : cam;
    std::memcpy(out->prev_view, p->view, sizeof(out->prev_view));
    out->viewport_orthographic[0] = cam->viewport[0];
    out->viewport_orthographic[1] = cam->viewport[1];
    out->viewport_orthographic[2] = cam->orthographic;
    out->viewport_orthographic[3] = p->orthographic;
    out->focal_clip[0] = cam->persp_focal;
    out->focal_clip[1] = cam->ortho_focal;
    out->focal_clip[2] = cam->clip_y_sign;
    out->focal_clip[3] = cam->clip_z_01;
    out->prev_focal_clip[0] = p->persp_focal;
    out->prev_focal_clip[1] = p->ortho_focal;
    out->prev_focal_clip[2] = p->clip_y_sign;
    out->proj_z[0] = cam->proj[10];
    out->proj_z[1] = cam->proj[14];
    out->history_valid = prev ? 1.0f : 0.0f;
    out->view_changed = view_changed ? 1.0f : 0.0f;
    if (!invertible) return fail(GS_ERR_INVALID, "view matrix is not invertible; history dropped");
    return ok();
}

int gs_taa_accumulate(gs_ctx* c, const GsTaaUniforms* u, int use_frame_average) {
    if (!c || !u) return fail(GS_ERR_INVALID, "bad arguments");
    if (!c->have_gps) return fail(GS_ERR_STATE, "gs_taa_accumulate needs a gs_stochastic_sample first");
    GS_CHECK(set_device(c));
    if (c->composite_kind == 1) GS_CHECK(settle_composite(c, true, true));
    const int W = c->res_w, H = c->res_h;
    if (use_frame_average && (!c->have_avg || c->avg_w != W || c->avg_h != H))
        return fail(GS_ERR_STATE, "no frame average of this size; call gs_frame_average");
    if (u->history_valid != 0.0f && (!c->have_taa || c->taa_w != W || c->taa_h != H))
        return fail(GS_ERR_STATE, "TAA history is %dx%d, frame is %dx%d; pass history_valid 0 to restart",
                    c->taa_w, c->taa_h, W, H);
    if (u->viewport_orthographic[0] != (float)W || u->viewport_orthographic[1] != (float)H)
        return fail(GS_ERR_INVALID, "uniform viewport %gx%g does not match the %dx%d sample",
                    (double)u->viewport_orthographic[0], (double)u->viewport_orthographic[1], W, H);
    const size_t bytes = (size_t)W * H * 16;
    for (int k = 0; k < 2; ++k) {
        GS_CHECK(ensure(c, c->taa_color[k], bytes, "TAA colour history"));
        GS_CHECK(ensure(c, c->taa_xyz[k], bytes, "TAA position history"));
    }
    const int rd = c->taa_last, wr = c->have_taa ? 1 - c->taa_last : 0;
    TaaLaunch t{};
    t.u = *u;
    t.width = W;
    t.height = H;
    t.color = use_frame_average ? ptr<float4>(c->frame_avg) : ptr<float4>(c->sample);
    t.depth = ptr<float>(c->depth);
    t.hist_color = ptr<float4>(c->taa_color[rd]);
    t.hist_xyz = ptr<float4>(c->taa_xyz[rd]);
    t.out_color = ptr<float4>(c->taa_color[wr]);
    t.out_xyz = ptr<float4>(c->taa_xyz[wr]);
    GS_HIP(launch_taa(t, c->stream));
    c->taa_last = wr;
    c->taa_w = W;
    c->taa_h = H;
    c->have_taa = true;
    return ok();
}

int gs_display_layers(gs_ctx* c, int source, float exposure_ev) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    if (!c->lut_size) return fail(GS_ERR_STATE, "no display LUT uploaded");
    if (!std::isfinite(exposure_ev)) return fail(GS_ERR_INVALID, "exposure is not finite");
    GS_CHECK(set_device(c));
    LayersDisplayLaunch d{};
    int w = 0, h = 0;
    switch (source) {
        case GS_IMAGE_ALPHA_RGBA16F:
            if (!c->have_composite) return fail(GS_ERR_STATE, "no composite yet");
            if (c->views != 1) return fail(GS_ERR_STATE, "display of a single-view composite only");
            GS_CHECK(settle_composite(c, true, true));
            d.source = c->image.p;
            d.half_source = 1;
            w = c->img_w;
            h = c->img_h;
            break;
        case GS_IMAGE_GPS_SAMPLE_RGBA32F:
            if (!c->have_gps) return fail(GS_ERR_STATE, "no sample yet");
            if (c->composite_kind == 1) GS_CHECK(settle_composite(c, true, true));
            d.source = c->sample.p;
            w = c->res_w;
            h = c->res_h;
            break;
        case GS_IMAGE_ACCUM_RGBA32F:
            if (!c->have_accum) return fail(GS_ERR_STATE, "no accumulation yet");
            d.source = c->accum.p;
            w = c->acc_w;
            h = c->acc_h;
            break;
        case GS_IMAGE_TAA_RGBA32F:
            if (!c->have_taa) return fail(GS_ERR_STATE, "no TAA pass yet");
            d.source = c->taa_color[c->taa_last].p;
            w = c->taa_w;
            h = c->taa_h;
            break;
        default:
            return fail(GS_ERR_INVALID, "image %d is not a linear colour image", source);
    }
    GS_CHECK(ensure(c, c->display, (size_t)w * h * 16, "display image"));
    d.count = (uint32_t)(w * h);
    d.lut = ptr<uint16_t>(c->lut);
    d.lut_size = (int)c->lut_size;
    d.log2_min = c->lut_log2_min;
    d.log2_max = c->lut_log2_max;
    d.exposure_ev = exposure_ev;
    d.out = ptr<float4>(c->display);
    GS_HIP(launch_layers_display(d, c->stream));
    c->disp_w = w;
    c->disp_h = h;
    c->have_display = true;
    return ok();
}

int gs_read_image(gs_ctx* c, int which, void* host, size_t bytes) {
    if (!c || !host) return fail(GS_ERR_INVALID, "bad arguments");
    GS_CHECK(set_device(c));
    switch (which) {
        case GS_IMAGE_ALPHA_RGBA16F:
            if (!c->have_composite) return fail(GS_ERR_STATE, "no composite yet");
            GS_CHECK(settle_composite(c, true, true));
            GS_CHECK(copy_image(c, c->image.p, host, c->img_w, c->img_h, 8, bytes, gl_rows(c)));
            return ok();
        case GS_IMAGE_GPS_SAMPLE_RGBA32F:
            if (!c->have_gps) return fail(GS_ERR_STATE, "no GPS sample yet");
            GS_CHECK(settle_composite(c, true, true));
            GS_CHECK(copy_image(c, c->sample.p, host, c->res_w, c->res_h, 16, bytes, gl_rows(c)));
            return ok();
        case GS_IMAGE_GPS_DEPTH_R32F:
            if (!c->have_gps) return fail(GS_ERR_STATE, "no GPS sample yet");
            GS_CHECK(settle_composite(c, true, true));
            GS_CHECK(copy_image(c, c->depth.p, host, c->res_w, c->res_h, 4, bytes, gl_rows(c)));
            return ok();
        case GS_IMAGE_ACCUM_RGBA32F:
            if (!c->have_accum) return fail(GS_ERR_STATE, "no accumulation yet");
            GS_CHECK(copy_image(c, c->accum.p, host, c->acc_w, c->acc_h, 16, bytes, gl_rows(c)));
            return ok();
        case GS_IMAGE_FRAME_AVG_RGBA32F:
            if (!c->have_avg) return fail(GS_ERR_STATE, "no frame average yet");
            GS_CHECK(copy_image(c, c->frame_avg.p, host, c->avg_w, c->avg_h, 16, bytes, gl_rows(c)));
            return ok();
        case GS_IMAGE_DISPLAY_RGBA32F:
            if (!c->have_display) return fail(GS_ERR_STATE, "no display pass yet");
            GS_CHECK(copy_image(c, c->display.p, host, c->disp_w, c->disp_h, 16, bytes, gl_rows(c)));
            return ok();
        case GS_IMAGE_TAA_RGBA32F:
        case GS_IMAGE_TAA_XYZ_RGBA32F: {
            if (!c->have_taa) return fail(GS_ERR_STATE, "no TAA pass yet");
            const DevBuf& b = which == GS_IMAGE_TAA_RGBA32F ? c->taa_color[c->taa_last] : c->taa_xyz[c->taa_last];
            GS_CHECK(copy_image(c, b.p, host, c->taa_w, c->taa_h, 16, bytes, gl_rows(c)));
            return ok();
        }
        default:
            return fail(GS_ERR_INVALID, "unknown image %d", which);
    }
}

int gs_read_buffer(gs_ctx* c, int which, void* host, size_t bytes) {
    if (!c || !host) return fail(GS_ERR_INVALID, "bad arguments");
    GS_CHECK(set_device(c));
    if (!c->have_cull) return fail(GS_ERR_STATE, "no cull yet");
    if (c->views != 1) return fail(GS_ERR_STATE, "buffers are readable after single-view frames only");
    uint32_t V = 0;
    GS_CHECK(read_u32(c, &c->dc()->cull.visible_count, &V));
    const size_t n = c->n;
    switch (which) {
        case GS_BUFFER_SORTED_KEYS:
        case GS_BUFFER_SORTED_IDS: {
            if (!c->have_sort) return fail(GS_ERR_STATE, "no sort yet");
            if (bytes != n * 4) return fail(GS_ERR_INVALID, "sorted buffer needs %zu bytes", n * 4);
            uint32_t* h = static_cast<uint32_t*>(host);
            GS_HIP(hipMemcpyAsync(host, c->ids[c->sort_slot].p, (size_t)V * 4, hipMemcpyDeviceToHost, c->stream));
            if (which == GS_BUFFER_SORTED_KEYS) {
                // the depth sort writes no keys: each sorted id (a dense
                // [view][n] slot) indexes the cull's dense keys
                std::vector<uint32_t> dense((size_t)c->views * c->n);
                GS_HIP(hipMemcpyAsync(dense.data(), c->dense_keys.p, dense.size() * 4, hipMemcpyDeviceToHost,
                                      c->stream));
                GS_HIP(hipStreamSynchronize(c->stream));
                for (size_t i = 0; i < V; ++i) {
                    if (h[i] >= dense.size()) return fail(GS_ERR_INVALID, "sorted id %u out of range", h[i]);
                    h[i] = dense[h[i]];
                }
            } else {
                GS_HIP(hipStreamSynchronize(c->stream));
            }
            for (size_t i = V; i < n; ++i) h[i] = 0xFFFFFFFFu;
            return ok();
        }
        case GS_BUFFER_VISIBLE_IDS:
        case GS_BUFFER_VISIBLE_KEYS: {
            if (bytes < (size_t)V * 4) return fail(GS_ERR_INVALID, "visible buffer needs %zu bytes", (size_t)V * 4);
            GS_CHECK(ensure_compacted(c));
            DevBuf& src = which == GS_BUFFER_VISIBLE_KEYS ? c->keys[0] : c->ids[0];
            GS_HIP(hipMemcpyAsync(host, src.p, (size_t)V * 4, hipMemcpyDeviceToHost, c->stream));
            GS_HIP(hipStreamSynchronize(c->stream));
            return ok();
        }
        case GS_BUFFER_PROJECTED: {
            if (bytes != n * 64) return fail(GS_ERR_INVALID, "projected buffer needs %zu bytes", n * 64);
            GS_CHECK(ensure_compacted(c));
            std::vector<uint32_t> vis(V);
            std::vector<float> rec(n * kRecordVec4 * 4);
            if (V) GS_HIP(hipMemcpyAsync(vis.data(), c->ids[0].p, (size_t)V * 4, hipMemcpyDeviceToHost, c->stream));
            GS_HIP(hipMemcpyAsync(rec.data(), c->records.p, rec.size() * 4, hipMemcpyDeviceToHost, c->stream));
            GS_HIP(hipStreamSynchronize(c->stream));
            float* out = static_cast<float*>(host);
            for (size_t i = 0; i < n; ++i) {
                float* o = out + i * 16;
                memset(o, 0, 64);
                o[8] = 1.0f; o[10] = 1.0f; o[11] = 1.0f;  // culled record (shaders/cull.glsl:222-227)
            }
            for (uint32_t k = 0; k < V; ++k) {
                uint32_t id = vis[k];
                const float* r = rec.data() + (size_t)id * 12;
                float* o = out + (size_t)id * 16;
                float a = r[4], b = r[5], cc = r[6];
                float det = a * cc - b * b;
                if (det < 1e-6f) det = 1e-6f;
                o[0] = r[8]; o[1] = r[9]; o[2] = r[10]; o[3] = r[2];
                o[4] = r[0]; o[5] = r[1]; o[6] = r[7]; o[7] = r[11];
                o[8] = cc / det; o[9] = -b / det; o[10] = a / det; o[11] = r[3];
                o[12] = a; o[13] = b; o[14] = cc; o[15] = det;
            }
            return ok();
        }
        case GS_BUFFER_GPS_PACKED: {
            if (!c->have_gps || !c->keep_cells) return fail(GS_ERR_STATE, "needs a GPS sample taken with GS_OPT_GPS_KEEP_CELLS");
            size_t need = (size_t)c->gps_w * c->gps_h * 8;
            if (bytes != need) return fail(GS_ERR_INVALID, "packed buffer needs %zu bytes", need);
            GS_HIP(hipMemcpyAsync(host, c->cells.p, need, hipMemcpyDeviceToHost, c->stream));
            GS_HIP(hipStreamSynchronize(c->stream));
            return ok();
        }
        default:
            return fail(GS_ERR_INVALID, "unknown buffer %d", which);
    }
}

int gs_stats(gs_ctx* c, GsStats* out) {
    if (!c || !out) return fail(GS_ERR_INVALID, "bad arguments");
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, true, true));
    DevCounters h;
    GS_HIP(hipMemcpyAsync(&h, c->counters.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    GS_HIP(hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof(*out));
    out->gaussian_count = c->n;
    out->visible_count = c->have_cull ? h.cull.visible_count : 0;
    out->tile_entries = c->have_composite ? c->last_dup : 0;
    out->gps_max_work_items = c->gps_budget;
    if (c->have_gps) {
        out->gps_requested_work_items = h.gps.requested;
        // reservations are contiguous from 0: some reservation crossed the
        // budget (the reference's overflow flag) exactly when the total did
        out->gps_work_budget_exhausted = h.gps.requested > c->gps_budget ? 1u : 0u;
        out->gps_dropped_work_items = h.gps.requested > c->gps_budget ? h.gps.requested - c->gps_budget : 0;
    }
    if (c->have_cull && c->stats_collected) {
        memcpy(out->diagnostics, h.cull.diag, sizeof(out->diagnostics));
        out->diagnostics_valid = 1;
    }
    return ok();
}

int gs_synchronize(gs_ctx* c) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, true, true));
    GS_HIP(hipStreamSynchronize(c->stream));
    return ok();
}

int gs_set_profiling(gs_ctx* c, int enabled) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    c->profiling = enabled != 0;
    return ok();
}

int gs_stage_times(gs_ctx* c, double* ms, uint32_t* counts) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    GS_CHECK(set_device(c));
    GS_HIP(hipStreamSynchronize(c->stream));
    for (auto& t : c->timed) {
        float e = 0.0f;
        GS_HIP(hipEventElapsedTime(&e, t.start, t.stop));
        c->stage_ms[t.stage] += e;
        c->stage_count[t.stage] += 1;
        c->free_events.push_back(t.start);
        c->free_events.push_back(t.stop);
    }
    c->timed.clear();
    for (int k = 0; k < GS_STAGE_COUNT; ++k) {
        if (ms) ms[k] = c->stage_ms[k];
        if (counts) counts[k] = c->stage_count[k];
    }
    return ok();
}

int gs_reset_stage_times(gs_ctx* c) {
    if (!c) return fail(GS_ERR_INVALID, "ctx is NULL");
    GS_CHECK(gs_stage_times(c, nullptr, nullptr));
    for (int k = 0; k < GS_STAGE_COUNT; ++k) {
        c->stage_ms[k] = 0.0;
        c->stage_count[k] = 0;
    }
    return ok();
}

void* gs_stream(gs_ctx* c) { return c ? (void*)c->stream : nullptr; }

}  // extern "C"

namespace gs {

int api_error(int code, const char* msg) { return fail(code, "%s", msg); }
int api_ok() { return ok(); }

int ctx_launch_target(gs_ctx* c, hipStream_t* stream, int* device) {
    if (!c) return GS_ERR_INVALID;
    *stream = c->stream;
    *device = c->device;
    return GS_OK;
}

int ctx_frames(gs_ctx* c, CtxFrames* out) {
    if (!c || !out) return fail(GS_ERR_INVALID, "bad arguments");
    GS_CHECK(set_device(c));
    GS_CHECK(settle_composite(c, true, true));
    memset(out, 0, sizeof(*out));
    out->stream = c->stream;
    out->device = c->device;
    if (c->have_accum && c->acc_w == c->res_w && c->acc_h == c->res_h) {
        out->accum = ptr<float4>(c->accum);
        out->depth = ptr<float>(c->depth);
        out->gps_w = c->acc_w;
        out->gps_h = c->acc_h;
    }
    if (c->have_composite && c->views == 1) {
        out->alpha = ptr<uint16_t>(c->image);
        out->alpha_w = c->img_w;
        out->alpha_h = c->img_h;
    }
    out->clip_y_sign = c->cam.clip_y_sign;
    return GS_OK;
}

}  // namespace gs

extern "C" {

}  // extern "C"
