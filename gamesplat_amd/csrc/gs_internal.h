// gs_internal.h - device data layout, launch interfaces and the wave64
// primitives shared by the kernels of libgamesplat_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/gs_api.h"

namespace gs {

// a[i]; with O32 (arrays under 4 GiB) the byte offset is formed in 32 bits,
// so a load or store takes the array base in scalar registers and one offset
// register rather than a 64-bit address per element.
template <bool O32, typename T>
__device__ __forceinline__ T& at(T* a, uint32_t i) {
    if constexpr (O32)
        return *reinterpret_cast<T*>(reinterpret_cast<char*>(const_cast<void*>(static_cast<const void*>(a))) +
                                     (uint32_t)(i * (uint32_t)sizeof(T)));
    else
        return a[i];
}

// ---------------------------------------------------------------------------
// HBM layout
//
// Scene: 15 float4 planes of N entries each (plane p, Gaussian i at p*N + i):
//   P0  pos.xyz, opacity          P1  rot w,x,y,z
//   P2  scale.xyz, dc.r           P3  dc.g, dc.b, rest[0], rest[1]
//   P4+j rest[2+4j .. 5+4j] (j = 0..9)      P14 rest[42..44], 0
// A lane of a wave reads plane p of its own Gaussian, so every load of the
// cull is one fully coalesced 128-bit load per lane.
constexpr int kPlanes = 15;

// Projected splat, 48 bytes (3 float4) per Gaussian, written only for
// visible Gaussians:
//   R0 centre.x, centre.y, opacity, ndc_z
//   R1 cov a, cov b * clip_y_sign, cov c, radius_x      (a, c include +0.3)
//   R2 colour r, g, b, radius_y
// The conic and the clamped determinant are recomputed from R1 with the
// reference's operations wherever they are needed.
constexpr int kRecordVec4 = 3;

// Views per batched launch (one bit each in the cull's visibility mask).
constexpr int kMaxViews = 64;

// One view of a batched frame.  Per-view arrays (records, pixel rectangles,
// dense keys) are laid out [view][n], so the compacted value view * n + id
// addresses them directly; tile ids are global (tile_base + local tile) and
// the RGBA16F images are packed one after the other (pixel_base).
struct ViewInfo {
    int width, height, tiles_x, tiles_y;
    uint32_t tile_base;
    uint32_t pixel_base;
};

// Device counters.  Each stage clears its own group with one fill launch
// before it runs, so a stage can be re-run without re-running earlier ones.
// Totals are written by the scan kernels, never accumulated with atomics.
struct CullCounters {
    uint32_t visible_count;
    uint32_t pad0;
    uint32_t key_min, key_max;   // range of the visible depth keys (all views), 8-byte aligned
    uint32_t diag[8];
    uint32_t pad[4];
};
static_assert(sizeof(CullCounters) == 64, "CullCounters layout");
struct SortCounters {
    uint32_t scratch[512];   // per pass: digit column totals
};
struct BinCounters {
    uint32_t dup_total;   // (tile, splat) pairs requested
    uint32_t dup_count;   // min(dup_total, capacity): pairs written and sorted
    uint32_t pad[6];
};
struct GpsCounters {
    uint32_t requested;
    uint32_t pad[7];
};
struct DevCounters {
    CullCounters cull;
    SortCounters depth_sort;
    BinCounters bin;
    SortCounters tile_sort;
    GpsCounters gps;
    uint32_t composite_queue[8];  // persistent composite: tiles taken per XCD band
};

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave64 prefix sum, all 64 lanes active, in registers: shifts by
// 1, 2, 4 and 8 lanes within each row of 16 (lanes shifted in from outside
// the row read 0), then row 0's total is broadcast into row 1 and row 2's
// into row 3 (row_bcast:15), and rows 0-1's total into rows 2 and 3
// (row_bcast:31).  Rows a broadcast does not write keep 0 as its result.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// Exclusive prefix over a block of WAVES wave64s; returns the block total
// through *total.  `scratch` holds WAVES words of LDS.  All threads call it.
template <int WAVES>
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t* scratch, uint32_t* total) {
    uint32_t inc = wave_inclusive_sum(v);
    uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 63) scratch[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < WAVES; ++k) {
        uint32_t s = scratch[k];
        if ((uint32_t)k < w) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// Index of the last entry of the non-decreasing LDS array start[0..n) that is
// <= i (threads with nothing to emit share the next thread's start, so this
// is the owner of output i).
__device__ __forceinline__ uint32_t owner_of(const uint32_t* start, uint32_t n, uint32_t i) {
    uint32_t lo = 0u, hi = n - 1u;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if (start[mid] <= i) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// Map a linear block index so that each of the 8 XCDs of an MI355X works on
// one contiguous range of work items (blocks are dealt to XCDs round-robin).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    const uint32_t xcds = 8u;
    uint32_t per = nblocks / xcds, rem = nblocks % xcds;
    uint32_t xcd = bid % xcds, local = bid / xcds;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// ---------------------------------------------------------------------------
// Kernel launch interfaces.

// gs_scan.hip: function attributes the launches below need (the dynamic
// LDS size of the scan), set once per process before any frame is captured.
hipError_t init_kernel_attributes();
// gs_scan.hip: exclusive scan of nb block counts into offsets; *total gets
// the sum, and *clamped (when given) min(sum, clamp_cap).  block_diag
// (nb x 8, may be NULL) is reduced into diag_out.
hipError_t launch_scan_counts(const uint32_t* counts, uint32_t nb, uint32_t* offsets, uint32_t* total,
                              const uint32_t* block_diag, uint32_t* diag_out, hipStream_t s,
                              uint32_t clamp_cap = 0, uint32_t* clamped = nullptr,
                              const uint2* block_range = nullptr, uint32_t* range_out = nullptr);
// block_range (nb x (min, max), may be NULL) is reduced into range_out[0..1].
// The same totals, diagnostics and range without the offsets.
hipError_t launch_reduce_counts(const uint32_t* counts, uint32_t nb, uint32_t* total, const uint32_t* block_diag,
                                uint32_t* diag_out, hipStream_t s, const uint2* block_range = nullptr,
                                uint32_t* range_out = nullptr);

// gs_cull.hip
struct CullLaunch {
    const float4* planes;
    uint32_t n;
    const GsCameraUniforms* cams;   // device, one per view
    uint32_t views;                 // 1 .. kMaxViews
    GsSplatEffectParams effect;
    int effect_active;
    int sh_degree;
    float sh_lod_distance;
    int collect_stats;
    float4* records;         // [view][n] x 3 float4
    uint32_t* dense_keys;    // [view][n]: depth key or 0xFFFFFFFF when culled
    uint32_t* block_counts;  // [view][block]: visible Gaussians per 256-block
    uint2* block_range;      // [view][block]: min and max of the visible depth keys (or NULL)
    uint32_t* block_diag;    // 8 diagnostics per block (collect_stats, one view)
    // optional [view][n]: covered pixel rectangle of each visible splat
    uint2* rect_px;
    int tight_rects;         // cut the rectangle to the opacity-aware ellipse box
    // optional [view][n]: the same rectangle grown by packed_pad pixels (1:
    // every StochasticSplats jitter up to half a pixel) as its tile
    // rectangle at tile size 1 << packed_shift (pack_tile_rect)
    uint32_t* packed_rects;
    int packed_shift, packed_pad;
    // GSAnim appearance: 0 none, 1 int8 deltas (48 per row and state, state
    // 0 excluded), 2 half coefficients (48 per row and layer)
    int anim;
    const uint4* anim_data;       // 16-byte aligned rows: 3 (deltas) or 6 (layers) uint4
    int anim_state_a, anim_state_b;
    float anim_weight;
    int anim_layers;              // layers summed (anim 2)
    const float* layer_weights;   // device, 3 per layer
    float anim_scales[48];
};
hipError_t launch_cull(const CullLaunch& a, hipStream_t s);
// With pack_tile (8, 16 or 32) it also writes, per visible splat, the tile
// rectangle of its pixel rectangle rect_px at that tile size as
// x0 | x1 << 8 | y0 << 16 | y1 << 24 (0 when empty; every view must span at
// most 255 tiles a side), which the depth sort carries to the binning; with
// packed (the cull's packed_rects) that packed rectangle is moved instead
// and rect_px is not read.
hipError_t launch_compact(const uint32_t* dense_keys, uint32_t n, uint32_t views, const uint32_t* offsets,
                          uint32_t* vis_keys, uint32_t* vis_ids, const uint2* rect_px, int pack_tile,
                          uint32_t* vis_rects, hipStream_t s, const uint32_t* packed = nullptr);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
// Partial boxes of the n Gaussian centres of plane P0: kCentreBoundBlocks
// groups of 8 floats (min x, y, z, max x, y, z, non-finite flag, 0).
constexpr uint32_t kCentreBoundBlocks = 256;
hipError_t launch_centre_bounds(const float4* p0, uint32_t n, float* partials, hipStream_t s);

// gs_probe.hip: HBM rate probe `which` (0 read, 1 fifteen planes, 2 write,
// 3 copy) over n float4 of a (read) and b (written); sink takes one float.
hipError_t launch_probe(int which, const float4* a, float4* b, size_t n, float* sink, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t value, size_t count, hipStream_t s);

// gs_sort.hip
struct SortLaunch {
    // keys[0] input; passes ping-pong keys[1] <-> keys[2]; with key_range the
    // last pass the range needs writes keys[3] instead
    uint32_t* keys[4];
    uint32_t* vals[4];
    uint32_t* vals2[4];      // optional second value moved with the same permutation (NULL: none)
    // optional device (min, max) of the keys: the passes sort key - min, and
    // those past the bits the range needs exit at once
    const uint32_t* key_range;
    const uint32_t* count;   // device element count
    uint32_t capacity;       // grid bound
    uint32_t* scratch;       // 512 words
    uint32_t* table;         // sort_table_words(capacity) words
    int passes;
    int bits;                // digit width (9, 8, 7, 6 or 4)
    int wide_index;          // 1: 64-bit element indexing even where 32-bit offsets suffice
    int key16;               // 1: keys are 16-bit (below 2^16; no dense pass, no key range)
    int skip_last_keys;      // 1 (with key_range): the last pass writes values only, not keys
    // optional dense first pass (digit widths 7 and 4): it reads dense_count
    // slots of dense_keys (0xFFFFFFFF: culled) instead of keys[0] and count,
    // keeps the others in slot order with the slot as value, and with vals2
    // their tile rectangles: dense_rect_packed[slot] as it is, else
    // dense_rect_px[slot] packed at 1 << dense_rect_shift pixels a tile
    const uint32_t* dense_keys;
    uint32_t dense_count;
    const uint2* dense_rect_px;
    const uint32_t* dense_rect_packed;
    int dense_rect_shift;
};
// Stable LSD radix sort.  Returns the index (1, 2, or 3 with key_range) of
// the buffer holding the result through *result_slot.
hipError_t launch_radix_sort(const SortLaunch& a, int* result_slot, hipStream_t s);
uint32_t sort_table_words(uint32_t capacity);

// gs_composite.hip
struct BinLaunch {
    const uint32_t* sorted_ids;  // view * n + id, depth order
    const uint32_t* count;   // visible count over all views (device)
    uint32_t capacity;       // views * n (grid bound)
    uint32_t n;              // per-view stride of the [view][n] arrays
    int tile;
    const ViewInfo* views;   // device
    uint32_t single_tiles_x; // one view: its tiles per row (tile base 0); 0: look the view up
    const uint2* rect_px;    // per Gaussian: covered pixel rectangle (from the cull)
    // optional: per depth-sorted position, the packed tile rectangle at this
    // bin's tile size, moved by the depth sort (replaces the rect_px gather)
    const uint32_t* sorted_rects;
    // jittered StochasticSplats samples: the rectangle is recomputed from
    // the record with the quad moved by (jitter_x, jitter_y) and grown by
    // pad pixels on each side
    const float4* records;
    int jittered;
    int tight;               // the cull's tight-rectangle setting
    int pad;
    float jitter_x, jitter_y;
    float clip_z_01;         // the view's depth range (quads outside it are clipped)
    uint2* rects;            // per front-to-back position: tile rectangle (without sorted_rects)
    uint32_t* block_counts;  // pairs per 256-block
    uint32_t* block_offsets;
    uint32_t* clear;         // optional: clear_words words zeroed by the count pass (the tile ranges)
    uint32_t clear_words;
    DevCounters* counters;
    uint32_t* dup_keys;
    uint32_t* dup_vals;
    uint32_t dup_capacity;   // pairs that fit in dup_keys / dup_vals
    int wide_index;          // 1: 64-bit pair indexing even where 32-bit offsets suffice
    int key16;               // 1: dup_keys holds 16-bit tile ids (every tile id below 2^16)
};
// Counts and emits (tile id, splat id) pairs; leaves the requested total in
// counters->bin.dup_total and min(total, capacity) in counters->bin.dup_count.
hipError_t launch_bin(const BinLaunch& a, hipStream_t s);
hipError_t launch_tile_ranges(const uint32_t* keys, int key16, const uint32_t* count, uint32_t capacity, uint2* ranges,
                              hipStream_t s);

struct CompositeLaunch {
    const float4* records;   // [view][n] x 3 float4
    const uint32_t* tile_vals;
    const uint2* ranges;     // per global tile id
    int tile;
    const ViewInfo* views;   // device
    uint32_t nviews;
    uint32_t total_tiles;
    uint16_t* image;         // RGBA16F per view, framebuffer row order, packed by pixel_base
    int quadrant_cull;       // 16x16 tiles: 8x8 quadrant kernel; 1 = rectangle + ellipse test, 2 = rectangle only
    int tight_rects;         // per-splat pixel rectangle cut to the opacity-aware ellipse box
    // 1: the reference's arithmetic (back to front, half-precision storage
    // after every blend; tile 16).  0: front to back in fp32.
    int reference_blend;
    // 0: one block per tile.  >0: that many persistent blocks pull tiles from
    // 8 per-XCD queues (queue: 8 counters, cleared before the launch).
    uint32_t persistent_blocks;
    uint32_t* queue;
};
hipError_t launch_composite(const CompositeLaunch& a, hipStream_t s);

// StochasticSplats sample over the 16x16 tile lists of one view.
struct StochasticLaunch {
    CompositeLaunch tiles;   // records, tile lists and ranges, tile = 16
    float frame_seed;
    float background;
    float clip_z_01;
    float jitter_x, jitter_y;  // TAA jitter of the quads and fragment centres (pixels)
    float4* sample;          // W x H RGBA32F, framebuffer rows
    float* depth;            // W x H window depth of the kept splat (1 where none)
};
hipError_t launch_stochastic(const StochasticLaunch& a, hipStream_t s);

// gs_gps.hip
struct GpsLaunch {
    const float4* records;
    const uint32_t* vis_ids;
    const uint32_t* vis_count;   // device
    uint32_t capacity;           // Gaussian capacity (grid bound)
    uint32_t max_work_items;
    float ss;
    float clip_z_01;
    int linear_color;
    uint4* prepared;             // 3 x uint4 per visible position
    uint32_t* work;              // owner position per work item
    uint32_t* item_count;        // per visible position
    uint32_t* item_base;         // per visible position
    uint32_t* block_counts;
    uint32_t* block_offsets;
    uint32_t* prepare_in_count;  // per count block: 1 = pass 1 writes the prepared records (the
                                 // block kept work last frame), 0 = pass 2 writes them for kept
                                 // splats; pass 2 updates it
    DevCounters* counters;
    // splat
    uint64_t* cells;
    int gps_width, gps_height;
    uint32_t frame_seed;
    // resolve
    int width, height;
    float background;
    int clear_cells;             // resolve writes the empty value back
    float4* sample;
    float* depth;
};
hipError_t launch_gps_expand(const GpsLaunch& a, hipStream_t s);
hipError_t launch_gps_splat(const GpsLaunch& a, hipStream_t s);
hipError_t launch_gps_resolve(const GpsLaunch& a, hipStream_t s);
hipError_t launch_accumulate(const float4* sample, float4* accum, uint32_t count, float n, hipStream_t s);

// gs_taa.hip: taa_accum_fs over W x H framebuffer-row images.
struct TaaLaunch {
    GsTaaUniforms u;
    int width, height;
    const float4* color;       // current colour (frame average or sample)
    const float* depth;        // current sample depth
    const float4* hist_color;  // history read buffer
    const float4* hist_xyz;
    float4* out_color;         // history write buffer
    float4* out_xyz;
};
hipError_t launch_taa(const TaaLaunch& a, hipStream_t s);

// gs_display.hip: layers_display_fs over count pixels.
struct LayersDisplayLaunch {
    const void* source;       // RGBA16F (half_source) or RGBA32F pixels
    int half_source;
    uint32_t count;
    const uint16_t* lut;      // lut_size^3 RGB uint16, red fastest
    int lut_size;
    float log2_min, log2_max, exposure_ev;
    float4* out;
};
hipError_t launch_layers_display(const LayersDisplayLaunch& a, hipStream_t s);

// gs_unet.hip: the denoiser's kernels.  Tensors are NHWC bf16 (raw uint16
// bits) viewed through a channel stride and a channel offset.
struct TensorView {
    uint16_t* p;
    int cs;   // channel stride of the buffer
    int co;   // first channel of the view
};
struct ConvLaunch {
    const uint16_t* x;
    int xs, xo;              // input channel stride / offset
    uint16_t* y;
    int ys, yo;              // output channel stride / offset
    const uint16_t* wt;      // bf16 [cout][kpad]
    const float* bias;       // [cout] or NULL
    int n, h, w, cin, cout, kpad, relu;
    // optional: input channels [0, split) are the nearest 2x upsample of x2
    // (n, h/2, w/2); channels [split, cin) come from x at the same index
    const uint16_t* x2;
    int x2s, x2o, split;
    // optional: 2x2 average pool of the (bf16) output into pool (n, h/2, w/2)
    uint16_t* pool;
    int ps, po;
    // optional (n = 1, 32 output channels, no x2 / pool): the 1x1 head fused
    // after the ReLU, out = b + W u + head_inp[0..3] with u the bf16 output,
    // W [4][32] then b [4] in head; written as RGBA16F at the pixels
    // (x, y) < (head_w, head_h), row stride head_w; y is then not written
    const float* head;
    const uint16_t* head_inp;  // 8 channels per pixel, h x w
    uint16_t* head_out;
    int head_w, head_h;
};
hipError_t launch_conv3x3(const ConvLaunch& a, int nt, hipStream_t s);  // nt = cout / 16
struct WGradLaunch {
    const uint16_t* x;
    int xs, xo;
    const uint16_t* dz;
    int dzs, dzo;
    float* partial;          // [chunks][9][cout][cin]
    float* bias_partial;     // [chunks][cout]
    int n, h, w, cin, cout;
    int chunk_pixels;        // multiple of 32
};
hipError_t launch_conv3x3_wgrad(const WGradLaunch& a, int chunks, float* gw, float* gb, hipStream_t s);
hipError_t launch_relu_mask(TensorView dy, TensorView y, long pixels, int c, hipStream_t s);
hipError_t launch_pool2(TensorView in, TensorView out, int n, int h, int w, int c, hipStream_t s);
hipError_t launch_pool2_bwd(TensorView dout, TensorView din, int n, int h, int w, int c, hipStream_t s);
hipError_t launch_up2(TensorView in, TensorView out, int n, int h, int w, int c, hipStream_t s);
hipError_t launch_up2_bwd(TensorView dout, TensorView din, int n, int h, int w, int c, hipStream_t s);
struct HeadLaunch {
    const uint16_t* u;       // [pixels][32]
    const uint16_t* inp;     // [pixels][8]
    const float* params;     // head.w (4 x 32) then head.b (4)
    long pixels;
    float* out;              // [pixels][4] fp32 or NULL
    uint16_t* image;         // [pixels][4] RGBA16F or NULL
    const float* target;     // [pixels][4] or NULL (inference)
    float inv_count;         // 1 / (pixels * 4)
    uint16_t* du;            // [pixels][32] bf16 gradient
    float* partial;          // head_partial_floats(pixels)
};
hipError_t launch_head(const HeadLaunch& a, float* loss, float* gw, hipStream_t s);
uint32_t head_partial_floats(long pixels);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint32_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s);
hipError_t launch_pack_weights(const float* w, int cin, int cout, int kpad, int kpad_t, uint16_t* wf, uint16_t* wb,
                               hipStream_t s);
struct PackLaunch {
    const float4* accum;     // width x height running mean
    const float* depth;      // width x height GPS depth
    const uint16_t* alpha;   // width x height RGBA16F sorted-alpha target (or NULL)
    int width, height;
    int x0, y0, hp, wp;      // region written (pixels beyond the frame are zero)
    uint16_t* inp;           // [hp][wp][8]
    float* target;           // [hp][wp][4] or NULL
};
hipError_t launch_pack_input(const PackLaunch& a, hipStream_t s);
// Squared-error sum and max abs difference of two half arrays, one float2
// (sum, max) per block.
hipError_t launch_image_error(const uint16_t* a, const uint16_t* b, uint32_t count, float2* partial, uint32_t blocks,
                              hipStream_t s);
hipError_t launch_gather(const uint16_t* pool_in, const float* pool_tgt, const uint32_t* idx, uint32_t pixels,
                         uint32_t batch, uint16_t* inp, float* tgt, hipStream_t s);

// gs_static.hip: device address of the static sort scratch on `device`
// (which must be current), looked up once per device by gs_create.
hipError_t static_scratch_init(int device);

// gs_api.cpp: the stream and device a context launches on (no other work).
int ctx_launch_target(gs_ctx* c, hipStream_t* stream, int* device);

// gs_api.cpp: what the denoiser reads from a render context.
struct CtxFrames {
    hipStream_t stream;
    int device;
    const float4* accum;     // running mean of GPS samples (NULL before gs_accumulate)
    const float* depth;      // last GPS depth
    int gps_w, gps_h;        // accumulation size (0 if none)
    const uint16_t* alpha;   // last sorted-alpha RGBA16F (NULL before a composite)
    int alpha_w, alpha_h;
    float clip_y_sign;       // +1: rows are bottom-first (GL)
};
int ctx_frames(gs_ctx* c, CtxFrames* out);
// Error reporting shared with the C ABI (gs_last_error).
int api_error(int code, const char* msg);
int api_ok();

}  // namespace gs
