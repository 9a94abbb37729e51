// gs_internal.h - device data layout, launch interfaces and the wave64
// primitives shared by the kernels of libgamesplat_hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/gs_api.h"

namespace gs {

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

// Device counters.  Each stage clears its own group with one fill launch
// before it runs, so a stage can be re-run without re-running earlier ones.
struct CullCounters {
    uint32_t visible_count;
    uint32_t ticket;
    uint32_t diag[8];
};
struct SortCounters {
    uint32_t ticket[8];
    uint32_t hist[8][256];
};
struct BinCounters {
    uint32_t ticket;
    uint32_t dup_total;
    uint32_t pad[6];
};
struct GpsCounters {
    uint32_t ticket;
    uint32_t requested;
    uint32_t overflow;
    uint32_t pad[5];
};
struct DevCounters {
    CullCounters cull;
    SortCounters depth_sort;
    BinCounters bin;
    SortCounters tile_sort;
    GpsCounters gps;
};

// Decoupled look-back status words: [63:62] flag, [61:32] epoch, [31:0] value.
// A word whose epoch differs from the launch's epoch reads as "not yet
// published", so status arrays never need clearing between launches.
constexpr uint64_t kFlagAggregate = 1ull;
constexpr uint64_t kFlagInclusive = 2ull;

__device__ __forceinline__ uint64_t make_status(uint64_t flag, uint32_t epoch, uint32_t value) {
    return (flag << 62) | ((uint64_t)(epoch & 0x3FFFFFFFu) << 32) | value;
}

__device__ __forceinline__ void store_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t load_status(uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of `aggregate` over all earlier tiles of a look-back chain
// with stride `stride` (status[tile * stride]).  Publishes this tile's
// aggregate first and its inclusive prefix last.  Values add modulo 2^32.
__device__ __forceinline__ uint32_t lookback(uint64_t* status, uint32_t tile, uint32_t stride,
                                             uint32_t epoch, uint32_t aggregate) {
    uint64_t* mine = status + (size_t)tile * stride;
    if (tile == 0) {
        store_status(mine, make_status(kFlagInclusive, epoch, aggregate));
        return 0u;
    }
    store_status(mine, make_status(kFlagAggregate, epoch, aggregate));
    uint32_t prefix = 0u;
    int64_t j = (int64_t)tile - 1;
    const uint32_t want = epoch & 0x3FFFFFFFu;
    while (j >= 0) {
        uint64_t s = load_status(status + (size_t)j * stride);
        uint32_t e = (uint32_t)(s >> 32) & 0x3FFFFFFFu;
        uint64_t flag = s >> 62;
        if (e != want || flag == 0) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        prefix += (uint32_t)s;
        if (flag == kFlagInclusive) break;
        --j;
    }
    store_status(mine, make_status(kFlagInclusive, epoch, prefix + aggregate));
    return prefix;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave64 prefix sum.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    uint32_t l = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (l >= (uint32_t)d) v += t;
    }
    return v;
}

// Exclusive prefix over a block of WAVES wave64s; returns the block total
// through *total.  `scratch` holds WAVES words of LDS.
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

// Map a linear block index so that each of the 8 XCDs of an MI355X works on
// one contiguous range of work items (blocks are dealt to XCDs round-robin).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    const uint32_t xcds = 8u;
    uint32_t per = nblocks / xcds, rem = nblocks % xcds;
    uint32_t xcd = bid % xcds, local = bid / xcds;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// ---------------------------------------------------------------------------
// Kernel launch interfaces (one .hip file each).

struct CullLaunch {
    const float4* planes;
    uint32_t n;
    GsCameraUniforms cam;
    GsSplatEffectParams effect;
    int effect_active;
    int sh_degree;
    float sh_lod_distance;
    int collect_stats;
    float4* records;
    uint32_t* vis_keys;
    uint32_t* vis_ids;
    DevCounters* counters;
    uint64_t* status;
    uint32_t epoch;
};
hipError_t launch_cull(const CullLaunch& a, hipStream_t s);
hipError_t launch_clear_counters(DevCounters* c, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, size_t count, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t value, size_t count, hipStream_t s);
hipError_t launch_transpose_upload(const float* aos, uint32_t count, uint32_t first, uint32_t n,
                                   float4* planes, hipStream_t s);

struct SortLaunch {
    uint32_t* keys[3];       // keys[0] input; passes ping-pong keys[1] <-> keys[2]
    uint32_t* vals[3];
    const uint32_t* count;   // device element count
    uint32_t capacity;       // grid bound
    uint32_t* hist;          // [passes][256]
    uint32_t* tickets;       // [passes]
    uint64_t* status;        // [tiles * 256]
    uint32_t first_epoch;
    int passes;
    int bits;                // digit width (8 or 4)
};
// Histograms of all digit passes.  hist must be zeroed.
hipError_t launch_radix_histogram(const uint32_t* keys, const uint32_t* count, uint32_t capacity,
                                  uint32_t* hist, int passes, int bits, hipStream_t s);
// Stable LSD onesweep.  Returns the index (1 or 2) of the buffer holding the
// result through *result_slot.
hipError_t launch_onesweep(const SortLaunch& a, int* result_slot, hipStream_t s);
uint32_t onesweep_tile_elems();

struct BinLaunch {
    const float4* records;
    const uint32_t* sorted_ids;
    const uint32_t* count;   // visible count (device)
    uint32_t capacity;
    int width, height, tile, tiles_x, tiles_y;
    uint32_t* offsets;       // per front-to-back position
    DevCounters* counters;
    uint64_t* status;
    uint32_t epoch;
    uint32_t* dup_keys;
    uint32_t* dup_vals;
};
hipError_t launch_bin_count(const BinLaunch& a, hipStream_t s);
hipError_t launch_bin_emit(const BinLaunch& a, hipStream_t s);
hipError_t launch_tile_ranges(const uint32_t* keys, uint32_t count, uint2* ranges, hipStream_t s);

struct CompositeLaunch {
    const float4* records;
    const uint32_t* tile_vals;
    const uint2* ranges;
    int width, height, tile, tiles_x, tiles_y;
    uint16_t* image;         // RGBA16F, framebuffer row order
    int quadrant_cull;       // 16x16 tiles: per-wave 8x8 quadrant culling kernel
};
hipError_t launch_composite(const CompositeLaunch& a, hipStream_t s);

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
    uint32_t* item_base;         // per visible position
    DevCounters* counters;
    uint64_t* status;
    uint32_t epoch;
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

}  // namespace gs
