// gs_sort.hip - stable LSD radix sort of 32-bit keys with 32-bit values for
// gfx950.
//
// Replaces radix_hist_cs / radix_prefix_cs / radix_scatter_cs
// (shaders/sort.glsl:4-164) and their 12 dispatches
// (src/renderer.cpp:1873-1962).  Differences from the reference:
//   - one up-front kernel histograms every digit of every pass at once;
//   - tiles are 3840 keys (256 lanes x 15), not 256, so per-tile tables are
//     15x smaller;
//   - the per-pass prefix is a parallel column scan (one block per digit)
//     instead of one workgroup looping over all groups twice;
//   - ranking inside a tile uses wave64 ballots (peer matching over the digit
//     bits), and keys are reordered by digit in LDS before the scatter, so
//     equal digits leave the CU as contiguous runs.
// Each pass is upsweep (tile digit counts) -> column scan -> downsweep.  No
// kernel waits on another block or takes a global ticket (device-scope
// atomics on one address serialize across XCDs on MI355X).  The element
// count lives in device memory; blocks past it exit at once.
#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kItems = 15;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kTile = kThreads * kItems;

template <int BITS>
__global__ void __launch_bounds__(kThreads) radix_histogram_kernel(const uint32_t* __restrict__ keys,
                                                                   const uint32_t* count_ptr,
                                                                   uint32_t* hist, int passes) {
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t s_hist[8 * RADIX];
    const uint32_t count = *count_ptr;
    for (int i = threadIdx.x; i < passes * RADIX; i += kThreads) s_hist[i] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < count; i += stride) {
        uint32_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&s_hist[p * RADIX + ((k >> (p * BITS)) & (RADIX - 1))], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * RADIX; i += kThreads) {
        uint32_t v = s_hist[i];
        if (v) atomicAdd(&hist[(i / RADIX) * 256 + (i % RADIX)], v);
    }
}

struct PassArgs {
    const uint32_t* keys_in;
    const uint32_t* vals_in;
    uint32_t* keys_out;
    uint32_t* vals_out;
    const uint32_t* count;
    const uint32_t* hist;   // this pass, RADIX bins
    uint32_t* table;        // [RADIX][tiles]: tile digit counts, then global offsets
    uint32_t tiles;         // table stride (grid size)
    int shift;
};

// Digit counts of one tile (same tile layout as the downsweep).
template <int BITS>
__global__ void __launch_bounds__(kThreads) upsweep_kernel(PassArgs a) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t MASK = RADIX - 1;
    __shared__ uint32_t s_hist[kWaves][RADIX];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t count = *a.count;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)kTile;
    if (base >= count) return;
    for (uint32_t i = tid; i < (uint32_t)(kWaves * RADIX); i += kThreads) (&s_hist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t wbase = base + wave * (64u * kItems);
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        if (g < count) atomicAdd(&s_hist[wave][(a.keys_in[g] >> a.shift) & MASK], 1u);
    }
    __syncthreads();
    if (tid < (uint32_t)RADIX) {
        uint32_t c = 0u;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) c += s_hist[w][tid];
        a.table[(size_t)tid * a.tiles + tile] = c;
    }
}

// One block per digit: exclusive scan of the digit's column over the tiles
// in use, offset by the number of keys with a smaller digit.
template <int BITS>
__global__ void __launch_bounds__(kThreads) colscan_kernel(PassArgs a) {
    constexpr int RADIX = 1 << BITS;
    __shared__ uint32_t s_scan[kWaves];
    const uint32_t tid = threadIdx.x, d = blockIdx.x;
    const uint32_t count = *a.count;
    const uint32_t used = (count + kTile - 1) / kTile;
    static_assert(RADIX <= kThreads, "one thread per digit");
    uint32_t below;  // keys with a smaller digit
    (void)block_exclusive_sum<kWaves>(tid < d ? a.hist[tid] : 0u, s_scan, &below);
    const uint32_t per = (used + kThreads - 1) / kThreads;
    const uint32_t t0 = min(used, tid * per), t1 = min(used, t0 + per);
    uint32_t* col = a.table + (size_t)d * a.tiles;
    uint32_t sum = 0u;
    for (uint32_t t = t0; t < t1; ++t) sum += col[t];
    uint32_t all;
    uint32_t run = below + block_exclusive_sum<kWaves>(sum, s_scan, &all);
    for (uint32_t t = t0; t < t1; ++t) {
        uint32_t c = col[t];
        col[t] = run;
        run += c;
    }
}

template <int BITS>
__global__ void __launch_bounds__(kThreads) downsweep_kernel(PassArgs a) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t MASK = RADIX - 1;
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_whist[kWaves][RADIX];
    __shared__ uint32_t s_local[RADIX];
    __shared__ uint32_t s_dest[RADIX];
    __shared__ uint32_t s_scan[kWaves];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t count = *a.count;
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * (uint32_t)kTile;
    if (base >= count) return;
    for (uint32_t i = tid; i < (uint32_t)(kWaves * RADIX); i += kThreads) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t valid = min((uint32_t)kTile, count - base);
    const int shift = a.shift;

    // 1. load, wave-contiguous (each load is a coalesced 256-byte access)
    const uint32_t wbase = base + wave * (64u * kItems);
    uint32_t keys[kItems], vals[kItems], ranks[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        bool in = g < count;
        keys[i] = in ? a.keys_in[g] : 0xFFFFFFFFu;
        vals[i] = in ? a.vals_in[g] : 0u;
    }

    // 2. stable rank within the wave: peers = lanes with the same digit,
    //    found with BITS ballots; the lowest peer updates the wave's counter
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        uint32_t d = g < count ? (keys[i] >> shift) & MASK : MASK;
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t below = (uint32_t)__popcll(peers & lt);
        uint32_t leader = (uint32_t)__ffsll((unsigned long long)peers) - 1u;
        uint32_t old = 0u;
        if (lane == leader) {
            old = s_whist[wave][d];
            s_whist[wave][d] = old + (uint32_t)__popcll(peers);
        }
        old = __shfl(old, (int)leader, 64);
        ranks[i] = old + below;
    }
    __syncthreads();

    // 3. per-digit offsets: within the tile (over waves, then digits) and global
    uint32_t bin_total = 0u;
    if (tid < (uint32_t)RADIX) {
        uint32_t sum = 0u;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            uint32_t c = s_whist[w][tid];
            s_whist[w][tid] = sum;
            sum += c;
        }
        bin_total = sum;
    }
    uint32_t total;
    uint32_t local = block_exclusive_sum<kWaves>(bin_total, s_scan, &total);
    if (tid < (uint32_t)RADIX) {
        s_local[tid] = local;
        s_dest[tid] = a.table[(size_t)tid * a.tiles + tile] - local;
    }
    __syncthreads();

    // 4. reorder in LDS by digit, then write runs out
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        uint32_t d = g < count ? (keys[i] >> shift) & MASK : MASK;
        uint32_t p = s_local[d] + s_whist[wave][d] + ranks[i];
        s_keys[p] = keys[i];
        s_vals[p] = vals[i];
    }
    __syncthreads();
    // invalid lanes carry the last digit and rank after every valid key, so
    // the valid keys occupy local positions [0, valid)
    for (uint32_t p = tid; p < valid; p += kThreads) {
        uint32_t k = s_keys[p];
        uint32_t d = (k >> shift) & MASK;
        uint32_t out = s_dest[d] + p;
        a.keys_out[out] = k;
        a.vals_out[out] = s_vals[p];
    }
}

template <int BITS>
hipError_t run_passes(const SortLaunch& a, int* result_slot, hipStream_t s) {
    uint32_t tiles = (a.capacity + kTile - 1) / kTile;
    if (tiles < 1) tiles = 1;
    int src = 0, dst = 1;
    for (int p = 0; p < a.passes; ++p) {
        PassArgs pa;
        pa.keys_in = a.keys[src];
        pa.vals_in = a.vals[src];
        pa.keys_out = a.keys[dst];
        pa.vals_out = a.vals[dst];
        pa.count = a.count;
        pa.hist = a.hist + p * 256;
        pa.table = a.table;
        pa.tiles = tiles;
        pa.shift = p * BITS;
        hipLaunchKernelGGL(upsweep_kernel<BITS>, dim3(tiles), dim3(kThreads), 0, s, pa);
        hipLaunchKernelGGL(colscan_kernel<BITS>, dim3(1 << BITS), dim3(kThreads), 0, s, pa);
        hipLaunchKernelGGL(downsweep_kernel<BITS>, dim3(tiles), dim3(kThreads), 0, s, pa);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src = dst;
        dst = dst == 1 ? 2 : 1;
    }
    *result_slot = src;
    return hipSuccess;
}

}  // namespace

uint32_t sort_table_words(uint32_t capacity) {
    uint32_t tiles = (capacity + kTile - 1) / kTile;
    return (tiles < 1 ? 1 : tiles) * 256u;
}

hipError_t launch_radix_histogram(const uint32_t* keys, const uint32_t* count, uint32_t capacity,
                                  uint32_t* hist, int passes, int bits, hipStream_t s) {
    if (passes < 1 || passes > 8) return hipErrorInvalidValue;
    uint32_t blocks = (capacity + 16383u) / 16384u;
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    if (bits == 8) {
        if (passes > 4) return hipErrorInvalidValue;
        hipLaunchKernelGGL(radix_histogram_kernel<8>, dim3(blocks), dim3(kThreads), 0, s, keys, count, hist, passes);
    } else if (bits == 4) {
        hipLaunchKernelGGL(radix_histogram_kernel<4>, dim3(blocks), dim3(kThreads), 0, s, keys, count, hist, passes);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_radix_sort(const SortLaunch& a, int* result_slot, hipStream_t s) {
    if (a.bits == 8) return run_passes<8>(a, result_slot, s);
    if (a.bits == 4) return run_passes<4>(a, result_slot, s);
    return hipErrorInvalidValue;
}

}  // namespace gs
