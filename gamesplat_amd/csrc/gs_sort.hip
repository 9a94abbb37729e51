// gs_sort.hip - stable LSD radix sort of 32-bit keys with 32-bit values for
// gfx950.
//
// Replaces radix_hist_cs / radix_prefix_cs / radix_scatter_cs
// (shaders/sort.glsl:4-164) and their 12 dispatches
// (src/renderer.cpp:1873-1962).  Each pass is
//   upsweep    digit counts of every 3840-key tile (256 lanes x 15 keys);
//   colscan    one block per digit: exclusive scan of that digit's column over
//              the tiles, plus the column total (the reference's single
//              workgroup loops over every group twice instead);
//   downsweep  digit bases from the column totals (every block scans them),
//              stable ranking inside the tile with wave64 ballots (peers =
//              lanes whose digit matches on every bit), reorder by digit in
//              LDS, then contiguous runs of equal digits are written out:
//              keys, values (and second values) in turn through one staging
//              array, which keeps the block's LDS small enough for seven
//              blocks per CU.
// An optional second 32-bit value array moves with the same permutation.
// Keys can also be 16-bit (K16: the tile ids of frames with at most 65536
// tiles), held as 32-bit values in registers and LDS.
// The first pass can also read the cull's dense [view][n] keys directly
// (DENSE): slots holding the culled key 0xFFFFFFFF take no digit count and no
// rank, so that pass also does the order-preserving compaction of the visible
// splats; their values are the slots themselves and their second values the
// splats' packed tile rectangles.
// No kernel waits on another block or takes a global ticket (device-scope
// atomics on one address serialize across XCDs on MI355X).  The element
// count lives in device memory; blocks past the tiles it needs exit at once,
// and the others take their tiles XCD by XCD (xcd_remap): the digit runs of
// neighbouring tiles, which end and start in the same cache lines of the
// output, are then written through the same L2.
#include "gs_internal.h"
#include "gs_raster.h"

namespace gs {

namespace {

constexpr int kItems = 15;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kTile = kThreads * kItems;

struct PassArgs {
    const uint32_t* keys_in;
    const uint32_t* vals_in;
    const uint32_t* vals2_in;   // second value array (P2 passes)
    uint32_t* keys_out;
    uint32_t* vals_out;
    uint32_t* vals2_out;
    uint32_t* keys_last;        // output of the last pass the key range needs (adaptive)
    uint32_t* vals_last;
    uint32_t* vals2_last;
    const uint32_t* count;
    const uint32_t* key_range;  // device (min, max) of the keys (read when adaptive)
    uint32_t* table;            // [RADIX][tiles]: tile digit counts, then column offsets
    uint32_t tiles;             // table stride (grid size)
    uint32_t* col_total;        // [RADIX]
    int adaptive;
    int skip_last_keys;         // the last pass the range needs writes no keys
    int pass, passes;           // this pass, passes launched
    int shift;
    // dense first pass: keys_in holds dense_count slots (0xFFFFFFFF: culled);
    // the second value is rect_packed[slot] or rect_px[slot] packed at tile
    // size 1 << rect_shift (neither: no second value)
    int dense;
    uint32_t dense_count;
    const uint2* rect_px;
    const uint32_t* rect_packed;
    int rect_shift;
};

// Whether pass `pass` runs and what it adds to or subtracts from the keys.
// Fixed passes all run and the buffers alternate.  Adaptive passes sort
// key - min: the first subtracts min as it loads the keys, the last one the
// range needs adds it back and writes the *_last buffers, the later ones exit.
struct PassPlan {
    bool active, last;
    uint32_t sub, add;
};

// Key i of an array of 32-bit or (K16) 16-bit keys.
template <bool O32, bool K16>
__device__ __forceinline__ uint32_t load_key(const uint32_t* keys, uint32_t i) {
    if constexpr (K16) return at<O32>(reinterpret_cast<const uint16_t*>(keys), i);
    else return at<O32>(keys, i);
}

template <bool O32, bool K16>
__device__ __forceinline__ void store_key(uint32_t* keys, uint32_t i, uint32_t k) {
    if constexpr (K16) at<O32>(reinterpret_cast<uint16_t*>(keys), i) = (uint16_t)k;
    else at<O32>(keys, i) = k;
}

template <int BITS>
__device__ __forceinline__ PassPlan pass_plan(const PassArgs& a, uint2 kr) {
    int needed = a.passes;
    if (a.adaptive) {
        const uint32_t range = kr.y > kr.x ? kr.y - kr.x : 0u;
        const int sig = range ? 32 - __clz((int)range) : 0;
        needed = min(a.passes, max(1, (sig + BITS - 1) / BITS));
    }
    PassPlan p;
    p.active = a.pass < needed;
    p.last = a.adaptive && a.pass == needed - 1;
    p.sub = a.adaptive && a.pass == 0 ? kr.x : 0u;
    p.add = p.last ? kr.x : 0u;
    return p;
}

// The element count (the slot count of a dense pass) and the key range,
// loaded together (one memory wait).
__device__ __forceinline__ uint32_t load_count_range(const PassArgs& a, uint2* kr) {
    const uint32_t count = a.dense ? a.dense_count : *a.count;
    *kr = *reinterpret_cast<const uint2*>(a.key_range);
    return count;
}

// Digit counts of one tile.  Each wave counts into kCopies histograms, lane
// l into copy l % kCopies, so a digit shared by many lanes (the top digit of
// clustered keys) serializes kCopies times fewer LDS atomics.
template <int BITS, bool O32, bool DENSE, bool K16>
__global__ void __launch_bounds__(kThreads) upsweep_kernel(PassArgs a) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t MASK = RADIX - 1;
    constexpr int kCopies = RADIX > 256 ? 2 : 4;
    __shared__ uint32_t s_hist[kWaves * kCopies][RADIX];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint2 kr;
    const uint32_t count = load_count_range(a, &kr);
    const uint32_t used = (count + (uint32_t)kTile - 1u) / (uint32_t)kTile;
    if (blockIdx.x >= used) return;
    const uint32_t tile = xcd_remap(blockIdx.x, used);
    const uint32_t base = tile * (uint32_t)kTile;
    const PassPlan plan = pass_plan<BITS>(a, kr);
    if (!plan.active) return;
    for (uint32_t i = tid; i < (uint32_t)(kWaves * kCopies * RADIX); i += kThreads) (&s_hist[0][0])[i] = 0u;
    __syncthreads();
    uint32_t* hist = s_hist[wave * kCopies + lane % kCopies];
    const uint32_t wbase = base + wave * (64u * kItems);
    const uint32_t* keys_in = a.keys_in;
    uint32_t keys[kItems];
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        keys[i] = g < count ? load_key<O32, K16>(keys_in, g) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        if (DENSE ? keys[i] != 0xFFFFFFFFu : g < count)
            atomicAdd(&hist[((keys[i] - plan.sub) >> a.shift) & MASK], 1u);
    }
    __syncthreads();
    for (uint32_t d = tid; d < (uint32_t)RADIX; d += kThreads) {
        uint32_t c = 0u;
#pragma unroll
        for (int w = 0; w < kWaves * kCopies; ++w) c += s_hist[w][d];
        a.table[(size_t)d * a.tiles + tile] = c;
    }
}

// One block per digit: exclusive scan of the digit's column over the tiles in
// use, column total to col_total[digit].  A thread's run of 8 counts is two
// 16-byte accesses (the column stride is a multiple of 4 words, and a group
// of 4 that starts below the tiles in use lies inside the column).
template <int BITS>
__global__ void __launch_bounds__(kThreads) colscan_kernel(PassArgs a) {
    constexpr int kRun = 8;
    __shared__ uint32_t s_scan[kWaves];
    const uint32_t tid = threadIdx.x, d = blockIdx.x;
    uint2 kr;
    const uint32_t count = load_count_range(a, &kr);
    if (!pass_plan<BITS>(a, kr).active) return;
    const uint32_t used = (count + kTile - 1) / kTile;
    uint32_t* col = a.table + (size_t)d * a.tiles;
    uint32_t carry = 0u;
    for (uint32_t first = 0; first < used; first += kThreads * kRun) {
        const uint32_t t0 = first + tid * kRun;
        uint32_t v[kRun];
        uint32_t sum = 0u;
#pragma unroll
        for (int k = 0; k < kRun; k += 4) {
            uint4 q = make_uint4(0u, 0u, 0u, 0u);
            if (t0 + k < used) q = *reinterpret_cast<const uint4*>(col + t0 + k);
            v[k] = t0 + k < used ? q.x : 0u;
            v[k + 1] = t0 + k + 1 < used ? q.y : 0u;
            v[k + 2] = t0 + k + 2 < used ? q.z : 0u;
            v[k + 3] = t0 + k + 3 < used ? q.w : 0u;
            sum += v[k] + v[k + 1] + v[k + 2] + v[k + 3];
        }
        uint32_t chunk;
        uint32_t run = carry + block_exclusive_sum<kWaves>(sum, s_scan, &chunk);
#pragma unroll
        for (int k = 0; k < kRun; k += 4) {
            uint4 o;
            o.x = run; run += v[k];
            o.y = run; run += v[k + 1];
            o.z = run; run += v[k + 2];
            o.w = run; run += v[k + 3];
            if (t0 + k < used) *reinterpret_cast<uint4*>(col + t0 + k) = o;
        }
        carry += chunk;
    }
    if (tid == 0) a.col_total[d] = carry;
}

template <int BITS, bool P2, bool O32, bool DENSE, bool K16>
__global__ void __launch_bounds__(kThreads) downsweep_kernel(PassArgs a) {
    constexpr int RADIX = 1 << BITS;
    constexpr uint32_t MASK = RADIX - 1;
    __shared__ uint32_t s_buf[kTile];   // keys, then values, then second values
    __shared__ uint32_t s_whist[kWaves][RADIX];
    __shared__ uint32_t s_local[RADIX];
    __shared__ uint32_t s_dest[RADIX];
    __shared__ uint32_t s_scan[kWaves];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint2 kr;
    const uint32_t count = load_count_range(a, &kr);
    const uint32_t used = (count + (uint32_t)kTile - 1u) / (uint32_t)kTile;
    if (blockIdx.x >= used) return;
    const uint32_t tile = xcd_remap(blockIdx.x, used);
    const uint32_t base = tile * (uint32_t)kTile;
    const PassPlan plan = pass_plan<BITS>(a, kr);
    if (!plan.active) return;
    const uint32_t* keys_in = a.keys_in;
    const uint32_t* vals_in = a.vals_in;
    const uint32_t* vals2_in = a.vals2_in;
    uint32_t* keys_out = plan.last ? a.keys_last : a.keys_out;
    uint32_t* vals_out = plan.last ? a.vals_last : a.vals_out;
    uint32_t* vals2_out = plan.last ? a.vals2_last : a.vals2_out;
    for (uint32_t i = tid; i < (uint32_t)(kWaves * RADIX); i += kThreads) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t valid = min((uint32_t)kTile, count - base);
    const int shift = a.shift;

    // 1. load, wave-contiguous (each load is a coalesced 256-byte access).
    //    Positions past the count read the last element (no branch) and take
    //    the key 0xFFFFFFFF: the last digit, ranked after every valid key.
    const uint32_t wbase = base + wave * (64u * kItems);
    const uint32_t last = count - 1u;
    // ranks (< kTile) two to a register and output digits four (BITS > 8:
    // two) to a register: fewer registers hold more waves in flight
    constexpr int kDigitBits = BITS <= 8 ? 8 : 16, kPerReg = 32 / kDigitBits;
    constexpr uint32_t kDigitMask = (1u << kDigitBits) - 1u;
    constexpr int kPairs = (kItems + 1) / 2, kDigitRegs = (kItems + kPerReg - 1) / kPerReg;
    static_assert(kTile <= 65536 && BITS <= 16, "packed ranks and digits");
    uint32_t keys[kItems], ranks[kPairs];
    uint32_t kept = 0u;  // DENSE: bit i set when item i holds a visible splat
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        uint32_t g = wbase + (uint32_t)i * 64u + lane;
        const uint32_t k = load_key<O32, K16>(keys_in, min(g, last));
        if constexpr (DENSE) {
            const bool keep = g < count && k != 0xFFFFFFFFu;
            kept |= (uint32_t)keep << i;
            keys[i] = keep ? k - plan.sub : 0xFFFFFFFFu;
        } else {
            keys[i] = g < count ? k - plan.sub : 0xFFFFFFFFu;
        }
    }

    // 2. stable rank within the wave: peers = lanes with the same digit,
    //    found with BITS ballots (a lane keeps the ballot where its bit is
    //    set, the complement where it is clear); the peers below a lane give
    //    its rank.  Every lane reads the wave's digit counter before the
    //    lowest peer (no peers below it) advances it: LDS accesses of one
    //    wave complete in program order, and no other wave uses the counter.
    //    In a dense pass only kept lanes are peers, and only they count.
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint32_t d = (keys[i] >> shift) & MASK;
        uint32_t plo = ~0u, phi = ~0u;
        if constexpr (DENSE) {
            const uint64_t kb = __ballot((kept >> i) & 1u);
            plo = (uint32_t)kb;
            phi = (uint32_t)(kb >> 32);
        }
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const uint32_t t = (uint32_t)((int32_t)(d << (31 - b)) >> 31);  // 0 or ~0: bit b of d
            const uint64_t bal = __ballot(t != 0u);
            plo &= ~((uint32_t)bal ^ t);
            phi &= ~((uint32_t)(bal >> 32) ^ t);
        }
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));
        const uint32_t old = s_whist[wave][d];
        if (below == 0u && (!DENSE || ((kept >> i) & 1u))) s_whist[wave][d] = old + (uint32_t)(__popc(plo) + __popc(phi));
        const uint32_t rank = old + below;
        ranks[i >> 1] = (i & 1) ? ranks[i >> 1] | (rank << 16) : rank;
    }
    __syncthreads();
    // the digits and their counter addresses are recomputed after the
    // barrier rather than held in registers across it (an empty asm
    // statement that the compiler must assume changes the keys)
#pragma unroll
    for (int i = 0; i < kItems; ++i) asm volatile("" : "+v"(keys[i]));
    if constexpr (DENSE) asm volatile("" : "+v"(kept));
    // likewise the lane's first position: the later loads and slots form
    // their positions from it instead of holding those of the key loads
    uint32_t lane_slot = wbase + lane;
    asm volatile("" : "+v"(lane_slot));

    // 3. per-digit offsets: within the tile (over waves, then digits) and
    //    global.  Thread t holds digits [t * kDPT, (t + 1) * kDPT).
    constexpr int kDPT = RADIX > kThreads ? RADIX / kThreads : 1;
    const bool has_digits = tid * kDPT < (uint32_t)RADIX;
    uint32_t dsum[kDPT], ctot[kDPT];
    uint32_t tsum = 0u, csum = 0u;
#pragma unroll
    for (int j = 0; j < kDPT; ++j) {
        const uint32_t d = tid * kDPT + (uint32_t)j;
        uint32_t sum = 0u;
        ctot[j] = 0u;
        if (has_digits) {
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                uint32_t c = s_whist[w][d];
                s_whist[w][d] = sum;
                sum += c;
            }
            // digit bases: every block scans the column totals itself (no
            // separate single-block launch between colscan and downsweep)
            ctot[j] = a.col_total[d];
        }
        dsum[j] = sum;
        tsum += sum;
        csum += ctot[j];
    }
    uint32_t total;
    uint32_t local = block_exclusive_sum<kWaves>(tsum, s_scan, &total);
    // a dense tile's kept splats, ranked first: they fill local positions
    // [0, total)
    const uint32_t out_count = DENSE ? total : valid;
    const bool write_keys = !(plan.last && a.skip_last_keys);
    uint32_t all;
    uint32_t dbase = block_exclusive_sum<kWaves>(csum, s_scan, &all);
    if (has_digits) {
#pragma unroll
        for (int j = 0; j < kDPT; ++j) {
            const uint32_t d = tid * kDPT + (uint32_t)j;
            s_local[d] = local;
            s_dest[d] = dbase + a.table[(size_t)d * a.tiles + tile] - local;
            local += dsum[j];
            dbase += ctot[j];
        }
    }
    __syncthreads();

    // 4. reorder in LDS by digit, then write runs out: the keys first (which
    //    also gives every local position its output position), then the
    //    values, then the second values, through one staging array.  Each
    //    value array is loaded while the previous round is written out, so
    //    only one array's items are held in registers at a time.
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const uint32_t d = (keys[i] >> shift) & MASK;
        const uint32_t rank = ((ranks[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu) + s_local[d] + s_whist[wave][d];
        ranks[i >> 1] = (i & 1) ? (ranks[i >> 1] & 0xFFFFu) | (rank << 16) : (ranks[i >> 1] & 0xFFFF0000u) | rank;
        // a culled slot of a dense pass has no rank (and its position may lie
        // past the tile)
        if (!DENSE || ((kept >> i) & 1u)) s_buf[rank] = keys[i];
    }
    // values of positions past the count are staged but never written out;
    // a dense pass's value is the slot (formed when it is staged)
    uint32_t v[kItems];
    if constexpr (!DENSE) {
#pragma unroll
        for (int i = 0; i < kItems; ++i) v[i] = at<O32>(vals_in, min(lane_slot + (uint32_t)i * 64u, last));
    }
    __syncthreads();
    // invalid lanes carry the last digit and rank after every valid key, so
    // the valid keys occupy local positions [0, valid)
    // the output position of local position p is s_dest[digit] + p
    uint32_t digits[kDigitRegs] = {};
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint32_t p = tid + (uint32_t)j * kThreads;
        if (p < out_count) {
            uint32_t k = s_buf[p];
            const uint32_t d = (k >> shift) & MASK;
            digits[j / kPerReg] |= d << ((j % kPerReg) * kDigitBits);
            if (write_keys) store_key<O32, K16>(keys_out, s_dest[d] + p, k + plan.add);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        if constexpr (DENSE) {
            if ((kept >> i) & 1u) s_buf[(ranks[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu] = lane_slot + (uint32_t)i * 64u;
        } else {
            s_buf[(ranks[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu] = v[i];
        }
    }
    if constexpr (P2) {
        // a dense pass packs each slot's rectangle (all loads issued before
        // any is used; culled slots load too and are never staged)
        if (DENSE && a.rect_packed) {
#pragma unroll
            for (int i = 0; i < kItems; ++i) v[i] = at<O32>(a.rect_packed, min(lane_slot + (uint32_t)i * 64u, last));
        } else if (DENSE) {
            // every rectangle is loaded whole before any is packed (the
            // packing tests the first word, and a branch on it would load
            // the second word only after the first arrived)
            uint2 px[kItems];
#pragma unroll
            for (int i = 0; i < kItems; ++i) px[i] = a.rect_px[min(lane_slot + (uint32_t)i * 64u, last)];
#pragma unroll
            for (int i = 0; i < kItems; ++i) asm volatile("" : "+v"(px[i].x), "+v"(px[i].y));
#pragma unroll
            for (int i = 0; i < kItems; ++i) v[i] = pack_tile_rect(px[i], a.rect_shift);
        } else {
#pragma unroll
            for (int i = 0; i < kItems; ++i) v[i] = at<O32>(vals2_in, min(lane_slot + (uint32_t)i * 64u, last));
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const uint32_t p = tid + (uint32_t)j * kThreads;
        if (p < out_count)
            at<O32>(vals_out, s_dest[(digits[j / kPerReg] >> ((j % kPerReg) * kDigitBits)) & kDigitMask] + p) = s_buf[p];
    }
    if constexpr (P2) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kItems; ++i)
            if (!DENSE || ((kept >> i) & 1u)) s_buf[(ranks[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu] = v[i];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            const uint32_t p = tid + (uint32_t)j * kThreads;
            if (p < out_count)
                at<O32>(vals2_out, s_dest[(digits[j / kPerReg] >> ((j % kPerReg) * kDigitBits)) & kDigitMask] + p) =
                    s_buf[p];
        }
    }
}

template <int BITS, bool O32, bool DENSE, bool K16 = false>
void launch_pass(const PassArgs& pa, bool p2, uint32_t tiles, hipStream_t s) {
    hipLaunchKernelGGL((upsweep_kernel<BITS, O32, DENSE, K16>), dim3(tiles), dim3(kThreads), 0, s, pa);
    hipLaunchKernelGGL(colscan_kernel<BITS>, dim3(1 << BITS), dim3(kThreads), 0, s, pa);
    if (p2)
        hipLaunchKernelGGL((downsweep_kernel<BITS, true, O32, DENSE, K16>), dim3(tiles), dim3(kThreads), 0, s, pa);
    else
        hipLaunchKernelGGL((downsweep_kernel<BITS, false, O32, DENSE, K16>), dim3(tiles), dim3(kThreads), 0, s, pa);
}

// DENSE_OK: the digit widths a dense first pass is built for (the depth
// sort's)
// Tiles of a sort over `capacity` elements, rounded up to a multiple of 4:
// the digit table's column stride.
uint32_t table_stride(uint32_t capacity) {
    uint32_t tiles = (capacity + kTile - 1) / kTile;
    if (tiles < 1) tiles = 1;
    return (tiles + 3u) & ~3u;
}

template <int BITS, bool DENSE_OK>
hipError_t run_passes(const SortLaunch& a, int* result_slot, hipStream_t s) {
    const uint32_t tiles = table_stride(a.capacity);
    PassArgs pa{};
    pa.count = a.count;
    // a readable placeholder when the passes are fixed (not used for them)
    pa.key_range = a.key_range ? a.key_range : a.scratch;
    pa.adaptive = a.key_range ? 1 : 0;
    pa.skip_last_keys = a.key_range && a.skip_last_keys ? 1 : 0;
    pa.keys_last = a.keys[3];
    pa.vals_last = a.vals[3];
    pa.vals2_last = a.vals2[3];
    pa.passes = a.passes;
    pa.table = a.table;
    pa.tiles = tiles;
    pa.col_total = a.scratch;
    // every array holds at most capacity 4-byte elements
    const bool o32 = (uint64_t)a.capacity * 4u <= 0xFFFFFFFFull && !a.wide_index;
    const bool p2 = a.vals2[0] != nullptr;
    int src = 0, dst = 1;
    for (int p = 0; p < a.passes; ++p) {
        const bool dense = p == 0 && a.dense_keys;
        pa.keys_in = dense ? a.dense_keys : a.keys[src];
        pa.vals_in = a.vals[src];
        pa.vals2_in = a.vals2[src];
        pa.keys_out = a.keys[dst];
        pa.vals_out = a.vals[dst];
        pa.vals2_out = a.vals2[dst];
        pa.pass = p;
        pa.shift = p * BITS;
        pa.dense = dense ? 1 : 0;
        pa.dense_count = dense ? a.dense_count : 0u;
        pa.rect_px = dense ? a.dense_rect_px : nullptr;
        pa.rect_packed = dense ? a.dense_rect_packed : nullptr;
        pa.rect_shift = dense ? a.dense_rect_shift : 0;
        if (a.key16 && o32)
            launch_pass<BITS, true, false, true>(pa, p2, tiles, s);
        else if (a.key16)
            launch_pass<BITS, false, false, true>(pa, p2, tiles, s);
        else if constexpr (DENSE_OK) {
            if (dense && o32)
                launch_pass<BITS, true, true>(pa, p2, tiles, s);
            else if (dense)
                launch_pass<BITS, false, true>(pa, p2, tiles, s);
            else if (o32)
                launch_pass<BITS, true, false>(pa, p2, tiles, s);
            else
                launch_pass<BITS, false, false>(pa, p2, tiles, s);
        } else {
            if (dense) return hipErrorInvalidValue;
            if (o32)
                launch_pass<BITS, true, false>(pa, p2, tiles, s);
            else
                launch_pass<BITS, false, false>(pa, p2, tiles, s);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src = dst;
        dst = dst == 1 ? 2 : 1;
    }
    *result_slot = a.key_range ? 3 : src;
    return hipSuccess;
}

}  // namespace

uint32_t sort_table_words(uint32_t capacity) {
    return table_stride(capacity) * 512u;  // up to 9-bit digits
}

hipError_t launch_radix_sort(const SortLaunch& a, int* result_slot, hipStream_t s) {
    // the last pass may reach past bit 31 (its top digits are then zero)
    if (a.passes < 1 || (a.passes - 1) * a.bits >= 32) return hipErrorInvalidValue;
    if (a.vals2[0] && (!a.vals2[1] || !a.vals2[2])) return hipErrorInvalidValue;
    if (a.key_range && (!a.keys[3] || !a.vals[3] || (a.vals2[0] && !a.vals2[3]))) return hipErrorInvalidValue;
    // 16-bit keys: no dense pass, no key range, and every key below 2^16
    if (a.key16 && (a.dense_keys || a.key_range || a.passes * a.bits > 32)) return hipErrorInvalidValue;
    if (a.dense_keys) {
        // the dense slots fit the grid; their second value is a packed
        // rectangle, from pixel rectangles at 8, 16 or 32 pixels a tile
        if (a.dense_count > a.capacity) return hipErrorInvalidValue;
        if (a.vals2[0] && !a.dense_rect_packed &&
            (!a.dense_rect_px || a.dense_rect_shift < 3 || a.dense_rect_shift > 5))
            return hipErrorInvalidValue;
    }
    if (a.bits == 9) return run_passes<9, true>(a, result_slot, s);
    if (a.bits == 8) return run_passes<8, false>(a, result_slot, s);
    if (a.bits == 7) return run_passes<7, true>(a, result_slot, s);
    if (a.bits == 6) return run_passes<6, false>(a, result_slot, s);
    if (a.bits == 4) return run_passes<4, true>(a, result_slot, s);
    return hipErrorInvalidValue;
}

}  // namespace gs
