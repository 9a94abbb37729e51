// gs_static.hip - the depth sort over caller buffers with no allocation.
//
// gs_sort_pairs_static runs the kernels of gs_sort (the stable LSD radix
// sort of gs_sort.hip) on (key, id) arrays the caller owns in device memory.
// Everything else the sort needs - one ping-pong buffer pair, the per-tile
// digit table, the column totals and the element count - lives in `scratch`,
// a fixed device array of this file, so the call makes no hipMalloc, hipFree
// or allocating library call: it fits code under hipGraph capture, serving
// loops and harnesses that forbid allocation.
//
// Passes: five 7-bit passes over the full 32-bit keys.  keys_out / ids_out
// serve as the other ping-pong buffer, so the fifth pass lands the result
// there and the inputs are only read.  Every byte of `scratch` the sort
// reads is written earlier in the same call (the count by a fill kernel,
// each pass's table and totals by its upsweep and column scan), so the
// array needs no clearing between calls.  Calls on one GPU share its copy
// of the array: calls that overlap in time on one GPU (two streams, two
// contexts) race on it.
#include <stdio.h>

#include <mutex>

#include "gs_internal.h"

namespace gs {

namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr uint32_t kMaxPairs = GS_SORT_STATIC_MAX_PAIRS;
constexpr uint32_t kTableWords = ((((kMaxPairs + 3839u) / 3840u) + 3u) & ~3u) * 512u;  // sort_table_words(kMaxPairs)
constexpr size_t kKeysOff = 0;
constexpr size_t kIdsOff = align256(kKeysOff + 4ull * kMaxPairs);
constexpr size_t kTableOff = align256(kIdsOff + 4ull * kMaxPairs);
constexpr size_t kColOff = align256(kTableOff + 4ull * kTableWords);
constexpr size_t kCountOff = align256(kColOff + 4ull * 512);
constexpr size_t kScratchBytes = align256(kCountOff + 16);

}  // namespace

// The sort's scratch: ping-pong keys, ping-pong ids, digit table, column
// totals, element count (about 34 MiB; every GPU the library runs on holds
// one copy until the process ends).
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint8_t* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device
hipError_t g_scratch_error[kMaxDevices] = {};  // why the lookup failed (hipSuccess: not tried or fine)

uint8_t* scratch_on(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace

// Called by gs_create with `device` current: looks up this device's address
// of `scratch` once.  Host code uses that address, never the array's name.
hipError_t static_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_error[device] = e;
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<uint8_t*>(p);
    return hipSuccess;
}

}  // namespace gs

extern "C" int gs_sort_pairs_static(gs_ctx* ctx, const uint32_t* keys_in, const uint32_t* ids_in, uint32_t* keys_out,
                                    uint32_t* ids_out, uint32_t count) {
    using namespace gs;
    hipStream_t stream = nullptr;
    int device = 0;
    if (ctx_launch_target(ctx, &stream, &device) != GS_OK) return api_error(GS_ERR_INVALID, "ctx is NULL");
    if (count > kMaxPairs) {
        char msg[160];
        snprintf(msg, sizeof(msg), "gs_sort_pairs_static: %u pairs exceed the static scratch's %u", count, kMaxPairs);
        return api_error(GS_ERR_LIMIT, msg);
    }
    if (count == 0) return api_ok();
    if (!keys_in || !ids_in || !keys_out || !ids_out)
        return api_error(GS_ERR_INVALID, "gs_sort_pairs_static: NULL buffer");
    if ((uintptr_t)keys_in % 4 || (uintptr_t)ids_in % 4 || (uintptr_t)keys_out % 4 || (uintptr_t)ids_out % 4)
        return api_error(GS_ERR_INVALID, "gs_sort_pairs_static: buffers must be 4-byte aligned");
    if (sort_table_words(kMaxPairs) != kTableWords) return api_error(GS_ERR_STATE, "sort table layout mismatch");
    uint8_t* base = scratch_on(device);
    if (!base) {
        char msg[200];
        snprintf(msg, sizeof(msg), "gs_sort_pairs_static: no scratch address on device %d (%s)", device,
                 hipGetErrorString(g_scratch_error[device < 0 || device >= kMaxDevices ? 0 : device]));
        return api_error(GS_ERR_STATE, msg);
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return api_error(GS_ERR_HIP, hipGetErrorString(e));

    uint32_t* dcount = reinterpret_cast<uint32_t*>(base + kCountOff);
    e = launch_fill_u32(dcount, count, 1, stream);
    SortLaunch s{};
    // pass p reads buffer p % 2 == 0 ? (p ? 2 : 0) : 1 and writes the other:
    // 0 -> 1 -> 2 -> 1 -> 2 -> 1, so buffer 1 (the caller's output) ends
    // with the sorted pairs
    s.keys[0] = const_cast<uint32_t*>(keys_in);
    s.vals[0] = const_cast<uint32_t*>(ids_in);
    s.keys[1] = keys_out;
    s.vals[1] = ids_out;
    s.keys[2] = reinterpret_cast<uint32_t*>(base + kKeysOff);
    s.vals[2] = reinterpret_cast<uint32_t*>(base + kIdsOff);
    s.count = dcount;
    s.capacity = count;
    s.scratch = reinterpret_cast<uint32_t*>(base + kColOff);
    s.table = reinterpret_cast<uint32_t*>(base + kTableOff);
    s.bits = 7;
    s.passes = 5;
    int slot = 0;
    if (e == hipSuccess) e = launch_radix_sort(s, &slot, stream);
    if (e != hipSuccess) return api_error(GS_ERR_HIP, hipGetErrorString(e));
    if (slot != 1) return api_error(GS_ERR_STATE, "gs_sort_pairs_static: result not in the output buffers");
    return api_ok();
}
