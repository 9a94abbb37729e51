// gs_scan.hip - exclusive scan of per-block counts (the middle step of every
// reduce-then-scan in this library: visible-Gaussian compaction, tile-pair
// offsets, GPS work reservation).
//
// The kernels that produce counts never wait on each other and never take a
// global ticket: on MI355X device-scope atomics on one address serialize
// across the XCDs, which made a per-block ticket or total counter cost
// ~25 ns per block.  One 1024-thread block scans up to 32768 counts staged
// in LDS; more counts take a reduce-then-scan over blocks of 8192 (below).
// The total goes to *total.
#include <mutex>

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kScanThreads = 1024;
constexpr int kRun = 32;             // consecutive counts per thread per round
constexpr uint32_t kLdsCounts = 32768;  // counts staged in LDS (128 KiB)

// The block's smallest lo and largest hi into out[0], out[1].
__device__ __forceinline__ void finish_range(uint32_t lo, uint32_t hi, uint32_t* out, uint32_t* s_lo, uint32_t* s_hi) {
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor((int)lo, d, 64));
        hi = max(hi, (uint32_t)__shfl_xor((int)hi, d, 64));
    }
    if ((tid & 63u) == 0) {
        s_lo[tid >> 6] = lo;
        s_hi[tid >> 6] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < kScanThreads / 64; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        out[0] = lo;
        out[1] = hi;
    }
}

// Smallest min and largest max of nb (min, max) pairs into out[0], out[1].
__device__ __forceinline__ void reduce_range(const uint2* __restrict__ block_range, uint32_t nb, uint32_t* out,
                                             uint32_t* s_lo, uint32_t* s_hi) {
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    // eight loads in flight per thread
    constexpr int kBatch = 8;
    for (uint32_t i0 = threadIdx.x; i0 < nb; i0 += kScanThreads * kBatch) {
        uint2 r[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t i = i0 + (uint32_t)j * kScanThreads;
            r[j] = i < nb ? block_range[i] : make_uint2(0xFFFFFFFFu, 0u);
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            lo = min(lo, r[j].x);
            hi = max(hi, r[j].y);
        }
    }
    finish_range(lo, hi, out, s_lo, s_hi);
}

__device__ __forceinline__ void reduce_diag(const uint32_t* __restrict__ block_diag, uint32_t nb, uint32_t* diag_out,
                                            uint32_t* s_diag) {
    const uint32_t tid = threadIdx.x;
    // SplatDiagnostics: sums except the two maxima (fields 1 and 5)
    if (tid < 8) s_diag[tid] = 0u;
    __syncthreads();
    for (uint32_t i = tid; i < nb * 8u; i += kScanThreads) {
        uint32_t f = i & 7u, v = block_diag[i];
        if (!v) continue;
        if (f == 1u || f == 5u) atomicMax(&s_diag[f], v);
        else atomicAdd(&s_diag[f], v);
    }
    __syncthreads();
    if (tid < 8) diag_out[tid] = s_diag[tid];
}

// Up to kLdsCounts counts: coalesced loads into LDS, each thread scans one
// contiguous run out of LDS (runs of odd length stay clear of bank
// conflicts), one block scan of the run sums, coalesced stores.
__global__ void __launch_bounds__(kScanThreads) scan_counts_lds_kernel(const uint32_t* __restrict__ counts,
                                                                       uint32_t nb, uint32_t* offsets,
                                                                       uint32_t* total,
                                                                       const uint32_t* __restrict__ block_diag,
                                                                       uint32_t* diag_out, uint32_t clamp_cap,
                                                                       uint32_t* clamped,
                                                                       const uint2* __restrict__ block_range,
                                                                       uint32_t* range_out) {
    extern __shared__ uint32_t s_c[];
    __shared__ uint32_t s_scan[kScanThreads / 64];
    __shared__ uint32_t s_diag[8];
    __shared__ uint32_t s_lo[kScanThreads / 64], s_hi[kScanThreads / 64];
    const uint32_t tid = threadIdx.x;
    // eight loads in flight per thread before their LDS stores; the blocks'
    // key ranges (when given) are loaded with their counts (a clamped
    // duplicate of the last one leaves the min and max unchanged)
    constexpr int kBatch = 8;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t i0 = tid; i0 < nb; i0 += kScanThreads * kBatch) {
        uint32_t v[kBatch];
        uint2 rg[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t i = i0 + (uint32_t)j * kScanThreads;
            v[j] = counts[min(i, nb - 1u)];
        }
        if (block_range) {
#pragma unroll
            for (int j = 0; j < kBatch; ++j) rg[j] = block_range[min(i0 + (uint32_t)j * kScanThreads, nb - 1u)];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t i = i0 + (uint32_t)j * kScanThreads;
            if (i < nb) s_c[i] = v[j];
        }
        if (block_range) {
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                lo = min(lo, rg[j].x);
                hi = max(hi, rg[j].y);
            }
        }
    }
    __syncthreads();
    const uint32_t run = (nb + kScanThreads - 1) / kScanThreads;
    const uint32_t r0 = tid * run, r1 = min(r0 + run, nb);
    uint32_t sum = 0u;
    for (uint32_t i = r0; i < r1; ++i) sum += s_c[i];
    uint32_t all;
    uint32_t acc = block_exclusive_sum<kScanThreads / 64>(sum, s_scan, &all);
    for (uint32_t i = r0; i < r1; ++i) {
        const uint32_t v = s_c[i];
        s_c[i] = acc;
        acc += v;
    }
    __syncthreads();
    for (uint32_t i = tid; i < nb; i += kScanThreads) offsets[i] = s_c[i];
    if (tid == 0) {
        *total = all;
        if (clamped) *clamped = min(all, clamp_cap);
    }
    if (block_diag) reduce_diag(block_diag, nb, diag_out, s_diag);
    if (block_range) finish_range(lo, hi, range_out, s_lo, s_hi);
}

__global__ void __launch_bounds__(kScanThreads) scan_counts_kernel(const uint32_t* __restrict__ counts,
                                                                   uint32_t nb, uint32_t* offsets,
                                                                   uint32_t* total,
                                                                   const uint32_t* __restrict__ block_diag,
                                                                   uint32_t* diag_out, uint32_t clamp_cap,
                                                                   uint32_t* clamped,
                                                                   const uint2* __restrict__ block_range,
                                                                   uint32_t* range_out) {
    __shared__ uint32_t s_scan[kScanThreads / 64];
    __shared__ uint32_t s_diag[8];
    __shared__ uint32_t s_lo[kScanThreads / 64], s_hi[kScanThreads / 64];
    const uint32_t tid = threadIdx.x;
    // Rounds of 1024 runs of kRun consecutive counts: every thread issues the
    // loads of its run together (one memory latency per round), sums it, one
    // block scan gives the run offsets, and the run is written back.
    uint32_t carry = 0u;
    for (uint32_t first = 0; first < nb; first += kRun * kScanThreads) {
        const uint32_t r0 = first + tid * kRun;
        uint32_t v[kRun];
        uint32_t sum = 0u;
#pragma unroll
        for (int k = 0; k < kRun; ++k) {
            v[k] = r0 + k < nb ? counts[r0 + k] : 0u;
            sum += v[k];
        }
        uint32_t chunk;
        uint32_t run = carry + block_exclusive_sum<kScanThreads / 64>(sum, s_scan, &chunk);
#pragma unroll
        for (int k = 0; k < kRun; ++k) {
            if (r0 + k < nb) offsets[r0 + k] = run;
            run += v[k];
        }
        carry += chunk;
    }
    if (tid == 0) {
        *total = carry;
        if (clamped) *clamped = min(carry, clamp_cap);
    }
    if (block_diag) reduce_diag(block_diag, nb, diag_out, s_diag);
    if (block_range) reduce_range(block_range, nb, range_out, s_lo, s_hi);
}

// Totals only (no offsets): the sum of nb counts into *total, and the
// diagnostics and key range as above.  Eight loads in flight per thread.
__global__ void __launch_bounds__(kScanThreads) reduce_counts_kernel(const uint32_t* __restrict__ counts,
                                                                     uint32_t nb, uint32_t* total,
                                                                     const uint32_t* __restrict__ block_diag,
                                                                     uint32_t* diag_out,
                                                                     const uint2* __restrict__ block_range,
                                                                     uint32_t* range_out) {
    __shared__ uint32_t s_sum[kScanThreads / 64];
    __shared__ uint32_t s_diag[8];
    __shared__ uint32_t s_lo[kScanThreads / 64], s_hi[kScanThreads / 64];
    const uint32_t tid = threadIdx.x;
    constexpr int kBatch = 8;
    uint32_t sum = 0u, lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t i0 = tid; i0 < nb; i0 += kScanThreads * kBatch) {
        uint32_t v[kBatch];
        uint2 rg[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const uint32_t i = i0 + (uint32_t)j * kScanThreads;
            v[j] = i < nb ? counts[i] : 0u;
        }
        if (block_range) {
#pragma unroll
            for (int j = 0; j < kBatch; ++j) rg[j] = block_range[min(i0 + (uint32_t)j * kScanThreads, nb - 1u)];
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) sum += v[j];
        if (block_range) {
#pragma unroll
            for (int j = 0; j < kBatch; ++j) {
                lo = min(lo, rg[j].x);
                hi = max(hi, rg[j].y);
            }
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sum += (uint32_t)__shfl_xor((int)sum, d, 64);
    if ((tid & 63u) == 0) s_sum[tid >> 6] = sum;
    __syncthreads();
    if (tid == 0) {
        uint32_t all = 0u;
        for (int w = 0; w < kScanThreads / 64; ++w) all += s_sum[w];
        *total = all;
    }
    if (block_diag) reduce_diag(block_diag, nb, diag_out, s_diag);
    if (block_range) finish_range(lo, hi, range_out, s_lo, s_hi);
}

// Counts past the LDS staging limit, without diagnostics or key ranges, are
// scanned by many blocks: each block of kChunkThreads threads owns kChunk
// consecutive counts (kChunkItems per thread, two 16-byte loads).
//   1. chunk_sum_kernel    each block's sum into offsets[block * kChunk]
//                          (the first slot of its own range);
//   2. chunk_scan_kernel   one block scans those sums in place (exclusive)
//                          and writes the total;
//   3. chunk_apply_kernel  each block reads its base from that slot, then
//                          writes the exclusive offsets of its counts.
// A slot is written in step 1 and read and overwritten in step 3 only by its
// own block, so no block waits on another.
constexpr int kChunkThreads = 1024;
constexpr int kChunkItems = 8;
constexpr uint32_t kChunk = kChunkThreads * kChunkItems;

// This thread's kChunkItems counts of block b's chunk (zero past nb).
__device__ __forceinline__ void load_chunk(const uint32_t* __restrict__ counts, uint32_t nb, uint32_t i0,
                                           uint32_t (&v)[kChunkItems]) {
    if (i0 + kChunkItems <= nb) {
        const uint4 a = *reinterpret_cast<const uint4*>(counts + i0);
        const uint4 b = *reinterpret_cast<const uint4*>(counts + i0 + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kChunkItems; ++k) v[k] = i0 + k < nb ? counts[i0 + k] : 0u;
    }
}

__global__ void __launch_bounds__(kChunkThreads) chunk_sum_kernel(const uint32_t* __restrict__ counts, uint32_t nb,
                                                                  uint32_t* offsets) {
    __shared__ uint32_t s_scan[kChunkThreads / 64];
    const uint32_t i0 = blockIdx.x * kChunk + threadIdx.x * kChunkItems;
    uint32_t v[kChunkItems];
    load_chunk(counts, nb, i0, v);
    uint32_t sum = 0u;
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) sum += v[k];
    uint32_t all;
    (void)block_exclusive_sum<kChunkThreads / 64>(sum, s_scan, &all);
    if (threadIdx.x == 0) offsets[blockIdx.x * kChunk] = all;
}

__global__ void __launch_bounds__(kScanThreads) chunk_scan_kernel(uint32_t* offsets, uint32_t chunks, uint32_t* total,
                                                                  uint32_t clamp_cap, uint32_t* clamped) {
    __shared__ uint32_t s_scan[kScanThreads / 64];
    uint32_t carry = 0u;
    for (uint32_t first = 0; first < chunks; first += kScanThreads) {
        const uint32_t c = first + threadIdx.x;
        const uint32_t v = c < chunks ? offsets[(size_t)c * kChunk] : 0u;
        uint32_t all;
        const uint32_t before = block_exclusive_sum<kScanThreads / 64>(v, s_scan, &all);
        if (c < chunks) offsets[(size_t)c * kChunk] = carry + before;
        carry += all;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        if (clamped) *clamped = min(carry, clamp_cap);
    }
}

__global__ void __launch_bounds__(kChunkThreads) chunk_apply_kernel(const uint32_t* __restrict__ counts, uint32_t nb,
                                                                    uint32_t* offsets) {
    __shared__ uint32_t s_scan[kChunkThreads / 64];
    __shared__ uint32_t s_base;
    const uint32_t i0 = blockIdx.x * kChunk + threadIdx.x * kChunkItems;
    if (threadIdx.x == 0) s_base = offsets[blockIdx.x * kChunk];
    uint32_t v[kChunkItems];
    load_chunk(counts, nb, i0, v);
    uint32_t sum = 0u;
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) sum += v[k];
    uint32_t all;
    // the barriers inside the block scan order the base's read before any
    // write of the block's range
    uint32_t run = block_exclusive_sum<kChunkThreads / 64>(sum, s_scan, &all) + s_base;
    uint32_t o[kChunkItems];
#pragma unroll
    for (int k = 0; k < kChunkItems; ++k) {
        o[k] = run;
        run += v[k];
    }
    if (i0 + kChunkItems <= nb) {
        *reinterpret_cast<uint4*>(offsets + i0) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(offsets + i0 + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kChunkItems; ++k)
            if (i0 + k < nb) offsets[i0 + k] = o[k];
    }
}

}  // namespace

hipError_t launch_reduce_counts(const uint32_t* counts, uint32_t nb, uint32_t* total, const uint32_t* block_diag,
                                uint32_t* diag_out, hipStream_t s, const uint2* block_range, uint32_t* range_out) {
    if (block_range && !range_out) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reduce_counts_kernel, dim3(1), dim3(kScanThreads), 0, s, counts, nb, total, block_diag,
                       diag_out, block_range, range_out);
    return hipGetLastError();
}

hipError_t init_kernel_attributes() {
    // once per process; the attribute belongs to the kernel on every device
    static std::once_flag once;
    static hipError_t status = hipSuccess;
    std::call_once(once, [] {
        status = hipFuncSetAttribute(reinterpret_cast<const void*>(&scan_counts_lds_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, kLdsCounts * 4);
    });
    return status;
}

hipError_t launch_scan_counts(const uint32_t* counts, uint32_t nb, uint32_t* offsets, uint32_t* total,
                              const uint32_t* block_diag, uint32_t* diag_out, hipStream_t s, uint32_t clamp_cap,
                              uint32_t* clamped, const uint2* block_range, uint32_t* range_out) {
    if (block_range && !range_out) return hipErrorInvalidValue;
    if (nb <= kLdsCounts) {
        hipError_t e = init_kernel_attributes();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(scan_counts_lds_kernel, dim3(1), dim3(kScanThreads), (size_t)(nb ? nb : 1) * 4, s, counts,
                           nb, offsets, total, block_diag, diag_out, clamp_cap, clamped, block_range, range_out);
    } else if (!block_diag && !block_range && reinterpret_cast<uintptr_t>(counts) % 16 == 0 &&
               reinterpret_cast<uintptr_t>(offsets) % 16 == 0) {
        const uint32_t chunks = (uint32_t)(((uint64_t)nb + kChunk - 1) / kChunk);
        hipLaunchKernelGGL(chunk_sum_kernel, dim3(chunks), dim3(kChunkThreads), 0, s, counts, nb, offsets);
        hipLaunchKernelGGL(chunk_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, offsets, chunks, total, clamp_cap,
                           clamped);
        hipLaunchKernelGGL(chunk_apply_kernel, dim3(chunks), dim3(kChunkThreads), 0, s, counts, nb, offsets);
    } else {
        hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(kScanThreads), 0, s, counts, nb, offsets, total,
                           block_diag, diag_out, clamp_cap, clamped, block_range, range_out);
    }
    return hipGetLastError();
}

}  // namespace gs
