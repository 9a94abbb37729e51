// gs_gps.hip - Gaussian Point Splatting sample and running-mean accumulation.
//
// Replaces gps_clear_cs, gps_expand_cs, gps_splat_cs, gps_resolve_fs
// (shaders/gps.glsl:6-386) as sequenced by renderer_draw_gps_sample
// (src/renderer.cpp:1964-2118), and accum_fs (shaders/accum.glsl:19-45).
//
// Work-budget semantics are the reference's: each visible splat asks for
// floor(E/ss^2) work items plus one with probability frac(E/ss^2) (frame
// invariant hash), items are reserved from a global budget, and a
// reservation that does not fit keeps only what fits.  Reservations are
// contiguous from 0, so the reference's overflow flag is exactly
// "requested > budget".  The reservation order is the visible-index order
// (count, scan, reserve instead of an atomicAdd race), so which items are
// dropped on overflow is deterministic.  Each work item stores only its owner (4 bytes instead of
// the reference's 8-byte (splat, item) pair); the item index is recovered
// from the owner's base offset.
//
// Depth and colour are one 64-bit word per supersample cell,
// depth_key << 32 | payload, updated with a single 64-bit atomicMin: the
// nearest point wins, ties go to the smaller payload, and the colour can no
// longer race with the depth as the reference's two 32-bit writes can
// (shaders/gps.glsl:279-282).  The resolve re-clears the cells it reads, so
// no separate clear pass runs per frame.
#include "gs_internal.h"
#include "gs_math.h"

namespace gs {

namespace {

constexpr uint64_t kEmptyCell = 0xFFFFFFFF00000000ull;

// Prepared record of visible position v (gps.glsl:151-169): mean, Cholesky
// factor of the twice-dilated covariance, opacity and its Li2, depth key,
// colour (or the splat id for linear colour) and the splat id.
__device__ void write_prepared(const GpsLaunch& a, uint32_t v, uint32_t splat_id, float4 r0, float4 r1) {
    const float4 r2 = a.records[(size_t)splat_id * kRecordVec4 + 2];
    const float cov_a = r1.x + 0.3f;
    const float cov_b = r1.y;
    const float cov_c = r1.z + 0.3f;
    float l00 = __builtin_sqrtf(maxf(cov_a, 1.0e-5f));
    float l10 = cov_b / l00;
    float l11 = __builtin_sqrtf(maxf(cov_c - l10 * l10, 1.0e-5f));
    float alpha = clampf(r0.z, 1.0e-6f, 1.0f);
    float dilog_alpha = dilog(alpha);
    float ndc_depth = r0.w;
    float z = mixf(ndc_depth * 0.5f + 0.5f, ndc_depth, a.clip_z_01);
    uint32_t depth_key = f2u_sat(clampf(z, 0.0f, 1.0f) * 4294967294.0f);
    uint32_t rr = (uint32_t)__builtin_rintf(clampf(r2.x, 0.0f, 1.0f) * 255.0f);
    uint32_t gg = (uint32_t)__builtin_rintf(clampf(r2.y, 0.0f, 1.0f) * 255.0f);
    uint32_t bb = (uint32_t)__builtin_rintf(clampf(r2.z, 0.0f, 1.0f) * 255.0f);
    uint32_t packed = rr | (gg << 8) | (bb << 16) | 0xFF000000u;
    if (a.linear_color) packed = splat_id;
    uint4* pr = a.prepared + (size_t)v * 3;
    pr[0] = make_uint4(fbits(r0.x), fbits(r0.y), fbits(l00), fbits(l10));
    pr[1] = make_uint4(fbits(l11), fbits(alpha), fbits(dilog_alpha), 0u);
    pr[2] = make_uint4(depth_key, packed, splat_id, 0u);
}

// Pass 1: work-item count (gps.glsl:127-142) per visible splat, item count
// per block.  The prepared record is needed only by splats whose
// reservation starts inside the budget (the reference writes it for those
// alone, gps.glsl:148-169), which only the scan decides: a block writes it
// here, for every splat with items, when its hint says that the block kept
// work in the frame before, and leaves it to pass 2 (kept splats only)
// otherwise.  Either way the same records reach the splat pass; the hint
// only saves writing records of blocks that lie past the budget.
__global__ void __launch_bounds__(256) gps_count_kernel(GpsLaunch a) {
    __shared__ uint32_t s_scan[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t V = *a.vis_count;
    const uint32_t v = blockIdx.x * 256u + tid;
    const float ss = maxf(a.ss, 1.0f);
    const bool prepare_here = a.prepare_in_count[blockIdx.x] != 0u;
    uint32_t count = 0u;
    if (v < V) {
        const uint32_t splat_id = a.vis_ids[v];
        const float4 r0 = a.records[(size_t)splat_id * kRecordVec4 + 0];
        const float4 r1 = a.records[(size_t)splat_id * kRecordVec4 + 1];
        // second dilation and expected point count
        const float cov_a = r1.x + 0.3f;
        const float cov_b = r1.y;
        const float cov_c = r1.z + 0.3f;
        float det = cov_a * cov_c - cov_b * cov_b;
        float expected_points = ss * ss * 6.28318530718f * __builtin_sqrtf(maxf(det, 0.0f)) *
            dilog(clampf(r0.z, 0.0f, 1.0f));
        float ppw = maxf(ss * ss, 1.0f);
        float scaled = maxf(expected_points / ppw, 0.0f);
        float fl = clampf(__builtin_floorf(scaled), 0.0f, 4294967295.0f);
        count = f2u_sat(fl);
        if (hash01(splat_id + 0x9e3779b9u) < fractf(scaled) && count < 0xFFFFFFFFu) count += 1u;
        a.item_count[v] = count;
        if (count && prepare_here) write_prepared(a, v, splat_id, r0, r1);
    }
    uint32_t total;
    (void)block_exclusive_sum<4>(count, s_scan, &total);
    if (tid == 0) a.block_counts[blockIdx.x] = total;
}

// Pass 2 (after the block counts are scanned): reservation offsets in
// visible-index order; the block's items form one contiguous range of the
// work list, written together (owner position per item) up to the budget.
__global__ void __launch_bounds__(256) gps_reserve_kernel(GpsLaunch a) {
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_start[256];
    const uint32_t tid = threadIdx.x;
    const uint32_t V = *a.vis_count;
    if (blockIdx.x * 256u >= V) return;
    const uint32_t v = blockIdx.x * 256u + tid;
    const uint32_t count = v < V ? a.item_count[v] : 0u;
    uint32_t total;
    const uint32_t local = block_exclusive_sum<4>(count, s_scan, &total);
    s_start[tid] = local;
    __syncthreads();
    const uint32_t capacity = a.max_work_items;
    const uint32_t block_base = a.block_offsets[blockIdx.x];
    const uint32_t room = block_base < capacity ? capacity - block_base : 0u;
    const uint32_t emit = min(total, room);
    for (uint32_t i = tid; i < emit; i += 256u) a.work[block_base + i] = blockIdx.x * 256u + owner_of(s_start, 256u, i);
    const bool prepared = a.prepare_in_count[blockIdx.x] != 0u;
    if (count && block_base + local < capacity) {
        a.item_base[v] = block_base + local;
        if (!prepared) {
            const uint32_t splat_id = a.vis_ids[v];
            write_prepared(a, v, splat_id, a.records[(size_t)splat_id * kRecordVec4 + 0],
                           a.records[(size_t)splat_id * kRecordVec4 + 1]);
        }
    }
    // hint for the next frame's pass 1: this block kept work
    __syncthreads();
    if (tid == 0) a.prepare_in_count[blockIdx.x] = block_base < capacity ? 1u : 0u;
}

__global__ void __launch_bounds__(256) gps_splat_kernel(GpsLaunch a) {
    const uint32_t requested = a.counters->gps.requested;
    const uint32_t n = min(requested, a.max_work_items);
    const uint32_t fs = (uint32_t)(float)a.frame_seed;
    const float ss = a.ss;
    int ppw = (int)__builtin_rintf(ss * ss);
    if (ppw < 1) ppw = 1;
    const float ssm = maxf(ss, 1.0f);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += stride) {
        const uint32_t v = a.work[w];
        const uint32_t sample_idx = w - a.item_base[v];
        const uint4* pr = a.prepared + (size_t)v * 3;
        const uint4 q0 = pr[0], q1 = pr[1], q2 = pr[2];
        const float mx = bitsf(q0.x), my = bitsf(q0.y), l00 = bitsf(q0.z), l10 = bitsf(q0.w);
        const float l11 = bitsf(q1.x), alpha = bitsf(q1.y), dla = bitsf(q1.z);
        const uint32_t splat_id = q2.z;
        const uint64_t value = ((uint64_t)q2.x << 32) | q2.y;
        const uint32_t seed0 = splat_id * 747796405u + fs * 2891336453u;
        const uint32_t seed1 = splat_id * 277803737u + 0x9E3779B9u;
        for (int p = 0; p < ppw; ++p) {
            uint32_t x = seed0 + (sample_idx * (uint32_t)ppw + (uint32_t)p) * 374761393u;
            uint32_t y = seed1;
            pcg2d(x, y);
            float r0 = (float)x * 2.32830643654e-10f;
            float r1 = (float)y * 2.32830643654e-10f;
            float arg = inv_dilog((1.0f - r0) * dla) / alpha;
            arg = clampf(arg, 1.0e-7f, 1.0f);
            float radius = __builtin_sqrtf(-2.0f * logf_fma(arg));
            float az = 6.28318530718f * r1;
            float sn, cs;
            gps_sincosf(az, &sn, &cs);
            float lx = cs * radius, ly = sn * radius;
            float ptx = (mx + l00 * lx) * ssm;
            float pty = (my + (l10 * lx + l11 * ly)) * ssm;
            int px = (int)__builtin_floorf(ptx + 0.5f), py = (int)__builtin_floorf(pty + 0.5f);
            if (px >= 0 && py >= 0 && px < a.gps_width && py < a.gps_height) {
                // A cell only decreases during this kernel: every value in its
                // modification order is at most the one before, and the later
                // atomicMin of this thread reads no earlier value than this
                // load.  A cell already at or below `value` makes the atomic a
                // no-op, so it is skipped (most are: a pixel receives several
                // points, and each new one beats the nearest so far ever less
                // often).
                auto* cp = reinterpret_cast<unsigned long long*>(a.cells + (size_t)py * a.gps_width + px);
                if (__hip_atomic_load(cp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > value)
                    atomicMin(cp, (unsigned long long)value);
            }
        }
    }
}

__global__ void __launch_bounds__(256) gps_resolve_kernel(GpsLaunch a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npx = (uint32_t)(a.width * a.height);
    if (i >= npx) return;
    const int x = (int)(i % (uint32_t)a.width), y = (int)(i / (uint32_t)a.width);
    int ss = (int)a.ss;
    ss = ss < 1 ? 1 : (ss > 4 ? 4 : ss);
    float sum0 = 0.0f, sum1 = 0.0f, sum2 = 0.0f, cov = 0.0f;
    uint32_t min_key = 0xFFFFFFFFu;
    for (int yy = 0; yy < ss; ++yy) {
        for (int xx = 0; xx < ss; ++xx) {
            int sx = x * ss + xx, sy = y * ss + yy;
            if (sx >= a.gps_width || sy >= a.gps_height) continue;
            uint64_t* cp = a.cells + (size_t)sy * a.gps_width + sx;
            uint64_t cell = *cp;
            if (a.clear_cells) *cp = kEmptyCell;
            uint32_t key = (uint32_t)(cell >> 32);
            if (key != 0xFFFFFFFFu) {
                uint32_t payload = (uint32_t)cell;
                if (a.linear_color) {
                    float4 c = a.records[(size_t)payload * kRecordVec4 + 2];
                    sum0 += c.x; sum1 += c.y; sum2 += c.z;
                } else {
                    sum0 += (float)(payload & 255u) / 255.0f;
                    sum1 += (float)((payload >> 8) & 255u) / 255.0f;
                    sum2 += (float)((payload >> 16) & 255u) / 255.0f;
                }
                cov += 1.0f;
                min_key = min(min_key, key);
            }
        }
    }
    if (cov <= 0.0f) {
        a.sample[i] = make_float4(a.background, a.background, a.background, 0.0f);
        a.depth[i] = 1.0f;
        return;
    }
    float sub = (float)(ss * ss);
    a.sample[i] = make_float4(sum0 / sub, sum1 / sub, sum2 / sub, cov / sub);
    a.depth[i] = (float)min_key / 4294967294.0f;
}

__global__ void accumulate_kernel(const float4* __restrict__ sample, float4* accum, uint32_t count, float n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 s = sample[i];
    float4 h = n <= 1.0f ? s : accum[i];
    float m = maxf(n, 1.0f);
    accum[i] = make_float4(h.x + (s.x - h.x) / m, h.y + (s.y - h.y) / m, h.z + (s.z - h.z) / m,
                           h.w + (s.w - h.w) / m);
}

}  // namespace

hipError_t launch_gps_expand(const GpsLaunch& a, hipStream_t s) {
    const uint32_t blocks = (a.capacity + 255u) / 256u;
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(gps_count_kernel, dim3(blocks), dim3(256), 0, s, a);
    hipError_t e = launch_scan_counts(a.block_counts, blocks, a.block_offsets, &a.counters->gps.requested, nullptr,
                                      nullptr, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gps_reserve_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gps_splat(const GpsLaunch& a, hipStream_t s) {
    uint32_t blocks = (a.max_work_items + 255u) / 256u;
    if (blocks > 8192u) blocks = 8192u;
    if (!blocks) blocks = 1;
    hipLaunchKernelGGL(gps_splat_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gps_resolve(const GpsLaunch& a, hipStream_t s) {
    uint32_t npx = (uint32_t)(a.width * a.height);
    if (!npx) return hipSuccess;
    hipLaunchKernelGGL(gps_resolve_kernel, dim3((npx + 255u) / 256u), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_accumulate(const float4* sample, float4* accum, uint32_t count, float n, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(accumulate_kernel, dim3((count + 255u) / 256u), dim3(256), 0, s, sample, accum, count, n);
    return hipGetLastError();
}

}  // namespace gs
