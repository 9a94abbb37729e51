// gs_composite.hip - tile binning and front-to-back sorted-alpha compositing.
//
// Replaces the instanced quad draw of the depth-sorted splats into the
// RGBA16F target with ONE / ONE_MINUS_SRC_ALPHA blending
// (shaders/splat.glsl:6-105, src/renderer.cpp:277-299,1687-1720,2534-2541).
//
// Coverage follows the rasteriser: a splat covers pixel (i, j) when its quad
// centre +- radius, taken through the same clip-space round trip as the
// vertex shader and snapped to 1/256 pixel, contains the pixel centre
// (i + 0.5, j + 0.5) with left/bottom edges inclusive.  Inside the quad the
// fragment test and alpha are the fragment shader's.
//
//   bin_count   per front-to-back position: tiles touched by the splat's
//               covered pixels (the tile rectangle the depth sort carried
//               along, else from the pixel rectangle the cull wrote), pair
//               count per block;
//   (scan)      block offsets;
//   bin_emit    (tile id, splat id) pairs in front-to-back order;
//   (sort)      stable LSD sort by tile id keeps the depth order in a tile;
//   tile_ranges [begin, end) of every tile;
//   composite   one block per tile, splat batches staged in LDS, all pixels
//               of the tile walk the batch front to back and stop once their
//               transmittance drops below 1e-4; the block stops when every
//               pixel has.
#include <hip/hip_fp16.h>

#include "gs_internal.h"
#include "gs_math.h"
#include "gs_raster.h"

namespace gs {

namespace {

constexpr float kTransmittanceMin = 1e-4f;
// The compositors carry each pixel's transmittance T as u = T / 255 and a
// splat's alpha as alpha' = 255 alpha = 2^(power - log2(1/255)), so the
// fragment test alpha >= 1/255 is power' >= 0 and the blend weight
// alpha T = alpha' u needs no extra scaling (u' = u - alpha' u / 255).
constexpr float kInv255 = 1.0f / 255.0f;
constexpr float kAlphaCap255 = 0.99f * 255.0f;  // alpha <= 0.99
constexpr float kUMin = kTransmittanceMin / 255.0f;

// Pass 1 over the depth-sorted splats, nearest first: tile rectangle and
// pair count per splat, pair count per block.
__global__ void __launch_bounds__(256) bin_count_kernel(BinLaunch a) {
    __shared__ uint32_t s_scan[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t V = *a.count;
    const uint32_t q = blockIdx.x * 256u + tid;
    for (uint32_t w = q; w < a.clear_words; w += gridDim.x * 256u) a.clear[w] = 0u;
    uint32_t cnt = 0u;
    if (q < V && a.sorted_rects) {
        const uint32_t r = a.sorted_rects[V - 1u - q];
        cnt = (((r >> 8) & 0xFFu) - (r & 0xFFu)) * ((r >> 24) - ((r >> 16) & 0xFFu));
    } else if (q < V) {
        const uint32_t val = a.sorted_ids[V - 1u - q];
        uint2 px;
        if (a.jittered) {
            const float4 r0 = a.records[(size_t)val * kRecordVec4 + 0];
            const float4 r1 = a.records[(size_t)val * kRecordVec4 + 1];
            const float4 r2 = a.records[(size_t)val * kRecordVec4 + 2];
            const ViewInfo& vi = a.views[val / a.n];
            px = depth_clipped(r0.w, a.clip_z_01)
                     ? make_uint2(0u, 0u)
                     : pixel_rect(r0.x, r0.y, r1.x, r1.y, r1.z, r1.w, r2.w, r0.z, vi.width, vi.height, a.tight != 0,
                                  a.jitter_x, a.jitter_y, a.pad);
        } else {
            px = a.rect_px[val];
        }
        uint2 t = make_uint2(0u, 0u);
        if (px.x) {
            const uint32_t T = (uint32_t)a.tile;
            const uint32_t x0 = (px.x & 0xFFFFu) / T, x1 = ((px.x >> 16) - 1u) / T + 1u;
            const uint32_t y0 = (px.y & 0xFFFFu) / T, y1 = ((px.y >> 16) - 1u) / T + 1u;
            cnt = (x1 - x0) * (y1 - y0);
            t = make_uint2(x0 | (x1 << 16), y0 | (y1 << 16));
        }
        a.rects[q] = t;
    }
    uint32_t total;
    (void)block_exclusive_sum<4>(cnt, s_scan, &total);
    if (tid == 0) a.block_counts[blockIdx.x] = total;
}

// Pass 2 (after the block counts are scanned): each block writes the
// (tile id, splat id) pairs of its 256 splats as one contiguous range.  A
// splat with at most kOwnPairs pairs writes them itself, row by row of its
// tile rectangle; the pairs of larger splats are written by the whole block,
// each output slot finding its splat by a binary search over the large
// splats' offsets.  Pairs past the capacity are not written.
constexpr uint32_t kOwnPairs = 16u;

// A tile key of the pairs: 32-bit, or 16-bit with K16.
template <bool O32, bool K16>
__device__ __forceinline__ void store_tile_key(uint32_t* keys, uint32_t i, uint32_t k) {
    if constexpr (K16) at<O32>(reinterpret_cast<uint16_t*>(keys), i) = (uint16_t)k;
    else at<O32>(keys, i) = k;
}

template <bool O32, bool K16>
__global__ void __launch_bounds__(256) bin_emit_kernel(BinLaunch a) {
    __shared__ uint32_t s_scan[4];
    __shared__ uint32_t s_lstart[256];   // large splats: first pair (block-relative) ...
    __shared__ uint32_t s_lrel[256];     // ... and their offset among the large splats' pairs
    __shared__ uint32_t s_id[256];
    __shared__ uint32_t s_tbase[256];
    __shared__ uint32_t s_tx[256];
    __shared__ uint2 s_rect[256];
    const uint32_t tid = threadIdx.x;
    const uint32_t V = *a.count;
    if (blockIdx.x * 256u >= V) return;
    const uint32_t q = blockIdx.x * 256u + tid;
    // the block's first pair, loaded with the splats (no dependence on them)
    const uint32_t base = a.block_offsets[blockIdx.x];
    uint32_t cnt = 0u, val = 0u, tbase = 0u, tx = a.single_tiles_x;
    uint2 r = make_uint2(0u, 0u);
    if (q < V) {
        if (a.sorted_rects) {
            const uint32_t p = a.sorted_rects[V - 1u - q];
            r = make_uint2((p & 0xFFu) | (((p >> 8) & 0xFFu) << 16), ((p >> 16) & 0xFFu) | ((p >> 24) << 16));
        } else {
            r = a.rects[q];
        }
        cnt = ((r.x >> 16) - (r.x & 0xFFFFu)) * ((r.y >> 16) - (r.y & 0xFFFFu));
        val = a.sorted_ids[V - 1u - q];
        if (cnt && !a.single_tiles_x) {
            const ViewInfo& vi = a.views[val / a.n];
            tbase = vi.tile_base;
            tx = (uint32_t)vi.tiles_x;
        }
    }
    uint32_t total;
    const uint32_t start = block_exclusive_sum<4>(cnt, s_scan, &total);
    const uint32_t room = base < a.dup_capacity ? a.dup_capacity - base : 0u;
    const uint32_t emit = min(total, room);
    const bool large = cnt > kOwnPairs;
    if (cnt && !large) {
        const uint32_t x0 = r.x & 0xFFFFu, x1 = r.x >> 16;
        uint32_t xx = x0, row = tbase + (r.y & 0xFFFFu) * tx;
        for (uint32_t i = start; i < start + cnt && i < emit; ++i) {
            store_tile_key<O32, K16>(a.dup_keys, base + i, row + xx);
            at<O32>(a.dup_vals, base + i) = val;
            if (++xx == x1) {
                xx = x0;
                row += tx;
            }
        }
    }
    // the large splats, compacted in order: their pair offsets among
    // themselves and their first pair in the block's range
    uint32_t lpairs;
    const uint32_t lrel = block_exclusive_sum<4>(large ? cnt : 0u, s_scan, &lpairs);
    if (lpairs == 0u) return;  // uniform: no splat of the block is large
    uint32_t nlarge;
    const uint32_t li = block_exclusive_sum<4>(large ? 1u : 0u, s_scan, &nlarge);
    if (large) {
        s_lstart[li] = start;
        s_lrel[li] = lrel;
        s_id[li] = val;
        s_tbase[li] = tbase;
        s_tx[li] = tx;
        s_rect[li] = r;
    }
    __syncthreads();
    for (uint32_t j = tid; j < lpairs; j += 256u) {
        const uint32_t o = owner_of(s_lrel, nlarge, j);
        const uint32_t k = j - s_lrel[o];
        const uint32_t i = s_lstart[o] + k;
        if (i >= emit) continue;
        const uint2 rr = s_rect[o];
        const uint32_t x0 = rr.x & 0xFFFFu, w = (rr.x >> 16) - x0, y0 = rr.y & 0xFFFFu;
        store_tile_key<O32, K16>(a.dup_keys, base + i, s_tbase[o] + (y0 + k / w) * s_tx[o] + x0 + k % w);
        at<O32>(a.dup_vals, base + i) = s_id[o];
    }
}

// [begin, end) of every tile's entries in the tile-sorted keys: each run
// start writes its own begin and the previous run's end.  A thread takes four
// consecutive keys (one 16-byte load, 8-byte with 16-bit keys, when they are
// all in range); the key before them is the last of the neighbouring lane's
// four.
template <bool K16>
__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys32, const uint32_t* count_ptr, uint2* ranges) {
    const uint16_t* keys16 = reinterpret_cast<const uint16_t*>(keys32);
    auto key = [&](uint32_t i) -> uint32_t { return K16 ? (uint32_t)keys16[i] : keys32[i]; };
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *count_ptr;
    const uint32_t i0 = t * 4u;
    uint32_t k[4];
    if (i0 + 3u < count) {
        if constexpr (K16) {
            const uint2 v = *reinterpret_cast<const uint2*>(keys16 + i0);
            k[0] = v.x & 0xFFFFu; k[1] = v.x >> 16; k[2] = v.y & 0xFFFFu; k[3] = v.y >> 16;
        } else {
            const uint4 v = *reinterpret_cast<const uint4*>(keys32 + i0);
            k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) k[j] = i0 + (uint32_t)j < count ? key(i0 + j) : 0xFFFFFFFFu;
    }
    uint32_t prev = __shfl_up(k[3], 1, 64);
    if (__lane_id() == 0) prev = (i0 > 0 && i0 - 1u < count) ? key(i0 - 1u) : 0xFFFFFFFFu;
    if (i0 >= count) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = i0 + (uint32_t)j;
        if (i >= count) break;
        const uint32_t pk = j ? k[j - 1] : prev;
        if (i == 0 || pk != k[j]) {
            ranges[k[j]].x = i;
            if (i > 0) ranges[pk].y = i;
        }
        if (i == count - 1u) ranges[k[j]].y = count;
    }
}

// One splat as the compositors keep it in LDS (4 x 16 bytes and 8 bytes,
// plus the quadrant mask).  The fragment power is kept as a quadratic
// polynomial in the pixel centre relative to the tile centre, (X, Y) = (x +
// 0.5 - xc, y + 0.5 - yc), in log2 units (scaled by log2 e) with the opacity
// folded in:
//   power' = A X^2 + B X Y + C Y^2 + E X + F Y + G
// (A, B, C from the conic; E, F, G from the splat centre relative to the
// tile centre and log2 o - log2(1/255)), so a pixel's blend step is five
// fused multiply-adds on per-pixel constants X^2, X Y, Y^2, X, Y, one
// v_exp_f32 for 255 alpha, and its fragment test a comparison with 0.  The
// covered pixels [i0, i1) x [j0, j1) ride in par.w and col.w as the packed
// corners (i0 | j0 << 16, (i1 - 1) | (j1 - 1) << 16), so one step reads
// three 16-byte records at one address.  The fourth record and the centre
// hold the record's covariance (a, b, c), opacity and centre, read only for
// the fragments whose fused test lies within par.z of the cutoff (under one
// wave step in a hundred): the reference's conic divisions and cutoff log
// are taken there, not at staging.
// (A native 4-float vector, so each record is one 128-bit LDS access.)
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
struct StagedSplats {
    f32x4_t* geo;   // A, B, C, E
    f32x4_t* par;   // F, G, decision margin (log2 units), rect (low corner bits)
    f32x4_t* col;   // r, g, b, rect (high corner bits)
    f32x4_t* exact; // a, b, c, opacity of the record: the exact test's inputs
    uint32_t* quads;  // 16x16 quadrant walk: bit q set when the splat can touch quadrant q; bit 4: slow step
    float2* ctr;    // cx, cy
};
// Inclusive 16-bit corners of a pixel rectangle [i0, i1) x [j0, j1) (image
// coordinates are below 2^16); an empty one becomes (0xFFFF, 0xFFFF) twice,
// which no pixel of an image equals.
__device__ __forceinline__ uint2 pack_rect16(int4 r) {
    if (r.x >= r.y || r.z >= r.w) return make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
    return make_uint2((uint32_t)r.x | ((uint32_t)r.z << 16), (uint32_t)(r.y - 1) | ((uint32_t)(r.w - 1) << 16));
}

__device__ __forceinline__ int4 unpack_rect16(uint2 r) {
    return make_int4((int)(r.x & 0xFFFFu), (int)(r.y & 0xFFFFu) + 1, (int)(r.x >> 16), (int)(r.y >> 16) + 1);
}

typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

// Whether the pixel p = x | y << 16 lies in the packed rectangle r: clamping
// both coordinates into it at once (packed 16-bit max and min) leaves p
// unchanged exactly when it does.
__device__ __forceinline__ bool in_rect16(uint32_t p, uint2 r) {
    const u16x2_t c = __builtin_elementwise_min(
        __builtin_elementwise_max(__builtin_bit_cast(u16x2_t, p), __builtin_bit_cast(u16x2_t, r.x)),
        __builtin_bit_cast(u16x2_t, r.y));
    return __builtin_bit_cast(uint32_t, c) == p;
}

constexpr float kLog2e = 1.4426950408889634f;
constexpr float kLog2AlphaMin = -7.99435343685886f;  // log2(1/255)
constexpr float kLog2AlphaCap255 = 7.97985386716375f;  // log2(0.99 * 255)

// The view and tile of a global tile id.
struct TileRef {
    int width, height, tx, ty;
    uint16_t* image;  // the view's RGBA16F image
    uint2 range;      // [begin, end) of the tile's entries
};

__device__ __forceinline__ TileRef tile_ref(const CompositeLaunch& a, uint32_t item) {
    uint32_t lo = 0u, hi = a.nviews - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (a.views[mid].tile_base <= item) lo = mid; else hi = mid - 1u;
    }
    const ViewInfo vi = a.views[lo];
    const uint32_t local = item - vi.tile_base;
    TileRef t;
    t.width = vi.width;
    t.height = vi.height;
    t.tx = (int)(local % (uint32_t)vi.tiles_x);
    t.ty = (int)(local / (uint32_t)vi.tiles_x);
    t.image = a.image + (size_t)vi.pixel_base * 4;
    t.range = a.ranges[item];
    return t;
}

// What the quadrant test needs of a staged splat: centre, covered pixels and
// the half widths of its alpha ellipse across the diagonals (diag: they
// apply).
struct StagedTerms {
    float cx, cy, wp, wm;
    int4 rect;
    bool slow, diag;
};

// Stage splat g into slot t of a T x T tile: conic from one reciprocal of the
// clamped determinant in log2 units, pixel rectangle, the covariance,
// opacity and centre the exact test needs (it forms the reference's conic
// (c, -b, a) / det, shaders/cull.glsl:273-275, and its fragment cutoff
// log((1/255) / max(o, 1e-8)), shaders/splat.glsl:70, itself), and the
// margin within which the fused test may decide differently from the
// reference's float test.
//
// Margin: with K the conic and kappa = lambda_max / lambda_min its condition
// number, |k0 dx^2| + |2 k1 dx dy| + |k2 dy^2| <= 2 kappa q at a pixel where
// the form is q; at the cutoff, q log2(e) / 2 = |log2((1/255) / o)|.  The
// float form with the staged coefficients A, B, C (log2 o folded in) and the
// reference's float power each carry about 8 roundings of those terms
// (staging, products, sums): together about 2e-6 kappa |cutoff|, plus a few
// 1e-6 from the log2 of o and the rounded cutoff; four times that is the
// first part.  The expanded polynomial adds its own rounding: forming E, F
// and G takes at most four roundings of a value no larger than S, the sum of
// the magnitudes of every product and partial sum involved (over the tile's
// pixel range), and its five fused multiply-adds one each, so it differs
// from the form above by at most about 9 eps S; 16 eps S is the second part.
// A splat whose conic is not positive definite in float always takes the
// reference's test.
//
// Slow step: the rectangle test and the 0.99 alpha cap can be left out of a
// splat's blend steps when neither can change a result.  Pixels outside the
// alpha ellipse's bounds fail the reference's fragment test, and the fused
// test decides as the reference does wherever it lies outside the margin
// (and takes the reference's test inside it), so when the quad's pixel
// bounds hold the ellipse's bounds, the rectangle excludes nothing the
// fragment test keeps.  The cap binds only where 255 alpha = 2^power'
// exceeds 0.99 * 255; power' is at most G' = log2 o - log2(1/255) at the
// centre, plus the margin.  Everything else takes the slow step.
struct SplatRecord {
    float4 r0, r1, r2;
};

__device__ __forceinline__ SplatRecord load_record(const CompositeLaunch& a, uint32_t g) {
    SplatRecord r;
    r.r0 = a.records[(size_t)g * kRecordVec4 + 0];
    r.r1 = a.records[(size_t)g * kRecordVec4 + 1];
    r.r2 = a.records[(size_t)g * kRecordVec4 + 2];
    return r;
}

template <int T>
__device__ __forceinline__ StagedTerms stage_splat(const CompositeLaunch& a, const TileRef& tr, const SplatRecord& rec,
                                                   uint32_t t, StagedSplats s) {
    const float4 r0 = rec.r0, r1 = rec.r1, r2 = rec.r2;
    float det = r1.x * r1.z - r1.y * r1.y;
    if (det < 1e-6f) det = 1e-6f;
    // the conic (c, -b, a) / det to about an ulp (the margin below covers
    // it); the reference's own divisions are taken in the exact test only
    const float rdet = __builtin_amdgcn_rcpf(det);
    const float k0 = r1.z * rdet, k1 = -r1.y * rdet, k2 = r1.x * rdet;
    const float log2o = __log2f(maxf(r0.z, 1e-8f));
    StagedTerms st;
    st.cx = r0.x;
    st.cy = r0.y;
    const float A = -0.5f * kLog2e * k0, B = -kLog2e * k1, C = -0.5f * kLog2e * k2;
    const float Dc = log2o - kLog2AlphaMin;
    const float cutoff = -Dc;  // log2((1/255) / o)
    const int4 quad = quad_pixel_bounds(r0.x, r0.y, r1.w, r2.w, tr.width, tr.height);
    float cgb;
    const bool bounded = ellipse_cutoff_bound(r1.x, r1.y, r1.z, r0.z, &cgb);
    int4 ell = quad;
    st.diag = bounded;
    st.wp = st.wm = 0.0f;
    if (bounded) {
        ell = ellipse_box(r0.x, r0.y, r1.x, r1.z, cgb);
        const float2 w = ellipse_diagonal_halfwidths(r1.x, r1.y, r1.z, cgb);
        st.wp = w.x;
        st.wm = w.y;
    }
    st.rect = quad;
    if (bounded && a.tight_rects != 0) {
        st.rect.x = max(quad.x, ell.x);
        st.rect.y = min(quad.y, ell.y);
        st.rect.z = max(quad.z, ell.z);
        st.rect.w = min(quad.w, ell.w);
    }
    st.rect = make_int4(max(st.rect.x, 0), min(st.rect.y, tr.width), max(st.rect.z, 0), min(st.rect.w, tr.height));
    // the expanded polynomial about the tile centre; S bounds the magnitudes
    // of every product and sum formed in its coefficients and its evaluation
    const float u = r0.x - (float)(tr.tx * T + T / 2), v = r0.y - (float)(tr.ty * T + T / 2);
    const float Au = A * u, Bv = B * v, Bu = B * u, Cv = C * v;
    const float E = -2.0f * Au - Bv, F = -2.0f * Cv - Bu;
    const float G = (Au + Bv) * u + Cv * v + Dc;
    constexpr float h = T / 2 - 0.5f;
    const float S = h * h * (fabsf(A) + fabsf(B) + fabsf(C)) +
                    h * (2.0f * fabsf(Au) + fabsf(Bv) + 2.0f * fabsf(Cv) + fabsf(Bu)) +
                    (fabsf(Au * u) + fabsf(Bv * u) + fabsf(Cv * v) + fabsf(Dc));
    const float ht = 0.5f * (k0 + k2), hd = 0.5f * (k0 - k2);
    const float sp = __builtin_sqrtf(hd * hd + k1 * k1);
    const float kappa = (ht + sp) * __builtin_amdgcn_rcpf(ht - sp);
    const float margin = ht - sp > 0.0f
                             ? 8e-6f * kappa * __builtin_fabsf(cutoff) + 3e-5f + 16.0f * 5.9604645e-8f * S
                             : __builtin_inff();
    const bool rect_free = bounded && ell.x >= quad.x && ell.y <= quad.y && ell.z >= quad.z && ell.w <= quad.w;
    const bool cap_free = Dc + margin + 1e-4f < kLog2AlphaCap255;
    st.slow = !(rect_free && cap_free);
    const uint2 rect = pack_rect16(st.rect);
    s.geo[t] = f32x4_t{A, B, C, E};
    s.par[t] = f32x4_t{F, G, margin, __uint_as_float(rect.x)};
    s.col[t] = f32x4_t{r2.x, r2.y, r2.z, __uint_as_float(rect.y)};
    s.exact[t] = f32x4_t{r1.x, r1.y, r1.z, r0.z};
    s.ctr[t] = make_float2(r0.x, r0.y);
    return st;
}

// A pixel's constants for the blend steps: its centre (x + 0.5, y + 0.5),
// its position x | y << 16, and the monomials of (X, Y) relative to the tile
// centre (all exact in float).
struct PixelTerms {
    float fx, fy, X, Y, XX, XY, YY;
    uint32_t pxy;
};

template <int T>
__device__ __forceinline__ PixelTerms pixel_terms(int px, int py, int tx, int ty) {
    PixelTerms p;
    p.fx = (float)px + 0.5f;
    p.fy = (float)py + 0.5f;
    p.X = (float)(px - tx * T) + (0.5f - (float)(T / 2));
    p.Y = (float)(py - ty * T) + (0.5f - (float)(T / 2));
    p.XX = p.X * p.X;
    p.XY = p.X * p.Y;
    p.YY = p.Y * p.Y;
    p.pxy = (uint32_t)px | ((uint32_t)py << 16);
    return p;
}

// One front-to-back blend step of a pixel: coverage (the pixel rectangle)
// and the fragment test select the alpha or 0, and a zero alpha leaves colour
// and transmittance bit-for-bit unchanged.  A saturated pixel (T < 1e-4)
// keeps blending until its whole wave is saturated; what it adds is below
// 1e-4 of a colour, and it only brings the result closer to blending every
// splat.  power' is the staged polynomial (log2 units relative to the
// cutoff), so 255 alpha = min(2^power', 255 * 0.99) and the fragment test
// alpha >= 1/255 is power' >= 0; tr holds T / 255.  Where power' lies within
// the splat's margin of 0 (a few fragments per thousand steps, on the rim of
// the alpha ellipse), the test is redone with the reference's float
// operations in source order (d = pixel centre - splat centre, power =
// -0.5 (k0 dx dx + 2 k1 dx dy + k2 dy dy), discard if power > 0 or power <
// log((1/255) / o); shaders/splat.glsl:84-96), so the set of blended
// fragments is the reference's.  Elsewhere the power > 0 discard cannot
// fire: the cull's +0.3 dilation keeps det >= 0.09, so the conic is positive
// definite, and its float evaluation stays within a relative 1e-2 of the
// exact form for every conditioning the dilation allows.  SLOW = false
// leaves out the rectangle test and the cap (see stage_splat).
template <bool SLOW>
__device__ __forceinline__ void blend_step(const PixelTerms& p, f32x4_t geo, f32x4_t par, f32x4_t col,
                                           const f32x4_t* exact, const float2* ctr, float& cr, float& cg, float& cb,
                                           float& tr) {
    const float power =
        fmaf(geo.x, p.XX, fmaf(geo.y, p.XY, fmaf(geo.z, p.YY, fmaf(geo.w, p.X, fmaf(par.x, p.Y, par.y)))));
    float alpha = __builtin_amdgcn_exp2f(power);
    if (SLOW) alpha = minf(alpha, kAlphaCap255);
    // power' >= margin: passes; below -margin: fails; between: the reference's test
    const bool lo = power >= -par.z;
    bool pass = lo;
    if (__builtin_expect(lo & !(power >= par.z), 0)) {
        // the reference's conic (shaders/cull.glsl:273-275) from the staged
        // covariance (a, b, c) and opacity, IEEE divisions
        const f32x4_t e = *exact;
        float det = e.x * e.z - e.y * e.y;
        if (det < 1e-6f) det = 1e-6f;
        const float k0 = e.z / det, k1 = -e.y / det, k2 = e.x / det, k1x2 = 2.0f * k1;
        const float arg = (1.0f / 255.0f) / maxf(e.w, 1e-8f);
        const float2 c = *ctr;
        const float dx = p.fx - c.x, dy = p.fy - c.y;
        const float pr = -0.5f * (k0 * dx * dx + k1x2 * dx * dy + k2 * dy * dy);
        // pr < cutoff, the cutoff being the deterministic float log of arg:
        // v_log_f32 (1 ulp of log2) times ln 2 lies within 4e-6 of it over
        // the cutoff's range (|cutoff| < 16), so where pr is 1e-5 or more
        // away from that estimate the two compare alike; closer, the
        // deterministic log decides
        const float est = __builtin_amdgcn_logf(arg) * 0.69314718f;
        bool below = pr < est;
        if (__builtin_expect(__builtin_fabsf(pr - est) < 1e-5f, 0)) below = pr < logf_(arg);
        pass = !(pr > 0.0f) & !below;
    }
    if (SLOW) pass &= in_rect16(p.pxy, make_uint2(__float_as_uint(par.w), __float_as_uint(col.w)));
    alpha = pass ? alpha : 0.0f;
    const float w = alpha * tr;
    cr = fmaf(col.x, w, cr);
    cg = fmaf(col.y, w, cg);
    cb = fmaf(col.z, w, cb);
    tr = fmaf(w, -kInv255, tr);
}

__device__ __forceinline__ void store_pixel(const TileRef& t, int px, int py, float cr, float cg, float cb,
                                            float tr) {
    size_t o = ((size_t)py * t.width + px) * 4;
    ushort4 h;
    h.x = __half_as_ushort(__float2half_rn(cr));
    h.y = __half_as_ushort(__float2half_rn(cg));
    h.z = __half_as_ushort(__float2half_rn(cb));
    h.w = __half_as_ushort(__float2half_rn(1.0f - tr * 255.0f));
    *reinterpret_cast<ushort4*>(t.image + o) = h;
}

template <int T>
struct TileShape {
    static constexpr int kPixels = T * T;
    static constexpr int kThreads = kPixels < 256 ? kPixels : 256;
    static constexpr int kPerThread = kPixels / kThreads;
    static constexpr int kBatch = kThreads;
};

// T x T tile, PPT pixels per thread; every pixel walks every staged splat of
// its tile (with the slow step).
template <int T>
__device__ __forceinline__ void composite_tile(const CompositeLaunch& a, uint32_t item, StagedSplats st) {
    using S = TileShape<T>;
    constexpr int NT = S::kThreads, PPT = S::kPerThread, BATCH = S::kBatch;
    const uint32_t tid = threadIdx.x;
    const TileRef t = tile_ref(a, item);

    int px[PPT], py[PPT];
    PixelTerms pt[PPT];
    float cr[PPT], cg[PPT], cb[PPT], tr[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        int p = (int)tid + k * NT;
        px[k] = t.tx * T + p % T;
        py[k] = t.ty * T + p / T;
        pt[k] = pixel_terms<T>(px[k], py[k], t.tx, t.ty);
        cr[k] = cg[k] = cb[k] = 0.0f;
        // pixels outside the image start saturated and are never written
        tr[k] = px[k] < t.width && py[k] < t.height ? kInv255 : 0.0f;
    }

    for (uint32_t base = t.range.x; base < t.range.y; base += BATCH) {
        bool live = false;
#pragma unroll
        for (int k = 0; k < PPT; ++k) live |= tr[k] >= kUMin;
        if (!__syncthreads_or(live)) break;
        uint32_t e = base + tid;
        if (e < t.range.y) stage_splat<T>(a, t, load_record(a, a.tile_vals[e]), tid, st);
        __syncthreads();
        const int cnt = (int)min((uint32_t)BATCH, t.range.y - base);
        if (__ballot(live)) {
            for (int j = 0; j < cnt; ++j) {
                const f32x4_t geo = st.geo[j], par = st.par[j], col = st.col[j];
#pragma unroll
                for (int k = 0; k < PPT; ++k)
                    blend_step<true>(pt[k], geo, par, col, st.exact + j, st.ctr + j, cr[k], cg[k], cb[k], tr[k]);
            }
        }
    }
    __syncthreads();  // the staged batch is free for the next tile

#pragma unroll
    for (int k = 0; k < PPT; ++k)
        if (px[k] < t.width && py[k] < t.height) store_pixel(t, px[k], py[k], cr[k], cg[k], cb[k], tr[k]);
}

// Whether the alpha ellipse of a staged splat can reach a pixel centre of
// the box [xa, xb] x [ya, yb] (pixel-centre offsets from the splat centre):
// the box must meet both diagonal slabs |dx + dy| <= wp and |dx - dy| <= wm
// that hold every pixel the fragment test can keep (with the rectangle,
// already cut to the ellipse's axis-aligned box, an octagon around the
// ellipse).  The sums and differences of the box corners range over [xa +
// ya, xb + yb] and [xa - yb, xb - ya].
__device__ __forceinline__ bool ellipse_meets_box(const StagedTerms& s, float xa, float xb, float ya, float yb) {
    return (xa + ya <= s.wp) & (xb + yb >= -s.wp) & (xa - yb <= s.wm) & (xb - ya >= -s.wm);
}

// T x T tile (16 or 32), (T / 8)^2 wave64s, each wave owning one 8x8
// quadrant.  Every splat of an LDS batch is tested against the quadrant's
// pixel range with one ballot per 64 splats; the wave then walks only the
// overlapping splats (scalar bit iteration) with one LDS wait per splat, and
// stops as soon as all 64 of its pixels are saturated.  Whether a splat takes
// the slow blend step is a second ballot, so the choice is a scalar branch.
template <int T>
struct QuadShape {
    static constexpr int kSide = T / 8;             // quadrants a side
    static constexpr int kQuads = kSide * kSide;    // waves
    static constexpr int kThreads = 64 * kQuads;    // also the batch
    static constexpr uint32_t kSlowBit = 1u << kQuads;
};

template <int T>
__device__ __forceinline__ void composite_tile_quad(const CompositeLaunch& a, uint32_t item, StagedSplats st) {
    using Q = QuadShape<T>;
    constexpr int BATCH = Q::kThreads, SIDE = Q::kSide;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const TileRef t = tile_ref(a, item);

    const int qx0 = t.tx * T + (int)(wave % SIDE) * 8, qy0 = t.ty * T + (int)(wave / SIDE) * 8;
    const int px = qx0 + (int)(lane & 7u), py = qy0 + (int)(lane >> 3);
    PixelTerms pt = pixel_terms<T>(px, py, t.tx, t.ty);
    // kept in registers through the walk rather than recomputed in it
    asm volatile("" : "+v"(pt.XX), "+v"(pt.XY), "+v"(pt.YY), "+v"(pt.X), "+v"(pt.Y));
    const bool valid = px < t.width && py < t.height;
    // pixels outside the image start saturated and are never written
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, tr = valid ? kInv255 : 0.0f;  // tr: T / 255

    // a batch's list entries are loaded before the previous batch's walk, so
    // its staging waits for the splat records only
    const uint32_t end = t.range.y;
    uint32_t g_next = t.range.x + tid < end ? a.tile_vals[t.range.x + tid] : 0u;
    for (uint32_t base = t.range.x; base < end; base += BATCH) {
        if (!__syncthreads_or(tr >= kUMin)) break;
        uint32_t e = base + tid;
        if (e < end) {
            const StagedTerms s = stage_splat<T>(a, t, load_record(a, g_next), tid, st);
            // the quadrants the splat can touch, tested once here for all
            // waves (the tests share the splat's terms): the rectangle's
            // pixel-centre offsets in each column and row of quadrants
            uint32_t quads = s.slow ? Q::kSlowBit : 0u;
            float xa[SIDE], xb[SIDE], ya[SIDE], yb[SIDE];
            bool ox[SIDE], oy[SIDE];
#pragma unroll
            for (int i = 0; i < SIDE; ++i) {
                const int x0 = t.tx * T + 8 * i, y0 = t.ty * T + 8 * i;
                ox[i] = s.rect.x < x0 + 8 && s.rect.y > x0;
                oy[i] = s.rect.z < y0 + 8 && s.rect.w > y0;
                xa[i] = (float)max(x0, s.rect.x) + 0.5f - s.cx;
                xb[i] = (float)min(x0 + 8, s.rect.y) - 0.5f - s.cx;
                ya[i] = (float)max(y0, s.rect.z) + 0.5f - s.cy;
                yb[i] = (float)min(y0 + 8, s.rect.w) - 0.5f - s.cy;
            }
            const bool ellipse_test = a.quadrant_cull == 1 && s.diag;
#pragma unroll
            for (int q = 0; q < Q::kQuads; ++q) {
                const int i = q % SIDE, j = q / SIDE;
                bool hit = ox[i] && oy[j];
                if (ellipse_test) hit = hit && ellipse_meets_box(s, xa[i], xb[i], ya[j], yb[j]);
                quads |= hit ? 1u << q : 0u;
            }
            st.quads[tid] = quads;
        }
        __syncthreads();
        if (base + BATCH + tid < end) g_next = a.tile_vals[base + BATCH + tid];
        const uint32_t cnt = min((uint32_t)BATCH, end - base);
        if (!__ballot(tr >= kUMin)) continue;
        for (uint32_t chunk = 0; chunk < cnt; chunk += 64) {
            // which of these 64 splats touch this wave's quadrant, and which
            // of them take the slow step
            const uint32_t j = chunk + lane;
            const uint32_t q = j < cnt ? st.quads[j] : 0u;
            uint64_t mask = __ballot((q >> wave) & 1u);
            const uint64_t slow = __ballot((q & Q::kSlowBit) != 0u);
            while (mask) {
                const uint32_t b = (uint32_t)__builtin_ctzll(mask);
                const uint32_t k = chunk + b;
                mask &= mask - 1ull;
                const f32x4_t geo = st.geo[k], par = st.par[k], col = st.col[k];
                if ((slow >> b) & 1ull)
                    blend_step<true>(pt, geo, par, col, st.exact + k, st.ctr + k, cr, cg, cb, tr);
                else
                    blend_step<false>(pt, geo, par, col, st.exact + k, st.ctr + k, cr, cg, cb, tr);
            }
            if (!__ballot(tr >= kUMin)) break;
        }
    }
    __syncthreads();  // the staged batch is free for the next tile

    if (valid) store_pixel(t, px, py, cr, cg, cb, tr);
}

// Reference blend arithmetic on 16x16 tiles.  The reference draws the
// depth-sorted quads far to near into an RGBA16F target with ONE /
// ONE_MINUS_SRC_ALPHA blending (src/renderer.cpp:292-299): each fragment's
// (rgb alpha, alpha) is added to (1 - alpha) times the stored value, and the
// sum is stored as half precision again (round to nearest even) before the
// next fragment reads it.  Here each pixel walks its tile list from the back
// (the list is front to back), evaluates the fragment shader
// (shaders/splat.glsl:84-104) with the reference's operations in source
// order (conic = (c, -b, a) / det from the record, power =
// -0.5 (k0 dx dx + 2 k1 dx dy + k2 dy dy), the deterministic log and exp) and
// rounds every channel to half after every blend.  No pixel stops early: a
// fragment behind a saturated one still changes the rounding of the stored
// value.  Staged per splat: geo (cx, cy, k0, 2 k1), par (k2, rect low
// corners, cutoff, opacity), col (r, g, b, rect high corners).
__device__ __forceinline__ float round_half(float v) { return __half2float(__float2half_rn(v)); }

__device__ __forceinline__ void composite_tile_ref(const CompositeLaunch& a, uint32_t item, StagedSplats st) {
    constexpr int BATCH = 256;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const TileRef t = tile_ref(a, item);

    const int qx0 = t.tx * 16 + (int)(wave & 1u) * 8, qy0 = t.ty * 16 + (int)(wave >> 1) * 8;
    const int px = qx0 + (int)(lane & 7u), py = qy0 + (int)(lane >> 3);
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const uint32_t pxy = (uint32_t)px | ((uint32_t)py << 16);
    const bool valid = px < t.width && py < t.height;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;  // the black (0, 0, 0, 0) clear

    for (uint32_t hi = t.range.y; hi > t.range.x;) {
        const uint32_t lo = hi - t.range.x > (uint32_t)BATCH ? hi - (uint32_t)BATCH : t.range.x;
        const uint32_t cnt = hi - lo;
        __syncthreads();  // the previous batch has been walked
        if (tid < cnt) {
            const uint32_t g = a.tile_vals[lo + tid];
            const float4 r0 = a.records[(size_t)g * kRecordVec4 + 0];
            const float4 r1 = a.records[(size_t)g * kRecordVec4 + 1];
            const float4 r2 = a.records[(size_t)g * kRecordVec4 + 2];
            float det = r1.x * r1.z - r1.y * r1.y;
            if (det < 1e-6f) det = 1e-6f;
            const float k0 = r1.z / det, k1 = -r1.y / det, k2 = r1.x / det;
            const uint2 rect = pack_rect16(splat_pixel_rect(r0.x, r0.y, r1.x, r1.y, r1.z, r1.w, r2.w, r0.z, t.width,
                                                            t.height, a.tight_rects != 0));
            st.geo[tid] = f32x4_t{r0.x, r0.y, k0, 2.0f * k1};
            st.par[tid] = f32x4_t{k2, __uint_as_float(rect.x), logf_((1.0f / 255.0f) / maxf(r0.z, 1e-8f)), r0.z};
            st.col[tid] = f32x4_t{r2.x, r2.y, r2.z, __uint_as_float(rect.y)};
            const int4 r = unpack_rect16(rect);
            uint32_t quads = 0u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int x0 = t.tx * 16 + (q & 1) * 8, y0 = t.ty * 16 + (q >> 1) * 8;
                quads |= (r.x < x0 + 8 && r.y > x0 && r.z < y0 + 8 && r.w > y0) ? 1u << q : 0u;
            }
            st.quads[tid] = quads;
        }
        __syncthreads();
        // chunks of 64 staged splats, last first; within a chunk the highest
        // set bit of the quadrant ballot first
        for (int chunk = (int)((cnt - 1u) & ~63u); chunk >= 0; chunk -= 64) {
            const uint32_t j = (uint32_t)chunk + lane;
            uint64_t mask = __ballot(j < cnt && ((st.quads[j] >> wave) & 1u));
            while (mask) {
                const int bit = 63 - __builtin_clzll(mask);
                mask &= ~(1ull << bit);
                const uint32_t k = (uint32_t)chunk + (uint32_t)bit;
                const f32x4_t geo = st.geo[k], par = st.par[k], col = st.col[k];
                const uint2 r = make_uint2(__float_as_uint(par.y), __float_as_uint(col.w));
                const float dx = fx - geo.x, dy = fy - geo.y;
                const float power = -0.5f * (geo.z * dx * dx + geo.w * dx * dy + par.x * dy * dy);
                const bool take = in_rect16(pxy, r) & !(power > 0.0f) & !(power < par.z);
                if (take) {
                    const float alpha = minf(par.w * expf_(power), 0.99f);
                    const float keep = 1.0f - alpha;
                    cr = round_half(col.x * alpha + cr * keep);
                    cg = round_half(col.y * alpha + cg * keep);
                    cb = round_half(col.z * alpha + cb * keep);
                    ca = round_half(alpha + ca * keep);
                }
            }
        }
        hi = lo;
    }
    __syncthreads();  // the staged batch is free for the next tile

    if (valid) {
        const size_t o = ((size_t)py * t.width + px) * 4;
        ushort4 h;
        h.x = __half_as_ushort(__float2half_rn(cr));
        h.y = __half_as_ushort(__float2half_rn(cg));
        h.z = __half_as_ushort(__float2half_rn(cb));
        h.w = __half_as_ushort(__float2half_rn(ca));
        *reinterpret_cast<ushort4*>(t.image + o) = h;
    }
}

// StochasticSplats sample (shaders/splat.glsl:107-146): every covered
// fragment that passes the alpha test is kept when hash13(pixel centre,
// id + frame_seed) < alpha, and the kept fragment with the smallest 24-bit
// window depth wins, ties to the lowest index (the reference draws the
// splats in index order with a LESS depth test).  Along a tile list the
// depth does not decrease, so a pixel is settled once the current splat lies
// more than a few D24 steps behind its winner.  The fragment arithmetic is
// the reference's, in source order, with the deterministic exp and log (the
// cutoff's log in the binary32 FMA form), so keep decisions are those of the
// CPU restatement.
__device__ __forceinline__ float fract_(float x) { return x - __builtin_floorf(x); }

__device__ __forceinline__ float hash13(float x, float y, float z) {
    float px = fract_(x * 0.1031f), py = fract_(y * 0.1031f), pz = fract_(z * 0.1031f);
    const float d = (px * (py + 33.33f) + py * (pz + 33.33f)) + pz * (px + 33.33f);
    px = px + d;
    py = py + d;
    pz = pz + d;
    return fract_((px + py) * pz);
}

__device__ __forceinline__ uint32_t depth24(float ndc_z, float clip_z_01) {
    float d = clip_z_01 > 0.5f ? ndc_z : ndc_z * 0.5f + 0.5f;
    d = minf(maxf(d, 0.0f), 1.0f);
    return (uint32_t)__builtin_rint((double)d * 16777215.0);
}

__global__ void __launch_bounds__(256) stochastic_kernel(StochasticLaunch sl) {
    constexpr int BATCH = 256;
    constexpr uint32_t kDepthSlack = 4u;  // D24 steps: float ndc_z is monotone in depth up to rounding
    // covered pixels as packed inclusive 16-bit corners (x, y; z, w unused:
    // the same 16-byte stride as the other records, so one address serves all)
    __shared__ u32x4_t s_rect[BATCH];
    __shared__ f32x4_t s_geo[BATCH];  // cx, cy, k0, 2 k1
    __shared__ f32x4_t s_par[BATCH];  // k2, opacity, cutoff, D24 depth (bits)
    __shared__ f32x4_t s_col[BATCH];  // r, g, b (half precision), id (bits)
    const CompositeLaunch& a = sl.tiles;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t item = xcd_remap(blockIdx.x, a.total_tiles);
    const TileRef t = tile_ref(a, item);

    const int qx0 = t.tx * 16 + (int)(wave & 1u) * 8, qy0 = t.ty * 16 + (int)(wave >> 1) * 8;
    const int px = qx0 + (int)(lane & 7u), py = qy0 + (int)(lane >> 3);
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const uint32_t pxy = (uint32_t)px | ((uint32_t)py << 16);
    const bool valid = px < t.width && py < t.height;
    uint32_t best_q = 16777215u, best_id = 0xFFFFFFFFu;
    float cr = sl.background, cg = sl.background, cb = sl.background;
    bool kept = false, open = valid;

    for (uint32_t base = t.range.x; base < t.range.y; base += BATCH) {
        if (!__syncthreads_or(open)) break;
        const uint32_t e = base + tid;
        if (e < t.range.y) {
            const uint32_t g = a.tile_vals[e];
            const float4 r0 = a.records[(size_t)g * kRecordVec4 + 0];
            const float4 r1 = a.records[(size_t)g * kRecordVec4 + 1];
            const float4 r2 = a.records[(size_t)g * kRecordVec4 + 2];
            float det = r1.x * r1.z - r1.y * r1.y;
            if (det < 1e-6f) det = 1e-6f;
            const float k0 = r1.z / det, k1 = -r1.y / det, k2 = r1.x / det;
            const uint2 rect = pack_rect16(splat_pixel_rect(r0.x, r0.y, r1.x, r1.y, r1.z, r1.w, r2.w, r0.z, t.width,
                                                            t.height, false, sl.jitter_x, sl.jitter_y));
            s_rect[tid] = u32x4_t{rect.x, rect.y, 0u, 0u};
            s_geo[tid] = f32x4_t{r0.x + sl.jitter_x, r0.y + sl.jitter_y, k0, 2.0f * k1};
            s_par[tid] = f32x4_t{k2, r0.z, logf_fma((1.0f / 255.0f) / maxf(r0.z, 1e-8f)),
                                 __uint_as_float(depth24(r0.w, sl.clip_z_01))};
            s_col[tid] = f32x4_t{__half2float(__float2half_rn(r2.x)), __half2float(__float2half_rn(r2.y)),
                                 __half2float(__float2half_rn(r2.z)), __uint_as_float(g)};
        }
        __syncthreads();
        const uint32_t cnt = min((uint32_t)BATCH, t.range.y - base);
        if (!__ballot(open)) continue;
        for (uint32_t chunk = 0; chunk < cnt; chunk += 64) {
            const uint32_t j = chunk + lane;
            bool hit = false;
            if (j < cnt) {
                const int4 r = unpack_rect16(make_uint2(s_rect[j].x, s_rect[j].y));
                hit = r.x < qx0 + 8 && r.y > qx0 && r.z < qy0 + 8 && r.w > qy0;
            }
            uint64_t mask = __ballot(hit);
            while (mask) {
                const uint32_t k = chunk + (uint32_t)__builtin_ctzll(mask);
                mask &= mask - 1ull;
                const uint2 r = make_uint2(s_rect[k].x, s_rect[k].y);
                const f32x4_t geo = s_geo[k], par = s_par[k], col = s_col[k];
                const uint32_t q = __float_as_uint(par.w), id = __float_as_uint(col.w);
                const float dx = fx - geo.x, dy = fy - geo.y;
                const float power = -0.5f * (geo.z * dx * dx + geo.w * dx * dy + par.x * dy * dy);
                const bool covered = in_rect16(pxy, r) & !(power > 0.0f) & !(power < par.z);
                const bool nearer = q < best_q || (q == best_q && id < best_id);
                const bool cand = covered & nearer;
                // the wave tests its candidate pixels together, without
                // per-lane branches (the scalar unit is shared by the CU)
                if (__ballot(cand)) {
                    // keep iff h < alpha = min(o exp(power), 0.99).  The
                    // hardware exp2 brackets alpha within 2e-5 relative: its
                    // 1-ulp error, the rounding of power * log2(e) (1.2e-7
                    // |power| relative) and the float rounding of exp stay
                    // below 5e-6 while power >= -40, and the fragment test
                    // keeps power >= log(1 / (255 o)) >= -5.6 for o <= 1.
                    // Only a hash inside the bracket needs the deterministic
                    // exp, so keep decisions match the CPU restatement.
                    const float h = hash13(fx, fy, (float)id + sl.frame_seed);
                    const float oe = par.y * __builtin_amdgcn_exp2f(power * 1.44269504f);
                    bool keep = cand & (h < minf(oe * (1.0f - 2e-5f), 0.99f));
                    const bool unsure = cand & !keep & (h < minf(oe * (1.0f + 2e-5f), 0.99f));
                    if (__ballot(unsure)) {
                        const bool exact = h < minf(par.y * expf_(power), 0.99f);
                        keep = unsure ? exact : keep;
                    }
                    best_q = keep ? q : best_q;
                    best_id = keep ? id : best_id;
                    cr = keep ? col.x : cr;
                    cg = keep ? col.y : cg;
                    cb = keep ? col.z : cb;
                    kept = kept | keep;
                }
                open = open & !(q > best_q + kDepthSlack);
            }
            if (!__ballot(open)) break;
        }
    }

    if (valid) {
        const size_t o = (size_t)py * t.width + px;
        sl.sample[o] = kept ? make_float4(cr, cg, cb, 1.0f) : make_float4(cr, cg, cb, 0.0f);
        sl.depth[o] = kept ? (float)((double)best_q / 16777215.0) : 1.0f;
    }
}

// Next tile of a persistent block: blocks are dealt to the 8 XCDs
// round-robin, each XCD drains its own contiguous band of tiles (its L2 sees
// neighbouring tiles), then takes tiles from the other bands.  A plain
// (agent-coherent) read skips exhausted bands without an atomic.
__device__ __forceinline__ uint32_t next_tile(const CompositeLaunch& a, uint32_t xcd) {
    const uint32_t total = a.total_tiles;
    for (uint32_t k = 0; k < 8u; ++k) {
        const uint32_t band = (xcd + k) & 7u;
        const uint32_t lo = (uint32_t)(((uint64_t)total * band) / 8u);
        const uint32_t hi = (uint32_t)(((uint64_t)total * (band + 1u)) / 8u);
        if (hi == lo) continue;
        if (__hip_atomic_load(&a.queue[band], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi - lo) continue;
        const uint32_t c = atomicAdd(&a.queue[band], 1u);
        if (c < hi - lo) return lo + c;
    }
    return total;
}

template <int T, bool QUAD>
struct CompositeKernelShape {
    static constexpr int kThreads = QUAD ? QuadShape<T>::kThreads : TileShape<T>::kThreads;
};

// The reference-arithmetic compositor over 16x16 tiles, launched like
// composite_kernel below.
__global__ void __launch_bounds__(256) composite_ref_kernel(CompositeLaunch a) {
    __shared__ f32x4_t s_geo[256];
    __shared__ f32x4_t s_par[256];
    __shared__ f32x4_t s_col[256];
    __shared__ uint32_t s_quads[256];
    __shared__ uint32_t s_item;
    const StagedSplats st{s_geo, s_par, s_col, nullptr, s_quads, nullptr};
    if (a.persistent_blocks == 0u) {
        composite_tile_ref(a, xcd_remap(blockIdx.x, a.total_tiles), st);
        return;
    }
    const uint32_t xcd = blockIdx.x & 7u;
    for (;;) {
        if (threadIdx.x == 0) s_item = next_tile(a, xcd);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= a.total_tiles) break;
        composite_tile_ref(a, item, st);
    }
}

// One block per tile (persistent_blocks == 0, XCD-aware tile order) or a
// persistent grid pulling tiles from the per-XCD queues.
template <int T, bool QUAD>
// The quadrant kernel at eight waves per SIMD (at most 64 VGPRs): its blend
// loop is VALU-issue bound and the other waves hide its LDS reads.
__global__ void __launch_bounds__((CompositeKernelShape<T, QUAD>::kThreads), (QUAD ? 8 : 1))
    composite_kernel(CompositeLaunch a) {
    constexpr int BATCH = CompositeKernelShape<T, QUAD>::kThreads;
    __shared__ f32x4_t s_geo[BATCH];
    __shared__ f32x4_t s_par[BATCH];
    __shared__ f32x4_t s_col[BATCH];
    __shared__ f32x4_t s_exact[BATCH];
    __shared__ float2 s_ctr[BATCH];
    __shared__ uint32_t s_quads[QUAD ? BATCH : 1];
    __shared__ uint32_t s_item;
    const StagedSplats st{s_geo, s_par, s_col, s_exact, s_quads, s_ctr};
    if (a.persistent_blocks == 0u) {
        const uint32_t item = xcd_remap(blockIdx.x, a.total_tiles);
        if constexpr (QUAD) composite_tile_quad<T>(a, item, st); else composite_tile<T>(a, item, st);
        return;
    }
    const uint32_t xcd = blockIdx.x & 7u;
    for (;;) {
        if (threadIdx.x == 0) s_item = next_tile(a, xcd);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= a.total_tiles) break;
        if constexpr (QUAD) composite_tile_quad<T>(a, item, st); else composite_tile<T>(a, item, st);
    }
}

}  // namespace

hipError_t launch_bin(const BinLaunch& a, hipStream_t s) {
    const uint32_t blocks = (a.capacity + 255u) / 256u;
    if (!blocks) return hipSuccess;
    if (a.sorted_rects && a.jittered) return hipErrorInvalidValue;
    if (a.clear_words && !a.clear) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bin_count_kernel, dim3(blocks), dim3(256), 0, s, a);
    // the tile sort works on min(total, capacity) entries
    hipError_t e = launch_scan_counts(a.block_counts, blocks, a.block_offsets, &a.counters->bin.dup_total, nullptr,
                                      nullptr, s, a.dup_capacity, &a.counters->bin.dup_count);
    if (e != hipSuccess) return e;
    // pair positions stay below the capacity
    const bool o32 = (uint64_t)a.dup_capacity * 4u <= 0xFFFFFFFFull && !a.wide_index;
    if (o32 && a.key16)
        hipLaunchKernelGGL((bin_emit_kernel<true, true>), dim3(blocks), dim3(256), 0, s, a);
    else if (o32)
        hipLaunchKernelGGL((bin_emit_kernel<true, false>), dim3(blocks), dim3(256), 0, s, a);
    else if (a.key16)
        hipLaunchKernelGGL((bin_emit_kernel<false, true>), dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((bin_emit_kernel<false, false>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tile_ranges(const uint32_t* keys, int key16, const uint32_t* count, uint32_t capacity, uint2* ranges,
                              hipStream_t s) {
    if (!capacity) return hipSuccess;
    if (reinterpret_cast<uintptr_t>(keys) % 16) return hipErrorInvalidValue;
    const uint32_t threads = (capacity + 3u) / 4u;
    if (key16)
        hipLaunchKernelGGL(tile_ranges_kernel<true>, dim3((threads + 255u) / 256u), dim3(256), 0, s, keys, count, ranges);
    else
        hipLaunchKernelGGL(tile_ranges_kernel<false>, dim3((threads + 255u) / 256u), dim3(256), 0, s, keys, count,
                           ranges);
    return hipGetLastError();
}

hipError_t launch_stochastic(const StochasticLaunch& a, hipStream_t s) {
    if (!a.tiles.total_tiles) return hipSuccess;
    if (a.tiles.tile != 16 || a.tiles.nviews != 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stochastic_kernel, dim3(a.tiles.total_tiles), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_composite(const CompositeLaunch& a, hipStream_t s) {
    if (!a.total_tiles || !a.nviews) return hipSuccess;
    const uint32_t grid = a.persistent_blocks ? a.persistent_blocks : a.total_tiles;
    if (a.reference_blend) {
        if (a.tile != 16) return hipErrorInvalidValue;
        hipLaunchKernelGGL(composite_ref_kernel, dim3(grid), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    switch (a.tile) {
        case 8:
            hipLaunchKernelGGL((composite_kernel<8, false>), dim3(grid), dim3(TileShape<8>::kThreads), 0, s, a);
            break;
        case 16:
            if (a.quadrant_cull)
                hipLaunchKernelGGL((composite_kernel<16, true>), dim3(grid), dim3(QuadShape<16>::kThreads), 0, s, a);
            else
                hipLaunchKernelGGL((composite_kernel<16, false>), dim3(grid), dim3(TileShape<16>::kThreads), 0, s, a);
            break;
        case 32:
            if (a.quadrant_cull)
                hipLaunchKernelGGL((composite_kernel<32, true>), dim3(grid), dim3(QuadShape<32>::kThreads), 0, s, a);
            else
                hipLaunchKernelGGL((composite_kernel<32, false>), dim3(grid), dim3(TileShape<32>::kThreads), 0, s, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
