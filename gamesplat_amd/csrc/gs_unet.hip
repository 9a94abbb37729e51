// gs_unet.hip - kernels of the learned denoiser: a 4-level U-Net in bf16 on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, fp32 accumulation).
//
// Tensors are NHWC bf16 views (pointer, channel stride, channel offset), so
// the decoder's concatenations [upsampled, skip] are two channel slices of
// one buffer that their producers write directly.  Every channel count and
// offset is a multiple of 8: a lane moves 8 channels (16 bytes) at a time.
//
// MFMA 16x16x32 operand layout (lane l of the wave64): A row l % 16 holds
// k = 8 (l / 16) .. +7; B column l % 16 holds the same k; the accumulator
// holds D[4 (l / 16) + i][l % 16], i = 0..3.
//
// 3x3 convolution as implicit GEMM: M = output pixels, N = output channels,
// K = 9 taps x input channels (tap = (dy + 1) * 3 + dx + 1, channels
// fastest, zero padded to a multiple of 32).  Each wave owns 32 pixels of
// one row (two 16-pixel row groups) and every output channel; a block is 4
// rows of 32 pixels.  The input gradient uses the same kernel with the
// weights transposed per tap and the taps reversed.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>

#include "gs_internal.h"

namespace gs {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4v_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// float -> bf16, round to nearest, ties to even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8_t as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8_t, u); }

// Block tile: 32 pixels wide, 4 * RG / 2 rows (RG 16-pixel row groups per
// wave).  The input tile with a one-pixel halo is staged in LDS once (channel
// pitch + 8 spreads the 16-byte reads of consecutive pixels over the banks).
// Channels taken from the nearest 2x upsample of a half-size tensor are
// staged at half size (the (TR / 2 + 1) x (TW / 2 + 1) half-size pixels the
// tile's halo covers, a quarter of the bytes), so a wide decoder input holds
// more blocks per CU.  A table maps each group of 8 k to its LDS offset from
// the lane's pixel: (tap row * TW + tap column) * pitch + channel at full
// size; at half size the offset depends on the parities of the lane's row and
// column, so each group has four entries (bit 31 marks the half-size part).
// The MFMA's A operand is the weights (M = output channel, through the L1/L2
// caches) and its B operand the input pixels (N = pixel, from LDS), so a
// lane's accumulator holds four consecutive output channels of one pixel: one
// packed 8-byte store each.  k beyond 9 * cin (the zero padding of K) reads
// the lane's own pixel: the padded weights are zero, so those products add
// nothing.
constexpr int kConvTileW = 32;

__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
    typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
    const bf16x2_t v = {(__bf16)lo, (__bf16)hi};  // round to nearest even
    return __builtin_bit_cast(uint32_t, v);
}

// Value of lane ^ 1 (the neighbouring pixel of a row group).
__device__ __forceinline__ uint32_t lane_xor1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
}

// Stages the tile pixels p0, p0 + ppass, ... (below TR * TW) of one channel
// group into LDS at pitch cp, eight loads in flight per thread.  src points at
// the group's channels of the image; pixel (yy, xx) of the tile is read at
// ((yy >> sh) * sw + (xx >> sh)) * ss.  With CHECK, pixels outside the h x w
// image are skipped (the caller has zeroed the tile).  A skipped element
// loads the image's first pixel and writes it to the 16-byte dummy slot
// s_dummy: every load and store is unconditional, so all loads of a batch
// are issued before the first store waits.
template <int TR, int TW, bool CHECK>
__device__ __forceinline__ void stage_tile(uint16_t* s_dst, int cp, uint16_t* s_dummy, const uint16_t* src, int sh,
                                           int sw, int ss, int y0, int x0, int h, int w, int p0, int ppass) {
    constexpr int kBatch = 8, kTilePix = TR * TW;
    int tx = p0 % TW, ty = p0 / TW;
    const int dtx = ppass % TW, dty = ppass / TW;
    for (int pb = p0; pb < kTilePix; pb += ppass * kBatch) {
        typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
        u32x4_t v[kBatch];
        uint16_t* dst[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int yy = y0 + ty - 1, xx = x0 + tx - 1, p = pb + j * ppass;
            bool ok = p < kTilePix;
            if (CHECK) ok = ok && (unsigned)yy < (unsigned)h && (unsigned)xx < (unsigned)w;
            dst[j] = ok ? s_dst + p * cp : s_dummy;
            v[j] = *reinterpret_cast<const u32x4_t*>(src + (ok ? ((yy >> sh) * sw + (xx >> sh)) * ss : 0));
            tx += dtx;
            ty += dty;
            if (tx >= TW) { tx -= TW; ++ty; }
        }
        // keep the scheduler from pairing each load with its store
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < kBatch; ++j) *reinterpret_cast<u32x4_t*>(dst[j]) = v[j];
    }
}

// Half-size rows and columns of a block's upsampled part: the full-size
// halo rows y0 - 1 .. y0 + BR (y0 even) lie on half-size rows y0 / 2 - 1 ..
// y0 / 2 + BR / 2.
template <int BR>
struct HalfTile {
    static constexpr int kRows = BR / 2 + 2, kCols = kConvTileW / 2 + 2;
};

// LDS element offset, from the lane's pixel, of k-group kgp (8 channels of one
// tap) for a lane whose output row and column have parities (ya, xb); bit 31
// set: the group lies in the half-size part.  k past 9 * cin: offset 0.
__device__ __forceinline__ uint32_t conv_tab_entry(int kgp, int ya, int xb, int cin, int split, int cpf, int cph,
                                                   int tw, int hw) {
    const int k = kgp * 8;
    if (k >= 9 * cin) return 0u;
    const int tap = k / cin, ci = k - tap * cin, dy = tap / 3, dx = tap % 3;
    if (ci >= split) return (uint32_t)((dy * tw + dx) * cpf + (ci - split));
    // half-size row of input row (lane row + dy - 1): (row >> 1) + ((ya + dy - 1) >> 1) + 1
    // relative to the half tile's first row, and likewise for the column
    const int oy = ((ya + dy + 1) >> 1), ox = ((xb + dx + 1) >> 1);
    return 0x80000000u | (uint32_t)((oy * hw + ox) * cph + ci);
}

// HALF: channels [0, split) come from the half-size tensor x2 (kept at half
// size in LDS); otherwise every channel comes from x and the table holds one
// entry per k-group.
// The 32-channel kernel without a half-size part (whose input tiles allow
// five blocks per CU) is held to 96 VGPRs, five waves per SIMD.
template <int NT, int RG, bool HALF, bool HEAD>
__global__ void __launch_bounds__(256, (NT == 2 && !HALF ? 5 : 1)) conv3x3_kernel(ConvLaunch a) {
    extern __shared__ __align__(16) uint16_t s_in[];
    constexpr int RPW = RG / 2;            // rows per wave
    constexpr int BR = 4 * RPW;            // rows per block
    constexpr int TW = kConvTileW + 2, TR = BR + 2;
    constexpr int HR = HalfTile<BR>::kRows, HW = HalfTile<BR>::kCols;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int y0 = (int)blockIdx.y * BR;
    const int x0 = (int)blockIdx.x * kConvTileW;
    const int n = (int)blockIdx.z;
    const int cin = a.cin;
    const int split = HALF ? a.split : 0;  // channels [0, split) at half size
    const int cpf = cin - split + 8, cph = split + 8;
    const int hbase = TR * TW * cpf;       // element offset of the half-size part
    const int words_in = hbase + (split ? HR * HW * cph : 0);
    uint16_t* s_dummy = s_in + words_in;   // 16 bytes
    uint32_t* s_tab = reinterpret_cast<uint32_t*>(s_dummy + 8);

    // k-group table: kpad / 8 <= 256 groups; with HALF four entries each,
    // ordered [group][column parity][row parity]
    constexpr int kEntries = HALF ? 4 : 1;
    for (int i = tid; i < a.kpad / 8 * kEntries; i += 256)
        s_tab[i] = HALF ? conv_tab_entry(i >> 2, i & 1, (i >> 1) & 1, cin, split, cpf, cph, TW, HW)
                        : conv_tab_entry(i, 0, 0, cin, 0, cpf, cph, TW, HW);

    // stage rows y0-1 .. y0+BR, columns x0-1 .. x0+32 (zero outside the
    // image) of the full-size channels, and the half-size pixels under them
    // of the upsampled ones.  Thread tid owns channel group tid % groups of
    // every ppass-th tile pixel from tid / groups on.  Offsets are 32-bit
    // within one image (the host checks h * w * stride < 2^31).  Blocks
    // clear of the image border (then the half-size part is too) skip the
    // bounds checks; the others zero the tile first.
    {
        const bool interior = y0 >= 1 && x0 >= 1 && y0 + BR < a.h && x0 + kConvTileW < a.w;
        if (!interior) {
            for (int i = tid; i < words_in / 8; i += 256) reinterpret_cast<uint4*>(s_in)[i] = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
        }
        const int gf = (cin - split) / 8;
        if (gf > 0) {
            const int g = tid % gf, p0 = tid / gf, ppass = 256 / gf;
            const uint16_t* src = a.x + (size_t)n * a.h * a.w * a.xs + a.xo + split + g * 8;
            if (p0 < ppass) {
                if (interior)
                    stage_tile<TR, TW, false>(s_in + g * 8, cpf, s_dummy, src, 0, a.w, a.xs, y0, x0, a.h, a.w, p0, ppass);
                else
                    stage_tile<TR, TW, true>(s_in + g * 8, cpf, s_dummy, src, 0, a.w, a.xs, y0, x0, a.h, a.w, p0, ppass);
            }
        }
        if (HALF) {
            const int gh = split / 8;
            const int g = tid % gh, p0 = tid / gh, ppass = 256 / gh;
            const int h2 = a.h / 2, w2 = a.w / 2;
            const uint16_t* src = a.x2 + (size_t)n * h2 * w2 * a.x2s + a.x2o + g * 8;
            if (p0 < ppass) {
                if (interior)
                    stage_tile<HR, HW, false>(s_in + hbase + g * 8, cph, s_dummy, src, 0, w2, a.x2s, y0 / 2, x0 / 2, h2,
                                              w2, p0, ppass);
                else
                    stage_tile<HR, HW, true>(s_in + hbase + g * 8, cph, s_dummy, src, 0, w2, a.x2s, y0 / 2, x0 / 2, h2,
                                             w2, p0, ppass);
            }
        }
    }
    __syncthreads();

    const int row = lane & 15, kg = lane >> 4;
    f32x4_t acc[RG][NT];
#pragma unroll
    for (int r = 0; r < RG; ++r)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[r][t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

    // LDS element offset of this lane's pixel in each row group (tap 0), at
    // full and at half size
    int pix[RG], pixh[HALF ? RG : 1];
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        const int ty = wave * RPW + r / 2, tx = (r & 1) * 16 + row;
        pix[r] = (ty * TW + tx) * cpf;
        if constexpr (HALF) pixh[r] = hbase + ((ty >> 1) * HW + (tx >> 1)) * cph;
    }
    const int xb = row & 1;
    const uint16_t* wl = a.wt + row * a.kpad + kg * 8;
    const int wtile = 16 * a.kpad;
    const int chunks = a.kpad / 32;
    auto load_w = [&](uint4 (&w)[NT], int c) {
#pragma unroll
        for (int t = 0; t < NT; ++t) w[t] = *reinterpret_cast<const uint4*>(wl + t * wtile + c * 32);
    };
    // with HALF the k-group's two entries for this lane's column parity (row
    // parity 0, 1)
    typedef std::conditional_t<HALF, uint2, uint32_t> TabEntry;
    auto tab = [&](int c) -> TabEntry {
        if constexpr (HALF) return *reinterpret_cast<const uint2*>(s_tab + ((c * 4 + kg) * 4 + xb * 2));
        else return s_tab[c * 4 + kg];
    };
    // a chunk's RG pixel fragments are all read before its first MFMA, so
    // one LDS latency (not RG) precedes the chunk's MFMAs
    auto chunk = [&](const uint4 (&w)[NT], TabEntry off) {
        uint4 pv[RG];
#pragma unroll
        for (int r = 0; r < RG; ++r) {
            int adr;
            if constexpr (HALF) {
                const uint32_t e = ((wave * RPW + r / 2) & 1) ? off.y : off.x;
                adr = (int)(e & 0x7FFFFFFFu) + ((e >> 31) ? pixh[r] : pix[r]);
            } else {
                adr = (int)off + pix[r];
            }
            pv[r] = *reinterpret_cast<const uint4*>(s_in + adr);
        }
        // (the register pressure would otherwise lead the scheduler to read
        // each fragment into one register tuple just before its MFMAs)
#pragma unroll
        for (int r = 0; r < RG; ++r) {
            u32x4v_t t = __builtin_bit_cast(u32x4v_t, pv[r]);
            asm volatile("" : "+v"(t));
            pv[r] = __builtin_bit_cast(uint4, t);
        }
#pragma unroll
        for (int r = 0; r < RG; ++r)
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(w[t]), as_bf16x8(pv[r]), acc[r][t], 0,
                                                                     0, 0);
    };
    // the bias is loaded before the main loop, which hides its latency
    float bias[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) bias[t][i] = a.bias ? a.bias[t * 16 + 4 * kg + i] : 0.0f;
    // weights PD chunks ahead: a chunk's MFMAs take ~16 NT RG cycles, too few
    // to cover an L2 load with the two-chunk lead of a ping-pong pair
    constexpr int PD = NT <= 4 ? 3 : 2;
    uint4 w[PD][NT];
    TabEntry off[PD];
#pragma unroll
    for (int j = 0; j < PD; ++j) {
        off[j] = TabEntry{};
        if (j < chunks) {
            load_w(w[j], j);
            off[j] = tab(j);
        }
    }
    for (int c = 0; c < chunks; c += PD) {
#pragma unroll
        for (int j = 0; j < PD; ++j) {
            if (c + j < chunks) {
                chunk(w[j], off[j]);
                if (c + j + PD < chunks) {
                    load_w(w[j], c + j + PD);
                    off[j] = tab(c + j + PD);
                }
            }
        }
    }
    // acc[r][t][i] = output channel t * 16 + 4 kg + i of pixel `row` of row group r
    uint16_t* yb = a.y + (size_t)n * a.h * a.w * a.ys + a.yo + 4 * kg;
    uint32_t q[RG][NT][2];
    // fused 1x1 head: the four lanes of a pixel (k-groups kg) each sum the
    // terms of their eight channels, two butterfly steps add the partial
    // sums, and the k-group-0 lane writes the pixel
    float4 hw[HEAD ? 4 : 1][NT];
    float4 hb = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (HEAD) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) hw[j][t] = *reinterpret_cast<const float4*>(a.head + j * 32 + t * 16 + 4 * kg);
        hb = *reinterpret_cast<const float4*>(a.head + 128);
    }
#pragma unroll
    for (int r = 0; r < RG; ++r) {
        const int y = y0 + wave * RPW + r / 2, xx = x0 + (r & 1) * 16 + row;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[i] = acc[r][t][i] + bias[t][i];
                if (a.relu) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
            }
            q[r][t][0] = pack_bf16x2(v[0], v[1]);
            q[r][t][1] = pack_bf16x2(v[2], v[3]);
        }
        if (!HEAD && y < a.h && xx < a.w) {
            uint16_t* yp = yb + (y * a.w + xx) * a.ys;
#pragma unroll
            for (int t = 0; t < NT; ++t)
                *reinterpret_cast<uint2*>(yp + t * 16) = make_uint2(q[r][t][0], q[r][t][1]);
        }
        if constexpr (HEAD) {
            float h[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const float u[4] = {__uint_as_float(q[r][t][0] << 16), __uint_as_float(q[r][t][0] & 0xFFFF0000u),
                                    __uint_as_float(q[r][t][1] << 16), __uint_as_float(q[r][t][1] & 0xFFFF0000u)};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    h[j] = __builtin_fmaf(hw[j][t].x, u[0], h[j]);
                    h[j] = __builtin_fmaf(hw[j][t].y, u[1], h[j]);
                    h[j] = __builtin_fmaf(hw[j][t].z, u[2], h[j]);
                    h[j] = __builtin_fmaf(hw[j][t].w, u[3], h[j]);
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                h[j] += __shfl_xor(h[j], 16, 64);
                h[j] += __shfl_xor(h[j], 32, 64);
            }
            if (kg == 0 && y < a.head_h && xx < a.head_w) {
                const uint2 iv = *reinterpret_cast<const uint2*>(a.head_inp + ((size_t)y * a.w + xx) * 8);
                ushort4 o;
                o.x = __half_as_ushort(__float2half_rn(h[0] + hb.x + __uint_as_float(iv.x << 16)));
                o.y = __half_as_ushort(__float2half_rn(h[1] + hb.y + __uint_as_float(iv.x & 0xFFFF0000u)));
                o.z = __half_as_ushort(__float2half_rn(h[2] + hb.z + __uint_as_float(iv.y << 16)));
                o.w = __half_as_ushort(__float2half_rn(h[3] + hb.w + __uint_as_float(iv.y & 0xFFFF0000u)));
                *reinterpret_cast<ushort4*>(a.head_out + ((size_t)y * a.head_w + xx) * 4) = o;
            }
        }
    }
    // 2x2 average pool: with two rows per wave, row groups r and r + 2 are
    // rows y and y + 1 of the same pixels, and the pixel x + 1 is the
    // neighbouring lane; summed in the order of pool2_kernel, so the result
    // is the same bits
    if constexpr (RG == 4 && !HEAD) {
      if (a.pool) {
        const int yp = (y0 + wave * RPW) / 2;
        uint16_t* pb = a.pool + (size_t)n * (a.h / 2) * (a.w / 2) * a.ps + a.po + 4 * kg;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int xx = x0 + r * 16 + row;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                uint32_t o[2];
#pragma unroll
                for (int h2 = 0; h2 < 2; ++h2) {
                    const uint32_t a0 = q[r][t][h2], b0 = lane_xor1(a0);
                    const uint32_t a1 = q[r + 2][t][h2], b1 = lane_xor1(a1);
                    float s[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int sh = e ? 0 : 16;
                        float sum = __uint_as_float((a0 << sh) & 0xFFFF0000u);
                        sum += __uint_as_float((b0 << sh) & 0xFFFF0000u);
                        sum += __uint_as_float((a1 << sh) & 0xFFFF0000u);
                        sum += __uint_as_float((b1 << sh) & 0xFFFF0000u);
                        s[e] = sum * 0.25f;
                    }
                    o[h2] = pack_bf16x2(s[0], s[1]);
                }
                if (!(row & 1) && 2 * yp + 1 < a.h && xx + 1 < a.w)
                    *reinterpret_cast<uint2*>(pb + (yp * (a.w / 2) + xx / 2) * a.ps + t * 16) = make_uint2(o[0], o[1]);
            }
        }
      }
    }
}

// Weight gradient of one tap over one chunk of pixels:
//   partial[chunk][tap][co][ci] = sum_p dz[p][co] * x[p + off(tap)][ci]
// as MFMA tiles (M = co, N = ci, K = 32 pixels per step); both operands are
// staged in LDS transposed ([channel][pixel]) so a lane's 8 pixels are one
// 16-byte read.  Tap 4 (the centre) also sums the bias gradient.
constexpr int kWgPix = 32;
constexpr int kWgMaxCh = 160;

template <int TPW>
__global__ void __launch_bounds__(256) conv3x3_wgrad_kernel(WGradLaunch a) {
    __shared__ __align__(16) uint16_t s_dz[kWgMaxCh][kWgPix];
    __shared__ __align__(16) uint16_t s_x[kWgMaxCh][kWgPix];
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int chunk = (int)blockIdx.x, tap = (int)blockIdx.y;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int cin16 = (a.cin + 15) / 16, ntiles = (a.cout / 16) * cin16;
    const int row = lane & 15, kg = lane >> 4;
    const long total = (long)a.n * a.h * a.w;

    f32x4_t acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;

    const long p_begin = (long)chunk * a.chunk_pixels;
    const long p_end = p_begin + a.chunk_pixels < total ? p_begin + a.chunk_pixels : total;
    for (long p0 = p_begin; p0 < p_end; p0 += kWgPix) {
        __syncthreads();
        // stage dz (cout channels) and the tap-shifted input (cin channels,
        // zero beyond cin up to a multiple of 16), 8 channels per load
        const int gz = a.cout / 8, gx = cin16 * 2;
        for (int i = tid; i < kWgPix * (gz + gx); i += 256) {
            const int p = i / (gz + gx), g = i - p * (gz + gx);
            const long q = p0 + p;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (g < gz) {
                if (q < p_end) v = *reinterpret_cast<const uint4*>(a.dz + q * a.dzs + a.dzo + g * 8);
                const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int c = 0; c < 8; ++c) s_dz[g * 8 + c][p] = e[c];
            } else {
                const int c0 = (g - gz) * 8;
                if (q < p_end && c0 < a.cin) {
                    const long xq = q % a.w, yq = (q / a.w) % a.h, nq = q / ((long)a.w * a.h);
                    const long yy = yq + dy, xx = xq + dx;
                    if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
                        v = *reinterpret_cast<const uint4*>(a.x + ((nq * a.h + yy) * a.w + xx) * a.xs + a.xo + c0);
                }
                const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int c = 0; c < 8; ++c) s_x[c0 + c][p] = e[c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
            const int tile = wave + 4 * t;
            if (tile < ntiles) {
                const int mt = tile / cin16, nt = tile - mt * cin16;
                const uint4 av = *reinterpret_cast<const uint4*>(&s_dz[mt * 16 + row][kg * 8]);
                const uint4 bv = *reinterpret_cast<const uint4*>(&s_x[nt * 16 + row][kg * 8]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(av), as_bf16x8(bv), acc[t], 0, 0, 0);
            }
        }
        if (tap == 4 && tid < a.cout)
            for (int p = 0; p < kWgPix; ++p) bsum += bf2f(s_dz[tid][p]);
    }

    float* out = a.partial + ((size_t)chunk * 9 + tap) * a.cout * a.cin;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave + 4 * t;
        if (tile >= ntiles) continue;
        const int mt = tile / cin16, nt = tile - mt * cin16;
        const int ci = nt * 16 + row;
        if (ci >= a.cin) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) out[(size_t)(mt * 16 + kg * 4 + i) * a.cin + ci] = acc[t][i];
    }
    if (tap == 4 && tid < a.cout) a.bias_partial[(size_t)chunk * a.cout + tid] = bsum;
}

// grad_w[co][tap][ci] = sum over chunks; grad_b[co] likewise.
__global__ void wgrad_reduce_kernel(const float* __restrict__ partial, const float* __restrict__ bias_partial,
                                    int chunks, int cout, int cin, float* gw, float* gb) {
    const int per = 9 * cout * cin;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < per) {
        const int tap = i / (cout * cin), rem = i - tap * cout * cin;
        const int co = rem / cin, ci = rem - co * cin;
        float s = 0.0f;
        for (int c = 0; c < chunks; ++c) s += partial[(size_t)c * per + i];
        gw[(size_t)co * 9 * cin + tap * cin + ci] = s;
    }
    if (i < cout) {
        float s = 0.0f;
        for (int c = 0; c < chunks; ++c) s += bias_partial[(size_t)c * cout + i];
        gb[i] = s;
    }
}

// dz = dy * (y > 0) in place (the ReLU derivative from the stored output).
__global__ void relu_mask_kernel(TensorView dy, TensorView y, long pixels, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int groups = c / 8;
    if (i >= pixels * groups) return;
    const long p = i / groups;
    const int g = (int)(i - p * groups);
    uint4* dp = reinterpret_cast<uint4*>(dy.p + p * dy.cs + dy.co + g * 8);
    const uint4 yv = *reinterpret_cast<const uint4*>(y.p + p * y.cs + y.co + g * 8);
    uint4 dv = *dp;
    uint16_t* d = reinterpret_cast<uint16_t*>(&dv);
    const uint16_t* yy = reinterpret_cast<const uint16_t*>(&yv);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (!(bf2f(yy[k]) > 0.0f)) d[k] = 0;
    *dp = dv;
}

// 2x2 average pool: out (n, h/2, w/2) <- in (n, h, w).
__global__ void pool2_kernel(TensorView in, TensorView out, int n, int h, int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ho = h / 2, wo = w / 2, groups = c / 8;
    if (i >= (long)n * ho * wo * groups) return;
    const int g = (int)(i % groups);
    const long p = i / groups;
    const long xo = p % wo, yo = (p / wo) % ho, nn = p / ((long)wo * ho);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) {
            const long q = (nn * h + 2 * yo + a2) * w + 2 * xo + b2;
            const uint4 v = *reinterpret_cast<const uint4*>(in.p + q * in.cs + in.co + g * 8);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += bf2f(e[k]);
        }
    uint4 o;
    uint16_t* e = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = f2bf(s[k] * 0.25f);
    *reinterpret_cast<uint4*>(out.p + p * out.cs + out.co + g * 8) = o;
}

// Adjoint of the pool, accumulated: din += 0.25 dout (in bf16).
__global__ void pool2_bwd_kernel(TensorView dout, TensorView din, int n, int h, int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int groups = c / 8;
    if (i >= (long)n * h * w * groups) return;
    const int g = (int)(i % groups);
    const long p = i / groups;
    const long x = p % w, y = (p / w) % h, nn = p / ((long)w * h);
    const long q = (nn * (h / 2) + y / 2) * (w / 2) + x / 2;
    const uint4 dv = *reinterpret_cast<const uint4*>(dout.p + q * dout.cs + dout.co + g * 8);
    uint4* ip = reinterpret_cast<uint4*>(din.p + p * din.cs + din.co + g * 8);
    uint4 iv = *ip;
    uint16_t* e = reinterpret_cast<uint16_t*>(&iv);
    const uint16_t* d = reinterpret_cast<const uint16_t*>(&dv);
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = f2bf(bf2f(e[k]) + 0.25f * bf2f(d[k]));
    *ip = iv;
}

// Nearest 2x upsample: out (n, 2h, 2w) <- in (n, h, w).
__global__ void up2_kernel(TensorView in, TensorView out, int n, int h, int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int H = 2 * h, W = 2 * w, groups = c / 8;
    if (i >= (long)n * H * W * groups) return;
    const int g = (int)(i % groups);
    const long p = i / groups;
    const long x = p % W, y = (p / W) % H, nn = p / ((long)W * H);
    const long q = (nn * h + y / 2) * w + x / 2;
    *reinterpret_cast<uint4*>(out.p + p * out.cs + out.co + g * 8) =
        *reinterpret_cast<const uint4*>(in.p + q * in.cs + in.co + g * 8);
}

// Adjoint of the upsample: din (n, h, w) = sum of its 2x2 block of dout.
__global__ void up2_bwd_kernel(TensorView dout, TensorView din, int n, int h, int w, int c) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int groups = c / 8;
    if (i >= (long)n * h * w * groups) return;
    const int g = (int)(i % groups);
    const long p = i / groups;
    const long x = p % w, y = (p / w) % h, nn = p / ((long)w * h);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) {
            const long q = (nn * 2 * h + 2 * y + a2) * (2 * w) + 2 * x + b2;
            const uint4 v = *reinterpret_cast<const uint4*>(dout.p + q * dout.cs + dout.co + g * 8);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += bf2f(e[k]);
        }
    uint4 o;
    uint16_t* e = reinterpret_cast<uint16_t*>(&o);
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = f2bf(s[k]);
    *reinterpret_cast<uint4*>(din.p + p * din.cs + din.co + g * 8) = o;
}

// 1x1 head (32 -> 4) plus the input's RGBA mean: out = b + W u + inp[0..3].
// With a target: dout = 2 (out - target) / count, the input-side gradient
// du = W^T dout (bf16), and per-block partial sums of the loss, dW and db.
__global__ void __launch_bounds__(256) head_kernel(HeadLaunch a) {
    __shared__ float s_w[4 * 32 + 4];
    __shared__ float s_red[256];
    const int tid = (int)threadIdx.x;
    if (tid < 132) s_w[tid] = a.params[tid];
    __syncthreads();
    const long p = (long)blockIdx.x * 256 + tid;
    const bool ok = p < a.pixels;
    float u[32];
    float o[4] = {0, 0, 0, 0}, d[4] = {0, 0, 0, 0};
    float lsum = 0.0f;
    if (ok) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const uint4 v = *reinterpret_cast<const uint4*>(a.u + p * 32 + g * 8);
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
            for (int k = 0; k < 8; ++k) u[g * 8 + k] = bf2f(e[k]);
        }
        const uint4 iv = *reinterpret_cast<const uint4*>(a.inp + p * 8);
        const uint16_t* ie = reinterpret_cast<const uint16_t*>(&iv);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float s = s_w[128 + j];
#pragma unroll
            for (int c = 0; c < 32; ++c) s += s_w[j * 32 + c] * u[c];
            o[j] = s + bf2f(ie[j]);
        }
        if (a.out) *reinterpret_cast<float4*>(a.out + p * 4) = make_float4(o[0], o[1], o[2], o[3]);
        if (a.image) {
            ushort4 h;
            h.x = __half_as_ushort(__float2half_rn(o[0]));
            h.y = __half_as_ushort(__float2half_rn(o[1]));
            h.z = __half_as_ushort(__float2half_rn(o[2]));
            h.w = __half_as_ushort(__float2half_rn(o[3]));
            *reinterpret_cast<ushort4*>(a.image + p * 4) = h;
        }
        if (a.target) {
            const float4 t = *reinterpret_cast<const float4*>(a.target + p * 4);
            const float tt[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float df = o[j] - tt[j];
                lsum += df * df;
                d[j] = 2.0f * df * a.inv_count;
            }
            float du[32];
#pragma unroll
            for (int c = 0; c < 32; ++c) du[c] = s_w[c] * d[0] + s_w[32 + c] * d[1] + s_w[64 + c] * d[2] + s_w[96 + c] * d[3];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                uint4 v;
                uint16_t* e = reinterpret_cast<uint16_t*>(&v);
#pragma unroll
                for (int k = 0; k < 8; ++k) e[k] = f2bf(du[g * 8 + k]);
                *reinterpret_cast<uint4*>(a.du + p * 32 + g * 8) = v;
            }
        }
    }
    if (!a.target) return;
    // per-block partials: loss, dW (4 x 32), db (4)
    float* part = a.partial + (size_t)blockIdx.x * 133;
    for (int k = 0; k < 133; ++k) {
        float v;
        if (k == 0) v = lsum;
        else if (k <= 128) v = d[(k - 1) / 32] * u[(k - 1) % 32];
        else v = d[k - 129];
        if (!ok) v = 0.0f;
        s_red[tid] = v;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) s_red[tid] += s_red[tid + s];
            __syncthreads();
        }
        if (tid == 0) part[k] = s_red[0];
        __syncthreads();
    }
}

__global__ void head_reduce_kernel(const float* __restrict__ partial, int blocks, float* loss, float* gw) {
    const int k = (int)threadIdx.x;
    if (k >= 133) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * 133 + k];
    if (k == 0) *loss = s;
    else gw[k - 1] = s;  // head.w (128) then head.b (4)
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, uint32_t count, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
}

// bf16 weight matrices of one layer from the fp32 master W[co][tap][ci]:
// forward wf[co][kpad] (k = tap * cin + ci) and input-gradient
// wb[ci][kpadT] (k = tap' * cout + co, value W[co][8 - tap'][ci]).
__global__ void pack_weights_kernel(const float* __restrict__ w, int cin, int cout, int kpad, int kpad_t,
                                    uint16_t* wf, uint16_t* wb) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (wf && i < cout * kpad) {
        const int co = i / kpad, k = i - co * kpad;
        wf[i] = k < 9 * cin ? f2bf(w[(size_t)co * 9 * cin + k]) : (uint16_t)0;
    }
    if (wb && i < cin * kpad_t) {
        const int ci = i / kpad_t, k = i - ci * kpad_t;
        uint16_t v = 0;
        if (k < 9 * cout) {
            const int tap = k / cout, co = k - tap * cout;
            v = f2bf(w[((size_t)co * 9 + (8 - tap)) * cin + ci]);
        }
        wb[i] = v;
    }
}

// Network input (8 bf16 channels: mean RGBA, depth, 0, 0, 0) from the GPS
// running mean and depth of a width x height frame, into an (hp, wp) tensor
// whose pixels beyond the frame are zero; optionally the crop (x0, y0) of
// size (hp, wp) only, with the matching sorted-alpha target (RGBA16F -> fp32).
__global__ void pack_input_kernel(PackLaunch a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)a.hp * a.wp) return;
    const int x = (int)(i % a.wp), y = (int)(i / a.wp);
    const int sx = a.x0 + x, sy = a.y0 + y;
    const bool in = sx < a.width && sy < a.height;
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float dep = 0.0f;
    if (in) {
        const size_t q = (size_t)sy * a.width + sx;
        c = a.accum[q];
        dep = a.depth[q];
    }
    uint4 v;
    uint16_t* e = reinterpret_cast<uint16_t*>(&v);
    e[0] = f2bf(c.x); e[1] = f2bf(c.y); e[2] = f2bf(c.z); e[3] = f2bf(c.w);
    e[4] = f2bf(dep); e[5] = 0; e[6] = 0; e[7] = 0;
    *reinterpret_cast<uint4*>(a.inp + i * 8) = v;
    if (a.target) {
        float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (in) {
            const ushort4 h = *reinterpret_cast<const ushort4*>(a.alpha + ((size_t)sy * a.width + sx) * 4);
            t = make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                            __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
        }
        reinterpret_cast<float4*>(a.target)[i] = t;
    }
}

// Gather batch crops from the pool: inp[b] = pool_in[idx[b]], tgt likewise.
__global__ void gather_kernel(const uint4* pool_in, const float4* pool_tgt, const uint32_t* idx, uint32_t per,
                              uint32_t batch, uint4* inp, float4* tgt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)per * batch) return;
    const size_t b = i / per, r = i - b * per;
    const size_t src = (size_t)idx[b] * per + r;
    inp[i] = pool_in[src];
    tgt[i] = pool_tgt[src];
}

inline uint32_t blocks_for(long n, int t = 256) { return (uint32_t)((n + t - 1) / t); }

}  // namespace

template <int NT, int RG, bool HALF, bool HEAD = false>
hipError_t conv3x3_launch_as(const ConvLaunch& a, hipStream_t s) {
    constexpr int BR = 4 * (RG / 2);
    if (HEAD && (a.n != 1 || a.cout != 32 || a.pool || a.x2 || !a.head_inp || !a.head_out || a.head_w > a.w ||
                 a.head_h > a.h || (reinterpret_cast<uintptr_t>(a.head) & 15)))
        return hipErrorInvalidValue;
    if (a.pool && (RG != 4 || a.h % 2 || a.w % 2)) return hipErrorInvalidValue;
    if (a.x2 && (a.split % 8 || a.split > a.cin || a.h % 2 || a.w % 2)) return hipErrorInvalidValue;
    if (a.kpad % 32 || a.kpad / 8 > 256 || a.kpad < 9 * a.cin) return hipErrorInvalidValue;
    // element offsets inside one image are 32-bit
    if ((size_t)a.h * a.w * std::max(std::max(a.xs, a.ys), a.x2 ? a.x2s : 0) >= (1u << 31)) return hipErrorInvalidValue;
    if (a.pool && (size_t)a.h * a.w * a.ps >= (1u << 31)) return hipErrorInvalidValue;
    const int split = HALF ? a.split : 0;
    if (HALF && (split == a.cin || split == 0)) return hipErrorInvalidValue;
    const size_t lds = ((size_t)(BR + 2) * (kConvTileW + 2) * (a.cin - split + 8) +
                        (HALF ? (size_t)HalfTile<BR>::kRows * HalfTile<BR>::kCols * (split + 8) : 0)) * 2 +
                       16 + (size_t)a.kpad / 8 * 4 * (HALF ? 4 : 1);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&conv3x3_kernel<NT, RG, HALF, HEAD>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((a.w + kConvTileW - 1) / kConvTileW, (a.h + BR - 1) / BR, a.n);
    hipLaunchKernelGGL((conv3x3_kernel<NT, RG, HALF, HEAD>), grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

template <int NT, int RG>
hipError_t conv3x3_launch(const ConvLaunch& a, hipStream_t s) {
    return a.x2 ? conv3x3_launch_as<NT, RG, true>(a, s) : conv3x3_launch_as<NT, RG, false>(a, s);
}

hipError_t launch_conv3x3(const ConvLaunch& a, int nt, hipStream_t s) {
    if (a.cin % 8 || a.cin > 160) return hipErrorInvalidValue;
    if (a.head) return nt == 2 ? conv3x3_launch_as<2, 4, false, true>(a, s) : hipErrorInvalidValue;
    switch (nt) {
        case 2: return conv3x3_launch<2, 4>(a, s);
        case 3: return conv3x3_launch<3, 4>(a, s);
        case 4: return conv3x3_launch<4, 4>(a, s);
        case 5: return conv3x3_launch<5, 2>(a, s);
        case 6: return conv3x3_launch<6, 2>(a, s);
        case 7: return conv3x3_launch<7, 2>(a, s);
        case 10: return conv3x3_launch<10, 2>(a, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_conv3x3_wgrad(const WGradLaunch& a, int chunks, float* gw, float* gb, hipStream_t s) {
    const int ntiles = (a.cout / 16) * ((a.cin + 15) / 16);
    const int tpw = (ntiles + 3) / 4;
    if (a.cin > kWgMaxCh || a.cout > kWgMaxCh || a.cout % 16) return hipErrorInvalidValue;
    dim3 grid(chunks, 9);
    if (tpw <= 2) hipLaunchKernelGGL(conv3x3_wgrad_kernel<2>, grid, dim3(256), 0, s, a);
    else if (tpw <= 4) hipLaunchKernelGGL(conv3x3_wgrad_kernel<4>, grid, dim3(256), 0, s, a);
    else if (tpw <= 6) hipLaunchKernelGGL(conv3x3_wgrad_kernel<6>, grid, dim3(256), 0, s, a);
    else if (tpw <= 10) hipLaunchKernelGGL(conv3x3_wgrad_kernel<10>, grid, dim3(256), 0, s, a);
    else return hipErrorInvalidValue;
    const int per = 9 * a.cout * a.cin;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(blocks_for(per)), dim3(256), 0, s, a.partial, a.bias_partial,
                       chunks, a.cout, a.cin, gw, gb);
    return hipGetLastError();
}

hipError_t launch_relu_mask(TensorView dy, TensorView y, long pixels, int c, hipStream_t s) {
    hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(pixels * (c / 8))), dim3(256), 0, s, dy, y, pixels, c);
    return hipGetLastError();
}

hipError_t launch_pool2(TensorView in, TensorView out, int n, int h, int w, int c, hipStream_t s) {
    hipLaunchKernelGGL(pool2_kernel, dim3(blocks_for((long)n * (h / 2) * (w / 2) * (c / 8))), dim3(256), 0, s, in,
                       out, n, h, w, c);
    return hipGetLastError();
}

hipError_t launch_pool2_bwd(TensorView dout, TensorView din, int n, int h, int w, int c, hipStream_t s) {
    hipLaunchKernelGGL(pool2_bwd_kernel, dim3(blocks_for((long)n * h * w * (c / 8))), dim3(256), 0, s, dout, din, n,
                       h, w, c);
    return hipGetLastError();
}

hipError_t launch_up2(TensorView in, TensorView out, int n, int h, int w, int c, hipStream_t s) {
    hipLaunchKernelGGL(up2_kernel, dim3(blocks_for((long)n * 4 * h * w * (c / 8))), dim3(256), 0, s, in, out, n, h,
                       w, c);
    return hipGetLastError();
}

hipError_t launch_up2_bwd(TensorView dout, TensorView din, int n, int h, int w, int c, hipStream_t s) {
    hipLaunchKernelGGL(up2_bwd_kernel, dim3(blocks_for((long)n * h * w * (c / 8))), dim3(256), 0, s, dout, din, n, h,
                       w, c);
    return hipGetLastError();
}

hipError_t launch_head(const HeadLaunch& a, float* loss, float* gw, hipStream_t s) {
    const uint32_t blocks = blocks_for(a.pixels);
    hipLaunchKernelGGL(head_kernel, dim3(blocks), dim3(256), 0, s, a);
    if (a.target) hipLaunchKernelGGL(head_reduce_kernel, dim3(1), dim3(192), 0, s, a.partial, (int)blocks, loss, gw);
    return hipGetLastError();
}

uint32_t head_partial_floats(long pixels) { return blocks_for(pixels) * 133u; }

hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint32_t count, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(count)), dim3(256), 0, s, p, g, m, v, count, lr, b1, b2, eps, c1,
                       c2);
    return hipGetLastError();
}

hipError_t launch_pack_weights(const float* w, int cin, int cout, int kpad, int kpad_t, uint16_t* wf, uint16_t* wb,
                               hipStream_t s) {
    const long n = (long)cout * kpad > (long)cin * kpad_t ? (long)cout * kpad : (long)cin * kpad_t;
    hipLaunchKernelGGL(pack_weights_kernel, dim3(blocks_for(n)), dim3(256), 0, s, w, cin, cout, kpad, kpad_t, wf, wb);
    return hipGetLastError();
}

hipError_t launch_pack_input(const PackLaunch& a, hipStream_t s) {
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for((long)a.hp * a.wp)), dim3(256), 0, s, a);
    return hipGetLastError();
}


// Squared-error sum and max abs difference of two half-precision arrays of
// `count` values (grid-stride, one (sum, max) per block; the caller adds the
// block sums in double).
__global__ void __launch_bounds__(256) image_error_kernel(const uint16_t* __restrict__ a,
                                                          const uint16_t* __restrict__ b, uint32_t count,
                                                          float2* partial) {
    __shared__ float s_sum[4], s_max[4];
    float sum = 0.0f, mx = 0.0f;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
        const float d = __half2float(__ushort_as_half(a[i])) - __half2float(__ushort_as_half(b[i]));
        sum = fmaf(d, d, sum);
        mx = fmaxf(mx, fabsf(d));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        sum += __shfl_xor(sum, o, 64);
        mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        s_sum[wave] = sum;
        s_max[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        partial[blockIdx.x] = make_float2((s_sum[0] + s_sum[1]) + (s_sum[2] + s_sum[3]),
                                          fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3])));
}

hipError_t launch_image_error(const uint16_t* a, const uint16_t* b, uint32_t count, float2* partial, uint32_t blocks,
                              hipStream_t s) {
    if (!blocks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(image_error_kernel, dim3(blocks), dim3(256), 0, s, a, b, count, partial);
    return hipGetLastError();
}

hipError_t launch_gather(const uint16_t* pool_in, const float* pool_tgt, const uint32_t* idx, uint32_t pixels,
                         uint32_t batch, uint16_t* inp, float* tgt, hipStream_t s) {
    hipLaunchKernelGGL(gather_kernel, dim3(blocks_for((long)pixels * batch)), dim3(256), 0, s,
                       reinterpret_cast<const uint4*>(pool_in), reinterpret_cast<const float4*>(pool_tgt), idx, pixels,
                       batch, reinterpret_cast<uint4*>(inp), reinterpret_cast<float4*>(tgt));
    return hipGetLastError();
}

}  // namespace gs
