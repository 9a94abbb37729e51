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

#include "gs_internal.h"

namespace gs {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

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
// pitch cin + 8 spreads the 16-byte reads of consecutive pixels over the
// banks); all 9 taps then read their A operands from LDS, and the weights
// (B operands) come through the L1/L2 caches.
constexpr int kConvTileW = 32;

template <int NT, int RG>
__global__ void __launch_bounds__(256) conv3x3_kernel(ConvLaunch a) {
    extern __shared__ __align__(16) uint16_t s_in[];
    constexpr int RPW = RG / 2;            // rows per wave
    constexpr int BR = 4 * RPW;            // rows per block
    constexpr int TW = kConvTileW + 2, TR = BR + 2;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int y0 = (int)blockIdx.y * BR;
    const int x0 = (int)blockIdx.x * kConvTileW;
    const int n = (int)blockIdx.z;
    const int cin = a.cin, cp = cin + 8, groups = cin / 8;
    const size_t img = (size_t)n * a.h;

    // stage rows y0-1 .. y0+BR, columns x0-1 .. x0+32 (zero outside the image);
    // with x2, channels below split are read from the half-size tensor.  Each
    // thread keeps kStageBatch loads in flight before it writes them to LDS.
    constexpr int kStageBatch = 8;
    const size_t img2 = (size_t)n * (a.h / 2);
    const int total = TR * TW * groups;
    for (int i0 = tid; i0 < total; i0 += 256 * kStageBatch) {
        uint4 v[kStageBatch];
        int dst[kStageBatch];
#pragma unroll
        for (int j = 0; j < kStageBatch; ++j) {
            const int i = i0 + j * 256;
            v[j] = make_uint4(0u, 0u, 0u, 0u);
            dst[j] = -1;
            if (i < total) {
                const int g = i % groups, pix = i / groups;
                const int tx = pix % TW, ty = pix / TW;
                const int yy = y0 + ty - 1, xx = x0 + tx - 1;
                dst[j] = pix * cp + g * 8;
                if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w) {
                    if (a.x2 && g * 8 < a.split)
                        v[j] = *reinterpret_cast<const uint4*>(a.x2 + ((img2 + yy / 2) * (a.w / 2) + xx / 2) * a.x2s +
                                                               a.x2o + g * 8);
                    else
                        v[j] = *reinterpret_cast<const uint4*>(a.x + ((img + yy) * a.w + xx) * a.xs + a.xo + g * 8);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kStageBatch; ++j)
            if (dst[j] >= 0) *reinterpret_cast<uint4*>(s_in + dst[j]) = v[j];
    }
    __syncthreads();

    const int row = lane & 15, kg = lane >> 4;
    f32x4_t acc[RG][NT];
#pragma unroll
    for (int g = 0; g < RG; ++g)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

    // this lane's k = kc + 8 kg as (tap, channel), advanced by 32 per chunk
    const int K = 9 * cin;
    int tap = 0, ci = kg * 8;
    while (ci >= cin) { ci -= cin; ++tap; }
    // weights of the next chunk are loaded while this chunk's MFMAs run
    const uint16_t* wrow = a.wt + (size_t)row * a.kpad + kg * 8;
    uint4 bnext[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bnext[t] = *reinterpret_cast<const uint4*>(wrow + (size_t)t * 16 * a.kpad);
    for (int kc = 0; kc < a.kpad; kc += 32) {
        const int kk = kc + kg * 8;
        uint4 bv[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[t] = bnext[t];
        if (kc + 32 < a.kpad) {
#pragma unroll
            for (int t = 0; t < NT; ++t)
                bnext[t] = *reinterpret_cast<const uint4*>(wrow + (size_t)t * 16 * a.kpad + kc + 32);
        }
        const bool live = kk < K;
        const int dy = tap / 3, dx = tap - dy * 3;  // 0..2: the halo offsets
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const int r = wave * RPW + g / 2, c = (g & 1) * 16 + row;
            uint4 av = make_uint4(0u, 0u, 0u, 0u);
            if (live) av = *reinterpret_cast<const uint4*>(s_in + (size_t)((r + dy) * TW + c + dx) * cp + ci);
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(av), as_bf16x8(bv[t]), acc[g][t], 0,
                                                                     0, 0);
        }
        ci += 32;
        while (ci >= cin && tap < 9) { ci -= cin; ++tap; }
    }

#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + row;
        const float b = a.bias ? a.bias[co] : 0.0f;
        uint16_t q[RG][4];
#pragma unroll
        for (int g = 0; g < RG; ++g) {
            const int y = y0 + wave * RPW + g / 2;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int xx = x0 + (g & 1) * 16 + kg * 4 + i;
                float v = acc[g][t][i] + b;
                if (a.relu) v = v > 0.0f ? v : 0.0f;
                q[g][i] = f2bf(v);
                if (y < a.h && xx < a.w) a.y[((img + y) * a.w + xx) * a.ys + a.yo + co] = q[g][i];
            }
        }
        // 2x2 average pool: with two rows per wave a lane holds whole 2x2
        // blocks (rows of groups g and g + 2, pixels i and i + 1); summed in
        // the order of pool2_kernel, so the result is the same bits
        if (RG == 4 && a.pool) {
            const int yp = (y0 + wave * RPW) / 2;
#pragma unroll
            for (int g = 0; g < 2; ++g)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int xx = x0 + g * 16 + kg * 4 + 2 * j;
                    if (2 * yp + 1 >= a.h || xx + 1 >= a.w) continue;
                    float sum = 0.0f;
                    sum += bf2f(q[g][2 * j]);
                    sum += bf2f(q[g][2 * j + 1]);
                    sum += bf2f(q[g + 2][2 * j]);
                    sum += bf2f(q[g + 2][2 * j + 1]);
                    a.pool[(((size_t)n * (a.h / 2) + yp) * (a.w / 2) + xx / 2) * a.ps + a.po + co] = f2bf(sum * 0.25f);
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

// RGBA16F (hp x wp, row pitch wp) -> the width x height image.
__global__ void crop_image_kernel(const uint16_t* src, int wp, uint16_t* dst, int width, int height) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)width * height) return;
    const int x = (int)(i % width), y = (int)(i / width);
    *reinterpret_cast<ushort4*>(dst + i * 4) = *reinterpret_cast<const ushort4*>(src + ((size_t)y * wp + x) * 4);
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

template <int NT, int RG>
hipError_t conv3x3_launch(const ConvLaunch& a, hipStream_t s) {
    constexpr int BR = 4 * (RG / 2);
    if (a.pool && (RG != 4 || a.h % 2 || a.w % 2)) return hipErrorInvalidValue;
    if (a.x2 && (a.split % 8 || a.split > a.cin || a.h % 2 || a.w % 2)) return hipErrorInvalidValue;
    const size_t lds = (size_t)(BR + 2) * (kConvTileW + 2) * (a.cin + 8) * 2;
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&conv3x3_kernel<NT, RG>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((a.w + kConvTileW - 1) / kConvTileW, (a.h + BR - 1) / BR, a.n);
    hipLaunchKernelGGL((conv3x3_kernel<NT, RG>), grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_conv3x3(const ConvLaunch& a, int nt, hipStream_t s) {
    if (a.cin % 8 || a.cin > 160) return hipErrorInvalidValue;
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

hipError_t launch_crop_image(const uint16_t* src, int wp, uint16_t* dst, int width, int height, hipStream_t s) {
    hipLaunchKernelGGL(crop_image_kernel, dim3(blocks_for((long)width * height)), dim3(256), 0, s, src, wp, dst, width,
                       height);
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
