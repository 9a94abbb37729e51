// gs_display.hip - display of a linear image through a GSAnim 0.3 LUT.
//
// Replaces layers_display_fs (shaders/accum.glsl:197-240): radiance times
// 2^exposure, mapped to [0, 1] in log2 between the asset's log2_min and
// log2_max, then an exact trilinear lookup of the uint16 RGB lattice
// (red fastest).  One thread per pixel; HBM-bound (16 or 8 B in, 16 B out,
// the lattice stays in the caches).
#include <hip/hip_fp16.h>

#include "gs_internal.h"
#include "gs_math.h"

namespace gs {

namespace {

__device__ __forceinline__ float lut_channel(const uint16_t* lut, int size, int x, int y, int z, int c) {
    const int i = ((z * size + y) * size + x) * 3 + c;
    return (float)lut[i] / 65535.0f;
}

__global__ void __launch_bounds__(256) layers_display_kernel(LayersDisplayLaunch a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= a.count) return;
    float rgb[3];
    if (a.half_source) {
        const uint2 h = reinterpret_cast<const uint2*>(a.source)[i];
        rgb[0] = __half2float(__ushort_as_half((unsigned short)(h.x & 65535u)));
        rgb[1] = __half2float(__ushort_as_half((unsigned short)(h.x >> 16)));
        rgb[2] = __half2float(__ushort_as_half((unsigned short)(h.y & 65535u)));
    } else {
        const float4 v = reinterpret_cast<const float4*>(a.source)[i];
        rgb[0] = v.x;
        rgb[1] = v.y;
        rgb[2] = v.z;
    }
    const float gain = exp2f(a.exposure_ev);
    const float floor_rad = exp2f(a.log2_min);
    const int size = a.lut_size;
    int lo[3], hi[3];
    float f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float rad = rgb[c] * gain;
        const float s = clampf((log2f(maxf(rad, floor_rad)) - a.log2_min) / (a.log2_max - a.log2_min), 0.0f, 1.0f);
        const float p = s * (float)(size - 1);
        const float fl = __builtin_floorf(p);
        lo[c] = (int)fl;
        hi[c] = min(lo[c] + 1, size - 1);
        f[c] = p - fl;
    }
    float out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float e000 = lut_channel(a.lut, size, lo[0], lo[1], lo[2], c);
        const float e100 = lut_channel(a.lut, size, hi[0], lo[1], lo[2], c);
        const float e010 = lut_channel(a.lut, size, lo[0], hi[1], lo[2], c);
        const float e110 = lut_channel(a.lut, size, hi[0], hi[1], lo[2], c);
        const float e001 = lut_channel(a.lut, size, lo[0], lo[1], hi[2], c);
        const float e101 = lut_channel(a.lut, size, hi[0], lo[1], hi[2], c);
        const float e011 = lut_channel(a.lut, size, lo[0], hi[1], hi[2], c);
        const float e111 = lut_channel(a.lut, size, hi[0], hi[1], hi[2], c);
        const float va = mixf(mixf(e000, e100, f[0]), mixf(e010, e110, f[0]), f[1]);
        const float vb = mixf(mixf(e001, e101, f[0]), mixf(e011, e111, f[0]), f[1]);
        out[c] = mixf(va, vb, f[2]);
    }
    a.out[i] = make_float4(out[0], out[1], out[2], 1.0f);
}

}  // namespace

hipError_t launch_layers_display(const LayersDisplayLaunch& a, hipStream_t s) {
    if (!a.count) return hipSuccess;
    hipLaunchKernelGGL(layers_display_kernel, dim3((a.count + 255u) / 256u), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gs
