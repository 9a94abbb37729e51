// gs_probe.hip - HBM rate probes: the read, write and copy rates a kernel of
// this library's access patterns can reach on the device, against which the
// HBM-bound kernels (cull, radix passes) are read.
//
//   read    one 16-byte load per thread (the cull's access: a lane reads one
//           float4 of a plane), grid of n / 256 blocks
//   planes  15 planes of m float4, thread i reads element i of each (the
//           cull's one-round-trip prefetch of all planes)
//   write   one 16-byte store per thread
//   copy    one 16-byte load and store per thread
// A result that could never be equal to the sum keeps the loads live.
#include "gs_internal.h"

namespace gs {

namespace {

__global__ void __launch_bounds__(256) probe_read_kernel(const float4* __restrict__ a, size_t n, float* sink) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 v = a[i];
    if (v.x + v.y + v.z + v.w == -1.5f) sink[0] = v.x;
}

__global__ void __launch_bounds__(256) probe_planes_kernel(const float4* __restrict__ a, size_t m, float* sink) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= m) return;
    float s = 0.0f;
#pragma unroll
    for (int p = 0; p < 15; ++p) {
        const float4 v = a[(size_t)p * m + i];
        s += (v.x + v.y) + (v.z + v.w);
    }
    if (s == -1.5f) sink[0] = s;
}

__global__ void __launch_bounds__(256) probe_write_kernel(float4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) b[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}

__global__ void __launch_bounds__(256) probe_copy_kernel(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) b[i] = a[i];
}

}  // namespace

hipError_t launch_probe(int which, const float4* a, float4* b, size_t n, float* sink, hipStream_t s) {
    switch (which) {
        case 0:
            hipLaunchKernelGGL(probe_read_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, a, n, sink);
            break;
        case 1: {
            const size_t m = n / 15;
            hipLaunchKernelGGL(probe_planes_kernel, dim3((uint32_t)((m + 255) / 256)), dim3(256), 0, s, a, m, sink);
            break;
        }
        case 2:
            hipLaunchKernelGGL(probe_write_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, b, n);
            break;
        case 3:
            hipLaunchKernelGGL(probe_copy_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
