// complex64 instantiation of the engine kernels
#include "gpe_kernels_impl.hpp"

namespace gpe {
GPE_INSTANTIATE(float)

__global__ __launch_bounds__(kThreads) void c128_to_c64_kernel(const double2* __restrict__ src,
                                                               float2* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const double2 v = src[i];
        dst[i] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ __launch_bounds__(kThreads) void c64_to_c128_kernel(const float2* __restrict__ src,
                                                               double2* __restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const float2 v = src[i];
        dst[i] = make_double2((double)v.x, (double)v.y);
    }
}

__global__ __launch_bounds__(kThreads) void f64_to_f32_kernel(const double* __restrict__ src, float* __restrict__ dst,
                                                             size_t n) {
    for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        dst[i] = (float)src[i];
}

static dim3 convert_grid(size_t n) {
    const size_t b = (n + kThreads - 1) / kThreads;
    return dim3((unsigned)(b < 65536 ? b : 65536));
}

void launch_c128_to_c64(const void* src, void* dst, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(c128_to_c64_kernel, convert_grid(n), dim3(kThreads), 0, s, (const double2*)src, (float2*)dst,
                       n);
}

void launch_f64_to_f32(const double* src, float* dst, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(f64_to_f32_kernel, convert_grid(n), dim3(kThreads), 0, s, src, dst, n);
}

void launch_c64_to_c128(const void* src, void* dst, size_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(c64_to_c128_kernel, convert_grid(n), dim3(kThreads), 0, s, (const float2*)src, (double2*)dst,
                       n);
}
}  // namespace gpe
