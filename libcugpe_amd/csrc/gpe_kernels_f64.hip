// complex128 instantiation of the engine kernels, plus the precision-free ones
#include "gpe_kernels_impl.hpp"

namespace gpe {
GPE_INSTANTIATE(double)

__global__ __launch_bounds__(kThreads) void finalize_kernel(const double* __restrict__ partials, int nb, int nq,
                                                            double* __restrict__ scal, int mode, int slot,
                                                            double npart, double dt, double div) {
    __shared__ double red[kThreads / 64];
    for (int q = 0; q < nq; q++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += kThreads) acc += partials[(size_t)q * nb + b];
        const double sum = block_sum(acc, red);
        if (threadIdx.x == 0) {
            if (mode == FIN_NORM) {
                scal[0] = sum;
                scal[1] = sqrt(npart / sum);
                scal[2] = -.5 * log(sum / npart) / dt;
            } else {
                scal[slot + q] = sum / div;
            }
        }
    }
}


bool fft_length_supported(int N) { return fft_tile_width<double>(N) > 0; }

// plain streaming copy (16-byte vectors); reference point for hardware byte counters
__global__ __launch_bounds__(kThreads) void copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                        size_t n16) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    size_t blocks = (n16 + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, (uint4*)dst, (const uint4*)src, n16);
}

// the step counter read by the passes of a run with user hooks: set once per
// run (outside any captured graph), advanced by one after every step
__global__ void set_counter_kernel(unsigned* c, unsigned v) { *c = v; }
__global__ void tick_kernel(unsigned* c) { *c = *c + 1u; }

void launch_set_counter(unsigned* c, unsigned v, hipStream_t s) {
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(1), 0, s, c, v);
}
void launch_tick(unsigned* c, hipStream_t s) { hipLaunchKernelGGL(tick_kernel, dim3(1), dim3(1), 0, s, c); }

void launch_finalize(const double* partials, int nb, int nq, double* scal, int mode, int slot, double npart,
                     double dt, double div, hipStream_t s) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, s, partials, nb, nq, scal, mode, slot, npart,
                       dt, div);
}
}  // namespace gpe
