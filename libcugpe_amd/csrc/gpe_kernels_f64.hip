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

void launch_finalize(const double* partials, int nb, int nq, double* scal, int mode, int slot, double npart,
                     double dt, double div, hipStream_t s) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, s, partials, nb, nq, scal, mode, slot, npart,
                       dt, div);
}
}  // namespace gpe
