// One FFT axis-pass instantiation: compiled once per precision and lattice
// length with -DGPE_REAL=float|double -DGPE_FFT_N=<length>.
#include "gpe_kernels.hpp"

#if !defined(GPE_REAL) || !defined(GPE_FFT_N)
#error "compile with -DGPE_REAL=<float|double> -DGPE_FFT_N=<length>"
#endif

namespace gpe {
template <typename T, int NN>
unsigned launch_fft_pass_n_impl(const PassArgs<T>& a, hipStream_t s) {
    static_assert(plan_ok(NN), "bad FFT plan");
    constexpr int W = tile_w<T, NN>();
    const unsigned blocks = (a.M + W - 1) / W;
    if (a.axis == 2)
        hipLaunchKernelGGL((fft_pass_kernel<T, NN, true>), dim3(blocks), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL((fft_pass_kernel<T, NN, false>), dim3(blocks), dim3(kThreads), 0, s, a);
    return blocks;
}

template <>
unsigned launch_fft_pass_n<GPE_REAL, GPE_FFT_N>(const PassArgs<GPE_REAL>& a, hipStream_t s) {
    return launch_fft_pass_n_impl<GPE_REAL, GPE_FFT_N>(a, s);
}
}  // namespace gpe
