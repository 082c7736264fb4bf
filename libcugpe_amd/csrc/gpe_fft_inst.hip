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
    if (a.axis == 2) {
        const unsigned zb = launch_zwave<T, NN>(a, s);
        if (zb) return zb;
    }
    // instantiated variants: z lean {F, I}, z full {F, I, IF}, y/x lean {F, I, FI}, y/x full {F}
    const bool lean = pass_is_lean(a);
    const int W = lean ? tile_w<T, NN>() : (a.axis == 2 ? full_tile_w<T, NN, true>() : full_tile_w<T, NN, false>());
    const unsigned blocks = (a.M + W - 1) / W;
    const int shape = pass_shape(a.dir1, a.dir2);
#define GPE_LAUNCH(CONTIG, FULL, SHAPE) \
    hipLaunchKernelGGL((fft_pass_kernel<T, NN, CONTIG, FULL, SHAPE>), dim3(blocks), dim3(kThreads), 0, s, a)
    if (a.axis == 2) {
        if (lean && shape == SHAPE_F) GPE_LAUNCH(true, false, SHAPE_F);
        else if (lean && shape == SHAPE_I) GPE_LAUNCH(true, false, SHAPE_I);
        else if (shape == SHAPE_F) GPE_LAUNCH(true, true, SHAPE_F);
        else if (shape == SHAPE_I) GPE_LAUNCH(true, true, SHAPE_I);
        else if (shape == SHAPE_IF) GPE_LAUNCH(true, true, SHAPE_IF);
        else return 0;
    } else {
        if (lean && shape == SHAPE_F) GPE_LAUNCH(false, false, SHAPE_F);
        else if (lean && shape == SHAPE_I) GPE_LAUNCH(false, false, SHAPE_I);
        else if (lean && shape == SHAPE_FI) GPE_LAUNCH(false, false, SHAPE_FI);
        else if (!lean && shape == SHAPE_F) GPE_LAUNCH(false, true, SHAPE_F);
        else return 0;
    }
#undef GPE_LAUNCH
    return blocks;
}

template <>
unsigned launch_fft_pass_n<GPE_REAL, GPE_FFT_N>(const PassArgs<GPE_REAL>& a, hipStream_t s) {
    return launch_fft_pass_n_impl<GPE_REAL, GPE_FFT_N>(a, s);
}
}  // namespace gpe
