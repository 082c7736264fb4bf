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
    // z passes: the wave kernels unless variant bit 0 asks for the LDS tiles
    if (a.axis == 2 && !(a.variant & 1)) {
        const unsigned zb = launch_zwave<T, NN>(a, s);
        if (zb) return zb;
    }
    // instantiated variants: z lean {F, I}, z full {F, I, IF}, y/x lean {F, I, FI} at full and
    // (variant bit 1) half tile width, y/x full {F}
    const bool lean = pass_is_lean(a);
    const bool half = lean && a.axis != 2 && (a.variant & 2) && lean_w<T, NN, 2>() < lean_w<T, NN, 1>();
    const int W = lean ? (half ? lean_w<T, NN, 2>() : lean_w<T, NN, 1>())
                       : (a.axis == 2 ? full_tile_w<T, NN, true>() : full_tile_w<T, NN, false>());
    const unsigned blocks = (a.M + W - 1) / W;
    const int shape = pass_shape(a.dir1, a.dir2);
#define GPE_LAUNCH(CONTIG, FULL, SHAPE, WD) \
    hipLaunchKernelGGL((fft_pass_kernel<T, NN, CONTIG, FULL, SHAPE, WD>), dim3(blocks), dim3(kThreads), 0, s, a)
    if (a.axis == 2) {
        if (lean && shape == SHAPE_F) GPE_LAUNCH(true, false, SHAPE_F, 1);
        else if (lean && shape == SHAPE_I) GPE_LAUNCH(true, false, SHAPE_I, 1);
        else if (shape == SHAPE_F) GPE_LAUNCH(true, true, SHAPE_F, 1);
        else if (shape == SHAPE_I) GPE_LAUNCH(true, true, SHAPE_I, 1);
        else if (shape == SHAPE_IF) GPE_LAUNCH(true, true, SHAPE_IF, 1);
        else return 0;
    } else if (half) {
        if (shape == SHAPE_F) GPE_LAUNCH(false, false, SHAPE_F, 2);
        else if (shape == SHAPE_I) GPE_LAUNCH(false, false, SHAPE_I, 2);
        else if (shape == SHAPE_FI) GPE_LAUNCH(false, false, SHAPE_FI, 2);
        else return 0;
    } else {
        if (lean && shape == SHAPE_F) GPE_LAUNCH(false, false, SHAPE_F, 1);
        else if (lean && shape == SHAPE_I) GPE_LAUNCH(false, false, SHAPE_I, 1);
        else if (lean && shape == SHAPE_FI) GPE_LAUNCH(false, false, SHAPE_FI, 1);
        else if (!lean && shape == SHAPE_F) GPE_LAUNCH(false, true, SHAPE_F, 1);
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
