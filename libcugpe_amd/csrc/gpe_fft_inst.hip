// One FFT axis-pass instantiation: compiled once per precision and lattice
// length with -DGPE_REAL=float|double -DGPE_FFT_N=<length>.
#include "gpe_kernels.hpp"

#if !defined(GPE_REAL) || !defined(GPE_FFT_N)
#error "compile with -DGPE_REAL=<float|double> -DGPE_FFT_N=<length>"
#endif

namespace gpe {
template <typename T, int NN>
unsigned launch_fft_pass_n_impl(const PassArgs<T>& a, hipStream_t s) {
    static_assert(plan_ok(NN, sizeof(T) == 4), "bad FFT plan");
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
// lean strided passes: the axis (0 x, 1 y) is a template argument
#define GPE_LAUNCH_S(SHAPE, WD, MK, CJ)                                                                        \
    do {                                                                                                       \
        if (a.axis == 0)                                                                                       \
            hipLaunchKernelGGL((fft_pass_kernel<T, NN, false, false, SHAPE, WD, MK, CJ, 0>), dim3(blocks),     \
                               dim3(kThreads), 0, s, a);                                                       \
        else                                                                                                   \
            hipLaunchKernelGGL((fft_pass_kernel<T, NN, false, false, SHAPE, WD, MK, CJ, 1>), dim3(blocks),     \
                               dim3(kThreads), 0, s, a);                                                       \
    } while (0)
// turn-arounds with a mid op: the x pass, and the y pass of the split kinetic factor
#define GPE_LAUNCH_MID(WD, MK, CJ)                                                                            \
    hipLaunchKernelGGL((fft_pass_kernel<T, NN, false, false, SHAPE_FI, WD, MK, CJ, 0>), dim3(blocks),          \
                       dim3(kThreads), 0, s, a)
// (PV: the alternative radix schedule of variant bit 6, for lengths that have one)
#define GPE_LAUNCH_SEP_PV(WD, PV)                                                                              \
    do {                                                                                                       \
        if (a.axis == 0)                                                                                       \
            hipLaunchKernelGGL((fft_pass_kernel<T, NN, false, false, SHAPE_FI, WD, MK_SEP, false, 0, false, PV>), \
                               dim3(blocks), dim3(kThreads), 0, s, a);                                         \
        else                                                                                                   \
            hipLaunchKernelGGL((fft_pass_kernel<T, NN, false, false, SHAPE_FI, WD, MK_SEP, false, 1, false, PV>), \
                               dim3(blocks), dim3(kThreads), 0, s, a);                                         \
    } while (0)
#define GPE_LAUNCH_SEP(WD)                                                                                     \
    do {                                                                                                       \
        if constexpr (has_alt_plan(NN, sizeof(T) == 4)) {                                                      \
            if (a.variant & 64) GPE_LAUNCH_SEP_PV(WD, 1);                                                      \
            else GPE_LAUNCH_SEP_PV(WD, 0);                                                                     \
        } else {                                                                                               \
            GPE_LAUNCH_SEP_PV(WD, 0);                                                                          \
        }                                                                                                      \
    } while (0)
    int mk = -1;
    bool cj = false;
    const bool has_mid = lean && mid_kind(a, mk, cj);
    // lean mid ops live in turn-arounds of the x axis, and of the y axis for
    // the split kinetic factor
    if (lean && has_mid && (shape != SHAPE_FI || a.axis == 2 || (a.axis == 1 && mk != MK_SEP))) return 0;
    // pipelined persistent strided passes (variant bit 2, complex64): every
    // tile must be W whole rows of W contiguous columns inside one lattice
    if constexpr (sizeof(T) == 4 && pipe_ok<NN>()) {
        constexpr int WP = lean_w<float, NN, 1>();
        const unsigned cols = a.axis == 1 ? (unsigned)a.ph.nz : (unsigned)(a.ph.ny * a.ph.nz);
        if (lean && (a.variant & 4) && a.axis != 2 && a.pipe_wgs > 0 && a.M % WP == 0 && cols % WP == 0) {
            const unsigned ntiles = a.M / WP;
            const unsigned pb = ntiles < a.pipe_wgs ? ntiles : a.pipe_wgs;
#define GPE_PIPE(SH, MKV, CJV, AXV) \
    hipLaunchKernelGGL((fft_pipe_kernel<NN, SH, MKV, CJV, AXV>), dim3(pb), dim3(kThreads), 0, s, a)
            if (shape == SHAPE_FI && has_mid && mk == MK_SEP) {
                if (a.axis == 0) GPE_PIPE(SHAPE_FI, MK_SEP, false, 0); else GPE_PIPE(SHAPE_FI, MK_SEP, false, 1);
                return pb;
            }
            if (shape == SHAPE_FI && a.axis == 0) {
                if (!has_mid) GPE_PIPE(SHAPE_FI, -1, false, 0);
                else switch (mk) {
                    case MK_EXPT_HL: if (cj) GPE_PIPE(SHAPE_FI, MK_EXPT_HL, true, 0); else GPE_PIPE(SHAPE_FI, MK_EXPT_HL, false, 0); break;
                    case MK_EXPT_64: if (cj) GPE_PIPE(SHAPE_FI, MK_EXPT_64, true, 0); else GPE_PIPE(SHAPE_FI, MK_EXPT_64, false, 0); break;
                    case MK_K2: GPE_PIPE(SHAPE_FI, MK_K2, false, 0); break;
                    case MK_VPP: GPE_PIPE(SHAPE_FI, MK_VPP, false, 0); break;
                    default: GPE_PIPE(SHAPE_FI, MK_KA, false, 0); break;
                }
                return pb;
            }
            if (!has_mid && (shape == SHAPE_F || shape == SHAPE_I)) {
                if (shape == SHAPE_F) {
                    if (a.axis == 0) GPE_PIPE(SHAPE_F, -1, false, 0); else GPE_PIPE(SHAPE_F, -1, false, 1);
                } else {
                    if (a.axis == 0) GPE_PIPE(SHAPE_I, -1, false, 0); else GPE_PIPE(SHAPE_I, -1, false, 1);
                }
                return pb;
            }
#undef GPE_PIPE
        }
    }
    if (a.axis == 2 && a.ph.modify && !lean) {
        // the physics z passes of engines with user hooks (INTERACTIONS / VEXT 100)
#define GPE_LAUNCH_UH(SHAPE)                                                                              \
    hipLaunchKernelGGL((fft_pass_kernel<T, NN, true, true, SHAPE, 1, -1, false, -1, true>), dim3(blocks), \
                       dim3(kThreads), 0, s, a)
        if (shape == SHAPE_F) GPE_LAUNCH_UH(SHAPE_F);
        else if (shape == SHAPE_I) GPE_LAUNCH_UH(SHAPE_I);
        else if (shape == SHAPE_IF) GPE_LAUNCH_UH(SHAPE_IF);
        else return 0;
#undef GPE_LAUNCH_UH
    } else if (a.axis == 2) {
        if (lean && shape == SHAPE_F) GPE_LAUNCH(true, false, SHAPE_F, 1);
        else if (lean && shape == SHAPE_I) GPE_LAUNCH(true, false, SHAPE_I, 1);
        else if (shape == SHAPE_F) GPE_LAUNCH(true, true, SHAPE_F, 1);
        else if (shape == SHAPE_I) GPE_LAUNCH(true, true, SHAPE_I, 1);
        else if (shape == SHAPE_IF) GPE_LAUNCH(true, true, SHAPE_IF, 1);
        else return 0;
    } else if (lean && shape == SHAPE_FI) {
        if (!has_mid) {
            if (half) GPE_LAUNCH_S(SHAPE_FI, 2, -1, false);
            else GPE_LAUNCH_S(SHAPE_FI, 1, -1, false);
        } else if (mk == MK_SEP) {
            if (half) GPE_LAUNCH_SEP(2);
            else GPE_LAUNCH_SEP(1);
        } else if (half) {
            switch (mk) {
            case MK_EXPT_HL: if (cj) GPE_LAUNCH_MID(2, MK_EXPT_HL, true); else GPE_LAUNCH_MID(2, MK_EXPT_HL, false); break;
            case MK_EXPT_64: if (cj) GPE_LAUNCH_MID(2, MK_EXPT_64, true); else GPE_LAUNCH_MID(2, MK_EXPT_64, false); break;
            case MK_K2: GPE_LAUNCH_MID(2, MK_K2, false); break;
            case MK_VPP: GPE_LAUNCH_MID(2, MK_VPP, false); break;
            default: GPE_LAUNCH_MID(2, MK_KA, false); break;
            }
        } else {
            switch (mk) {
            case MK_EXPT_HL: if (cj) GPE_LAUNCH_MID(1, MK_EXPT_HL, true); else GPE_LAUNCH_MID(1, MK_EXPT_HL, false); break;
            case MK_EXPT_64: if (cj) GPE_LAUNCH_MID(1, MK_EXPT_64, true); else GPE_LAUNCH_MID(1, MK_EXPT_64, false); break;
            case MK_K2: GPE_LAUNCH_MID(1, MK_K2, false); break;
            case MK_VPP: GPE_LAUNCH_MID(1, MK_VPP, false); break;
            default: GPE_LAUNCH_MID(1, MK_KA, false); break;
            }
        }
    } else if (lean) {
        if (shape == SHAPE_F) {
            if (half) GPE_LAUNCH_S(SHAPE_F, 2, -1, false);
            else GPE_LAUNCH_S(SHAPE_F, 1, -1, false);
        } else if (shape == SHAPE_I) {
            if (half) GPE_LAUNCH_S(SHAPE_I, 2, -1, false);
            else GPE_LAUNCH_S(SHAPE_I, 1, -1, false);
        } else {
            return 0;
        }
    } else {
        if (shape == SHAPE_F) GPE_LAUNCH(false, true, SHAPE_F, 1);
        else return 0;
    }
#undef GPE_LAUNCH_S
#undef GPE_LAUNCH_SEP
#undef GPE_LAUNCH_SEP_PV
#undef GPE_LAUNCH_MID
#undef GPE_LAUNCH
    return blocks;
}

template <>
unsigned launch_fft_pass_n<GPE_REAL, GPE_FFT_N>(const PassArgs<GPE_REAL>& a, hipStream_t s) {
    return launch_fft_pass_n_impl<GPE_REAL, GPE_FFT_N>(a, s);
}
}  // namespace gpe
