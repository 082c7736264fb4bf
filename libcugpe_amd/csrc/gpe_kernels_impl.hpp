// Kernel bodies and launchers, instantiated once per precision by
// gpe_kernels_f32.hip / gpe_kernels_f64.hip.
#pragma once

#include "gpe_kernels.hpp"

namespace gpe {

template <typename T>
unsigned launch_fft_pass_impl(int N, const PassArgs<T>& a, hipStream_t s) {
#define GPE_CASE(NN) \
    case NN: return launch_fft_pass_n<T, NN>(a, s);
    switch (N) {
        GPE_FFT_LENGTHS(GPE_CASE)
    default:
        return 0;
    }
#undef GPE_CASE
}

template <typename T, int NN>
constexpr int min_tile() {
    constexpr int w = full_tile_w<T, NN, true>();
    if constexpr (zw_lanes<T>(NN) != 0) return ZW<T, NN>::PENCILS < w ? ZW<T, NN>::PENCILS : w;
    return w;
}

template <typename T>
int fft_tile_width_impl(int N) {
#define GPE_CASE(NN) \
    case NN: return min_tile<T, NN>();  /* the narrowest tile of any pass */
    switch (N) {
        GPE_FFT_LENGTHS(GPE_CASE)
    default:
        return 0;
    }
#undef GPE_CASE
}

template <typename T>
int fft_lean_width_impl(int N) {
#define GPE_CASE(NN) \
    case NN: return lean_w<T, NN, 1>();
    switch (N) {
        GPE_FFT_LENGTHS(GPE_CASE)
    default:
        return 0;
    }
#undef GPE_CASE
}

// ---------------------------------------------------------- reductions

template <typename T>
__global__ __launch_bounds__(kThreads) void norm_partials_kernel(const cplx<T>* __restrict__ A, unsigned n,
                                                                 double gamma, double* __restrict__ partials) {
    __shared__ double red[kThreads / 64];
    double acc = 0.0;
    const unsigned stride = gridDim.x * kThreads;
    for (unsigned i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const cplx<T> v = A[i];
        acc += gamma * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scale_kernel(cplx<T>* __restrict__ A, unsigned n,
                                                         const double* __restrict__ scal) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) A[i] = cscale(A[i], (T)scal[1]);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void density_kernel(const cplx<T>* __restrict__ A, unsigned n, double gamma,
                                                           double* __restrict__ out) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const cplx<T> v = A[i];
        out[i] = gamma * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void mean_density_kernel(const cplx<T>* __restrict__ A, unsigned n,
                                                                unsigned batch, double gamma,
                                                                double* __restrict__ out) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        double acc = 0.0;
        for (unsigned b = 0; b < batch; b++) {
            const cplx<T> v = A[(size_t)b * n + i];
            acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        }
        out[i] = gamma * acc / (double)batch;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void phase_kernel(const cplx<T>* __restrict__ A, unsigned n,
                                                         double* __restrict__ out) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const cplx<T> v = A[i];
        out[i] = atan2((double)v.y, (double)v.x);
    }
}

// dev/gpe_engine.cu:584-591: psi = |psi| exp(i phase)
template <typename T>
__global__ __launch_bounds__(kThreads) void enforce_phase_kernel(cplx<T>* __restrict__ A, unsigned n,
                                                                 const double* __restrict__ phase) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const cplx<T> v = A[i];
        const T a = g_sqrt(v.x * v.x + v.y * v.y);
        T s, c;
        g_sincos((T)phase[i], &s, &c);
        A[i] = mk(a * c, a * s);
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void from_c128_kernel(const cplx<double>* __restrict__ in,
                                                             cplx<T>* __restrict__ out, unsigned n) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = mk((T)in[i].x, (T)in[i].y);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void to_c128_kernel(const cplx<T>* __restrict__ in,
                                                           cplx<double>* __restrict__ out, unsigned n) {
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = mk((double)in[i].x, (double)in[i].y);
}

// -------------------------------------------------- plaquette windings
//
// Winding of the phase of psi around each elementary plaquette, counted in
// quarter turns: the quadrant of every corner is compared with the next
// corner's; opposite quadrants are resolved by the sign of the cross product.
// All arithmetic is on FP64 values with separately rounded products (no fused
// multiply-add), so the result is bit-identical to the CPU oracle's on the same
// input, in either precision.

__device__ __forceinline__ int quadrant(double x, double y) {
    if (y >= 0.0 && x > 0.0) return 0;
    if (x <= 0.0 && y > 0.0) return 1;
    if (y <= 0.0 && x < 0.0) return 2;
    return 3;
}

__device__ __forceinline__ int quarter_turns(double ax, double ay, double bx, double by) {
    const int d = (quadrant(bx, by) - quadrant(ax, ay)) & 3;
    if (d == 1) return 1;
    if (d == 3) return -1;
    if (d == 0) return 0;
    const double cr = __dsub_rn(__dmul_rn(ax, by), __dmul_rn(ay, bx));
    return (cr < 0.0) ? -2 : 2;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void windings_kernel(const cplx<T>* __restrict__ A, int nx, int ny, int nz,
                                                            double rho_min, int8_t* __restrict__ w,
                                                            double* __restrict__ partials) {
    __shared__ double red[kThreads / 64];
    const unsigned n = (unsigned)nx * ny * nz;
    double cnt = 0.0;
    const unsigned stride = gridDim.x * kThreads;
    for (unsigned i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const unsigned ix = i / ((unsigned)ny * nz);
        const unsigned r = i - ix * ((unsigned)ny * nz);
        const unsigned iy = r / (unsigned)nz;
        const unsigned iz = r - iy * (unsigned)nz;
        const unsigned jx = (ix + 1 == (unsigned)nx) ? 0 : ix + 1;
        const unsigned jy = (iy + 1 == (unsigned)ny) ? 0 : iy + 1;
        const unsigned jz = (iz + 1 == (unsigned)nz) ? 0 : iz + 1;
        auto at = [&](unsigned a, unsigned b, unsigned c) { return (a * (unsigned)ny + b) * (unsigned)nz + c; };
        const unsigned corners[3][4] = {{at(ix, iy, iz), at(jx, iy, iz), at(jx, jy, iz), at(ix, jy, iz)},
                                        {at(ix, iy, iz), at(ix, jy, iz), at(ix, jy, jz), at(ix, iy, jz)},
                                        {at(ix, iy, iz), at(ix, iy, jz), at(jx, iy, jz), at(jx, iy, iz)}};
#pragma unroll
        for (int o = 0; o < 3; o++) {
            double re[4], im[4];
            bool masked = false;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const cplx<T> v = A[corners[o][c]];
                re[c] = (double)v.x;
                im[c] = (double)v.y;
                const double r2 = __dadd_rn(__dmul_rn(re[c], re[c]), __dmul_rn(im[c], im[c]));
                if (!(r2 >= rho_min)) masked = true;
            }
            int s = 0;
            if (!masked) {
#pragma unroll
                for (int c = 0; c < 4; c++) s += quarter_turns(re[c], im[c], re[(c + 1) & 3], im[(c + 1) & 3]);
            }
            const int wind = s / 4;
            w[(size_t)o * n + i] = (int8_t)wind;
            if (wind != 0) cnt += 1.0;
        }
    }
    const double s = block_sum(cnt, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

static inline int reduce_grid(unsigned n) {
    unsigned b = (n + kThreads - 1) / kThreads;
    if (b > (unsigned)kMaxReduceBlocks) b = kMaxReduceBlocks;
    if (b < 1) b = 1;
    return (int)b;
}

// Cloud parameters (dev/gpe_cloud_params.h:60-119) on the device: the
// largest gamma |psi|^2 of each workgroup into partials[blockIdx], and from
// workgroup 0 the densities along the three lines through the lattice centre
// (nx/2, ny/2, nz/2) into lines[nx | ny | nz].
template <typename T>
__global__ __launch_bounds__(kThreads) void cloud_kernel(const cplx<T>* __restrict__ A, unsigned n, int nx, int ny,
                                                         int nz, double gamma, double* __restrict__ partials,
                                                         double* __restrict__ lines) {
    __shared__ double red[kThreads / 64];
    double m = 0.0;
    const unsigned stride = gridDim.x * kThreads;
    for (unsigned i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const cplx<T> v = A[i];
        m = fmax(m, gamma * ((double)v.x * (double)v.x + (double)v.y * (double)v.y));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; w++) m = fmax(m, red[w]);
        partials[blockIdx.x] = m;
    }
    if (blockIdx.x == 0) {
        const unsigned cx = (unsigned)(nx / 2), cy = (unsigned)(ny / 2), cz = (unsigned)(nz / 2);
        const unsigned nl = (unsigned)(nx + ny + nz);
        for (unsigned i = threadIdx.x; i < nl; i += kThreads) {
            unsigned ix = cx, iy = cy, iz = cz;
            if (i < (unsigned)nx) ix = i;
            else if (i < (unsigned)(nx + ny)) iy = i - (unsigned)nx;
            else iz = i - (unsigned)(nx + ny);
            const cplx<T> v = A[((size_t)ix * (unsigned)ny + iy) * (unsigned)nz + iz];
            lines[i] = gamma * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
        }
    }
}

template <typename T>
int launch_norm_partials_impl(const cplx<T>* A, unsigned n, double gamma, double* partials, hipStream_t s) {
    const int nb = reduce_grid(n);
    hipLaunchKernelGGL((norm_partials_kernel<T>), dim3(nb), dim3(kThreads), 0, s, A, n, gamma, partials);
    return nb;
}

template <typename T>
int launch_windings_impl(const cplx<T>* A, int nx, int ny, int nz, double rho_min, int8_t* w, double* partials,
                         hipStream_t s) {
    const int nb = reduce_grid((unsigned)nx * ny * nz);
    hipLaunchKernelGGL((windings_kernel<T>), dim3(nb), dim3(kThreads), 0, s, A, nx, ny, nz, rho_min, w, partials);
    return nb;
}

static inline dim3 ew_grid(unsigned n) { return dim3((n + kThreads - 1) / kThreads); }

#define GPE_INSTANTIATE(T)                                                                                        \
    template <>                                                                                                   \
    unsigned launch_fft_pass<T>(int N, const PassArgs<T>& a, hipStream_t s) {                                     \
        return launch_fft_pass_impl<T>(N, a, s);                                                                  \
    }                                                                                                             \
    template <>                                                                                                   \
    int fft_tile_width<T>(int N) {                                                                                \
        return fft_tile_width_impl<T>(N);                                                                         \
    }                                                                                                             \
    template <>                                                                                                   \
    int fft_lean_width<T>(int N) {                                                                                \
        return fft_lean_width_impl<T>(N);                                                                         \
    }                                                                                                             \
    template <>                                                                                                   \
    int launch_norm_partials<T>(const cplx<T>* A, unsigned n, double gamma, double* partials, hipStream_t s) {    \
        return launch_norm_partials_impl<T>(A, n, gamma, partials, s);                                            \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_scale<T>(cplx<T> * A, unsigned n, const double* scal, hipStream_t s) {                            \
        hipLaunchKernelGGL((scale_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, A, n, scal);                      \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_density<T>(const cplx<T>* A, unsigned n, double gamma, double* out, hipStream_t s) {              \
        hipLaunchKernelGGL((density_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, A, n, gamma, out);              \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_mean_density<T>(const cplx<T>* A, unsigned n, unsigned batch, double gamma, double* out,          \
                                hipStream_t s) {                                                                  \
        hipLaunchKernelGGL((mean_density_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, A, n, batch, gamma, out);  \
    }                                                                                                             \
    template <>                                                                                                   \
    int launch_windings<T>(const cplx<T>* A, int nx, int ny, int nz, double rho_min, int8_t* w, double* partials, \
                           hipStream_t s) {                                                                       \
        return launch_windings_impl<T>(A, nx, ny, nz, rho_min, w, partials, s);                                   \
    }                                                                                                             \
    template <>                                                                                                   \
    int launch_cloud<T>(const cplx<T>* A, int nx, int ny, int nz, double gamma, double* partials, double* lines,  \
                        hipStream_t s) {                                                                          \
        const unsigned n = (unsigned)nx * (unsigned)ny * (unsigned)nz;                                            \
        unsigned nb = (n + kThreads - 1) / kThreads;                                                              \
        if (nb > (unsigned)kMaxReduceBlocks) nb = kMaxReduceBlocks;                                               \
        hipLaunchKernelGGL((cloud_kernel<T>), dim3(nb), dim3(kThreads), 0, s, A, n, nx, ny, nz, gamma, partials,  \
                           lines);                                                                                \
        return (int)nb;                                                                                           \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_phase<T>(const cplx<T>* A, unsigned n, double* out, hipStream_t s) {                              \
        hipLaunchKernelGGL((phase_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, A, n, out);                       \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_enforce_phase<T>(cplx<T> * A, unsigned n, const double* phase, hipStream_t s) {                   \
        hipLaunchKernelGGL((enforce_phase_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, A, n, phase);             \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_from_c128<T>(const cplx<double>* in, cplx<T>* out, unsigned n, hipStream_t s) {                   \
        hipLaunchKernelGGL((from_c128_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, in, out, n);                  \
    }                                                                                                             \
    template <>                                                                                                   \
    void launch_to_c128<T>(const cplx<T>* in, cplx<double>* out, unsigned n, hipStream_t s) {                     \
        hipLaunchKernelGGL((to_c128_kernel<T>), ew_grid(n), dim3(kThreads), 0, s, in, out, n);                    \
    }

}  // namespace gpe
