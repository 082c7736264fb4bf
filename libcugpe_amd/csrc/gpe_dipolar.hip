// Pointwise kernels of the dipolar path (INTERACTIONS -2): the density as a
// complex array for the convolution transform (__gpe_compute_density2C__,
// dev/gpe_engine.cu:618-626), the Strang-Euler potential step with the
// convolution potential (__gpe_eulerstrang_exp_Vstep__, dev/gpe_engine.cu:
// 984-1008) and the partial sums of the real-space dipolar energy
// (__gpe_compute_vdip_energy2__, dev/gpe_engine.cu:1768-1782).  The
// convolution itself runs in the FFT passes (MID_VPP).
#include "gpe_kernels.hpp"

namespace gpe {

namespace {

template <typename T>
__global__ __launch_bounds__(kThreads) void density_c_kernel(const cplx<T>* __restrict__ A, cplx<T>* __restrict__ B,
                                                             double gamma, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const cplx<T> v = A[i];
    B[i] = mk((T)(gamma * ((double)v.x * (double)v.x + (double)v.y * (double)v.y)), (T)0);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void dipolar_vstep_kernel(cplx<T>* __restrict__ A, const cplx<T>* __restrict__ B,
                                                                 Phys<T> ph, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const unsigned nyz = (unsigned)(ph.ny * ph.nz);
    const unsigned ix = (unsigned)(i / nyz), r = (unsigned)(i % nyz);
    const unsigned iy = r / (unsigned)ph.nz, iz = r % (unsigned)ph.nz;
    const cplx<T> l = modify_psi<true>(ph, A[i], ix, iy, iz, (unsigned)i, ph.it);
    const T lv = vext<true>(ph, ix, iy, iz, (unsigned)i, ph.it) + dedfdn<true>(ph, density(ph, l), ph.it, (unsigned)i) +
                 B[i].x;
    A[i] = mul_exp(l, ph.c.x * lv, ph.c.y * lv);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void dipolar_e2_kernel(const cplx<T>* __restrict__ A,
                                                              const cplx<T>* __restrict__ B, double gamma, size_t n,
                                                              double* __restrict__ partials) {
    __shared__ double red[kThreads / 64];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const cplx<T> v = A[i];
        s += gamma * ((double)v.x * (double)v.x + (double)v.y * (double)v.y) * (double)B[i].x;
    }
    const double b = block_sum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = b;
}

inline dim3 grid_of(size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

}  // namespace

template <typename T>
void launch_density_c(const cplx<T>* A, cplx<T>* B, double gamma, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((density_c_kernel<T>), grid_of(n), dim3(kThreads), 0, s, A, B, gamma, n);
}

template <typename T>
void launch_dipolar_vstep(cplx<T>* A, const cplx<T>* B, const Phys<T>& ph, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((dipolar_vstep_kernel<T>), grid_of(n), dim3(kThreads), 0, s, A, B, ph, n);
}

template <typename T>
int launch_dipolar_e2(const cplx<T>* A, const cplx<T>* B, double gamma, size_t n, double* partials, hipStream_t s) {
    size_t b = (n + kThreads - 1) / kThreads;
    if (b > (size_t)kMaxReduceBlocks) b = kMaxReduceBlocks;
    hipLaunchKernelGGL((dipolar_e2_kernel<T>), dim3((unsigned)b), dim3(kThreads), 0, s, A, B, gamma, n, partials);
    return (int)b;
}

#define GPE_DIP_INST(T)                                                                              \
    template void launch_density_c<T>(const cplx<T>*, cplx<T>*, double, size_t, hipStream_t);       \
    template void launch_dipolar_vstep<T>(cplx<T>*, const cplx<T>*, const Phys<T>&, size_t, hipStream_t); \
    template int launch_dipolar_e2<T>(const cplx<T>*, const cplx<T>*, double, size_t, double*, hipStream_t);
GPE_DIP_INST(float)
GPE_DIP_INST(double)
#undef GPE_DIP_INST

}  // namespace gpe
