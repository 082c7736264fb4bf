// Vortex detection that allocates no device memory: gpe_detect_vortices_noalloc.
//
// The windings kernel and the FP64 count reduction of gpe_engine_detect_vortices,
// run on a wave function the caller holds in device memory, writing into the
// caller's output buffers.  The only scratch memory, the per-workgroup partial
// counts of the reduction, is the file-scope device array `scratch` below: one
// copy per GPU, held until the process ends, shared by every call on that GPU
// (calls that overlap on one GPU race on it).  Every partial the reduction
// reads is written by the same call's windings kernel first, so nothing in it
// needs clearing between calls.
#include <mutex>

#include "gpe_engine.h"
#include "gpe_kernels.hpp"

namespace gpe {
namespace {

// reduction partials (one per workgroup of the windings kernel, at most
// kMaxReduceBlocks) followed by one slot for the count when the caller passes
// no count buffer
constexpr int kScratchDoubles = kMaxReduceBlocks + 32;
static __device__ __align__(256) double scratch[kScratchDoubles];

// largest lattice: the windings kernel indexes points with 32-bit unsigned
// arithmetic (three windings per point are addressed as o * n + i)
constexpr unsigned long long kMaxPoints = (1ull << 31) - 1;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on the current GPU, looked up once per GPU
int scratch_on_current_device(double** out) {
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return (int)err;
    if (dev < 0 || dev >= kMaxDevices) return GPE_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return (int)err;
        if (!p) return GPE_ERR_ARG;
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return 0;
}

template <typename T>
int detect_noalloc_t(const void* psi, int nx, int ny, int nz, double rho_min, int8_t* w, double* count,
                     hipStream_t s, double* sp) {
    const int nb = launch_windings<T>((const cplx<T>*)psi, nx, ny, nz, rho_min, w, sp, s);
    if (nb > kMaxReduceBlocks) return GPE_ERR_ARG;
    // the count goes to the caller's buffer, else to the slot after the partials
    double* dst = count ? count : sp + kMaxReduceBlocks;
    launch_finalize(sp, nb, 1, dst, FIN_STORE, 0, 1.0, 1.0, 1.0, s);
    const hipError_t err = hipGetLastError();
    return err == hipSuccess ? 0 : (int)err;
}

}  // namespace
}  // namespace gpe

extern "C" int gpe_detect_vortices_noalloc(const void* d_psi, int nx, int ny, int nz, int precision,
                                           double rho_min, int8_t* d_windings, double* d_count, void* stream) {
    using namespace gpe;
    if (!d_psi || !d_windings || nx < 1 || ny < 1 || nz < 1) return GPE_ERR_ARG;
    if (precision != GPE_PRECISION_C64 && precision != GPE_PRECISION_C128) return GPE_ERR_ARG;
    const unsigned long long n = (unsigned long long)nx * (unsigned long long)ny * (unsigned long long)nz;
    if (n > kMaxPoints) return GPE_ERR_TOO_LARGE;
    double* sp = nullptr;
    const int r = scratch_on_current_device(&sp);
    if (r) return r;
    const hipStream_t s = (hipStream_t)stream;
    return precision == GPE_PRECISION_C64 ? detect_noalloc_t<float>(d_psi, nx, ny, nz, rho_min, d_windings, d_count, s, sp)
                                          : detect_noalloc_t<double>(d_psi, nx, ny, nz, rho_min, d_windings, d_count, s, sp);
}
