// Pointwise kernels of the discrete adjoint of the real-time split step.
//
// The objective is L = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2 |psi(0)|^2) and the
// gradient is taken with respect to V_ext at every lattice point.  With
// dL = 2 Re sum conj(lambda) dpsi, one step backwards (forward quantities
// recomputed from the step's input psi0 and its output psi1) is
//
//   P1  phi = psi0 E0                       E0 = exp(c V0), V0 = Vext + g(rho(psi0))
//       phi <- K phi                        (kinetic step, FFT passes)
//   P2  phi3 = phi E0, V1 = (V0 + Vext + g(rho(phi3)))/2, W = exp(c V1)
//       vb = 2 Re(conj(lam) psi1 c), lam <- conj(W) lam, phi <- phi3
//       lam <- K^+ lam                      (conj kinetic factors)
//   P3  chi = psi0 W, lam0 = conj(W) lam, Vb1 = vb + 2 Re(conj(lam) chi c)
//       grad += Vb1/2, lp3 = (Vb1/2) g'(rho3) gamma phi3,
//       vb <- Vb1/2 + 2 Re(conj(lp3) phi3 c), lam <- conj(E0) lp3
//       lam <- K^+ lam
//   P4  lam <- lam0 + conj(E0) lam + gamma g'(rho0) Vb0 psi0,
//       Vb0 = vb + 2 Re(conj(lam) psi0 E0 c), grad += Vb0
//
// Every pointwise quantity is formed in FP64 from the stored states (complex64
// or complex128), so the factors exp(c V) are exact rotations and the gradient
// accumulates in FP64.  V_ext is read from a precomputed FP64 array holding the
// forward kernels' own values.
#include "gpe_kernels.hpp"

namespace gpe {

namespace {

struct AdjPhys {
    int interactions;
    double gamma;
    double up[kMaxUserParams];
    double cx, cy;  // step coefficient
};

__device__ __forceinline__ double adj_g(const AdjPhys& p, double rho) {
    const double PI = 3.14159265358979323846;
    switch (p.interactions) {
    case INT_UFG: return 0.37 * pow(3.0 * PI * PI * rho, 2.0 / 3.0) / 2.0;
    case INT_CONTACT: return 4 * PI * p.up[3] * rho;
    default: return 0.0;
    }
}
// derivative of dEDF/drho in rho (UFG: (2/3) g / rho, 0 at rho = 0)
__device__ __forceinline__ double adj_dg(const AdjPhys& p, double rho) {
    const double PI = 3.14159265358979323846;
    switch (p.interactions) {
    case INT_UFG: return rho > 0.0 ? (2.0 / 3.0) * adj_g(p, rho) / rho : 0.0;
    case INT_CONTACT: return 4 * PI * p.up[3];
    default: return 0.0;
    }
}

__device__ __forceinline__ cplx<double> ld(const cplx<float>* a, size_t i) {
    const cplx<float> v = a[i];
    return mk((double)v.x, (double)v.y);
}
__device__ __forceinline__ cplx<double> ld(const cplx<double>* a, size_t i) { return a[i]; }
__device__ __forceinline__ void st(cplx<float>* a, size_t i, cplx<double> v) { a[i] = mk((float)v.x, (float)v.y); }
__device__ __forceinline__ void st(cplx<double>* a, size_t i, cplx<double> v) { a[i] = v; }

__device__ __forceinline__ cplx<double> conj_(cplx<double> a) { return mk(a.x, -a.y); }
__device__ __forceinline__ cplx<double> ecv(const AdjPhys& p, double V) { return cexp_(p.cx * V, p.cy * V); }
// 2 Re(conj(l) y c)
__device__ __forceinline__ double sens(const AdjPhys& p, cplx<double> l, cplx<double> y) {
    const cplx<double> t = cmulc(y, l);  // y conj(l)
    return 2.0 * (t.x * p.cx - t.y * p.cy);
}
__device__ __forceinline__ double rho_of(const AdjPhys& p, cplx<double> v) { return p.gamma * (v.x * v.x + v.y * v.y); }

template <typename T>
__global__ __launch_bounds__(kThreads) void adj_p1(AdjPhys p, const double* __restrict__ vx,
                                                   const cplx<T>* __restrict__ psi0, cplx<T>* __restrict__ phi,
                                                   size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const cplx<double> a = ld(psi0, i);
    const double V0 = vx[i] + adj_g(p, rho_of(p, a));
    st(phi, i, cmul(a, ecv(p, V0)));
}

template <typename T>
__global__ __launch_bounds__(kThreads) void adj_p2(AdjPhys p, const double* __restrict__ vx,
                                                   const cplx<T>* __restrict__ psi0, const cplx<T>* __restrict__ psi1,
                                                   cplx<T>* __restrict__ phi, cplx<T>* __restrict__ lam,
                                                   double* __restrict__ vb, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const cplx<double> a = ld(psi0, i);
    const double V0 = vx[i] + adj_g(p, rho_of(p, a));
    const cplx<double> phi3 = cmul(ld(phi, i), ecv(p, V0));
    const double V1 = 0.5 * (V0 + vx[i] + adj_g(p, rho_of(p, phi3)));
    const cplx<double> W = ecv(p, V1);
    const cplx<double> l = ld(lam, i);
    vb[i] = sens(p, l, ld(psi1, i));
    st(lam, i, cmul(conj_(W), l));
    st(phi, i, phi3);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void adj_p3(AdjPhys p, const double* __restrict__ vx,
                                                   const cplx<T>* __restrict__ psi0, const cplx<T>* __restrict__ phi,
                                                   cplx<T>* __restrict__ lam, cplx<T>* __restrict__ lam0,
                                                   double* __restrict__ vb, double* __restrict__ grad, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const cplx<double> a = ld(psi0, i);
    const double V0 = vx[i] + adj_g(p, rho_of(p, a));
    const cplx<double> E0 = ecv(p, V0);
    const cplx<double> phi3 = ld(phi, i);
    const double r3 = rho_of(p, phi3);
    const double V1 = 0.5 * (V0 + vx[i] + adj_g(p, r3));
    const cplx<double> W = ecv(p, V1);
    const cplx<double> chi = cmul(a, W);
    const cplx<double> l = ld(lam, i);
    st(lam0, i, cmul(conj_(W), l));
    const double V1b = vb[i] + sens(p, l, chi);
    grad[i] += 0.5 * V1b;
    const double f = 0.5 * V1b * adj_dg(p, r3) * p.gamma;
    const cplx<double> lp3 = mk(f * phi3.x, f * phi3.y);
    vb[i] = 0.5 * V1b + sens(p, lp3, phi3);
    st(lam, i, cmul(conj_(E0), lp3));
}

template <typename T>
__global__ __launch_bounds__(kThreads) void adj_p4(AdjPhys p, const double* __restrict__ vx,
                                                   const cplx<T>* __restrict__ psi0, cplx<T>* __restrict__ lam,
                                                   const cplx<T>* __restrict__ lam0, const double* __restrict__ vb,
                                                   double* __restrict__ grad, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const cplx<double> a = ld(psi0, i);
    const double r0 = rho_of(p, a);
    const double V0 = vx[i] + adj_g(p, r0);
    const cplx<double> E0 = ecv(p, V0);
    const cplx<double> l = ld(lam, i);
    const double V0b = vb[i] + sens(p, l, cmul(a, E0));
    grad[i] += V0b;
    const double f = p.gamma * adj_dg(p, r0) * V0b;
    cplx<double> r = cadd(ld(lam0, i), cmul(conj_(E0), l));
    r = mk(r.x + f * a.x, r.y + f * a.y);
    st(lam, i, r);
}

// V_ext at every point exactly as the forward kernels evaluate it
template <typename T>
__global__ __launch_bounds__(kThreads) void adj_vfill(Phys<T> ph, double* __restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const unsigned nyz = (unsigned)(ph.ny * ph.nz);
    const unsigned ix = (unsigned)(i / nyz), r = (unsigned)(i % nyz);
    const unsigned iy = r / (unsigned)ph.nz, iz = r % (unsigned)ph.nz;
    out[i] = (double)vext(ph, ix, iy, iz, (unsigned)i, ph.it);
}

// per workgroup: Re and Im of sum conj(t) psi, sum |a|^2, sum |t|^2
template <typename T>
__global__ __launch_bounds__(kThreads) void adj_overlap(const cplx<T>* __restrict__ psi, const cplx<T>* __restrict__ tgt,
                                                        const cplx<T>* __restrict__ a0, size_t n,
                                                        double* __restrict__ partials) {
    __shared__ double red[kThreads / 64];
    double s[4] = {0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads) {
        const cplx<double> p = ld(psi, i), t = ld(tgt, i), q = ld(a0, i);
        s[0] += t.x * p.x + t.y * p.y;
        s[1] += t.x * p.y - t.y * p.x;
        s[2] += q.x * q.x + q.y * q.y;
        s[3] += t.x * t.x + t.y * t.y;
    }
    for (int k = 0; k < 4; k++) {
        const double b = block_sum(s[k], red);
        if (threadIdx.x == 0) partials[(size_t)k * gridDim.x + blockIdx.x] = b;
    }
}

// lam = (fr + i fi) tgt
template <typename T>
__global__ __launch_bounds__(kThreads) void adj_lam_init(cplx<T>* __restrict__ lam, const cplx<T>* __restrict__ tgt,
                                                         double fr, double fi, size_t n) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    st(lam, i, cmul(ld(tgt, i), mk(fr, fi)));
}

inline dim3 grid_of(size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

AdjPhys adj_phys(const AdjArgs& a) {
    AdjPhys p;
    p.interactions = a.interactions;
    p.gamma = a.gamma;
    for (int i = 0; i < kMaxUserParams; i++) p.up[i] = a.up[i];
    p.cx = a.cx;
    p.cy = a.cy;
    return p;
}

}  // namespace

template <typename T>
void launch_adjoint_step(const AdjArgs& a, int part, hipStream_t s) {
    const AdjPhys p = adj_phys(a);
    const cplx<T>* psi0 = (const cplx<T>*)a.psi0;
    const cplx<T>* psi1 = (const cplx<T>*)a.psi1;
    cplx<T>* phi = (cplx<T>*)a.phi;
    cplx<T>* lam = (cplx<T>*)a.lam;
    cplx<T>* lam0 = (cplx<T>*)a.lam0;
    switch (part) {
    case 1: hipLaunchKernelGGL((adj_p1<T>), grid_of(a.n), dim3(kThreads), 0, s, p, a.vx, psi0, phi, a.n); break;
    case 2:
        hipLaunchKernelGGL((adj_p2<T>), grid_of(a.n), dim3(kThreads), 0, s, p, a.vx, psi0, psi1, phi, lam, a.vb, a.n);
        break;
    case 3:
        hipLaunchKernelGGL((adj_p3<T>), grid_of(a.n), dim3(kThreads), 0, s, p, a.vx, psi0, (const cplx<T>*)phi, lam,
                           lam0, a.vb, a.grad, a.n);
        break;
    default:
        hipLaunchKernelGGL((adj_p4<T>), grid_of(a.n), dim3(kThreads), 0, s, p, a.vx, psi0, lam, (const cplx<T>*)lam0,
                           (const double*)a.vb, a.grad, a.n);
        break;
    }
}

template <typename T>
void launch_adjoint_vfill(const Phys<T>& ph, double* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((adj_vfill<T>), grid_of(n), dim3(kThreads), 0, s, ph, out, n);
}

template <typename T>
int launch_adjoint_overlap(const void* psi, const void* tgt, const void* a0, size_t n, double* partials,
                           hipStream_t s) {
    size_t b = (n + kThreads - 1) / kThreads;
    if (b > (size_t)kMaxReduceBlocks / 2) b = kMaxReduceBlocks / 2;  // 4 sums fit the engine's partials
    hipLaunchKernelGGL((adj_overlap<T>), dim3((unsigned)b), dim3(kThreads), 0, s, (const cplx<T>*)psi,
                       (const cplx<T>*)tgt, (const cplx<T>*)a0, n, partials);
    return (int)b;
}

template <typename T>
void launch_adjoint_lam_init(void* lam, const void* tgt, double fr, double fi, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((adj_lam_init<T>), grid_of(n), dim3(kThreads), 0, s, (cplx<T>*)lam, (const cplx<T>*)tgt, fr,
                       fi, n);
}

#define GPE_ADJ_INST(T)                                                                                      \
    template void launch_adjoint_step<T>(const AdjArgs&, int, hipStream_t);                                 \
    template void launch_adjoint_vfill<T>(const Phys<T>&, double*, size_t, hipStream_t);                     \
    template int launch_adjoint_overlap<T>(const void*, const void*, const void*, size_t, double*, hipStream_t); \
    template void launch_adjoint_lam_init<T>(void*, const void*, double, double, size_t, hipStream_t);
GPE_ADJ_INST(float)
GPE_ADJ_INST(double)
#undef GPE_ADJ_INST

}  // namespace gpe
