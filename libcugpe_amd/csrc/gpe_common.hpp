// Shared device types and physics for the MI355X split-step GPE engine.
//
// Complex numbers are {re, im} pairs of float (complex64 mode) or double
// (complex128 mode) with the natural 8- or 16-byte alignment, laid out like
// cuDoubleComplex / std::complex so host buffers cross the C ABI unchanged.
//
// The physics functions restate the hooks of the reference's
// gpe_user_defined.h (orginal/gpe_user_defined.h:66-126,
// dev/gpe_user_defined.h:81-554) with the INTERACTIONS / VEXT choice made at
// run time from a uniform kernel argument instead of at compile time, so one
// library serves every lattice and physics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// the user hook file compiled into the library (build.py passes the file named
// by GPE_USER_HOOKS; include/ is on the include path)
#ifndef GPE_USER_HOOKS_FILE
#define GPE_USER_HOOKS_FILE "gpe_user_hooks.h"
#endif

namespace gpe {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cplx<R> mk(R a, R b) {
    cplx<R> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cmulc(cplx<R> a, cplx<R> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) { return mk(s * a.x, s * a.y); }
template <typename R>
__host__ __device__ __forceinline__ R cnorm2(cplx<R> a) { return a.x * a.x + a.y * a.y; }

// precision-generic math wrappers (full-accuracy library versions, never the
// reduced-precision intrinsics)
__device__ __forceinline__ float g_exp(float a) { return expf(a); }
__device__ __forceinline__ double g_exp(double a) { return exp(a); }
__device__ __forceinline__ void g_sincos(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ __forceinline__ void g_sincos(double a, double* s, double* c) { sincos(a, s, c); }
__device__ __forceinline__ float g_cos(float a) { return cosf(a); }
__device__ __forceinline__ double g_cos(double a) { return cos(a); }
__device__ __forceinline__ float g_pow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double g_pow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float g_sqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double g_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float g_atan2(float a, float b) { return atan2f(a, b); }
__device__ __forceinline__ double g_atan2(double a, double b) { return atan2(a, b); }
__device__ __forceinline__ float g_hypot(float a, float b) { return hypotf(a, b); }
__device__ __forceinline__ double g_hypot(double a, double b) { return hypot(a, b); }

// exp(z) = e^Re (cos Im, sin Im)  (dev/gpe_complex.cuh:167-175)
template <typename R>
__device__ __forceinline__ cplx<R> cexp_(R re, R im) {
    R t = g_exp(re);
    R s, c;
    g_sincos(im, &s, &c);
    return mk(t * c, t * s);
}

constexpr int kMaxUserParams = 32;

// Physics codes: INTERACTIONS (dev/gpe_user_defined.h:90,169,276,356) and
// VEXT (dev/gpe_user_defined.h:476,520) of the reference, plus VEXT 2 (cosine
// lattice) and VEXT 3 (caller-supplied array) of this engine, and 100 for the
// user hook file compiled into the library (gpe_user_hooks.h by default).
enum { INT_UFG = 0, INT_ETF = 1, INT_CONTACT = -1, INT_DIPOLAR = -2, INT_USER = 100 };
enum { VEXT_HARMONIC = 0, VEXT_WOODS_SAXON = 1, VEXT_COSINE = 2, VEXT_ARRAY = 3, VEXT_USER = 100 };

// Per-launch physics / step constants (kernel argument, uniform per launch).
template <typename R>
struct Phys {
    int nx, ny, nz;
    int interactions, vext;
    unsigned it;             // time index of the launch (host value)
    const unsigned* itp;     // time index in device memory (step passes with user hooks), or null
    unsigned it_off;         // added to the time index (the next step's Vstep1 in its own pass)
    int modify;              // apply the user gpe_modify_psi
    R gamma;                 // 1 particles, 2 dimers
    cplx<R> c;               // step coefficient gamma*dt/2/(-beta + i alpha)
    R up[kMaxUserParams];    // user parameters (d_user_param)
    const double* uparam;    // user parameters in FP64 per lattice [batch][kMaxUserParams] (user hooks)
    unsigned n1;             // points per lattice
    double dt, t0, npart;    // d_dt, d_t0, d_npart of the user hooks
    const R* varr;           // VEXT_ARRAY potential
    const R* vtab;           // separable potentials: per-axis terms [nx | ny | nz]
    R vx0, vy0;              // vortex line position
    R vQ;                    // vortex charge
};

// time index of a launch: the device counter of a captured step, else the host value
template <typename R>
__device__ __forceinline__ unsigned time_index(const Phys<R>& p) {
    return (p.itp ? *p.itp : p.it) + p.it_off;
}

// The user hook file, compiled as member functions of a context that supplies
// the names the reference's hooks use (d_user_param, d_nx .. d_npart, NX .. NZ,
// cuDoubleComplex) for one lattice.
struct UserHooks {
    typedef unsigned int uint;
    typedef cplx<double> cuDoubleComplex;
    typedef cplx<double> Complex;
    const double* d_user_param;
    uint d_nx, d_ny, d_nz;
    double d_dt, d_t0, d_npart;
    static __device__ __forceinline__ cuDoubleComplex make_cuDoubleComplex(double r, double i) { return mk(r, i); }
    static __device__ __forceinline__ double cuCreal(cuDoubleComplex a) { return a.x; }
    static __device__ __forceinline__ double cuCimag(cuDoubleComplex a) { return a.y; }
    static __device__ __forceinline__ cuDoubleComplex cuCadd(cuDoubleComplex a, cuDoubleComplex b) { return cadd(a, b); }
    static __device__ __forceinline__ cuDoubleComplex cuCsub(cuDoubleComplex a, cuDoubleComplex b) { return csub(a, b); }
    static __device__ __forceinline__ cuDoubleComplex cuCmul(cuDoubleComplex a, cuDoubleComplex b) { return cmul(a, b); }
    static __device__ __forceinline__ cuDoubleComplex cuConj(cuDoubleComplex a) { return mk(a.x, -a.y); }
    static __device__ __forceinline__ double cuCabs(cuDoubleComplex a) { return hypot(a.x, a.y); }
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define NX d_nx
#define NY d_ny
#define NZ d_nz
#include GPE_USER_HOOKS_FILE
#undef NX
#undef NY
#undef NZ
};

// the hook context of the lattice holding global offset `off`
template <typename R>
__device__ __forceinline__ UserHooks user_hooks(const Phys<R>& p, unsigned off) {
    UserHooks h;
    h.d_user_param = p.uparam + (size_t)(off / p.n1) * kMaxUserParams;
    h.d_nx = (unsigned)p.nx;
    h.d_ny = (unsigned)p.ny;
    h.d_nz = (unsigned)p.nz;
    h.d_dt = p.dt;
    h.d_t0 = p.t0;
    h.d_npart = p.npart;
    return h;
}

// gpe_density: orginal/gpe_engine.cu:384-387
template <typename R>
__device__ __forceinline__ R density(const Phys<R>& p, cplx<R> v) {
    return p.gamma * (v.x * v.x + v.y * v.y);
}

// The physics functions below take the user hook file only in kernels
// instantiated with UH = true (the passes of engines with INTERACTIONS or
// VEXT 100); every other kernel is compiled without the hook code.

// gpe_external_potential (hooks listed above) at time index it; ix, iy, iz
// lattice indices, off the global offset of the point
template <bool UH = false, typename R>
__device__ __forceinline__ R vext(const Phys<R>& p, unsigned ix, unsigned iy, unsigned iz, unsigned off, unsigned it) {
    if constexpr (UH) {
        if (p.vext == VEXT_USER) return (R)user_hooks(p, off).gpe_external_potential(ix, iy, iz, it);
    }
    switch (p.vext) {
    case VEXT_HARMONIC:
        // 0.5 x^2 wx^2 + 0.5 y^2 wy^2 + 0.5 z^2 wz^2, each term tabulated per axis
        return p.vtab[ix] + p.vtab[p.nx + iy] + p.vtab[p.nx + p.ny + iz];
    case VEXT_WOODS_SAXON: {
        const R x = (R)ix - (R)1.0 * (R)(p.nx / 2);
        const R y = (R)iy - (R)1.0 * (R)(p.ny / 2);
        const R r = g_sqrt(x * x + y * y);
        return p.up[0] * ((R)1. - (R)1. / ((R)1 + g_exp((r - p.up[2]) / p.up[1])));
    }
    case VEXT_COSINE:
        // A (cos(2 pi P ix/nx) + cos(2 pi P iy/ny) + cos(2 pi P iz/nz)), A = up[0], P = up[1],
        // the cosines tabulated per axis
        return p.up[0] * (p.vtab[ix] + p.vtab[p.nx + iy] + p.vtab[p.nx + p.ny + iz]);
    case VEXT_ARRAY:
        return p.varr[off];
    default:
        return (R)0;
    }
}

// gpe_dEDFdn at time index it (off: the point's global offset, for the user
// hooks' per-lattice parameters)
template <bool UH = false, typename R>
__device__ __forceinline__ R dedfdn(const Phys<R>& p, R rho, unsigned it, unsigned off) {
    const R PI = (R)3.14159265358979323846;
    if constexpr (UH) {
        if (p.interactions == INT_USER) return (R)user_hooks(p, off).gpe_dEDFdn((double)rho, it);
    }
    switch (p.interactions) {
    case INT_UFG:
        return (R)0.37 * g_pow((R)3.0 * PI * PI * rho, (R)(2.0 / 3.0)) / (R)2.0;
    case INT_ETF: {
        const R XI = (R)0.37, CT = (R)0.901;
        const R a = p.up[3];
        if (rho > (R)1e-15) {
            const R kF = g_pow((R)3.0 * PI * PI * rho, (R)(1.0 / 3.0));
            const R eF = (R)0.5 * kF * kF;
            const R x = (R)1.0 / (a * kF);
            const R D = (XI + x * ((R)1.0 + CT) + (R)3.0 * PI * XI * x * x);
            return XI * eF * (XI + (R)0.8 * x) / D +
                   (R)0.2 * XI * eF * (XI + x) * x * (((R)1.0 + CT) + (R)6.0 * PI * XI * x) / (D * D) -
                   (R)1.0 / ((R)2.0 * a * a);
        }
        return -(R)1.0 / ((R)2.0 * a * a);
    }
    case INT_CONTACT:
        return (R)4 * PI * p.up[3] * rho;
    default:
        return (R)0;
    }
}

// gpe_EDF (energy density) at time index it; evaluated in FP64 by the energy
// reductions (up: the user parameters in FP64)
template <bool UH = false, typename R>
__device__ __forceinline__ double edf(const Phys<R>& p, const double* up, double rho, unsigned it, unsigned off) {
    const double PI = 3.14159265358979323846;
    if constexpr (UH) {
        if (p.interactions == INT_USER) return user_hooks(p, off).gpe_EDF(rho, it);
    }
    switch (p.interactions) {
    case INT_UFG:
        return 0.37 * 0.6 * rho * pow(3.0 * PI * PI * rho, 2.0 / 3.0) / 2.;
    case INT_ETF: {
        const double XI = 0.37, CT = 0.901;
        const double a = up[3];
        const double kF = pow(3.0 * PI * PI * rho, 1.0 / 3.0);
        const double eF = 0.5 * kF * kF;
        const double x = (kF > 2.2250738585072014e-308) ? 1.0 / (a * kF) : 1.7976931348623157e308;
        return 0.6 * eF * rho * XI * (XI + x) / (XI + x * (1.0 + CT) + 3.0 * PI * XI * x * x) - rho / (2.0 * a * a);
    }
    case INT_CONTACT:
    case INT_DIPOLAR:
        return 2 * PI * up[3] * rho * rho;
    default:
        return 0.;
    }
}

// gpe_modify_psi of the user hooks at time index it; the identity otherwise
template <bool UH = false, typename R>
__device__ __forceinline__ cplx<R> modify_psi(const Phys<R>& p, cplx<R> v, unsigned ix, unsigned iy, unsigned iz,
                                              unsigned off, unsigned it) {
    if constexpr (UH) {
        if (p.modify) {
            const cplx<double> r = user_hooks(p, off).gpe_modify_psi(ix, iy, iz, it, mk((double)v.x, (double)v.y));
            return mk((R)r.x, (R)r.y);
        }
    }
    return v;
}

// Vortex-line imprint (dev/gpe_engine.cu:485-511): phase Q*atan2(x, y) with
// x = ix - nx/2 - x0, y = iy - ny/2 - y0, modulus kept.
template <typename R>
__device__ __forceinline__ cplx<R> imprint(const Phys<R>& p, cplx<R> v, unsigned ix, unsigned iy) {
    const R x = (R)ix - (R)1.0 * (R)(p.nx / 2) - p.vx0;
    const R y = (R)iy - (R)1.0 * (R)(p.ny / 2) - p.vy0;
    const R a = g_hypot(v.x, v.y);
    const R ph = g_atan2(x, y) * p.vQ;
    R s, c;
    g_sincos(ph, &s, &c);
    return mk(a * c, a * s);
}

}  // namespace gpe
