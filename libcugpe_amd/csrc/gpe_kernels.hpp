// Launchers of the engine's device kernels.  Each precision is compiled in its
// own translation unit (gpe_kernels_f32.hip, gpe_kernels_f64.hip).
#pragma once

#include "gpe_fft.hpp"
#include "gpe_zwave.hpp"
#include "gpe_fused.hpp"
#include "gpe_plane.hpp"

namespace gpe {

// lattice lengths with a compiled FFT pass
bool fft_length_supported(int N);
// pencils per workgroup for a compiled length
template <typename T>
int fft_tile_width(int N);
// pencils per workgroup of the full-width lean strided passes
template <typename T>
int fft_lean_width(int N);

// the complex128 host layout of psi to and from a complex64 state, element by
// element with round-to-nearest (the host buffers of gpe_set_psi / gpe_get_psi)
void launch_c128_to_c64(const void* src, void* dst, size_t n, hipStream_t s);
void launch_c64_to_c128(const void* src, void* dst, size_t n, hipStream_t s);
void launch_f64_to_f32(const double* src, float* dst, size_t n, hipStream_t s);

// one FFT axis pass; returns the number of workgroups launched (0 on error)
template <typename T>
unsigned launch_fft_pass(int N, const PassArgs<T>& a, hipStream_t s);
// the same for one compiled length (one object file per precision and length)
template <typename T, int N>
unsigned launch_fft_pass_n(const PassArgs<T>& a, hipStream_t s);

#define GPE_FFT_LENGTHS(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(512)
#define GPE_DECLARE_PASS(NN)                                                        \
    template <>                                                                     \
    unsigned launch_fft_pass_n<float, NN>(const PassArgs<float>& a, hipStream_t s); \
    template <>                                                                     \
    unsigned launch_fft_pass_n<double, NN>(const PassArgs<double>& a, hipStream_t s);
GPE_FFT_LENGTHS(GPE_DECLARE_PASS)
#undef GPE_DECLARE_PASS

enum FinalizeMode {
    FIN_NORM = 0,   // scal[0] = sum, scal[1] = sqrt(npart/sum), scal[2] = -log(sum/npart)/(2 dt)
    FIN_STORE = 1,  // scal[slot + q] = sum_q / div
};

void launch_finalize(const double* partials, int nb, int nq, double* scal, int mode, int slot, double npart,
                     double dt, double div, hipStream_t s);

// sum of gamma |psi|^2 over the lattice into partials[0..nb), nb returned
template <typename T>
int launch_norm_partials(const cplx<T>* A, unsigned n, double gamma, double* partials, hipStream_t s);

// A *= scal[1]
template <typename T>
void launch_scale(cplx<T>* A, unsigned n, const double* scal, hipStream_t s);

// out = gamma |psi|^2 (FP64)
template <typename T>
void launch_density(const cplx<T>* A, unsigned n, double gamma, double* out, hipStream_t s);

// out = (1/batch) sum_b gamma |psi_b|^2 over `batch` lattices of n points
// stored one after another (FP64, summed in lattice order)
template <typename T>
void launch_mean_density(const cplx<T>* A, unsigned n, unsigned batch, double gamma, double* out, hipStream_t s);

// plaquette windings w[3][n]; count of non-zero windings summed into partials, nb returned
template <typename T>
int launch_windings(const cplx<T>* A, int nx, int ny, int nz, double rho_min, int8_t* w, double* partials,
                    hipStream_t s);

// per-workgroup maxima of gamma |psi|^2 into partials[0..nb) (nb returned, at
// most kMaxReduceBlocks) and the densities along the three lines through the
// lattice centre into lines[nx | ny | nz] (FP64)
template <typename T>
int launch_cloud(const cplx<T>* A, int nx, int ny, int nz, double gamma, double* partials, double* lines,
                 hipStream_t s);

// phase (arg psi) into a FP64 array, and enforcing a stored phase: psi = |psi| exp(i phase)
template <typename T>
void launch_phase(const cplx<T>* A, unsigned n, double* out, hipStream_t s);
template <typename T>
void launch_enforce_phase(cplx<T>* A, unsigned n, const double* phase, hipStream_t s);

// complex128 <-> working precision
template <typename T>
void launch_from_c128(const cplx<double>* in, cplx<T>* out, unsigned n, hipStream_t s);
template <typename T>
void launch_to_c128(const cplx<T>* in, cplx<double>* out, unsigned n, hipStream_t s);

// the device time index of the steps of a run with user hooks: set it, advance it by one
void launch_set_counter(unsigned* c, unsigned v, hipStream_t s);
void launch_tick(unsigned* c, hipStream_t s);

// plain device-to-device copy kernel of `bytes` (a multiple of 16)
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

constexpr int kMaxReduceBlocks = 1024;

// whole real-time steps of L^3 lattices held in one workgroup's LDS
// (gpe_fused.hip): the cube lengths compiled per precision, and nt steps of
// `batch` lattices stored one after another (uh: the user hook kernels);
// returns non-zero for a length without a kernel
template <typename T>
bool fused_length_ok(int L);
template <typename T>
int launch_fused_steps(const FusedArgs<T>& a, int L, unsigned batch, bool uh, hipStream_t s);

// y-z plane passes (gpe_plane.hip, modes of PlaneMode in gpe_plane.hpp): the
// plane lengths compiled (ny = nz = L), and one pass over `planes` planes
// (lattices x nx) of L x L points; returns non-zero for a length without a kernel
bool plane_length_ok(int L);
template <typename T>
int launch_plane(const PassArgs<T>& a, int L, int mode, unsigned planes, bool uh, hipStream_t s);

// dipolar path (gpe_dipolar.hip)
template <typename T>
void launch_density_c(const cplx<T>* A, cplx<T>* B, double gamma, size_t n, hipStream_t s);
template <typename T>
void launch_dipolar_vstep(cplx<T>* A, const cplx<T>* B, const Phys<T>& ph, size_t n, hipStream_t s);
// partials of sum gamma |A|^2 Re B; nb returned
template <typename T>
int launch_dipolar_e2(const cplx<T>* A, const cplx<T>* B, double gamma, size_t n, double* partials, hipStream_t s);

// discrete adjoint of the real-time split step (gpe_adjoint.hip)
struct AdjArgs {
    int interactions;
    double gamma;
    double up[kMaxUserParams];
    double cx, cy;          // step coefficient c
    size_t n;
    const double* vx;       // V_ext per point (FP64)
    const void* psi0;       // step input and output states
    const void* psi1;
    void* phi;              // work: phi1 / phi2 / phi3
    void* lam;              // adjoint state
    void* lam0;             // work: the part of the new adjoint state carried to P4
    double* vb;             // potential sensitivities (FP64)
    double* grad;           // dL/dV_ext accumulated (FP64)
};
// part 1..4: the pointwise kernels P1..P4 of one backward step
template <typename T>
void launch_adjoint_step(const AdjArgs& a, int part, hipStream_t s);
template <typename T>
void launch_adjoint_vfill(const Phys<T>& ph, double* out, size_t n, hipStream_t s);
// partials [4][nb]: Re, Im of sum conj(tgt) psi, sum |a0|^2, sum |tgt|^2; nb returned
template <typename T>
int launch_adjoint_overlap(const void* psi, const void* tgt, const void* a0, size_t n, double* partials,
                           hipStream_t s);
template <typename T>
void launch_adjoint_lam_init(void* lam, const void* tgt, double fr, double fi, size_t n, hipStream_t s);

}  // namespace gpe
