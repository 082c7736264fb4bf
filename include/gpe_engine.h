/*
 * MI355X-native split-step Gross-Pitaevskii engine: the C ABI.
 *
 * Drop-in boundary for the reference's gpe_engine.h (dev/gpe_engine.h:74-558,
 * orginal/gpe_engine.h:42-192).  Every function below is exported with C
 * linkage from libcugpe_amd/libgpe_hip.so; arguments are plain pointers,
 * sizes and doubles.  Two call surfaces share one implementation:
 *
 *  - the reference's singleton surface (gpe_create_engine ... gpe_destroy_engine),
 *    same names, argument meaning, defaults and error behaviour.  The lattice
 *    and physics the reference fixes at library build time (-DNX -DNY -DNZ
 *    -DINTERACTIONS -DVEXT) are registered at program start by this header when
 *    the driver is compiled with -DNX=.. -DNY=.. -DNZ=.., or by an explicit
 *    gpe_configure_lattice() call;
 *  - a handle-based surface (gpe_engine_*) for any number of engines of any
 *    lattice, physics and precision in one process.
 *
 * Layout: psi is complex128 {double x, y} (cuDoubleComplex) in C order
 * [nx][ny][nz], iz fastest.  Host buffers are owned by the caller and copied.
 *
 * Return codes (gpe_result_t): 0 success; > 0 a hipError_t from the HIP
 * runtime; < 0 an engine error, see GPE_ERR_* (the reference returns raw
 * cudaError_t / cufftResult values and -9 / -99 the same way).
 */
#ifndef GPE_ENGINE_H
#define GPE_ENGINE_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "gpe_timing.h"
#include "gpe_user_defined.h"

#ifdef __cplusplus
#define GPE_DEFAULT(v) = v
extern "C" {
#else
#define GPE_DEFAULT(v)
#endif

typedef int gpe_result_t;
typedef unsigned int uint;

/* complex128 with the layout and alignment of cuDoubleComplex */
typedef struct __attribute__((aligned(16))) {
    double x, y;
} cuCplx;
typedef cuCplx Complex; /* orginal/gpe_engine.h:171 */
/* the name the orginal drivers use (test_compare/gpe_imag.c:33-35 typedefs
 * cuCplx from it when built without -DDEV) */
#ifndef GPE_NO_CUDOUBLECOMPLEX
typedef cuCplx cuDoubleComplex;
#endif
#ifndef GPE_NO_CPLX
typedef double _Complex cplx; /* C99 complex128, dev/gpe_complex.cuh:36 */
#endif

#define GPE_SUCCESS 0
#define GPE_ERR_TOO_MANY_PARAMS (-9) /* size > MAX_USER_PARAMS (reference code)      */
#define GPE_ERR_NO_ENGINE (-10)      /* no engine created / NULL handle               */
#define GPE_ERR_LATTICE (-11)        /* lattice not configured or length unsupported  */
#define GPE_ERR_ARG (-12)            /* invalid argument                              */
#define GPE_ERR_VORTEX_NOT_SET (-13) /* gpe_evolve_vortex before gpe_set_vortex       */
#define GPE_ERR_PHASE_NOT_SET (-15)  /* enforced-phase evolution without a phase      */
#define GPE_ERR_TOO_LARGE (-16)      /* lattice above gpe_detect_vortices_noalloc's limit */
#define GPE_ERR_GAMMA (-99)          /* GPE_FOR not PARTICLES / DIMERS (reference code) */

#define GPE_PRECISION_C128 128
#define GPE_PRECISION_C64 64

/* ------------------------------------------------------------------------
 * Reference singleton surface (dev/gpe_engine.h:299-558).
 * --------------------------------------------------------------------- */

/* Lattice, physics and precision of the singleton engine; replaces the
 * reference's library build flags (dev/Makefile:37-51).  gpe_for 0 selects
 * the default for the interactions; precision GPE_PRECISION_C128 (default)
 * or GPE_PRECISION_C64. */
int gpe_configure_lattice(int nx, int ny, int nz, int interactions, int vext, int gpe_for, int precision);

/* dev/gpe_engine.h:299 */
void gpe_get_lattice(int *_nx, int *_ny, int *_nz);
/* dev/gpe_engine.h:310; nthreads is accepted and ignored (launch shapes are
 * the engine's own) */
int gpe_create_engine(double alpha, double beta, double dt, double npart, int nthreads GPE_DEFAULT(1024));
/* dev/gpe_engine.h:319 */
int gpe_change_alpha_beta(double alpha, double beta);
/* dev/gpe_engine.h:326, 333 */
int gpe_set_rte_evolution(void);
int gpe_set_ite_evolution(void);
/* dev/gpe_engine.h:341 */
int gpe_set_time(double t0);
/* dev/gpe_engine.h:352; copies `size` values, the rest are zero */
int gpe_set_user_params(int size, double *params);
/* dev/gpe_engine.h:361 */
int gpe_set_quantum_friction_coeff(double qfcoeff);
/* dev/gpe_engine.h:370 */
int gpe_set_vortex(double vortex_x0, double vortex_y0, int8_t Q GPE_DEFAULT(1));
/* dev/gpe_engine.h:378, 385 */
int gpe_get_phase(double *h_phase GPE_DEFAULT(NULL));
int gpe_set_phase(double *h_phase);
/* dev/gpe_engine.h:393, 402 */
int gpe_set_psi(double t, cuCplx *psi);
int gpe_get_psi(double *t, cuCplx *psi);
/* dev/gpe_engine.h:412, 425 */
int gpe_get_density(double *t, double *density);
int gpe_get_currents(double *t, double *jx, double *jy, double *jz);
/* dev/gpe_engine.h:432, 439 */
int gpe_normalize_psi(void);
int gpe_get_norm(double *norm);
/* dev/gpe_engine.h:446: evolves nt split steps (the quantum-friction
 * potential is applied when its coefficient is non-zero) */
int gpe_evolve(int nt);
/* dev/gpe_engine.h:455 */
int gpe_evolve_qf(int nt, double *chemical_potential GPE_DEFAULT(NULL));
/* dev/gpe_engine.h:465 */
int gpe_evolve_vortex(int nt, double *chemical_potential GPE_DEFAULT(NULL));
/* dev/gpe_engine.h:475: imprint only where |psi| > 1e-15 */
int gpe_evolve_vortex2(int nt, double *chemical_potential GPE_DEFAULT(NULL));
/* dev/gpe_engine.h:485: imprint in the predictor only */
int gpe_evolve_vortex3(int nt, double *chemical_potential GPE_DEFAULT(NULL));
/* dev/gpe_engine.h:494: phase from gpe_set_phase / gpe_get_phase enforced */
int gpe_evolve_enforced_phase(int nt, double *chemical_potential GPE_DEFAULT(NULL));
/* dev/gpe_engine.h:504, 514: dipolar bosons (INTERACTIONS -2), Strang-Euler
 * step with the convolution potential F^-1[Vpp(k) F[rho]]/N.  The reference's
 * (int, double*) overload is gpe_evolve_dipolar_mu at the C level; C++
 * callers get the overload below. */
int gpe_evolve_dipolar(int nt);
int gpe_evolve_dipolar_mu(int nt, double *chemical_potential);
/* dev/gpe_engine.h:524, 534: dipolar energy in k space (sum_k Vpp |rho_k|^2 / N)
 * and in real space ((1/2) sum rho V_dip) */
int gpe_energy_dipolar(double *t, double *edip);
int gpe_energy_dipolar2(double *t, double *edip);
/* dev/gpe_engine.h:545 */
int gpe_energy(double *t, double *ekin, double *eint, double *eext);
/* dev/gpe_engine.h:552 */
int gpe_destroy_engine(void);
/* dev/gpe_engine.h:558 */
void gpe_print_potential_type(void);
/* dev/gpe_engine.h:158-195 */
const char *gpe_get_error_string(gpe_result_t res);

/* Device selection and page-locked host memory for psi buffers: the drivers'
 * cudaSetDevice / cudaHostAlloc / cudaFreeHost, served by the library so a
 * driver needs no GPU runtime header (no counterpart in the reference's
 * engine). */
int gpe_set_device(int device);
int gpe_host_alloc(void **ptr, size_t bytes);
int gpe_host_free(void *ptr);

/* External potential array for VEXT 3 (no counterpart in the reference,
 * whose potentials are compiled in). */
int gpe_set_external_potential(const double *v);
/* Plaquette vortex detection on the current psi (no counterpart in the
 * reference; replaces its host tracker dev/gpe_vortex_detect.h:153-174):
 * windings[3][N] (NULL allowed), number of non-zero windings in *count. */
int gpe_detect_vortices(double rho_min, int8_t *windings, long *count);

/* ------------------------------------------------------------------------
 * Handle-based surface.
 * --------------------------------------------------------------------- */

typedef struct gpe_engine_s *gpe_engine_t;

typedef struct {
    int nx, ny, nz;
    int interactions; /* INTERACTIONS code */
    int vext;         /* VEXT code */
    int gpe_for;      /* PARTICLES, DIMERS or 0 for the default */
    int precision;    /* GPE_PRECISION_C128 or GPE_PRECISION_C64 */
} gpe_config_t;

int gpe_engine_create(gpe_engine_t *out, const gpe_config_t *cfg, double alpha, double beta, double dt, double npart);
int gpe_engine_destroy(gpe_engine_t e);
int gpe_engine_change_alpha_beta(gpe_engine_t e, double alpha, double beta);
int gpe_engine_set_time(gpe_engine_t e, double t0);
int gpe_engine_set_user_params(gpe_engine_t e, int size, const double *params);
int gpe_engine_set_quantum_friction_coeff(gpe_engine_t e, double qfcoeff);
int gpe_engine_set_vortex(gpe_engine_t e, double x0, double y0, int Q);
int gpe_engine_set_external_potential(gpe_engine_t e, const double *v);
/* The same from FP64 device memory (nx*ny*nz doubles; no host copy): the
 * caller's writes to d_v must be complete when the call starts (synchronize
 * the stream that produced it); the call returns after the engine has read
 * it.  No counterpart in the reference, whose potentials are host arrays. */
int gpe_engine_set_external_potential_device(gpe_engine_t e, const double *d_v);
int gpe_engine_set_psi(gpe_engine_t e, double t, const cuCplx *psi);
int gpe_engine_get_psi(gpe_engine_t e, double *t, cuCplx *psi);
int gpe_engine_normalize_psi(gpe_engine_t e);
int gpe_engine_get_norm(gpe_engine_t e, double *norm);
int gpe_engine_evolve_qf(gpe_engine_t e, int nt, double *chemical_potential);
int gpe_engine_evolve_vortex(gpe_engine_t e, int nt, double *chemical_potential);
int gpe_engine_evolve_vortex2(gpe_engine_t e, int nt, double *chemical_potential);
int gpe_engine_evolve_vortex3(gpe_engine_t e, int nt, double *chemical_potential);
int gpe_engine_evolve_enforced_phase(gpe_engine_t e, int nt, double *chemical_potential);
int gpe_engine_energy(gpe_engine_t e, double *t, double *ekin, double *eint, double *eext);
int gpe_engine_get_density(gpe_engine_t e, double *t, double *density);
int gpe_engine_get_currents(gpe_engine_t e, double *t, double *jx, double *jy, double *jz);
int gpe_engine_get_phase(gpe_engine_t e, double *h_phase);
int gpe_engine_set_phase(gpe_engine_t e, const double *h_phase);
int gpe_engine_detect_vortices(gpe_engine_t e, double rho_min, int8_t *windings, long *count);
/* Cloud parameters computed on the device (the host-array forms are in
 * gpe_cloud_params.h, dev/gpe_cloud_params.h:60-119): out[0] density at the
 * lattice centre, out[1] largest density, out[2..4] cloud range along x, y, z
 * (first distance from the centre where the density is below epsilon, NaN
 * when none is). */
int gpe_engine_cloud_params(gpe_engine_t e, double epsilon, double *out);
int gpe_engine_evolve_dipolar(gpe_engine_t e, int nt, double *chemical_potential);
/* method 1: gpe_energy_dipolar, 2: gpe_energy_dipolar2 */
int gpe_engine_energy_dipolar(gpe_engine_t e, int method, double *t, double *edip);
int gpe_engine_get_time(gpe_engine_t e, double *t);
/* Unnormalised 3-D DFT of psi in place, sign -1 forward / +1 inverse, by the
 * engine's own FFT passes (no counterpart in the reference, whose transform
 * is cuFFT's; exposed so the transform can be checked on its own). */
int gpe_engine_fft(gpe_engine_t e, int sign);
int gpe_engine_get_config(gpe_engine_t e, gpe_config_t *cfg);
/* device address of psi (complex64 or complex128 per the precision) and the
 * HIP stream the engine launches on, for interop and timing */
int gpe_engine_device_psi(gpe_engine_t e, void **ptr);
int gpe_engine_stream(gpe_engine_t e, void **stream);
int gpe_engine_synchronize(gpe_engine_t e);
/* Kernel timing with HIP events on the engine's stream: while enabled, every
 * FFT pass launched on the selected axis class (0: strided y and x passes,
 * 1: contiguous z passes, 2: y passes, 3: x passes) is bracketed by an event
 * pair; read returns the summed milliseconds and the number of timed launches. */
int gpe_engine_profile(gpe_engine_t e, int axis_class, int enable);
/* one plain copy kernel of the state array (psi -> work buffer), timed with
 * HIP events: the streaming-bandwidth reference of the bench and the
 * calibration point of hardware byte counters */
int gpe_engine_copy_benchmark(gpe_engine_t e, double *ms);
int gpe_engine_profile_read(gpe_engine_t e, double *total_ms, long *launches);
/* library build information (GPU target, compiled FFT lengths) */
const char *gpe_build_info(void);

/* Launcher choice of an engine, picked by timing (no counterpart in the
 * reference, whose transforms are cuFFT plans): bit 0 runs the z passes in
 * LDS tiles instead of wave registers, bit 1 halves the tile width of the
 * strided passes, bit 2 runs the strided passes of complex64 engines in
 * persistent workgroups that copy the next tile into LDS while transforming
 * the current one, bit 3 runs the kinetic factor in one x turn-around pass
 * between forward and inverse y passes, and bit 4 runs whole real-time steps
 * of cubes that fit in one workgroup's LDS (complex64 8^3 to 24^3,
 * complex128 8^3 to 16^3; no quantum friction, no pinning) in one kernel per
 * lattice, the lower bits serving the runs it does not cover, and bit 5
 * runs y-z planes of ny = nz = 32, 48 or 64 points in one pass (z and y
 * transforms with the physics between them: four passes per real-time step
 * instead of six; no quantum friction, no pinning), and bit 6 runs the
 * strided turn-around passes with the alternative radix schedule of their
 * length (the radices in the other order, or the first one split).  autotune
 * times
 * `steps` real-time steps per choice on the current state (restored
 * afterwards) and keeps the fastest. */
int gpe_engine_autotune(gpe_engine_t e, int steps, int *variant);
int gpe_engine_set_variant(gpe_engine_t e, int variant);
/* 1: an evolve call of 3 or more steps replays HIP graphs of the open pass,
 * a middle step and the last step, captured once per step plan and recaptured
 * after any parameter change; 0 (default, measured faster on one stream):
 * launch every pass directly.  The results are the same bit for bit (no
 * counterpart in the reference). */
int gpe_engine_set_graph(gpe_engine_t e, int on);

/* Vortex detection that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The windings and
 * count of gpe_engine_detect_vortices, computed by the same kernels from a
 * wave function in the caller's device memory: d_psi holds nx*ny*nz values in
 * C order (complex64 for precision GPE_PRECISION_C64, complex128 for
 * GPE_PRECISION_C128); d_windings (device memory, 3*nx*ny*nz int8) receives
 * windings[3][N]; *d_count (device memory, NULL allowed) receives the number
 * of non-zero windings as a double.  The work is enqueued on `stream` (a
 * hipStream_t, NULL for the null stream) and not waited for.  The reduction's
 * partial sums live in a fixed 8.2 KiB device array of the library, one per
 * GPU, which every call on that GPU shares: calls that overlap on one GPU
 * race.  nx*ny*nz above 2^31 - 1 returns GPE_ERR_TOO_LARGE, before any launch. */
int gpe_detect_vortices_noalloc(const void *d_psi, int nx, int ny, int nz, int precision, double rho_min,
                                int8_t *d_windings, double *d_count, void *stream);

/* Adjoint of a real-time run (no counterpart in the reference).  forward:
 * nt steps of gpe_evolve_qf from the current psi, keeping every ckpt-th state
 * (0: ceil(sqrt(nt))), and *loss = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2
 * |psi(0)|^2).  backward: dL/dV_ext at every point (N doubles) by the
 * discrete adjoint of the split step, recomputing one segment of states at a
 * time; psi stays psi(T).  Real time without quantum friction, INTERACTIONS
 * 0, -1 or -2 (GPE_ERR_ARG otherwise); V_ext any VEXT (gradient with respect
 * to the per-point values the step uses).  backward returns GPE_ERR_ARG when
 * a parameter of the engine (user parameters, potential, alpha/beta, quantum
 * friction, launcher variant, time origin) changed after the forward run. */
int gpe_engine_adjoint_forward(gpe_engine_t e, int nt, const cuCplx *psi_target, int ckpt, double *loss);
int gpe_engine_adjoint_backward(gpe_engine_t e, double *grad_vext);
/* The same gradient written to FP64 device memory (N doubles), complete when
 * the call returns. */
int gpe_engine_adjoint_backward_device(gpe_engine_t e, double *d_grad_vext);

/* ------------------------------------------------------------------------
 * Ensembles: many independent lattices of possibly different shapes
 * (truncated-Wigner realisations, parameter sweeps) advanced together in
 * real time.  No counterpart in the reference, which holds one lattice per
 * process (dev/gpe_engine.cu:43-44); every lattice follows the same split
 * step as gpe_evolve_qf.  Lattices of one shape share one batched engine;
 * a step is one HIP graph launch covering every group.
 * --------------------------------------------------------------------- */

typedef struct gpe_ensemble_s *gpe_ensemble_t;

/* dims: ninst triples (nx, ny, nz); alpha/beta/dt/npart as gpe_create_engine.
 * Evolution is real time only (beta must be 0 when gpe_ensemble_evolve runs). */
int gpe_ensemble_create(gpe_ensemble_t *out, int ninst, const int *dims, int interactions, int vext, int gpe_for,
                        int precision, double alpha, double beta, double dt, double npart);
int gpe_ensemble_destroy(gpe_ensemble_t s);
int gpe_ensemble_set_user_params(gpe_ensemble_t s, int size, const double *params);
int gpe_ensemble_set_quantum_friction_coeff(gpe_ensemble_t s, double qfcoeff);
/* Parameter sweeps: the user parameters of lattice inst alone, read by the
 * user hooks as its d_user_param (INTERACTIONS or VEXT 100; GPE_ERR_ARG for
 * built-in physics, which reads the ensemble's gpe_ensemble_set_user_params).
 * Copies `size` values and zeroes the rest, like gpe_set_user_params. */
int gpe_ensemble_set_instance_params(gpe_ensemble_t s, int inst, int size, const double *params);
/* VEXT 3 ensembles: the potential array of lattice inst (nx*ny*nz doubles of
 * its shape); every lattice needs one before gpe_ensemble_evolve, which
 * otherwise returns GPE_ERR_ARG without launching anything */
int gpe_ensemble_set_external_potential(gpe_ensemble_t s, int inst, const double *v);
int gpe_ensemble_set_psi(gpe_ensemble_t s, int inst, const cuCplx *psi);
int gpe_ensemble_get_psi(gpe_ensemble_t s, int inst, double *t, cuCplx *psi);
int gpe_ensemble_evolve(gpe_ensemble_t s, int nt);
/* 1: replay captured HIP graphs (default), 0: launch the passes directly */
int gpe_ensemble_set_graph(gpe_ensemble_t s, int on);
int gpe_ensemble_autotune(gpe_ensemble_t s, int steps);
/* launcher variant of one group (the bits of gpe_engine_set_variant) */
int gpe_ensemble_set_variant(gpe_ensemble_t s, int group, int variant);
int gpe_ensemble_groups(gpe_ensemble_t s, int *ngroups);
int gpe_ensemble_group_info(gpe_ensemble_t s, int group, int *nx, int *ny, int *nz, int *count, int *variant);
int gpe_ensemble_instance(gpe_ensemble_t s, int inst, int *group, int *slot);
/* mean of gamma |psi|^2 over the lattices of one group (FP64, nx*ny*nz values) */
int gpe_ensemble_mean_density(gpe_ensemble_t s, int group, double *out);

#ifdef __cplusplus
}
#endif

/* ------------------------------------------------------------------------
 * Driver-side helpers of the reference header (dev/gpe_engine.h:119-133,
 * 252-273): index macros, checked malloc and the abort-on-error wrapper.
 * --------------------------------------------------------------------- */
#define ixyz2ixiyiz(ixyz, _ix, _iy, _iz, i) \
    i = ixyz;                               \
    _ix = i / (ny * nz);                    \
    i = i - _ix * ny * nz;                  \
    _iy = i / nz;                           \
    _iz = i - _iy * nz;

#define ixiyiz2ixyz(_ixyz, _ix, _iy, _iz, ny, nz) _ixyz = _iz + nz * _iy + nz * ny * _ix

#define gpemalloc(pointer, size, type)                                              \
    if ((pointer = (type *)malloc((size) * sizeof(type))) == NULL) {                \
        fprintf(stderr, "error: cannot malloc()! file=`%s`, line=%d\n", __FILE__, __LINE__); \
        exit(EXIT_FAILURE);                                                         \
    }

#define gpe_exec(cmd, ierr)                                                                      \
    {                                                                                            \
        ierr = cmd;                                                                              \
        if (ierr) {                                                                              \
            fprintf(stderr, "error: cannot execute: %s\tfile=`%s`, line=%d\t", #cmd, __FILE__, __LINE__); \
            fprintf(stderr, "Error=%d (%s)\nExiting!\n", ierr, gpe_get_error_string(ierr));     \
            exit(EXIT_FAILURE);                                                                  \
        }                                                                                        \
    }

#ifdef __cplusplus
/* the reference's C++ overload gpe_evolve_dipolar(int, double*) (dev/gpe_engine.h:514) */
static inline int gpe_evolve_dipolar(int nt, double *chemical_potential)
{
    return gpe_evolve_dipolar_mu(nt, chemical_potential);
}
#endif

/* A driver compiled with -DNX=.. -DNY=.. -DNZ=.. (the reference's build flags)
 * registers its lattice and physics before main(), so the unchanged
 * gpe_create_engine(alpha, beta, dt, npart) call creates that engine. */
#if defined(NX) && defined(NY) && defined(NZ)
#ifndef GPE_PRECISION
#define GPE_PRECISION GPE_PRECISION_C128
#endif
static void __attribute__((constructor, used)) gpe_register_compiled_lattice_(void)
{
    gpe_configure_lattice(NX, NY, NZ, INTERACTIONS, VEXT, GPE_FOR, GPE_PRECISION);
}
#endif

#endif
