/*
 * FP64 CPU oracle of the split-step Gross-Pitaevskii engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing that is measured or shipped.  The product path
 * (libcugpe_amd/) never links or calls it.
 *
 * It restates, in plain C + OpenMP and FP64 throughout, the algorithm of the
 * reference's orginal/gpe_engine.cu (split step, energy, normalisation, quantum
 * friction, density, currents) and the vortex additions of dev/gpe_engine.cu
 * (imprint, evolve_vortex, chemical potential).  Each function in
 * gpe_oracle.c cites the reference file:line it follows.
 *
 * Where the reference calls cuFFT (closed source), the oracle uses its own
 * mixed-radix Stockham FFT (fft_oracle.c) pinned against the DFT definition:
 * FORWARD = exp(-2 pi i jk/n) per axis, INVERSE = exp(+...), unnormalised.
 *
 * Complex arrays are interleaved (re, im) doubles in C order [nx][ny][nz]
 * (iz fastest), i.e. the layout of cuDoubleComplex / std::complex<double>.
 */
#ifndef GPE_ORACLE_H
#define GPE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpeo gpeo_t;

/* physics selectors; values follow the reference's INTERACTIONS / VEXT macros
 * (dev/gpe_user_defined.h:90,169,276,356 and 476,520) plus two additions:
 * VEXT 2 = separable cosine lattice  V = A (cos(kx ix) + cos(ky iy) + cos(kz iz)),
 * VEXT 3 = caller-supplied array V[ixyz]. */
gpeo_t *gpeo_create(int nx, int ny, int nz, int interactions, int vext,
                    int gpe_for, double alpha, double beta, double dt,
                    double npart);
void gpeo_destroy(gpeo_t *e);

int gpeo_change_alpha_beta(gpeo_t *e, double alpha, double beta);
int gpeo_set_time(gpeo_t *e, double t0);
int gpeo_set_user_params(gpeo_t *e, int size, const double *params);
int gpeo_set_quantum_friction_coeff(gpeo_t *e, double qf);
int gpeo_set_vortex(gpeo_t *e, double x0, double y0, int Q);
int gpeo_set_vext_array(gpeo_t *e, const double *v);

int gpeo_set_psi(gpeo_t *e, double t, const double *psi);
int gpeo_get_psi(gpeo_t *e, double *t, double *psi);
int gpeo_normalize_psi(gpeo_t *e);
int gpeo_get_norm(gpeo_t *e, double *norm);

int gpeo_evolve_qf(gpeo_t *e, int nt, double *mu);
int gpeo_evolve_vortex(gpeo_t *e, int nt, double *mu);
int gpeo_evolve_vortex2(gpeo_t *e, int nt, double *mu);
int gpeo_evolve_vortex3(gpeo_t *e, int nt, double *mu);
int gpeo_evolve_enforced_phase(gpeo_t *e, int nt, double *mu);
int gpeo_set_phase(gpeo_t *e, const double *phase);
int gpeo_get_phase(gpeo_t *e, double *phase);
int gpeo_energy(gpeo_t *e, double *t, double *ekin, double *eint, double *eext);
int gpeo_get_density(gpeo_t *e, double *t, double *rho);
int gpeo_get_currents(gpeo_t *e, double *t, double *jx, double *jy, double *jz);

/* engine state introspection used by the pinning tests */
double gpeo_gamma(const gpeo_t *e);
int gpeo_tables(const gpeo_t *e, double *kx, double *ky, double *kz,
                double *Kx, double *Ky, double *Kz, double *step_coeff);
int gpeo_vext_value(const gpeo_t *e, int ix, int iy, int iz, double *v);
int gpeo_dedfdn_value(const gpeo_t *e, double rho, double *dv, double *edf);

/* stand-alone 3-D FFT in place: sign -1 forward, +1 inverse, unnormalised */
int gpeo_fft3d(int nx, int ny, int nz, double *data, int sign);
/* stand-alone 1-D FFT in place (length n), same conventions */
int gpeo_fft1d(int n, double *data, int sign);
/* exp(-2 pi i m / n) with exact octant reduction (re, im) */
void gpeo_unit_root(long m, long n, double *re, double *im);

/* Plaquette phase-winding vortex detection.  For every elementary plaquette of
 * the periodic lattice and each orientation o (0: xy plane, normal +z;
 * 1: yz plane, normal +x; 2: zx plane, normal +y) writes the integer winding
 * of psi around the plaquette's corners taken counter-clockwise about the
 * normal, into w[o*N + ixyz] where ixyz is the plaquette's lowest corner.
 * Plaquettes with a corner of |psi|^2 < rho_min get 0.  Returns the number
 * of non-zero windings.  psi may be given in FP64 (psi64) or FP32 (psi32,
 * pass psi64 = NULL). */
long gpeo_plaquette_windings(int nx, int ny, int nz, const double *psi64,
                             const float *psi32, double rho_min, int8_t *w);

/* Gradient dL/dV_ext[ixyz] (grad, N doubles) of
 *   L = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2 |psi(0)|^2)
 * over nt real-time steps from the current psi (the discrete adjoint of the
 * split step, states checkpointed every ckpt steps, 0: ceil(sqrt(nt))).
 * Real time without quantum friction and INTERACTIONS 0, -1 or -2 only
 * (otherwise -12).  Leaves psi at psi(T). */
int gpeo_adjoint_overlap(gpeo_t *e, int nt, const double *psi_target, int ckpt, double *loss, double *grad);

/* Dipolar bosons (INTERACTIONS -2): the Strang-Euler step of
 * gpe_evolve_dipolar, the k-space and real-space dipolar energies of
 * gpe_energy_dipolar / gpe_energy_dipolar2, and the convolution potential
 * V_dip = Re F^-1[Vpp(k) F[rho]]/N on its own (a_con = up[3], a_dip = up[4]). */
int gpeo_evolve_dipolar(gpeo_t *e, int nt, double *mu);
int gpeo_energy_dipolar(gpeo_t *e, double *t, double *edip);
int gpeo_energy_dipolar2(gpeo_t *e, double *t, double *edip);
int gpeo_dipolar_potential(gpeo_t *e, double *vdip);

int gpeo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
