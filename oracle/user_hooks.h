/*
 * Oracle-side entry points of the user hook file (TEST INFRASTRUCTURE ONLY):
 * gpe_external_potential, gpe_dEDFdn, gpe_EDF and gpe_modify_psi evaluated
 * with the parameters and lattice of one oracle engine.
 */
#ifndef GPEO_USER_HOOKS_H
#define GPEO_USER_HOOKS_H

typedef struct {
    const double *up; /* d_user_param */
    int nx, ny, nz;
    double dt, t0, npart;
} gpeo_uh_ctx;

double gpeo_uh_vext(const gpeo_uh_ctx *c, unsigned ix, unsigned iy, unsigned iz, unsigned it);
double gpeo_uh_dedfdn(const gpeo_uh_ctx *c, double rho, unsigned it);
double gpeo_uh_edf(const gpeo_uh_ctx *c, double rho, unsigned it);
void gpeo_uh_modify(const gpeo_uh_ctx *c, unsigned ix, unsigned iy, unsigned iz, unsigned it, double *re, double *im);

#endif
