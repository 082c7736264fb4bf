/*
 * The user hook file (INTERACTIONS / VEXT 100) compiled for the FP64 CPU
 * oracle (TEST INFRASTRUCTURE ONLY, like the rest of oracle/).
 *
 * The hook file is the one the GPU library is built with: a list of the
 * reference's device hook functions (dev/gpe_user_defined.h:81-84, 118-135,
 * 488-513) in their names and signatures, reading d_user_param, d_nx .. d_npart
 * and NX .. NZ.  Here those names are thread-local variables bound before
 * every call, so OpenMP threads evaluating hooks of different engines (or
 * lattices) never share them, and the functions become static C functions.
 */
#include <math.h>

#include "user_hooks.h"

typedef unsigned int uint;
typedef struct {
    double x, y;
} cuDoubleComplex;
typedef cuDoubleComplex Complex;

static inline cuDoubleComplex make_cuDoubleComplex(double r, double i)
{
    cuDoubleComplex c = {r, i};
    return c;
}

static __thread const double *d_user_param;
static __thread uint d_nx, d_ny, d_nz;
static __thread double d_dt, d_t0, d_npart;

#ifndef GPE_USER_HOOKS_FILE
#define GPE_USER_HOOKS_FILE "../include/gpe_user_hooks.h"
#endif

#define __device__
#define inline static inline
#define NX d_nx
#define NY d_ny
#define NZ d_nz
#include GPE_USER_HOOKS_FILE
#undef NX
#undef NY
#undef NZ
#undef inline
#undef __device__

static void bind(const gpeo_uh_ctx *c)
{
    d_user_param = c->up;
    d_nx = (uint)c->nx;
    d_ny = (uint)c->ny;
    d_nz = (uint)c->nz;
    d_dt = c->dt;
    d_t0 = c->t0;
    d_npart = c->npart;
}

double gpeo_uh_vext(const gpeo_uh_ctx *c, unsigned ix, unsigned iy, unsigned iz, unsigned it)
{
    bind(c);
    return gpe_external_potential(ix, iy, iz, it);
}

double gpeo_uh_dedfdn(const gpeo_uh_ctx *c, double rho, unsigned it)
{
    bind(c);
    return gpe_dEDFdn(rho, it);
}

double gpeo_uh_edf(const gpeo_uh_ctx *c, double rho, unsigned it)
{
    bind(c);
    return gpe_EDF(rho, it);
}

void gpeo_uh_modify(const gpeo_uh_ctx *c, unsigned ix, unsigned iy, unsigned iz, unsigned it, double *re, double *im)
{
    bind(c);
    cuDoubleComplex p = {*re, *im};
    p = gpe_modify_psi(ix, iy, iz, it, p);
    *re = p.x;
    *im = p.y;
}
