/*
 * FP64 CPU oracle of the split-step GPE engine (TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product path).
 *
 * Restatement of orginal/gpe_engine.cu (step, energy, normalisation, quantum
 * friction, density, currents) and of dev/gpe_engine.cu (chemical potential,
 * vortex imprint, evolve_vortex), with the physics hooks of
 * orginal/gpe_user_defined.h and dev/gpe_user_defined.h.  Every array is FP64
 * and every sum is taken in a fixed order that does not depend on the thread
 * count, so results are reproducible run to run.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

#include "gpe_oracle.h"
#include "user_hooks.h"

typedef struct { double re, im; } cd;

static inline cd cmul(cd a, cd b) { cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cd cscale(cd a, double s) { cd r = {s * a.re, s * a.im}; return r; }
/* exp(z) = e^Re (cos Im, sin Im): dev/gpe_complex.cuh:167-175, orginal/gpe_engine.cu:358-366 */
static inline cd cexp_(cd z) { double t = exp(z.re); cd r = {t * cos(z.im), t * sin(z.im)}; return r; }

#define QF_EPSILON 1.0e-12 /* orginal/gpe_engine.cu:129 */
#define MAXUP 32           /* orginal/gpe_user_defined.h:45 */

struct gpeo {
    int nx, ny, nz;
    size_t n;
    int interactions, vext;
    double gamma;
    double alpha, beta, dt, npart, t0;
    unsigned it;
    double qf;            /* stored as qf/(gamma N): orginal/gpe_engine.cu:642 */
    double up[MAXUP];     /* d_user_param */
    cd c;                 /* d_step_coeff = gamma dt/2 / (-beta + i alpha) */
    double *kx, *ky, *kz;
    cd *Kx, *Ky, *Kz;     /* Kz includes 1/N (d_exp_kkz2_over_nxyz) */
    cd *psi, *psi2;
    double *wrkR;         /* V0 of the predictor (d_wrk2R) */
    double *qfpot;        /* d_wrk3R */
    cd *wrk;              /* FFT scratch (d_wrk3C / d_wrk2 in the reference) */
    double *varr;         /* VEXT 3 array */
    double vx0, vy0;
    int vQ;
    double *phase;        /* enforced phase (d_phase) */
};

/* ------------------------------------------------------------------ hooks */

#define USER_CODE 100 /* INTERACTIONS / VEXT of the user hook file */

/* the user hook file's view of this engine */
static inline gpeo_uh_ctx uh_ctx(const gpeo_t *e)
{
    gpeo_uh_ctx c = {e->up, e->nx, e->ny, e->nz, e->dt, e->t0, e->npart};
    return c;
}

/* gpe_external_potential at time index it: VEXT 0 harmonic
 * (orginal/gpe_user_defined.h:79-101, dev/gpe_user_defined.h:488-513), VEXT 1
 * Woods-Saxon tube (dev/gpe_user_defined.h:533-554), VEXT 2 cosine lattice and
 * VEXT 3 array (additions of this engine), VEXT 100 the user hook file. */
static inline double vext_val(const gpeo_t *e, unsigned ix, unsigned iy, unsigned iz, size_t ixyz, unsigned it)
{
    const double *p = e->up;
    switch (e->vext) {
    case USER_CODE: {
        gpeo_uh_ctx c = uh_ctx(e);
        return gpeo_uh_vext(&c, ix, iy, iz, it);
    }
    case 0: {
        double x = (double)ix - 1.0 * (e->nx / 2);
        double y = (double)iy - 1.0 * (e->ny / 2);
        double z = (double)iz - 1.0 * (e->nz / 2);
        return 0.5 * x * x * p[0] * p[0] + 0.5 * y * y * p[1] * p[1] + 0.5 * z * z * p[2] * p[2];
    }
    case 1: {
        double x = (double)ix - 1.0 * (e->nx / 2);
        double y = (double)iy - 1.0 * (e->ny / 2);
        double r = sqrt(x * x + y * y);
        return p[0] * (1. - 1. / (1 + exp((r - p[2]) / p[1])));
    }
    case 2: {
        const double tp = 2.0 * M_PI * p[1];
        return p[0] * (cos(tp * (double)ix / (double)e->nx) + cos(tp * (double)iy / (double)e->ny) +
                       cos(tp * (double)iz / (double)e->nz));
    }
    case 3:
        return e->varr ? e->varr[ixyz] : 0.0;
    default:
        return 0.0;
    }
}

#define XI_ 0.37
#define CONTACT_ 0.901

/* gpe_dEDFdn at time index it: INTERACTIONS 0 unitary Fermi gas
 * (orginal/gpe_user_defined.h:122-126, dev/gpe_user_defined.h:131-135), 1 ETF
 * BEC side (dev 214-243), -1 contact bosons (dev 317-322), -2 dipolar local
 * part (dev 400-404), 100 the user hook file. */
static inline double dedfdn(const gpeo_t *e, double rho, unsigned it)
{
    switch (e->interactions) {
    case USER_CODE: {
        gpeo_uh_ctx c = uh_ctx(e);
        return gpeo_uh_dedfdn(&c, rho, it);
    }
    case 0:
        return 0.37 * pow(3.0 * M_PI * M_PI * rho, 2.0 / 3.0) / 2.0;
    case 1: {
        const double a = e->up[3];
        if (rho > 1e-15) {
            const double kF = pow(3.0 * M_PI * M_PI * rho, 1.0 / 3.0);
            const double eF = 0.5 * kF * kF;
            const double x = (kF > DBL_MIN) ? 1.0 / (a * kF) : DBL_MAX;
            const double D = (XI_ + x * (1.0 + CONTACT_) + 3.0 * M_PI * XI_ * x * x);
            return XI_ * eF * (XI_ + 0.8 * x) / D +
                   0.2 * XI_ * eF * (XI_ + x) * x * ((1.0 + CONTACT_) + 6.0 * M_PI * XI_ * x) / (D * D) -
                   1.0 / (2.0 * a * a);
        }
        return -1.0 / (2.0 * a * a);
    }
    case -1:
        return 4 * M_PI * e->up[3] * rho;
    case -2:
        return 0.;
    default:
        return 0.;
    }
}

/* gpe_EDF: orginal/gpe_user_defined.h:109-114, dev/gpe_user_defined.h:118-123,
 * 196-206, 305-309, 387-392 */
static inline double edf(const gpeo_t *e, double rho, unsigned it)
{
    switch (e->interactions) {
    case USER_CODE: {
        gpeo_uh_ctx c = uh_ctx(e);
        return gpeo_uh_edf(&c, rho, it);
    }
    case 0:
        return 0.37 * 0.6 * rho * pow(3.0 * M_PI * M_PI * rho, 2.0 / 3.0) / 2.;
    case 1: {
        const double a = e->up[3];
        const double kF = pow(3.0 * M_PI * M_PI * rho, 1.0 / 3.0);
        const double eF = 0.5 * kF * kF;
        const double x = (kF > DBL_MIN) ? 1.0 / (a * kF) : DBL_MAX;
        return 0.6 * eF * rho * XI_ * (XI_ + x) / (XI_ + x * (1.0 + CONTACT_) + 3.0 * M_PI * XI_ * x * x) -
               rho / (2.0 * a * a);
    }
    case -1:
        return 2 * M_PI * e->up[3] * rho * rho;
    case -2:
        return 2 * M_PI * e->up[3] * rho * rho;
    default:
        return 0.;
    }
}

/* gpe_density: orginal/gpe_engine.cu:384-387 */
static inline double density(const gpeo_t *e, cd p) { return e->gamma * (p.re * p.re + p.im * p.im); }

/* gpe_modify_psi at time index it: the identity of orginal/gpe_user_defined.h:66-69,
 * or the user hook file's when the physics comes from it */
static inline cd modify_psi(const gpeo_t *e, unsigned ix, unsigned iy, unsigned iz, unsigned it, cd p)
{
    if (e->interactions != USER_CODE && e->vext != USER_CODE) return p;
    gpeo_uh_ctx c = uh_ctx(e);
    gpeo_uh_modify(&c, ix, iy, iz, it, &p.re, &p.im);
    return p;
}

/* ---------------------------------------------------------- utilities */

#define IDX3(ix, iy, iz) (((size_t)(ix) * e->ny + (iy)) * e->nz + (iz))

/* sum of f over the lattice in a fixed order: per-x-plane partial sums,
 * then the planes in order (independent of the thread count) */
static double *plane_buf(const gpeo_t *e) { return (double *)malloc(sizeof(double) * (size_t)e->nx); }

static double ordered_sum_planes(const double *planes, int nx)
{
    double s = 0.0;
    for (int i = 0; i < nx; i++) s += planes[i];
    return s;
}

static double sum_density(const gpeo_t *e, const cd *p)
{
    double *pl = plane_buf(e);
    const size_t pn = (size_t)e->ny * e->nz;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++) {
        double s = 0.0;
        const cd *q = p + (size_t)ix * pn;
        for (size_t j = 0; j < pn; j++) s += density(e, q[j]);
        pl[ix] = s;
    }
    double r = ordered_sum_planes(pl, e->nx);
    free(pl);
    return r;
}

/* gpe_normalize: orginal/gpe_engine.cu:684-692 (density, tree sum, scale by
 * sqrt(npart/sum)); returns the sum before scaling */
static double normalize(gpeo_t *e, cd *p)
{
    double sum = sum_density(e, p);
    double f = sqrt(e->npart / sum);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) p[i] = cscale(p[i], f);
    return sum;
}

static void fft3(gpeo_t *e, cd *a, int sign) { gpeo_fft3d(e->nx, e->ny, e->nz, (double *)a, sign); }

/* reciprocal lattice and kinetic factors: orginal/gpe_engine.cu:446-517 and
 * gpe_change_alpha_beta 572-610; dev/gpe_engine.cu:126-281 */
static void fill_k(double *k, int n)
{
    int i, j;
    for (i = 0; i <= n / 2 - 1; i++) k[i] = 2. * M_PI / n * (double)i;
    j = -i;
    for (i = n / 2; i < n; i++) { k[i] = 2. * M_PI / n * (double)j; j++; }
}

/* complex exp of a complex argument, as std::exp(std::complex<double>) */
static cd cexp_host(double re, double im) { double t = exp(re); cd r = {t * cos(im), t * sin(im)}; return r; }

static void build_kinetic(gpeo_t *e)
{
    /* c3 = (gamma*0.5*dt) / (-beta + i alpha)  (std::complex division) */
    double a = e->gamma * 0.5 * e->dt;
    double br = -1.0 * e->beta, bi = e->alpha;
    double den = br * br + bi * bi;
    e->c.re = (a * br) / den;
    e->c.im = (-a * bi) / den;
    const double g2 = e->gamma * e->gamma;
    for (int i = 0; i < e->nx; i++) {
        double k2 = e->kx[i] * e->kx[i];
        e->Kx[i] = cexp_host(e->c.re * k2 / g2, e->c.im * k2 / g2);
    }
    for (int i = 0; i < e->ny; i++) {
        double k2 = e->ky[i] * e->ky[i];
        e->Ky[i] = cexp_host(e->c.re * k2 / g2, e->c.im * k2 / g2);
    }
    for (int i = 0; i < e->nz; i++) {
        double k2 = e->kz[i] * e->kz[i];
        cd t = cexp_host(e->c.re * k2 / g2, e->c.im * k2 / g2);
        e->Kz[i] = cscale(t, 1.0 / (double)e->n);
    }
}

/* ------------------------------------------------------------- lifecycle */

gpeo_t *gpeo_create(int nx, int ny, int nz, int interactions, int vext, int gpe_for, double alpha,
                    double beta, double dt, double npart)
{
    if (nx < 2 || ny < 2 || nz < 2) return NULL;
    gpeo_t *e = (gpeo_t *)calloc(1, sizeof(gpeo_t));
    e->nx = nx; e->ny = ny; e->nz = nz;
    e->n = (size_t)nx * ny * nz;
    e->interactions = interactions;
    e->vext = vext;
    if (gpe_for == 1) e->gamma = 1.0;
    else if (gpe_for == 2) e->gamma = 2.0;
    else e->gamma = (interactions == 0 || interactions == 1) ? 2.0 : 1.0;
    e->alpha = alpha; e->beta = beta; e->dt = dt; e->npart = npart;
    e->t0 = 0.0; e->it = 0; e->qf = 0.0;
    e->kx = (double *)malloc(sizeof(double) * nx);
    e->ky = (double *)malloc(sizeof(double) * ny);
    e->kz = (double *)malloc(sizeof(double) * nz);
    fill_k(e->kx, nx); fill_k(e->ky, ny); fill_k(e->kz, nz);
    e->Kx = (cd *)malloc(sizeof(cd) * nx);
    e->Ky = (cd *)malloc(sizeof(cd) * ny);
    e->Kz = (cd *)malloc(sizeof(cd) * nz);
    build_kinetic(e);
    e->psi = (cd *)calloc(e->n, sizeof(cd));
    e->psi2 = (cd *)calloc(e->n, sizeof(cd));
    e->wrkR = (double *)calloc(e->n, sizeof(double));
    e->wrk = (cd *)calloc(e->n, sizeof(cd));
    e->qfpot = NULL;
    e->varr = NULL;
    e->vQ = 1;
    return e;
}

void gpeo_destroy(gpeo_t *e)
{
    if (!e) return;
    free(e->kx); free(e->ky); free(e->kz);
    free(e->Kx); free(e->Ky); free(e->Kz);
    free(e->psi); free(e->psi2); free(e->wrkR); free(e->wrk);
    free(e->qfpot); free(e->varr); free(e->phase);
    free(e);
}

/* orginal/gpe_engine.cu:560-613 */
int gpeo_change_alpha_beta(gpeo_t *e, double alpha, double beta)
{
    e->alpha = alpha; e->beta = beta;
    build_kinetic(e);
    return 0;
}

/* orginal/gpe_engine.cu:615-624 */
int gpeo_set_time(gpeo_t *e, double t0) { e->t0 = t0; e->it = 0; return 0; }

/* orginal/gpe_engine.cu:626-635; copies `size` values and zeroes the rest
 * (the reference copies MAX_USER_PARAMS regardless of size) */
int gpeo_set_user_params(gpeo_t *e, int size, const double *params)
{
    if (size > MAXUP) return -9;
    if (size < 0) return -1;
    memset(e->up, 0, sizeof(e->up));
    for (int i = 0; i < size; i++) e->up[i] = params[i];
    return 0;
}

/* orginal/gpe_engine.cu:637-662 */
int gpeo_set_quantum_friction_coeff(gpeo_t *e, double qf)
{
    if (qf != 0.0) {
        e->qf = qf / (e->gamma * (double)e->n);
        if (!e->qfpot) e->qfpot = (double *)calloc(e->n, sizeof(double));
    } else {
        e->qf = 0.0;
        free(e->qfpot);
        e->qfpot = NULL;
    }
    return 0;
}

/* dev/gpe_engine.cu:465-476 */
int gpeo_set_vortex(gpeo_t *e, double x0, double y0, int Q) { e->vx0 = x0; e->vy0 = y0; e->vQ = Q; return 0; }

int gpeo_set_vext_array(gpeo_t *e, const double *v)
{
    if (!e->varr) e->varr = (double *)malloc(sizeof(double) * e->n);
    memcpy(e->varr, v, sizeof(double) * e->n);
    return 0;
}

/* orginal/gpe_engine.cu:699-717 */
int gpeo_set_psi(gpeo_t *e, double t, const double *psi)
{
    e->t0 = t; e->it = 0;
    memcpy(e->psi, psi, sizeof(cd) * e->n);
    return 0;
}

int gpeo_get_psi(gpeo_t *e, double *t, double *psi)
{
    memcpy(psi, e->psi, sizeof(cd) * e->n);
    *t = e->t0 + e->dt * e->it;
    return 0;
}

int gpeo_normalize_psi(gpeo_t *e) { normalize(e, e->psi); return 0; }

/* dev/gpe_engine.cu:679-689 */
int gpeo_get_norm(gpeo_t *e, double *norm) { *norm = sum_density(e, e->psi); return 0; }

/* ----------------------------------------------------------- step parts */

/* __gpe_multiply_by_expT__: orginal/gpe_engine.cu:880-895 */
static void multiply_by_expT(gpeo_t *e, cd *a)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd w = a[i];
                w = cmul(w, e->Kx[ix]);
                w = cmul(w, e->Ky[iy]);
                w = cmul(w, e->Kz[iz]);
                a[i] = w;
            }
}

/* gpe_compute_qf_potential: orginal/gpe_engine.cu:897-937 */
static void compute_qf_potential(gpeo_t *e, const cd *psi, cd *wrk, double *qfpot)
{
    memcpy(wrk, psi, sizeof(cd) * e->n);
    fft3(e, wrk, -1);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double f = e->qf * (e->kx[ix] * e->kx[ix] + e->ky[iy] * e->ky[iy] + e->kz[iz] * e->kz[iz]);
                wrk[i] = cscale(wrk[i], f);
            }
    fft3(e, wrk, +1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) {
        cd l = psi[i];
        double r = density(e, l);
        /* cplxMulI(cplxConj(l), wrk) = l.re*w.im + (-l.im)*w.re */
        qfpot[i] = (l.re * wrk[i].im + (-l.im) * wrk[i].re) / (r + QF_EPSILON);
    }
}

/* __gpe_exp_Vstep1_ / _qf_: orginal/gpe_engine.cu:723-772 */
static void vstep1(gpeo_t *e, unsigned it, const cd *psi_in, cd *psi_out, double *wrkR, const double *qf)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd l = modify_psi(e, ix, iy, iz, it, psi_in[i]);
                double r = density(e, l);
                double lv = vext_val(e, ix, iy, iz, i, it) + dedfdn(e, r, it);
                if (qf) lv = lv + qf[i];
                wrkR[i] = lv;
                cd ex = cexp_(cscale(e->c, lv));
                psi_out[i] = cmul(l, ex);
            }
}

/* __gpe_exp_Vstep2_: orginal/gpe_engine.cu:774-801 (wrkC aliases psi_in) */
static void vstep2(gpeo_t *e, unsigned it, cd *psi_in, cd *psi_out, const double *wrkR, cd *wrkC)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd l = psi_in[i];
                double lv = wrkR[i];
                cd ex = cexp_(cscale(e->c, lv));
                l = cmul(l, ex);
                double r = density(e, l);
                lv = 0.5 * (lv + vext_val(e, ix, iy, iz, i, it + 1) + dedfdn(e, r, it + 1));
                ex = cexp_(cscale(e->c, lv));
                wrkC[i] = ex;
                l = modify_psi(e, ix, iy, iz, it, psi_out[i]);
                psi_out[i] = cmul(l, ex);
            }
}

/* __gpe_exp_Vstep2_part1_: orginal/gpe_engine.cu:803-815 */
static void vstep2_part1(gpeo_t *e, cd *psi, const double *wrkR)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) psi[i] = cmul(psi[i], cexp_(cscale(e->c, wrkR[i])));
}

/* __gpe_exp_Vstep2_part2_ / _qf_: orginal/gpe_engine.cu:817-868 */
static void vstep2_part2(gpeo_t *e, unsigned it, cd *psi_in, cd *psi_out, const double *wrkR, cd *wrkC,
                         const double *qf)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd l = psi_in[i];
                double lv = wrkR[i];
                double r = density(e, l);
                if (qf) lv = 0.5 * (lv + vext_val(e, ix, iy, iz, i, it + 1) + dedfdn(e, r, it + 1) + qf[i]);
                else lv = 0.5 * (lv + vext_val(e, ix, iy, iz, i, it + 1) + dedfdn(e, r, it + 1));
                cd ex = cexp_(cscale(e->c, lv));
                wrkC[i] = ex;
                l = modify_psi(e, ix, iy, iz, it, psi_out[i]);
                psi_out[i] = cmul(l, ex);
            }
}

/* __gpe_exp_Vstep3_: orginal/gpe_engine.cu:870-878 */
static void vstep3(gpeo_t *e, cd *psi, const cd *W)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) psi[i] = cmul(psi[i], W[i]);
}

/* __gpe_imprint_vortexline_zdir_: dev/gpe_engine.cu:485-511.  Phase
 * Q*atan2(x, y) (arguments in this order), modulus kept. */
static void imprint_vortex(gpeo_t *e, cd *psi)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++) {
            double x = (double)ix - 1.0 * (e->nx / 2) - e->vx0;
            double y = (double)iy - 1.0 * (e->ny / 2) - e->vy0;
            double ph = atan2(x, y) * (double)e->vQ;
            double cp = cos(ph), sp = sin(ph);
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double a = hypot(psi[i].re, psi[i].im);
                psi[i].re = a * cp;
                psi[i].im = a * sp;
            }
        }
}

/* __gpe_imprint2_vortexline_zdir_: dev/gpe_engine.cu:520-548, the imprint
 * only where |psi| > 1e-15 */
static void imprint_vortex2(gpeo_t *e, cd *psi)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++) {
            double x = (double)ix - 1.0 * (e->nx / 2) - e->vx0;
            double y = (double)iy - 1.0 * (e->ny / 2) - e->vy0;
            double ph = atan2(x, y) * (double)e->vQ;
            double cp = cos(ph), sp = sin(ph);
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double a = hypot(psi[i].re, psi[i].im);
                if (a > 1e-15) {
                    psi[i].re = a * cp;
                    psi[i].im = a * sp;
                }
            }
        }
}

/* __gpe_enforce_phase__: dev/gpe_engine.cu:584-591, psi = |psi| exp(i phase) */
static void enforce_phase(gpeo_t *e, cd *psi)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) {
        double a = sqrt(psi[i].re * psi[i].re + psi[i].im * psi[i].im);
        psi[i].re = a * cos(e->phase[i]);
        psi[i].im = a * sin(e->phase[i]);
    }
}

enum { PIN_NONE = 0, PIN_VORTEX = 1, PIN_VORTEX2 = 2, PIN_VORTEX3 = 3, PIN_PHASE = 4 };

static void pin_start(gpeo_t *e, int pin)
{
    if (pin == PIN_VORTEX || pin == PIN_VORTEX3) imprint_vortex(e, e->psi);
    else if (pin == PIN_VORTEX2) imprint_vortex2(e, e->psi);
    else if (pin == PIN_PHASE) enforce_phase(e, e->psi);
}

/* re-applied after the predictor normalisation in imaginary time, except by
 * gpe_evolve_vortex3 (dev/gpe_engine.cu:1160-1162, 1252-1254, 1344-1346, 1436-1438) */
static void pin_mid(gpeo_t *e, int pin)
{
    if (pin == PIN_VORTEX) imprint_vortex(e, e->psi);
    else if (pin == PIN_VORTEX2) imprint_vortex2(e, e->psi);
    else if (pin == PIN_PHASE) enforce_phase(e, e->psi);
}

/* One split step shared by gpe_evolve_qf (orginal/gpe_engine.cu:948-1013,
 * dev/gpe_engine.cu:1035-1106), gpe_evolve_vortex / 2 / 3 (dev/gpe_engine.cu:
 * 1120-1199, 1207-1294, 1299-1386) and gpe_evolve_enforced_phase (1390-1477). */
static void one_step(gpeo_t *e, int pin, double *mu)
{
    pin_start(e, pin);

    if (e->qf != 0.0) {
        compute_qf_potential(e, e->psi, e->wrk, e->qfpot);
        vstep1(e, e->it, e->psi, e->psi2, e->wrkR, e->qfpot);
    } else {
        vstep1(e, e->it, e->psi, e->psi2, e->wrkR, NULL);
    }

    fft3(e, e->psi2, -1);
    multiply_by_expT(e, e->psi2);
    fft3(e, e->psi2, +1);

    if (e->beta == 0.0 && e->qf == 0.0) {
        vstep2(e, e->it, e->psi2, e->psi, e->wrkR, e->psi2);
    } else {
        vstep2_part1(e, e->psi2, e->wrkR);
        if (e->beta != 0.0) {
            normalize(e, e->psi2);
            pin_mid(e, pin);
        }
        if (e->qf != 0.0) {
            compute_qf_potential(e, e->psi, e->wrk, e->qfpot);
            vstep2_part2(e, e->it, e->psi2, e->psi, e->wrkR, e->psi2, e->qfpot);
        } else {
            vstep2_part2(e, e->it, e->psi2, e->psi, e->wrkR, e->psi2, NULL);
        }
    }

    fft3(e, e->psi, -1);
    multiply_by_expT(e, e->psi);
    fft3(e, e->psi, +1);

    vstep3(e, e->psi, e->psi2);

    if (e->beta != 0.0) {
        double norm = normalize(e, e->psi);
        if (mu) *mu = -.5 * log(norm / e->npart) / e->dt;  /* dev/gpe_engine.cu:1097-1102 */
    }
    e->it = e->it + 1;
}

int gpeo_evolve_qf(gpeo_t *e, int nt, double *mu)
{
    for (int i = 0; i < nt; i++) one_step(e, PIN_NONE, mu);
    return 0;
}

int gpeo_evolve_vortex(gpeo_t *e, int nt, double *mu)
{
    for (int i = 0; i < nt; i++) one_step(e, PIN_VORTEX, mu);
    return 0;
}

int gpeo_evolve_vortex2(gpeo_t *e, int nt, double *mu)
{
    for (int i = 0; i < nt; i++) one_step(e, PIN_VORTEX2, mu);
    return 0;
}

int gpeo_evolve_vortex3(gpeo_t *e, int nt, double *mu)
{
    for (int i = 0; i < nt; i++) one_step(e, PIN_VORTEX3, mu);
    return 0;
}

int gpeo_evolve_enforced_phase(gpeo_t *e, int nt, double *mu)
{
    if (!e->phase) return -15;
    for (int i = 0; i < nt; i++) one_step(e, PIN_PHASE, mu);
    return 0;
}

/* dev/gpe_engine.cu:561-570 */
int gpeo_set_phase(gpeo_t *e, const double *phase)
{
    if (!e->phase) e->phase = (double *)malloc(sizeof(double) * e->n);
    memcpy(e->phase, phase, sizeof(double) * e->n);
    return 0;
}

/* dev/gpe_engine.cu:552-559, 572-582: phase = arg psi, kept as the enforced phase */
int gpeo_get_phase(gpeo_t *e, double *phase)
{
    if (!e->phase) e->phase = (double *)malloc(sizeof(double) * e->n);
    for (size_t i = 0; i < e->n; i++) e->phase[i] = atan2(e->psi[i].im, e->psi[i].re);
    if (phase) memcpy(phase, e->phase, sizeof(double) * e->n);
    return 0;
}

/* gpe_energy: orginal/gpe_engine.cu:1076-1113 */
int gpeo_energy(gpeo_t *e, double *t, double *ekin, double *eint, double *eext)
{
    *t = e->t0 + e->dt * e->it;
    if (e->beta == 0.0) normalize(e, e->psi);

    double *pl_ext = plane_buf(e), *pl_int = plane_buf(e), *pl_kin = plane_buf(e);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++) {
        double se = 0.0, si = 0.0;
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double r = density(e, e->psi[i]);
                se += r * vext_val(e, ix, iy, iz, i, e->it);
                si += edf(e, r, e->it);
            }
        pl_ext[ix] = se;
        pl_int[ix] = si;
    }
    *eext = ordered_sum_planes(pl_ext, e->nx);
    *eint = ordered_sum_planes(pl_int, e->nx);

    memcpy(e->psi2, e->psi, sizeof(cd) * e->n);
    fft3(e, e->psi2, -1);
    const double norm = 2.0 * e->gamma * (double)e->n;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double f = (e->kx[ix] * e->kx[ix] + e->ky[iy] * e->ky[iy] + e->kz[iz] * e->kz[iz]) / norm;
                e->psi2[i] = cscale(e->psi2[i], f);
            }
    fft3(e, e->psi2, +1);
    const size_t pn = (size_t)e->ny * e->nz;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++) {
        double s = 0.0;
        for (size_t j = 0; j < pn; j++) {
            size_t i = (size_t)ix * pn + j;
            s += e->psi[i].re * e->psi2[i].re + e->psi[i].im * e->psi2[i].im;
        }
        pl_kin[ix] = s;
    }
    *ekin = ordered_sum_planes(pl_kin, e->nx);
    free(pl_ext); free(pl_int); free(pl_kin);
    return 0;
}

/* gpe_get_density: orginal/gpe_engine.cu:1115-1123 */
int gpeo_get_density(gpeo_t *e, double *t, double *rho)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) rho[i] = density(e, e->psi[i]);
    *t = e->t0 + e->dt * e->it;
    return 0;
}

/* gpe_get_currents: orginal/gpe_engine.cu:1125-1212 */
int gpeo_get_currents(gpeo_t *e, double *t, double *jx, double *jy, double *jz)
{
    *t = e->t0 + e->dt * e->it;
    cd *F = e->psi2;
    memcpy(F, e->psi, sizeof(cd) * e->n);
    fft3(e, F, -1);
    double *outs[3] = {jx, jy, jz};
    for (int a = 0; a < 3; a++) {
#pragma omp parallel for schedule(static)
        for (int ix = 0; ix < e->nx; ix++)
            for (int iy = 0; iy < e->ny; iy++)
                for (int iz = 0; iz < e->nz; iz++) {
                    size_t i = IDX3(ix, iy, iz);
                    double k = (a == 0) ? e->kx[ix] : (a == 1) ? e->ky[iy] : e->kz[iz];
                    e->wrk[i] = cscale(F[i], k / (e->gamma * (double)e->n));
                }
        fft3(e, e->wrk, +1);
        double *o = outs[a];
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < e->n; i++) o[i] = e->psi[i].re * e->wrk[i].re + e->psi[i].im * e->wrk[i].im;
    }
    return 0;
}

/* ------------------------------------------------------- introspection */

double gpeo_gamma(const gpeo_t *e) { return e->gamma; }

int gpeo_tables(const gpeo_t *e, double *kx, double *ky, double *kz, double *Kx, double *Ky, double *Kz,
                double *step_coeff)
{
    if (kx) memcpy(kx, e->kx, sizeof(double) * e->nx);
    if (ky) memcpy(ky, e->ky, sizeof(double) * e->ny);
    if (kz) memcpy(kz, e->kz, sizeof(double) * e->nz);
    if (Kx) memcpy(Kx, e->Kx, sizeof(cd) * e->nx);
    if (Ky) memcpy(Ky, e->Ky, sizeof(cd) * e->ny);
    if (Kz) memcpy(Kz, e->Kz, sizeof(cd) * e->nz);
    if (step_coeff) { step_coeff[0] = e->c.re; step_coeff[1] = e->c.im; }
    return 0;
}

int gpeo_vext_value(const gpeo_t *e, int ix, int iy, int iz, double *v)
{
    *v = vext_val(e, (unsigned)ix, (unsigned)iy, (unsigned)iz, IDX3(ix, iy, iz), e->it);
    return 0;
}

int gpeo_dedfdn_value(const gpeo_t *e, double rho, double *dv, double *edfv)
{
    if (dv) *dv = dedfdn(e, rho, e->it);
    if (edfv) *edfv = edf(e, rho, e->it);
    return 0;
}

/* ------------------------------------------------- plaquette windings */

/* quadrant of a non-zero complex number: 0 for arg in [0, pi/2), 1 for
 * [pi/2, pi), 2 for [pi, 3pi/2), 3 for [3pi/2, 2pi) */
static inline int quadrant(double x, double y)
{
    if (y >= 0.0 && x > 0.0) return 0;
    if (x <= 0.0 && y > 0.0) return 1;
    if (y <= 0.0 && x < 0.0) return 2;
    return 3;
}

/* signed quarter turns of the principal phase change from a to b */
static inline int quarter_turns(double ax, double ay, double bx, double by)
{
    int d = (quadrant(bx, by) - quadrant(ax, ay)) & 3;
    if (d == 1) return 1;
    if (d == 3) return -1;
    if (d == 0) return 0;
    double cr = ax * by - ay * bx; /* sign of sin of the phase change */
    return (cr < 0.0) ? -2 : 2;
}

long gpeo_plaquette_windings(int nx, int ny, int nz, const double *psi64, const float *psi32, double rho_min,
                             int8_t *w)
{
    const size_t n = (size_t)nx * ny * nz;
    long count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int ix = 0; ix < nx; ix++) {
        for (int iy = 0; iy < ny; iy++)
            for (int iz = 0; iz < nz; iz++) {
                const int jx = (ix + 1) % nx, jy = (iy + 1) % ny, jz = (iz + 1) % nz;
                size_t cidx[3][4] = {
                    {((size_t)ix * ny + iy) * nz + iz, ((size_t)jx * ny + iy) * nz + iz,
                     ((size_t)jx * ny + jy) * nz + iz, ((size_t)ix * ny + jy) * nz + iz},
                    {((size_t)ix * ny + iy) * nz + iz, ((size_t)ix * ny + jy) * nz + iz,
                     ((size_t)ix * ny + jy) * nz + jz, ((size_t)ix * ny + iy) * nz + jz},
                    {((size_t)ix * ny + iy) * nz + iz, ((size_t)ix * ny + iy) * nz + jz,
                     ((size_t)jx * ny + iy) * nz + jz, ((size_t)jx * ny + iy) * nz + iz}};
                const size_t here = ((size_t)ix * ny + iy) * nz + iz;
                for (int o = 0; o < 3; o++) {
                    double re[4], im[4];
                    int masked = 0;
                    for (int c = 0; c < 4; c++) {
                        size_t q = cidx[o][c];
                        re[c] = psi64 ? psi64[2 * q] : (double)psi32[2 * q];
                        im[c] = psi64 ? psi64[2 * q + 1] : (double)psi32[2 * q + 1];
                        double r2 = re[c] * re[c];
                        r2 = r2 + im[c] * im[c];
                        if (!(r2 >= rho_min)) masked = 1;
                    }
                    int s = 0;
                    if (!masked)
                        for (int c = 0; c < 4; c++) {
                            int d = (c + 1) & 3;
                            s += quarter_turns(re[c], im[c], re[d], im[d]);
                        }
                    int wind = s / 4;
                    w[(size_t)o * n + here] = (int8_t)wind;
                    if (wind != 0) count++;
                }
            }
    }
    return count;
}

/* ------------------------------------------------------------------ adjoint
 *
 * Gradient of the final-time overlap objective
 *     L = 1 - |<psi_t|psi(T)>|^2 / (|psi_t|^2 |psi(0)|^2)
 * with respect to V_ext at every lattice point, for T real-time split steps
 * of one_step above (no quantum friction, no phase pinning).  No counterpart
 * in the reference.  The discrete adjoint of one step, with the convention
 * dL = 2 Re sum conj(lambda) dpsi and c the step coefficient, is
 *
 *   forward:  V0 = Vext + g(rho(psi0)), E0 = exp(c V0), phi1 = psi0 E0,
 *             phi3 = K(phi1) E0, V1 = (V0 + Vext + g(rho(phi3)))/2,
 *             W = exp(c V1), chi = psi0 W, psi1 = K(chi) W
 *   backward: lambda_xi = conj(W) lambda1,  Vb1 = 2 Re(conj(lambda1) psi1 c)
 *             lambda_chi = K^+(lambda_xi),  Vb1 += 2 Re(conj(lambda_chi) chi c)
 *             lambda0 = conj(W) lambda_chi
 *             dL/dVext += Vb1/2,  Vb0 = Vb1/2,
 *             lambda_phi3 = (Vb1/2) g'(rho3) gamma phi3,  Vb0 += 2 Re(conj(lambda_phi3) phi3 c)
 *             lambda_phi1 = K^+(conj(E0) lambda_phi3), Vb0 += 2 Re(conj(lambda_phi1) phi1 c)
 *             lambda0 += conj(E0) lambda_phi1 + gamma g'(rho0) Vb0 psi0,  dL/dVext += Vb0
 * where K = F^-1 diag(Kx Ky Kz) F is the kinetic step and K^+ the same with
 * conj(Kx Ky Kz) (the unnormalised F and F^-1 are each other's adjoints), and
 * g = dEDF/drho with g' its derivative.  States are kept every `ckpt` steps
 * and each segment is recomputed once on the way back.
 */

static double dgdrho(const gpeo_t *e, double rho)
{
    switch (e->interactions) {
    case 0: /* g = 0.37 (3 pi^2 rho)^(2/3) / 2  ->  g' = (2/3) g / rho */
        return rho > 0.0 ? (2.0 / 3.0) * dedfdn(e, rho, 0) / rho : 0.0;
    case -1:
        return 4 * M_PI * e->up[3];
    case -2:
        return 0.0;
    default:
        return NAN;
    }
}

static void kinetic_step(gpeo_t *e, cd *a, int adjoint)
{
    fft3(e, a, -1);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd w = a[i];
                cd kx = e->Kx[ix], ky = e->Ky[iy], kz = e->Kz[iz];
                if (adjoint) { kx.im = -kx.im; ky.im = -ky.im; kz.im = -kz.im; }
                w = cmul(w, kx);
                w = cmul(w, ky);
                w = cmul(w, kz);
                a[i] = w;
            }
    fft3(e, a, +1);
}

static inline cd cconj(cd a) { cd r = {a.re, -a.im}; return r; }
/* 2 Re(conj(l) y c) */
static inline double sens(cd l, cd y, cd c) { return 2.0 * cmul(cmul(cconj(l), y), c).re; }

static void backward_step(gpeo_t *e, const cd *psi0, const cd *psi1, cd *lam, double *grad, cd *phi, cd *lam0,
                          double *vb)
{
    const cd c = e->c;
    const double G = e->gamma;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double V0 = vext_val(e, ix, iy, iz, i, 0) + dedfdn(e, density(e, psi0[i]), 0);
                phi[i] = cmul(psi0[i], cexp_(cscale(c, V0)));
            }
    kinetic_step(e, phi, 0); /* phi2 */
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double vx = vext_val(e, ix, iy, iz, i, 0);
                double V0 = vx + dedfdn(e, density(e, psi0[i]), 0);
                cd phi3 = cmul(phi[i], cexp_(cscale(c, V0)));
                double V1 = 0.5 * (V0 + vx + dedfdn(e, density(e, phi3), 0));
                cd W = cexp_(cscale(c, V1));
                vb[i] = sens(lam[i], psi1[i], c);
                lam[i] = cmul(cconj(W), lam[i]);
                phi[i] = phi3;
            }
    kinetic_step(e, lam, 1); /* lambda_chi */
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double vx = vext_val(e, ix, iy, iz, i, 0);
                double V0 = vx + dedfdn(e, density(e, psi0[i]), 0);
                cd E0 = cexp_(cscale(c, V0));
                cd phi3 = phi[i];
                double r3 = density(e, phi3);
                double V1 = 0.5 * (V0 + vx + dedfdn(e, r3, 0));
                cd W = cexp_(cscale(c, V1));
                cd chi = cmul(psi0[i], W);
                lam0[i] = cmul(cconj(W), lam[i]);
                double V1b = vb[i] + sens(lam[i], chi, c);
                grad[i] += 0.5 * V1b;
                double V0b = 0.5 * V1b;
                cd lp3 = cscale(phi3, 0.5 * V1b * dgdrho(e, r3) * G);
                V0b += sens(lp3, phi3, c);
                lam[i] = cmul(cconj(E0), lp3);
                vb[i] = V0b;
            }
    kinetic_step(e, lam, 1); /* lambda_phi1 */
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                double r0 = density(e, psi0[i]);
                double V0 = vext_val(e, ix, iy, iz, i, 0) + dedfdn(e, r0, 0);
                cd E0 = cexp_(cscale(c, V0));
                cd phi1 = cmul(psi0[i], E0);
                cd l0 = lam0[i];
                cd t = cmul(cconj(E0), lam[i]);
                l0.re += t.re;
                l0.im += t.im;
                double V0b = vb[i] + sens(lam[i], phi1, c);
                grad[i] += V0b;
                double f = G * dgdrho(e, r0) * V0b;
                l0.re += f * psi0[i].re;
                l0.im += f * psi0[i].im;
                lam[i] = l0;
            }
}

int gpeo_adjoint_overlap(gpeo_t *e, int nt, const double *psi_target, int ckpt, double *loss, double *grad)
{
    if (nt < 1 || !psi_target || !grad) return -12;
    if (e->beta != 0.0 || e->qf != 0.0 || e->vext == USER_CODE || isnan(dgdrho(e, 1.0))) return -12;
    const size_t n = e->n;
    const int C = ckpt > 0 ? ckpt : (int)ceil(sqrt((double)nt));
    const int nck = (nt + C - 1) / C;
    const cd *pt = (const cd *)psi_target;
    cd *ck = (cd *)malloc(sizeof(cd) * n * (size_t)nck);
    cd *seg = (cd *)malloc(sizeof(cd) * n * (size_t)(C + 1));
    cd *lam = (cd *)malloc(sizeof(cd) * n);
    cd *phi = (cd *)malloc(sizeof(cd) * n);
    cd *lam0 = (cd *)malloc(sizeof(cd) * n);
    double *vb = (double *)malloc(sizeof(double) * n);
    double *pl = plane_buf(e);
    if (!ck || !seg || !lam || !phi || !lam0 || !vb || !pl) {
        free(ck); free(seg); free(lam); free(phi); free(lam0); free(vb); free(pl);
        return -12;
    }
    const unsigned it0 = e->it;
    for (int k = 0; k < nt; k++) {
        if (k % C == 0) memcpy(ck + (size_t)(k / C) * n, e->psi, sizeof(cd) * n);
        one_step(e, PIN_NONE, NULL);
    }
    /* overlap and norms, summed per x plane in a fixed order */
    double ore = 0, oim = 0, n0 = 0, ntg = 0;
    for (int ix = 0; ix < e->nx; ix++) {
        double a = 0, b = 0, c0 = 0, d = 0;
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd p = e->psi[i], t = pt[i], q = ck[i];
                a += t.re * p.re + t.im * p.im;
                b += t.re * p.im - t.im * p.re;
                c0 += q.re * q.re + q.im * q.im;
                d += t.re * t.re + t.im * t.im;
            }
        ore += a;
        oim += b;
        n0 += c0;
        ntg += d;
    }
    const double s = 1.0 / (n0 * ntg);
    *loss = 1.0 - (ore * ore + oim * oim) * s;
    /* lambda_T = -s O psi_t */
    for (size_t i = 0; i < n; i++) {
        cd t = pt[i];
        lam[i].re = -s * (ore * t.re - oim * t.im);
        lam[i].im = -s * (ore * t.im + oim * t.re);
        grad[i] = 0.0;
    }
    cd *final_psi = (cd *)malloc(sizeof(cd) * n);
    memcpy(final_psi, e->psi, sizeof(cd) * n);
    for (int sgi = nck - 1; sgi >= 0; sgi--) {
        const int k0 = sgi * C;
        const int k1 = (k0 + C < nt) ? k0 + C : nt;
        memcpy(e->psi, ck + (size_t)sgi * n, sizeof(cd) * n);
        memcpy(seg, e->psi, sizeof(cd) * n);
        for (int j = 1; j <= k1 - k0; j++) {
            one_step(e, PIN_NONE, NULL);
            memcpy(seg + (size_t)j * n, e->psi, sizeof(cd) * n);
        }
        for (int k = k1 - 1; k >= k0; k--)
            backward_step(e, seg + (size_t)(k - k0) * n, seg + (size_t)(k - k0 + 1) * n, lam, grad, phi, lam0, vb);
    }
    memcpy(e->psi, final_psi, sizeof(cd) * n);
    e->it = it0 + (unsigned)nt;
    free(final_psi); free(ck); free(seg); free(lam); free(phi); free(lam0); free(vb); free(pl);
    return 0;
}

/* ------------------------------------------------------------------ dipolar
 *
 * INTERACTIONS -2: bosons with contact and z-polarised dipole-dipole
 * interactions (dev/gpe_user_defined.h:356-438).  The mean field is the
 * convolution V_dip = Re F^-1[ Vpp(k) F[rho] ] / N with
 *   Vpp(k) = 4 pi a_con                                   for |k|^2 < DBL_EPSILON,
 *            4 pi (a_con + a_dip (3 kz^2 / |k|^2 - 1))   otherwise
 * (gpe_Vpp_k, dev/gpe_user_defined.h:422-438; a_con = up[3], a_dip = up[4]);
 * the local dEDF/dn is 0 and the step is the Strang-Euler one of
 * gpe_evolve_dipolar (dev/gpe_engine.cu:1490-1605).
 */
static double vpp_k(const gpeo_t *e, double kx, double ky, double kz)
{
    const double ac = e->up[3], ad = e->up[4];
    const double k2 = kx * kx + ky * ky + kz * kz;
    if (k2 < DBL_EPSILON) return 4. * M_PI * ac;
    return 4. * M_PI * (ac + ad * (3. * kz * kz / k2 - 1.));
}

/* wrk = (rho, 0) -> F -> x Vpp/N -> F^-1: __gpe_compute_density2C__ (dev 618-626),
 * __gpe_compute_vint_k__ (dev 960-975) between the two transforms */
static void dipolar_potential(gpeo_t *e, const cd *psi, cd *wrk)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) {
        wrk[i].re = density(e, psi[i]);
        wrk[i].im = 0.0;
    }
    fft3(e, wrk, -1);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                wrk[i] = cscale(wrk[i], vpp_k(e, e->kx[ix], e->ky[iy], e->kz[iz]) / (double)e->n);
            }
    fft3(e, wrk, +1);
}

/* __gpe_eulerstrang_exp_Vstep__: dev/gpe_engine.cu:984-1008 */
static void eulerstrang_vstep(gpeo_t *e, cd *psi, const cd *vdip)
{
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++)
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                cd l = modify_psi(e, ix, iy, iz, e->it, psi[i]);
                double r = density(e, l);
                double lv = vext_val(e, ix, iy, iz, i, e->it) + dedfdn(e, r, e->it) + vdip[i].re;
                psi[i] = cmul(l, cexp_(cscale(e->c, lv)));
            }
}

/* gpe_evolve_dipolar(int) / (int, double*): dev/gpe_engine.cu:1490-1605 */
int gpeo_evolve_dipolar(gpeo_t *e, int nt, double *mu)
{
    for (int s = 0; s < nt; s++) {
        dipolar_potential(e, e->psi, e->wrk);
        eulerstrang_vstep(e, e->psi, e->wrk);
        fft3(e, e->psi, -1);
        multiply_by_expT(e, e->psi);
        fft3(e, e->psi, +1);
        dipolar_potential(e, e->psi, e->wrk);
        eulerstrang_vstep(e, e->psi, e->wrk);
        if (e->beta != 0.0) {
            double norm = normalize(e, e->psi);
            if (mu) *mu = -.5 * log(norm / e->npart) / e->dt;
        }
        e->it = e->it + 1;
    }
    return 0;
}

/* gpe_energy_dipolar: dev/gpe_engine.cu:1716-1760.  sum_k Re(rho_k(-k) Vpp(k)
 * rho_k(k)) / N, with -k taken modulo the lattice (the reference indexes
 * n - i, which leaves the lattice for i = 0). */
int gpeo_energy_dipolar(gpeo_t *e, double *t, double *edip)
{
    *t = e->t0 + e->dt * e->it;
    cd *rk = e->wrk;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < e->n; i++) {
        rk[i].re = density(e, e->psi[i]);
        rk[i].im = 0.0;
    }
    fft3(e, rk, -1);
    double *pl = plane_buf(e);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++) {
        double s = 0.0;
        const int rx = (e->nx - ix) % e->nx;
        for (int iy = 0; iy < e->ny; iy++) {
            const int ry = (e->ny - iy) % e->ny;
            for (int iz = 0; iz < e->nz; iz++) {
                const int rz = (e->nz - iz) % e->nz;
                size_t i = IDX3(ix, iy, iz), j = IDX3(rx, ry, rz);
                cd en = cscale(rk[i], vpp_k(e, e->kx[ix], e->ky[iy], e->kz[iz]));
                s += rk[j].re * en.re - rk[j].im * en.im;
            }
        }
        pl[ix] = s;
    }
    *edip = ordered_sum_planes(pl, e->nx) / (double)e->n;
    free(pl);
    return 0;
}

/* gpe_energy_dipolar2: dev/gpe_engine.cu:1768-1810.  (1/2) sum rho Re V_dip */
int gpeo_energy_dipolar2(gpeo_t *e, double *t, double *edip)
{
    *t = e->t0 + e->dt * e->it;
    dipolar_potential(e, e->psi, e->wrk);
    double *pl = plane_buf(e);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < e->nx; ix++) {
        double s = 0.0;
        for (int iy = 0; iy < e->ny; iy++)
            for (int iz = 0; iz < e->nz; iz++) {
                size_t i = IDX3(ix, iy, iz);
                s += density(e, e->psi[i]) * e->wrk[i].re;
            }
        pl[ix] = s;
    }
    *edip = 0.5 * ordered_sum_planes(pl, e->nx);
    free(pl);
    return 0;
}

/* the convolution potential alone (for the pinning tests) */
int gpeo_dipolar_potential(gpeo_t *e, double *vdip)
{
    dipolar_potential(e, e->psi, e->wrk);
    for (size_t i = 0; i < e->n; i++) vdip[i] = e->wrk[i].re;
    return 0;
}
