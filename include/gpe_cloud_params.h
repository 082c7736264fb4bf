/*
 * Cloud parameters of a wave function held on the host (a psi array copied
 * back with gpe_get_psi): density at the lattice centre, largest density and
 * the cloud range along each axis.  Same names, arguments and results as the
 * reference's dev/gpe_cloud_params.h:60-119:
 *   find_n0(psi)                     gamma |psi|^2 at (nx/2, ny/2, nz/2)
 *   find_max_density(psi)            largest gamma |psi|^2
 *   cloud_range_ix/iy/iz(psi, eps)   scanning from the centre index down to 1
 *                                    along the line through the centre, the
 *                                    first distance at which the density is
 *                                    below eps (default 1e-8); NaN if none is
 * Like the reference, the forms without sizes read the lattice from the
 * globals `int nx, ny, nz, nxyz` that the driver defines, and gamma from
 * GAMMA (gpe_user_defined.h).  psi is any pointer to complex128 values
 * (cuCplx, cuDoubleComplex, double complex).
 *
 * The gpe_cloud_* forms take the sizes and gamma as arguments.  The engine
 * computes the same five numbers on the device without a host copy of psi:
 * gpe_engine_cloud_params (gpe_engine.h).
 */
#ifndef GPE_CLOUD_PARAMS_H
#define GPE_CLOUD_PARAMS_H

#include <math.h>
#include <stddef.h>

#include "gpe_user_defined.h"

static inline double gpe_cloud_density_(const void *psi, size_t i, double gamma)
{
    const double *p = (const double *)psi + 2 * i;
    return gamma * (p[0] * p[0] + p[1] * p[1]);
}

static inline double gpe_cloud_n0(const void *psi, int nx, int ny, int nz, double gamma)
{
    return gpe_cloud_density_(psi, ((size_t)(nx / 2) * ny + (size_t)(ny / 2)) * nz + (size_t)(nz / 2), gamma);
}

static inline double gpe_cloud_max_density(const void *psi, size_t n, double gamma)
{
    double m = 0.0;
#if defined(_OPENMP)
#pragma omp parallel for reduction(max : m)
#endif
    for (long i = 0; i < (long)n; i++) {
        const double d = gpe_cloud_density_(psi, (size_t)i, gamma);
        if (d > m) m = d;
    }
    return m;
}

/* axis 0, 1, 2: x, y, z */
static inline double gpe_cloud_range(const void *psi, int nx, int ny, int nz, double gamma, int axis, double epsilon)
{
    const int len = axis == 0 ? nx : axis == 1 ? ny : nz;
    for (int i = len / 2; i > 0; i--) {
        const size_t ix = axis == 0 ? (size_t)i : (size_t)(nx / 2);
        const size_t iy = axis == 1 ? (size_t)i : (size_t)(ny / 2);
        const size_t iz = axis == 2 ? (size_t)i : (size_t)(nz / 2);
        if (gpe_cloud_density_(psi, (ix * ny + iy) * nz + iz, gamma) < epsilon) return (double)(len / 2 - i);
    }
    return NAN;
}

#ifndef GPE_CLOUD_NO_GLOBALS
extern int nx, ny, nz, nxyz;

static inline double find_n0(const void *psi) { return gpe_cloud_n0(psi, nx, ny, nz, GAMMA); }

static inline double find_max_density(const void *psi) { return gpe_cloud_max_density(psi, (size_t)nxyz, GAMMA); }

#ifdef __cplusplus
static inline double cloud_range_ix(const void *psi, const double epsilon = 1e-08)
{
    return gpe_cloud_range(psi, nx, ny, nz, GAMMA, 0, epsilon);
}
static inline double cloud_range_iy(const void *psi, const double epsilon = 1e-08)
{
    return gpe_cloud_range(psi, nx, ny, nz, GAMMA, 1, epsilon);
}
static inline double cloud_range_iz(const void *psi, const double epsilon = 1e-08)
{
    return gpe_cloud_range(psi, nx, ny, nz, GAMMA, 2, epsilon);
}
#else
static inline double cloud_range_ix(const void *psi, const double epsilon)
{
    return gpe_cloud_range(psi, nx, ny, nz, GAMMA, 0, epsilon);
}
static inline double cloud_range_iy(const void *psi, const double epsilon)
{
    return gpe_cloud_range(psi, nx, ny, nz, GAMMA, 1, epsilon);
}
static inline double cloud_range_iz(const void *psi, const double epsilon)
{
    return gpe_cloud_range(psi, nx, ny, nz, GAMMA, 2, epsilon);
}
#endif
#endif /* GPE_CLOUD_NO_GLOBALS */

#endif /* GPE_CLOUD_PARAMS_H */
