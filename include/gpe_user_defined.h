/*
 * Physics selection of the MI355X GPE engine, in the reference's vocabulary.
 *
 * Replaces the compile-time hook header of the reference
 * (dev/gpe_user_defined.h:35-561, orginal/gpe_user_defined.h:36-126).  The
 * engine evaluates the hooks (external potential, energy-density functional
 * and its derivative, wave-function modifier) on the device; which of the
 * built-in hooks runs is a run-time choice of the engine, and this header
 * carries the compile-time defaults a driver selects with
 * -DINTERACTIONS=<n> -DVEXT=<n> (exactly as the reference's Makefile does) and
 * the user-parameter enums that drivers index with.
 *
 *   INTERACTIONS  0  unitary Fermi gas, EDF = (3/5) xi (3 pi^2 n)^(2/3) n / 2,   xi = 0.37
 *                 1  ETF functional on the BEC side of the crossover (A_SCAT)
 *                -1  contact bosons, EDF = 2 pi a n^2                          (A_SCAT)
 *                -2  dipolar bosons, local contact part 2 pi a_con n^2         (A_CON, A_DIP)
 *   VEXT          0  harmonic trap centred at n/2: 0.5 sum_a omega_a^2 (i_a - n_a/2)^2
 *                 1  Woods-Saxon tube V0 (1 - 1/(1 + exp((r - R)/a)))          (V_0, A, R)
 *                 2  cosine lattice A (cos(2 pi P ix/nx) + cos(2 pi P iy/ny) + cos(2 pi P iz/nz))
 *                                                                              (V_AMP, V_PERIODS)
 *                 3  caller-supplied array (gpe_set_external_potential)
 *   INTERACTIONS / VEXT 100 (GPE_USER): the user hook file compiled into the
 *                 library (include/gpe_user_hooks.h, or the file named by
 *                 GPE_USER_HOOKS at build time): gpe_EDF / gpe_dEDFdn,
 *                 gpe_external_potential and gpe_modify_psi with the time index
 *   GPE_FOR       PARTICLES (gamma = 1) or DIMERS (gamma = 2); defaults as in the
 *                 reference: DIMERS for INTERACTIONS 0 and 1, PARTICLES otherwise.
 */
#ifndef GPE_USER_DEFINED_H
#define GPE_USER_DEFINED_H

#include <stdio.h>

#define PARTICLES 1
#define DIMERS 2
#define GPE_USER 100

#ifndef INTERACTIONS
#define INTERACTIONS 0
#endif
#ifndef VEXT
#define VEXT 0
#endif

/* The engine accepts up to 32 user parameters for every physics choice
 * (the orginal header's limit; dev/gpe_user_defined.h lowers it to 4 or 5). */
#ifndef MAX_USER_PARAMS
#define MAX_USER_PARAMS 32
#endif

#if (INTERACTIONS == -2)
typedef enum { OMEGA_X, OMEGA_Y, OMEGA_Z, A_CON, A_DIP } user_params_t;
#else
typedef enum { OMEGA_X, OMEGA_Y, OMEGA_Z, A_SCAT } user_params_t;
#endif

#if (VEXT == 1)
typedef enum { V_0, A, R } potential_params_t;
#elif (VEXT == 2)
typedef enum { V_AMP, V_PERIODS } potential_params_t;
#endif

#ifndef GPE_FOR
#if (INTERACTIONS == 0 || INTERACTIONS == 1)
#define GPE_FOR DIMERS
#else
#define GPE_FOR PARTICLES
#endif
#endif

#if (GPE_FOR == PARTICLES)
#define GAMMA 1.0
#elif (GPE_FOR == DIMERS)
#define GAMMA 2.0
#endif

#if (INTERACTIONS == 0)
#define UNITARY
#elif (INTERACTIONS == -1)
#define CONTACT
#elif (INTERACTIONS == -2)
#define DIPOLAR
#endif

static inline void gpe_print_interactions_type(void)
{
#if (INTERACTIONS == 0)
    printf("# GPE FOR FERMIONIC DIMERS IN UNITARY LIMIT\n");
#elif (INTERACTIONS == 1)
    printf("# GPE FOR FERMIONIC DIMERS IN BEC LIMIT\n");
#elif (INTERACTIONS == -1)
    printf("# GPE FOR BOSONS IN BEC\n");
#elif (INTERACTIONS == -2)
    printf("# GPE FOR BOSONS IN BEC WITH DIPOLAR INTERACTIONS\n");
#endif
}

#endif
