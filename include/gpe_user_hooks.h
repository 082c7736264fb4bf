/*
 * User hooks compiled into the engine: the device functions of the
 * reference's gpe_user_defined.h, in its names and signatures
 * (dev/gpe_user_defined.h:81-84 gpe_modify_psi, 118-135 gpe_EDF / gpe_dEDFdn,
 * 488-513 gpe_external_potential; orginal/gpe_user_defined.h:66-126).
 *
 * Build the library with your own file in place of this one:
 *
 *     GPE_USER_HOOKS=/path/to/my_hooks.h python -m libcugpe_amd.build --force
 *
 * and select the hooks with INTERACTIONS 100 and/or VEXT 100 (GPE_USER in
 * include/gpe_user_defined.h).  INTERACTIONS 100 takes gpe_EDF / gpe_dEDFdn
 * from this file, VEXT 100 takes gpe_external_potential; with either,
 * gpe_modify_psi is applied where the reference applies it (to psi0 at the
 * start of each step and to the corrector input, orginal/gpe_engine.cu:737,
 * 798, 838, 865).  GPE_FOR chooses gamma as usual (0: PARTICLES).
 *
 * The file is a list of function definitions (no #include, no extern
 * declarations); the same file is compiled into the GPU library, where the
 * functions run in FP64 at every lattice point, and into the FP64 CPU oracle
 * of the tests.  In scope, as in the reference:
 *
 *   d_user_param[]          this lattice's user parameters (gpe_set_user_params;
 *                           per lattice in an ensemble), FP64, 32 entries
 *   d_nx, d_ny, d_nz        lattice size; NX, NY, NZ are the same values
 *   d_dt, d_t0, d_npart     step, time origin (time = d_t0 + it*d_dt), particle number
 *   uint, cuDoubleComplex   {double x, y} and make_cuDoubleComplex(re, im)
 *
 * Every function must be a pure function of its arguments and these values:
 * the engine re-evaluates it where the reference stored its result.
 * gpe_modify_psi must keep |psi| (the reference's note: "this function assumes
 * that norm is not changed after modification").
 *
 * This example: a contact Bose gas with a three-body term and a Feshbach-like
 * ramp of the scattering length, in a harmonic trap whose strength is
 * modulated in time and whose offset is ramped, with an optional phase kick of
 * a moving lattice.  With slots 4..9 zero it is the reference's contact gas
 * (INTERACTIONS -1) in its harmonic trap (VEXT 0).
 */

/* d_user_param slots of this example (OMEGA_X .. A_SCAT as in the reference) */
enum {
    UH_OMEGA_X = 0,
    UH_OMEGA_Y = 1,
    UH_OMEGA_Z = 2,
    UH_A_SCAT = 3,
    UH_TRAP_EPS = 4, /* relative modulation depth of the trap */
    UH_TRAP_W = 5,   /* angular frequency of the modulation and of the kick */
    UH_G3 = 6,       /* three-body coefficient: EDF += g3 rho^3 / 3 */
    UH_KICK = 7,     /* phase kick kappa of gpe_modify_psi */
    UH_A_RAMP = 8,   /* a(t) = A_SCAT (1 + A_RAMP t) */
    UH_V_RAMP = 9    /* uniform potential offset V_RAMP t */
};

/* psi exp(i kappa sin(2 pi ix/nx - W t)): a moving lattice phase kick */
inline __device__ cuDoubleComplex gpe_modify_psi(uint ix, uint iy, uint iz, uint it, cuDoubleComplex psi)
{
    const double kappa = d_user_param[UH_KICK];
    if (kappa == 0.0) return psi;
    const double t = d_t0 + it * d_dt;
    const double ph = kappa * sin(2.0 * M_PI * ix / d_nx - d_user_param[UH_TRAP_W] * t);
    const double c = cos(ph), s = sin(ph);
    return make_cuDoubleComplex(psi.x * c - psi.y * s, psi.x * s + psi.y * c);
}

/* harmonic trap centred at n/2, strength modulated as 1 + eps sin(W t),
 * plus a uniform offset ramped linearly in time */
inline __device__ double gpe_external_potential(uint ix, uint iy, uint iz, uint it)
{
    const double t = d_t0 + it * d_dt;
    const double x = (double)(ix)-1.0 * (NX / 2);
    const double y = (double)(iy)-1.0 * (NY / 2);
    const double z = (double)(iz)-1.0 * (NZ / 2);
    const double wx = d_user_param[UH_OMEGA_X], wy = d_user_param[UH_OMEGA_Y], wz = d_user_param[UH_OMEGA_Z];
    const double trap = 0.5 * x * x * wx * wx + 0.5 * y * y * wy * wy + 0.5 * z * z * wz * wz;
    return trap * (1.0 + d_user_param[UH_TRAP_EPS] * sin(d_user_param[UH_TRAP_W] * t)) + d_user_param[UH_V_RAMP] * t;
}

/* contact interaction with a ramped scattering length plus a three-body term */
inline __device__ double gpe_EDF(double rho, uint it)
{
    const double a = d_user_param[UH_A_SCAT] * (1.0 + d_user_param[UH_A_RAMP] * (d_t0 + it * d_dt));
    return 2.0 * M_PI * a * rho * rho + d_user_param[UH_G3] * rho * rho * rho / 3.0;
}

inline __device__ double gpe_dEDFdn(double rho, uint it)
{
    const double a = d_user_param[UH_A_SCAT] * (1.0 + d_user_param[UH_A_RAMP] * (d_t0 + it * d_dt));
    return 4.0 * M_PI * a * rho + d_user_param[UH_G3] * rho * rho;
}
