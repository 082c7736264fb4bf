// Drop-in check of the singleton C ABI: the call sequence of the reference's
// differential drivers (test_compare/gpe_imag.c:58-163, gpe_real.c:66-186 and
// test_vortex/ite_vortex.c:137-310) written against include/gpe_engine.h and
// built with the reference's own flags (-DNX -DNY -DNZ), so gpe_create_engine
// takes the lattice registered by the header.  Imaginary time to a tolerance,
// then real time from the resulting psi, then a vortex imprint; writes
// psi_ite.bin, psi_rte.bin, psi_vortex.bin (raw complex128) and energies.txt.
//
//   ite_rte_driver <outdir> [max_ite_blocks]
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gpe_engine.h"
#include "gpe_io.h"

static void write_psi(const char* dir, const char* name, const cuCplx* psi, size_t n)
{
    char path[512];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE* f = fopen(path, "wb");
    if (!f || fwrite(psi, sizeof(cuCplx) * n, 1, f) != 1) {
        fprintf(stderr, "cannot write %s\n", path);
        exit(EXIT_FAILURE);
    }
    fclose(f);
}

int main(int argc, char** argv)
{
    const char* dir = argc > 1 ? argv[1] : ".";
    const int max_blocks = argc > 2 ? atoi(argv[2]) : 50;
    int ierr;
    gpe_exec(gpe_set_device(0), ierr);

    int nx, ny, nz;
    gpe_get_lattice(&nx, &ny, &nz);
    printf("# GPE engine compiled for lattice: %d x %d x %d\n", nx, ny, nz);
    const size_t nxyz = (size_t)nx * ny * nz;

    cuCplx* psi;
    gpe_exec(gpe_host_alloc((void**)&psi, sizeof(cuCplx) * nxyz), ierr);
    for (size_t i = 0; i < nxyz; i++) { psi[i].x = 1.0; psi[i].y = 0.0; }

    const double npart = 1000.0, dt = 0.025;
    double params[3] = {0.01, 0.10, 0.11};
    char epath[512];
    snprintf(epath, sizeof(epath), "%s/energies.txt", dir);
    FILE* fe = fopen(epath, "w");

    // imaginary time (alpha = 0, beta = 1)
    gpe_exec(gpe_create_engine(0.0, 1.0, dt, npart), ierr);
    gpe_exec(gpe_set_user_params(3, params), ierr);
    gpe_exec(gpe_set_psi(0.0, psi), ierr);
    gpe_exec(gpe_normalize_psi(), ierr);
    double t, ekin, eint, eext, etot, eprev, mu = 0.0;
    gpe_exec(gpe_energy(&t, &ekin, &eint, &eext), ierr);
    etot = ekin + eint + eext;
    fprintf(fe, "ite %.3f %.17g %.17g %.17g %.17g\n", t, etot, ekin, eint, eext);
    for (int b = 0; b < max_blocks; b++) {
        b_t();
        gpe_exec(gpe_evolve_qf(20, &mu), ierr);
        gpe_exec(gpe_energy(&t, &ekin, &eint, &eext), ierr);
        double rt = e_t(0);
        eprev = etot;
        etot = ekin + eint + eext;
        fprintf(fe, "ite %.3f %.17g %.17g %.17g %.17g %.17g %.4f\n", t, etot, ekin, eint, eext, mu, rt);
        if (fabs((eprev - etot) / npart) < 1.0e-9) break;
    }
    gpe_exec(gpe_get_psi(&t, psi), ierr);
    write_psi(dir, "psi_ite.bin", psi, nxyz);
    char info[512];
    snprintf(info, sizeof(info), "%s/psi_ite.bin", dir);
    save_info(info, nx, ny, nz, dt, 1, 0.0, 1.0, params[0], 0.0, npart);
    gpe_exec(gpe_destroy_engine(), ierr);

    // real time from the ground state (alpha = 1, beta = 0), as test_compare/gpe_real.c
    gpe_exec(gpe_create_engine(1.0, 0.0, dt, npart), ierr);
    gpe_exec(gpe_set_user_params(3, params), ierr);
    gpe_exec(gpe_set_psi(0.0, psi), ierr);
    for (int b = 0; b < 3; b++) {
        gpe_exec(gpe_evolve_qf(20), ierr);
        gpe_exec(gpe_energy(&t, &ekin, &eint, &eext), ierr);
        fprintf(fe, "rte %.3f %.17g %.17g %.17g %.17g\n", t, ekin + eint + eext, ekin, eint, eext);
    }
    gpe_exec(gpe_get_psi(&t, psi), ierr);
    write_psi(dir, "psi_rte.bin", psi, nxyz);

    // vortex imprint in imaginary time (test_vortex/ite_vortex.c:245-286)
    gpe_exec(gpe_set_ite_evolution(), ierr);
    gpe_exec(gpe_set_vortex(0.0, 2.5), ierr);
    gpe_exec(gpe_evolve_vortex(20), ierr);
    gpe_exec(gpe_energy(&t, &ekin, &eint, &eext), ierr);
    fprintf(fe, "vortex %.3f %.17g %.17g %.17g %.17g\n", t, ekin + eint + eext, ekin, eint, eext);
    gpe_exec(gpe_get_psi(&t, psi), ierr);
    write_psi(dir, "psi_vortex.bin", psi, nxyz);
    gpe_exec(gpe_destroy_engine(), ierr);
    fclose(fe);
    gpe_host_free(psi);
    printf("driver: OK\n");
    return 0;
}
