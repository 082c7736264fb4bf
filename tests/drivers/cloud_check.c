/* Reads a raw complex128 psi of nx*ny*nz points (C order) and prints the cloud
 * parameters of include/gpe_cloud_params.h, through the reference's names
 * (globals nx, ny, nz, nxyz and GAMMA from INTERACTIONS) and the sized forms:
 *   cloud_check <file> nx ny nz epsilon */
#include <stdio.h>
#include <stdlib.h>

#include "gpe_cloud_params.h"

int nx, ny, nz, nxyz;

int main(int argc, char **argv)
{
    if (argc != 6) return 2;
    nx = atoi(argv[2]);
    ny = atoi(argv[3]);
    nz = atoi(argv[4]);
    nxyz = nx * ny * nz;
    const double eps = atof(argv[5]);
    double *psi = (double *)malloc(sizeof(double) * 2 * (size_t)nxyz);
    FILE *f = fopen(argv[1], "rb");
    if (!f || !psi || fread(psi, sizeof(double) * 2, (size_t)nxyz, f) != (size_t)nxyz) return 3;
    fclose(f);
    printf("reference %.17g %.17g %.17g %.17g %.17g\n", find_n0(psi), find_max_density(psi),
           cloud_range_ix(psi, eps), cloud_range_iy(psi, eps), cloud_range_iz(psi, eps));
    printf("sized %.17g %.17g %.17g %.17g %.17g\n", gpe_cloud_n0(psi, nx, ny, nz, GAMMA),
           gpe_cloud_max_density(psi, (size_t)nxyz, GAMMA), gpe_cloud_range(psi, nx, ny, nz, GAMMA, 0, eps),
           gpe_cloud_range(psi, nx, ny, nz, GAMMA, 1, eps), gpe_cloud_range(psi, nx, ny, nz, GAMMA, 2, eps));
    free(psi);
    return 0;
}
