/* Writes and reads .info sidecars through include/gpe_io.h for the format
 * tests: "w <dftc>" writes the 10-line and header forms (<dftc>.info and
 * <dftc>_h.info), "r <info>" prints the 10 values read back by get_info. */
#include <stdio.h>
#include <string.h>

#include "gpe_io.h"

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    if (argv[1][0] == 'w') {
        char name[1024], hname[1024];
        snprintf(name, sizeof(name), "%s", argv[2]);
        snprintf(hname, sizeof(hname), "%s_h", argv[2]);
        save_info(name, 24, 16, 32, 0.0025, 7, 0.125, 2.5, 0.390625, 1.5, 10000.0);
        gpe_save_info(hname, 24, 16, 32, 0.0025, 7, 0.125, 2.5, 0.390625, 1.5, 10000.0, 17.5);
        return 0;
    }
    int nx, ny, nz, nom;
    double dt, a, asp, om, r0, np;
    get_info(argv[2], &nx, &ny, &nz, &dt, &nom, &a, &asp, &om, &r0, &np);
    printf("%d %d %d %.17g %d %.17g %.17g %.17g %.17g %.17g\n", nx, ny, nz, dt, nom, a, asp, om, r0, np);
    return 0;
}
