/*
 * ".dftc.info" sidecar of a wave-function file: the text file "<file>.info"
 * next to a raw complex128 .dftc frame file (C order [nom][nx][ny][nz]).
 * Same functions, line order and formats as the reference's dev/gpe_io.h:
 *   save_info / get_info          10 lines: nx ny nz dt nom a aspect omega_x r0 npart  (dev/gpe_io.h:16-44, 99-128)
 *   gpe_save_info / gpe_get_info  2 header lines (system type, mode), the same 10
 *                                 values in %.15e, and the total time              (dev/gpe_io.h:46-96, 131-168)
 * Each line is "<value>   # <name>"; readers take the first token of each line.
 */
#ifndef GPE_IO_H
#define GPE_IO_H

#include <stdio.h>
#include <stdlib.h>

#include "gpe_user_defined.h"

#define MAX_REC_LEN 256

static inline FILE *gpe_io_open_info_(const char *dftc_filename, const char *mode)
{
    char name[1024];
    snprintf(name, sizeof(name), "%s.info", dftc_filename);
    return fopen(name, mode);
}

static inline void save_info(char *dftc_filename, const int nx, const int ny, const int nz, const double dt,
                             const int nom, const double scat_lenght, const double aspect_ratio,
                             const double omega_x, const double r0, const double npart)
{
    FILE *f = gpe_io_open_info_(dftc_filename, "w");
    if (!f) {
        printf("Error: Cannot open file %s.info!\n", dftc_filename);
        return;
    }
    fprintf(f, "%d   # nx\n%d   # ny\n%d   # nz\n", nx, ny, nz);
    fprintf(f, "%lf   # dt\n%d   # nom\n", dt, nom);
    fprintf(f, "%lf   # a_scattering\n%lf   # aspect_ratio\n%lf   # omega_x\n", scat_lenght, aspect_ratio, omega_x);
    fprintf(f, "%lf   # r0\n%lf   # npart\n", r0, npart);
    fclose(f);
}

#ifdef __cplusplus
static inline void gpe_save_info(char *dftc_filename, const int nx, const int ny, const int nz, const double dt,
                                 const int nom, const double scat_lenght, const double aspect_ratio,
                                 const double omega_x, const double r0, const double npart,
                                 const double total_time = 0)
#else
static inline void gpe_save_info(char *dftc_filename, const int nx, const int ny, const int nz, const double dt,
                                 const int nom, const double scat_lenght, const double aspect_ratio,
                                 const double omega_x, const double r0, const double npart, const double total_time)
#endif
{
    FILE *f = gpe_io_open_info_(dftc_filename, "w");
    if (!f) {
        printf("Error: Cannot open file %s.info!\n", dftc_filename);
        return;
    }
#if (GPE_FOR == PARTICLES)
    fprintf(f, "%s   # system type\n", "bosons");
#if (INTERACTIONS == -2)
    fprintf(f, "%s   # mode\n", "dipolar");
#else
    fprintf(f, "%s   # mode\n", "contact");
#endif
#else
    fprintf(f, "%s   # system type\n", "fermions");
#if (INTERACTIONS == 0)
    fprintf(f, "%s   # mode\n", "unitary");
#else
    fprintf(f, "%s   # mode\n", "bec");
#endif
#endif
    fprintf(f, "%d   # nx\n%d   # ny\n%d   # nz\n", nx, ny, nz);
    fprintf(f, "%.15e   # dt\n%d   # nom\n", dt, nom);
    fprintf(f, "%.15e   # a_scattering\n%.15e   # aspect_ratio\n%.15e   # omega_x\n", scat_lenght, aspect_ratio,
            omega_x);
    fprintf(f, "%.15e   # r0\n%.15e   # npart\n%.15e   # total time\n", r0, npart, total_time);
    fclose(f);
}

static inline int gpe_io_read_values_(FILE *fp, int *nx, int *ny, int *nz, double *dt, int *nom, double *scat_lenght,
                                      double *aspect_ratio, double *omega_x, double *r0, double *npart)
{
    char s[MAX_REC_LEN];
    int ok = 0;
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%d", nx);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%d", ny);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%d", nz);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%lf", dt);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%d", nom);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%lf", scat_lenght);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%lf", aspect_ratio);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%lf", omega_x);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%lf", r0);
    if (fgets(s, MAX_REC_LEN, fp)) ok += sscanf(s, "%lf", npart);
    return ok;
}

/* reads the 10-line format; the argument is the .info file itself */
static inline void get_info(char *info_filename, int *nx, int *ny, int *nz, double *dt, int *nom,
                            double *scat_lenght, double *aspect_ratio, double *omega_x, double *r0, double *npart)
{
    FILE *fp = fopen(info_filename, "r");
    if (fp == NULL) {
        printf("Error: Cannot open file %s!\n", info_filename);
        exit(EXIT_FAILURE);
    }
    gpe_io_read_values_(fp, nx, ny, nz, dt, nom, scat_lenght, aspect_ratio, omega_x, r0, npart);
    fclose(fp);
}

/* reads the header format of "<dftc_filename>.info" */
#ifdef __cplusplus
static inline void gpe_get_info(char *dftc_filename, int *nx, int *ny, int *nz, double *dt, int *nom,
                                double *scat_lenght, double *aspect_ratio, double *omega_x, double *r0,
                                double *npart, double *total_time = NULL)
#else
static inline void gpe_get_info(char *dftc_filename, int *nx, int *ny, int *nz, double *dt, int *nom,
                                double *scat_lenght, double *aspect_ratio, double *omega_x, double *r0,
                                double *npart, double *total_time)
#endif
{
    FILE *fp = gpe_io_open_info_(dftc_filename, "r");
    if (fp == NULL) {
        printf("Error: Cannot open file %s.info!\n", dftc_filename);
        exit(EXIT_FAILURE);
    }
    char s[MAX_REC_LEN];
    if (!fgets(s, MAX_REC_LEN, fp) || !fgets(s, MAX_REC_LEN, fp)) { /* system type, mode */
        fclose(fp);
        return;
    }
    gpe_io_read_values_(fp, nx, ny, nz, dt, nom, scat_lenght, aspect_ratio, omega_x, r0, npart);
    if (total_time && fgets(s, MAX_REC_LEN, fp)) sscanf(s, "%lf", total_time);
    fclose(fp);
}

#endif
