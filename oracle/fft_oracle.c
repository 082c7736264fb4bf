/*
 * FP64 mixed-radix Stockham FFT for the CPU oracle (test infrastructure only).
 *
 * The reference calls cuFFT Z2Z (orginal/gpe_engine.cu:519-527 plan,
 * 963/966/997/1000/1103/1106 exec), which is closed source and not vendored,
 * so this file implements the transform from its definition:
 *   X[k] = sum_j x[j] exp(sign * 2 pi i j k / n),  sign = -1 forward, +1 inverse,
 * unnormalised, applied along z (contiguous), then y, then x.
 *
 * Algorithm: self-sorting Stockham stages.  Stage s with radix R and span Ns
 * (product of the previous radices) maps butterfly j (0 <= j < n/R):
 *   v[r]  = in[j + r n/R] * w^(r (j mod Ns) n/(Ns R))      (w = exp(sign 2 pi i/n))
 *   V     = DFT_R(v)
 *   out[(j/Ns) Ns R + (j mod Ns) + r Ns] = V[r]
 * which leaves the result in natural order after the last stage.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "gpe_oracle.h"

typedef struct { double re, im; } cd;

static inline cd cmul(cd a, cd b) { cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cd cadd(cd a, cd b) { cd r = {a.re + b.re, a.im + b.im}; return r; }
static inline cd csub(cd a, cd b) { cd r = {a.re - b.re, a.im - b.im}; return r; }

/* exp(-2 pi i m/n) with the angle reduced to the first octant in exact
 * integer arithmetic, so that large m keeps full accuracy. */
void gpeo_unit_root(long m, long n, double *re, double *im)
{
    m %= n;
    if (m < 0) m += n;
    long o = (8 * m) / n;               /* octant 0..7 */
    long num = 8 * m - o * n;           /* residual numerator, 0 <= num < n */
    double phi = (M_PI / 4.0) * ((double)num / (double)n);          /* in [0, pi/4) */
    double phc = (M_PI / 4.0) * ((double)(n - num) / (double)n);    /* pi/4 - phi */
    double c, s;
    switch (o) {
    case 0: c = cos(phi);  s = sin(phi);  break;
    case 1: c = sin(phc);  s = cos(phc);  break;
    case 2: c = -sin(phi); s = cos(phi);  break;
    case 3: c = -cos(phc); s = sin(phc);  break;
    case 4: c = -cos(phi); s = -sin(phi); break;
    case 5: c = -sin(phc); s = -cos(phc); break;
    case 6: c = sin(phi);  s = -cos(phi); break;
    default: c = cos(phc); s = -sin(phc); break;
    }
    *re = c;
    *im = -s;
}

#define MAXFAC 48
typedef struct {
    int n, nf;
    int fac[MAXFAC];
    cd *tw;              /* tw[m] = exp(-2 pi i m / n) */
} plan1d;

#define MAXPLANS 64
static plan1d g_plans[MAXPLANS];
static int g_nplans = 0;

static const plan1d *get_plan(int n)
{
    const plan1d *found = NULL;
#pragma omp critical(gpeo_fft_plan)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i].n == n) { found = &g_plans[i]; break; }
        if (!found && g_nplans < MAXPLANS) {
            plan1d *p = &g_plans[g_nplans];
            p->n = n;
            p->nf = 0;
            int r = n;
            while (r % 4 == 0 && p->nf < MAXFAC) { p->fac[p->nf++] = 4; r /= 4; }
            while (r % 2 == 0 && p->nf < MAXFAC) { p->fac[p->nf++] = 2; r /= 2; }
            for (int f = 3; r > 1 && p->nf < MAXFAC; f += 2)
                while (r % f == 0) { p->fac[p->nf++] = f; r /= f; }
            p->tw = (cd *)malloc(sizeof(cd) * (size_t)n);
            for (int m = 0; m < n; m++) gpeo_unit_root(m, n, &p->tw[m].re, &p->tw[m].im);
            g_nplans++;
            found = p;
        }
    }
    return found;
}

/* w^e for the plan's root, inverse direction conjugates */
static inline cd twid(const plan1d *p, long e, int sign)
{
    cd w = p->tw[e % p->n];
    if (sign > 0) w.im = -w.im;
    return w;
}

/* R-point DFT in place, v[q] <- sum_m v[m] exp(sign 2 pi i m q / R) */
static void dft_small(cd *v, int R, int sign, const plan1d *p)
{
    if (R == 2) {
        cd a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        cd t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        cd t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
        /* sign * i * t3 */
        cd it3 = {-sign * t3.im, sign * t3.re};
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, it3);
        v[3] = csub(t1, it3);
    } else if (R == 3) {
        const double h = 0.86602540378443864676; /* sqrt(3)/2 */
        cd t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        cd m = {v[0].re - 0.5 * t1.re, v[0].im - 0.5 * t1.im};
        cd it2 = {-sign * h * t2.im, sign * h * t2.re};
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, it2);
        v[2] = csub(m, it2);
    } else {
        /* generic O(R^2) DFT using the plan's n-th roots: exp(2 pi i/R) = w^(n/R) */
        cd tmp[64];
        cd *t = tmp;
        cd *heap = NULL;
        if (R > 64) { heap = (cd *)malloc(sizeof(cd) * (size_t)R); t = heap; }
        long step = p->n / R;
        for (int q = 0; q < R; q++) {
            cd acc = {0.0, 0.0};
            for (int m = 0; m < R; m++) acc = cadd(acc, cmul(v[m], twid(p, (long)((long)m * q % R) * step, sign)));
            t[q] = acc;
        }
        memcpy(v, t, sizeof(cd) * (size_t)R);
        free(heap);
    }
}

/* one 1-D transform of x (length n) using work (length n) */
static void fft1d_run(const plan1d *p, cd *x, cd *work, int sign)
{
    const int n = p->n;
    cd *in = x, *out = work;
    long Ns = 1;
    cd vbuf[64];
    for (int s = 0; s < p->nf; s++) {
        const int R = p->fac[s];
        const int M = n / R;
        cd *v = vbuf;
        cd *heap = NULL;
        if (R > 64) { heap = (cd *)malloc(sizeof(cd) * (size_t)R); v = heap; }
        const long tstep = n / (Ns * R);
        for (int j = 0; j < M; j++) {
            long k = j % Ns;
            for (int r = 0; r < R; r++) v[r] = in[j + r * M];
            if (k)
                for (int r = 1; r < R; r++) v[r] = cmul(v[r], twid(p, k * r * tstep, sign));
            dft_small(v, R, sign, p);
            long idxD = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) out[idxD + r * Ns] = v[r];
        }
        free(heap);
        cd *tmp = in; in = out; out = tmp;
        Ns *= R;
    }
    if (in != x) memcpy(x, in, sizeof(cd) * (size_t)n);
}

int gpeo_fft1d(int n, double *data, int sign)
{
    if (n < 1) return -1;
    const plan1d *p = get_plan(n);
    if (!p) return -2;
    cd *work = (cd *)malloc(sizeof(cd) * (size_t)n);
    fft1d_run(p, (cd *)data, work, sign);
    free(work);
    return 0;
}

/* transform all pencils of one axis; len = axis length, inner = stride */
static void fft_axis(cd *a, size_t total, int len, size_t inner, int sign)
{
    const plan1d *p = get_plan(len);
    const long npen = (long)(total / (size_t)len);
    if (inner == 1) {
#pragma omp parallel
        {
            cd *work = (cd *)malloc(sizeof(cd) * (size_t)len);
#pragma omp for schedule(static)
            for (long q = 0; q < npen; q++) fft1d_run(p, a + (size_t)q * len, work, sign);
            free(work);
        }
        return;
    }
    /* strided pencils, processed in groups of G consecutive columns */
    const int G = 8;
    const long ngroups = (npen + G - 1) / G;
#pragma omp parallel
    {
        cd *buf = (cd *)malloc(sizeof(cd) * (size_t)len * G);
        cd *work = (cd *)malloc(sizeof(cd) * (size_t)len);
#pragma omp for schedule(static)
        for (long g = 0; g < ngroups; g++) {
            long q0 = g * G;
            int cnt = (int)((q0 + G <= npen) ? G : (npen - q0));
            for (int c = 0; c < cnt; c++) {
                long q = q0 + c;
                size_t base = (size_t)(q / (long)inner) * (size_t)len * inner + (size_t)(q % (long)inner);
                for (int i = 0; i < len; i++) buf[(size_t)c * len + i] = a[base + (size_t)i * inner];
            }
            for (int c = 0; c < cnt; c++) fft1d_run(p, buf + (size_t)c * len, work, sign);
            for (int c = 0; c < cnt; c++) {
                long q = q0 + c;
                size_t base = (size_t)(q / (long)inner) * (size_t)len * inner + (size_t)(q % (long)inner);
                for (int i = 0; i < len; i++) a[base + (size_t)i * inner] = buf[(size_t)c * len + i];
            }
        }
        free(buf);
        free(work);
    }
}

int gpeo_fft3d(int nx, int ny, int nz, double *data, int sign)
{
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    size_t total = (size_t)nx * ny * nz;
    cd *a = (cd *)data;
    fft_axis(a, total, nz, 1, sign);
    fft_axis(a, total, ny, (size_t)nz, sign);
    fft_axis(a, total, nx, (size_t)ny * nz, sign);
    return 0;
}

int gpeo_num_threads(void)
{
    return omp_get_max_threads();
}
