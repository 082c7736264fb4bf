/*
 * Wall-clock / CPU-clock stopwatch used by the drivers around each block of
 * steps: b_t() starts, e_t(0) stops and returns the elapsed wall time in
 * seconds, e_t(1) returns the last wall time, e_t(2) the last CPU time.
 * Same calls and meaning as the reference's dev/gpe_timing.h:39-64 and
 * orginal/gpe_timing.h:31-51.
 */
#ifndef GPE_TIMING_H
#define GPE_TIMING_H

#include <sys/time.h>
#include <time.h>

static double gpe_timing_wall_;
static clock_t gpe_timing_cpu_;

static inline void b_t(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    gpe_timing_cpu_ = clock();
    gpe_timing_wall_ = (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

#ifdef __cplusplus
static inline double e_t(int type = 0)
#else
static inline double e_t(int type)
#endif
{
    if (type == 0) {
        struct timeval tv;
        clock_t c = clock();
        gettimeofday(&tv, NULL);
        gpe_timing_cpu_ = c - gpe_timing_cpu_;
        gpe_timing_wall_ = (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec - gpe_timing_wall_;
        return gpe_timing_wall_;
    }
    if (type == 2) return (double)gpe_timing_cpu_ / (double)CLOCKS_PER_SEC;
    return gpe_timing_wall_;
}

#endif
