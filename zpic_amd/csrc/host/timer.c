/*
 * timer.c - wall-clock microsecond timers (reference em2d/timer.c:24-84).
 */
#include <stdint.h>
#include <sys/time.h>

#include "timer.h"

uint64_t timer_ticks(void)
{
    struct timeval tv;
    gettimeofday(&tv, 0);
    return (uint64_t)tv.tv_sec * 1000000u + (uint64_t)tv.tv_usec;
}

double timer_interval_seconds(uint64_t start, uint64_t end) { return (double)(end - start) * 1.0e-6; }

double timer_cpu_seconds(void)
{
    struct timeval tv;
    gettimeofday(&tv, 0);
    return (double)tv.tv_sec + 1.0e-6 * (double)tv.tv_usec;
}

double timer_resolution(void)
{
    struct timeval a, b;
    gettimeofday(&a, 0);
    do {
        gettimeofday(&b, 0);
    } while (a.tv_usec == b.tv_usec);
    if (a.tv_sec < b.tv_sec) b.tv_usec += 1000000;
    return (double)(b.tv_usec - a.tv_usec) * 1.0e-6;
}
