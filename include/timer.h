/*
 * timer.h - wall-clock microsecond timers (reference em2d/timer.h:21-50).
 */
#ifndef TIMER_H_DROPIN
#define TIMER_H_DROPIN

#include <stdint.h>

uint64_t timer_ticks(void);
double timer_interval_seconds(uint64_t start, uint64_t end);
double timer_cpu_seconds(void);
double timer_resolution(void);

#endif
