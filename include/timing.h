#pragma once
/* Monotonic wall-clock ticks; ABI counterpart of timing.h:6-8. */

#include <stdint.h>

int64_t get_time_ticks(void);

int64_t get_tick_resolution(void);
