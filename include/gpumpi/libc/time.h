// Device time.h / sys/time.h for converted programs: wall time from the GPU's constant
// 100 MHz clock (the clock MPI_Wtime reads), counted from an arbitrary origin.
#pragma once
#include <hip/hip_runtime.h>
#include <time.h>
#include <sys/time.h>

namespace gpumpi_libc {
__device__ inline int get_time_of_day(struct timeval* tv, void*) {
    const unsigned long long t = wall_clock64();  // 100 MHz ticks
    tv->tv_sec = (time_t)(t / 100000000ull);
    tv->tv_usec = (suseconds_t)((t % 100000000ull) / 100ull);
    return 0;
}
__device__ inline time_t seconds(time_t* out) {
    const time_t s = (time_t)(wall_clock64() / 100000000ull);
    if (out) *out = s;
    return s;
}
}  // namespace gpumpi_libc

#define gettimeofday(tv, tz) ::gpumpi_libc::get_time_of_day((tv), (tz))
#define time(out) ::gpumpi_libc::seconds(out)
#define clock() ((clock_t)(wall_clock64() / (100000000ull / CLOCKS_PER_SEC)))
