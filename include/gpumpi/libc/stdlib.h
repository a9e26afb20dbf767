// Device stdlib for converted programs (gpu_libs/gpu_libc/stdlib.cuh:1-49,
// stdlib.h.cuh:23-72): allocation through the gpumpi heap -- the coalescing variant
// chosen by GPUMPI_MALLOC_V<n> as the reference dispatches it --, exit/abort ending the
// calling rank, number parsing, a per-rank rand(), qsort.  All functions are called
// wave-uniformly (every lane of the rank's wavefront with the same arguments).
#pragma once
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include "../gpu_malloc.h"
#include "../global_vars.h"
#include "../mpi.h"
#include "parse.h"

namespace gpumpi_libc {
__device__ inline void* alloc(size_t n, bool coalesced = false) {
#if defined(GPUMPI_MALLOC_V1)
    return __gpu_malloc_v1(n, coalesced);
#elif defined(GPUMPI_MALLOC_V2)
    return __gpu_malloc_v2(n, coalesced);
#elif defined(GPUMPI_MALLOC_V3)
    return __gpu_malloc_v3(n, coalesced);
#elif defined(GPUMPI_MALLOC_V4)
    return __gpu_malloc_v4(n, coalesced);
#elif defined(GPUMPI_MALLOC_V5)
    return __gpu_malloc_v5(n, coalesced);
#elif defined(GPUMPI_MALLOC_V6) || defined(GPUMPI_MALLOC_V7)
    return __gpu_malloc_v6(n, coalesced);
#else
    (void)coalesced;
    return __gpu_malloc(n);
#endif
}
__device__ inline void release(void* p) {
#if defined(GPUMPI_MALLOC_V1)
    __gpu_free_v1(p);
#elif defined(GPUMPI_MALLOC_V2)
    __gpu_free_v2(p);
#elif defined(GPUMPI_MALLOC_V3)
    __gpu_free_v3(p);
#elif defined(GPUMPI_MALLOC_V4)
    __gpu_free_v4(p);
#elif defined(GPUMPI_MALLOC_V5)
    __gpu_free_v5(p);
#elif defined(GPUMPI_MALLOC_V6) || defined(GPUMPI_MALLOC_V7)
    __gpu_free_v6(p);
#else
    __gpu_free(p);
#endif
}

// rand(): a per-rank 64-bit LCG (Knuth's MMIX constants), upper 31 bits returned.
__device__ unsigned long long rand_state[65536];
__device__ inline unsigned long long& rand_word() {
    return rand_state[::gpumpi::gpu_global_rank() & 65535];
}
__device__ inline int next_rand() {
    unsigned long long& s = rand_word();
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (int)(s >> 33);
}

// qsort: heapsort over byte-swapped elements (O(n log n), no recursion, no scratch).
__device__ inline void swap_bytes(char* a, char* b, size_t n) {
    for (size_t i = 0; i < n; i++) {
        const char t = a[i];
        a[i] = b[i];
        b[i] = t;
    }
}
__device__ inline void sift(char* base, size_t root, size_t n, size_t sz,
                            int (*cmp)(const void*, const void*)) {
    for (;;) {
        size_t child = 2 * root + 1;
        if (child >= n) return;
        if (child + 1 < n && cmp(base + child * sz, base + (child + 1) * sz) < 0) child++;
        if (cmp(base + root * sz, base + child * sz) >= 0) return;
        swap_bytes(base + root * sz, base + child * sz, sz);
        root = child;
    }
}
__device__ inline void heap_sort(void* v, size_t n, size_t sz, int (*cmp)(const void*, const void*)) {
    char* base = (char*)v;
    if (n < 2) return;
    for (size_t i = n / 2; i-- > 0;) sift(base, i, n, sz, cmp);
    for (size_t end = n - 1; end > 0; end--) {
        swap_bytes(base, base + end * sz, sz);
        sift(base, 0, end, sz, cmp);
    }
}
}  // namespace gpumpi_libc

#define GPUMPI_LIBC_STDLIB 1
#define malloc(...) ::gpumpi_libc::alloc(__VA_ARGS__)
#define calloc(n, size) __gpu_calloc((n), (size))
#define realloc(p, size) __gpu_realloc((p), (size))
#define free(p) ::gpumpi_libc::release(p)
#define exit(status) gpumpi_exit(status)
#define abort() gpumpi_exit(134)
#define atoi(s) ((int)::gpumpi_libc::parse_int((s), nullptr, 10))
#define atol(s) ((long)::gpumpi_libc::parse_int((s), nullptr, 10))
#define atoll(s) ((long long)::gpumpi_libc::parse_int((s), nullptr, 10))
#define strtol(s, end, base) ((long)::gpumpi_libc::parse_int((s), (end), (base)))
#define strtoll(s, end, base) ((long long)::gpumpi_libc::parse_int((s), (end), (base)))
#define strtoul(s, end, base) ((unsigned long)::gpumpi_libc::parse_int((s), (end), (base)))
#define atof(s) ::gpumpi_libc::parse_double((s), nullptr)
#define strtod(s, end) ::gpumpi_libc::parse_double((s), (end))
#define rand() ::gpumpi_libc::next_rand()
#define srand(seed) ((void)(::gpumpi_libc::rand_word() = (unsigned long long)(seed)))
#define qsort(base, n, size, cmp) ::gpumpi_libc::heap_sort((base), (n), (size), (cmp))
#define getenv(name) ((char*)nullptr)
