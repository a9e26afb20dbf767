// Device string.h for converted programs (gpu_libs/gpu_libc/string.cuh): memcpy and
// memset are the HIP device builtins; memcmp, memchr, memmove and the str* functions are
// written here (strdup / strndup allocate from the runtime's device heap).
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>
#include "../gpu_malloc.h"

namespace gpumpi_libc {
__device__ inline size_t str_len(const char* s) {
    size_t n = 0;
    while (s[n]) n++;
    return n;
}
__device__ inline char* str_cpy(char* d, const char* s) {
    size_t i = 0;
    for (; s[i]; i++) d[i] = s[i];
    d[i] = 0;
    return d;
}
__device__ inline char* str_ncpy(char* d, const char* s, size_t n) {
    size_t i = 0;
    for (; i < n && s[i]; i++) d[i] = s[i];
    for (; i < n; i++) d[i] = 0;
    return d;
}
__device__ inline int str_ncmp(const char* a, const char* b, size_t n) {
    for (size_t i = 0; i < n; i++) {
        const unsigned char x = (unsigned char)a[i], y = (unsigned char)b[i];
        if (x != y) return x < y ? -1 : 1;
        if (!x) return 0;
    }
    return 0;
}
__device__ inline int str_cmp(const char* a, const char* b) { return str_ncmp(a, b, (size_t)-1); }
__device__ inline char* str_cat(char* d, const char* s) { str_cpy(d + str_len(d), s); return d; }
__device__ inline char* str_ncat(char* d, const char* s, size_t n) {
    char* e = d + str_len(d);
    size_t i = 0;
    for (; i < n && s[i]; i++) e[i] = s[i];
    e[i] = 0;
    return d;
}
__device__ inline char* str_chr(const char* s, int c) {
    for (;; s++) {
        if (*s == (char)c) return (char*)s;
        if (!*s) return nullptr;
    }
}
__device__ inline char* str_rchr(const char* s, int c) {
    const char* r = nullptr;
    for (;; s++) {
        if (*s == (char)c) r = s;
        if (!*s) return (char*)r;
    }
}
__device__ inline char* str_str(const char* h, const char* n) {
    const size_t m = str_len(n);
    for (; *h; h++)
        if (str_ncmp(h, n, m) == 0) return (char*)h;
    return m == 0 ? (char*)h : nullptr;
}
__device__ inline int mem_cmp(const void* a, const void* b, size_t n) {
    const unsigned char* x = (const unsigned char*)a;
    const unsigned char* y = (const unsigned char*)b;
    for (size_t i = 0; i < n; i++)
        if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1;
    return 0;
}
__device__ inline void* mem_chr(const void* s, int c, size_t n) {
    const unsigned char* p = (const unsigned char*)s;
    for (size_t i = 0; i < n; i++)
        if (p[i] == (unsigned char)c) return (void*)(p + i);
    return nullptr;
}
__device__ inline void* mem_move(void* d, const void* s, size_t n) {
    unsigned char* dp = (unsigned char*)d;
    const unsigned char* sp = (const unsigned char*)s;
    if (dp < sp) {
        for (size_t i = 0; i < n; i++) dp[i] = sp[i];
    } else if (dp > sp) {
        for (size_t i = n; i-- > 0;) dp[i] = sp[i];
    }
    return d;
}
__device__ inline char* str_ndup(const char* s, size_t n) {
    size_t m = 0;
    while (m < n && s[m]) m++;
    char* d = (char*)__gpu_malloc(m + 1);
    if (!d) return nullptr;
    for (size_t i = 0; i < m; i++) d[i] = s[i];
    d[m] = 0;
    return d;
}
__device__ inline char* str_dup(const char* s) { return str_ndup(s, (size_t)-1); }
}  // namespace gpumpi_libc

#define GPUMPI_LIBC_STRING 1
#define strdup(s) ::gpumpi_libc::str_dup(s)
#define strndup(s, n) ::gpumpi_libc::str_ndup((s), (n))
#define memcmp(a, b, n) ::gpumpi_libc::mem_cmp((a), (b), (n))
#define memchr(s, c, n) ::gpumpi_libc::mem_chr((s), (c), (n))
#define memmove(d, s, n) ::gpumpi_libc::mem_move((d), (s), (n))
#define strlen(s) ::gpumpi_libc::str_len(s)
#define strcpy(d, s) ::gpumpi_libc::str_cpy((d), (s))
#define strncpy(d, s, n) ::gpumpi_libc::str_ncpy((d), (s), (n))
#define strcmp(a, b) ::gpumpi_libc::str_cmp((a), (b))
#define strncmp(a, b, n) ::gpumpi_libc::str_ncmp((a), (b), (n))
#define strcat(d, s) ::gpumpi_libc::str_cat((d), (s))
#define strncat(d, s, n) ::gpumpi_libc::str_ncat((d), (s), (n))
#define strchr(s, c) ::gpumpi_libc::str_chr((s), (c))
#define strrchr(s, c) ::gpumpi_libc::str_rchr((s), (c))
#define strstr(h, n) ::gpumpi_libc::str_str((h), (n))
