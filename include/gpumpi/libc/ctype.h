// Device ctype.h for converted programs (gpu_libs/gpu_libc/ctype.cuh): the C locale's
// classification, without glibc's host-side tables.
#pragma once
#include <hip/hip_runtime.h>
#include <ctype.h>

namespace gpumpi_libc {
__device__ inline int c_isdigit(int c) { return c >= '0' && c <= '9'; }
__device__ inline int c_isupper(int c) { return c >= 'A' && c <= 'Z'; }
__device__ inline int c_islower(int c) { return c >= 'a' && c <= 'z'; }
__device__ inline int c_isalpha(int c) { return c_isupper(c) || c_islower(c); }
__device__ inline int c_isalnum(int c) { return c_isalpha(c) || c_isdigit(c); }
__device__ inline int c_isspace(int c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
__device__ inline int c_isxdigit(int c) { return c_isdigit(c) || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'); }
__device__ inline int c_isprint(int c) { return c >= 32 && c < 127; }
__device__ inline int c_ispunct(int c) { return c_isprint(c) && !c_isalnum(c) && c != ' '; }
__device__ inline int c_toupper(int c) { return c_islower(c) ? c - 32 : c; }
__device__ inline int c_tolower(int c) { return c_isupper(c) ? c + 32 : c; }
}  // namespace gpumpi_libc

#undef isdigit
#undef isupper
#undef islower
#undef isalpha
#undef isalnum
#undef isspace
#undef isxdigit
#undef isprint
#undef ispunct
#undef toupper
#undef tolower
#define isdigit(c) ::gpumpi_libc::c_isdigit(c)
#define isupper(c) ::gpumpi_libc::c_isupper(c)
#define islower(c) ::gpumpi_libc::c_islower(c)
#define isalpha(c) ::gpumpi_libc::c_isalpha(c)
#define isalnum(c) ::gpumpi_libc::c_isalnum(c)
#define isspace(c) ::gpumpi_libc::c_isspace(c)
#define isxdigit(c) ::gpumpi_libc::c_isxdigit(c)
#define isprint(c) ::gpumpi_libc::c_isprint(c)
#define ispunct(c) ::gpumpi_libc::c_ispunct(c)
#define toupper(c) ::gpumpi_libc::c_toupper(c)
#define tolower(c) ::gpumpi_libc::c_tolower(c)
