// gpumpi device heap API (gfx950).
//
// Same entry points as the reference's device allocators so converted programs link
// unchanged; all variants are served by one lock-free size-class heap:
//   __gpu_malloc / __gpu_free / __gpu_calloc ........ gpu_libs/gpu_libc/stdlib.cuh:33-35
//   __gpu_malloc_coalesce / _v1 .. _v6 (+ frees) .... gpu_libs/gpu_libc/stdlib.cuh:37-65
//   __gpu_init_malloc / __gpu_clean_malloc (+ _v6) .. gpu_libs/gpu_libc/stdlib.cuh:49-50, 66-67
//   malloc / calloc / free remapping ................ gpu_libs/gpu_libc/stdlib.h.cuh:23-72
// The `coalesced` flag is accepted and ignored: a rank is a whole wavefront, so one call
// is already one request.  Blocks of 16 bytes or more are 16-byte aligned.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

__device__ void* __gpu_malloc(size_t size);
__device__ void* __gpu_calloc(size_t nmemb, size_t size);
__device__ void* __gpu_realloc(void* ptr, size_t size);
__device__ void __gpu_free(void* memptr);

__device__ void* __gpu_malloc_coalesce(size_t size, bool coalesced = false);
__device__ void __gpu_free_coalesce(void* memptr);
__device__ void* __gpu_malloc_v1(size_t size, bool coalesced = false);
__device__ void __gpu_free_v1(void* memptr);
__device__ void* __gpu_malloc_v2(size_t size, bool coalesced = false);
__device__ void __gpu_free_v2(void* memptr);
__device__ void* __gpu_malloc_v3(size_t size, bool coalesced = false);
__device__ void __gpu_free_v3(void* memptr);
__device__ void* __gpu_malloc_v4(size_t size, bool coalesced = false);
__device__ void __gpu_free_v4(void* memptr);
__device__ void* __gpu_malloc_v5(size_t size, bool coalesced = false);
__device__ void __gpu_free_v5(void* memptr);
__device__ void* __gpu_malloc_v6(size_t size, bool coalesced = false);
__device__ void __gpu_free_v6(void* memptr);
__device__ void __gpu_init_malloc();
__device__ void __gpu_clean_malloc();
__device__ void __gpu_init_malloc_v6();
__device__ void __gpu_clean_malloc_v6();

// Heap statistics of the running rank grid (bytes of live blocks, class-rounded).  A
// running rank adds its own operations in batches (every 64 operations or 1 MiB net), so
// while ranks run the value lags by at most that much per rank; it is exact once they end.
__device__ long long __gpu_heap_in_use();
__device__ long long __gpu_heap_capacity();
