// Wave-level primitives, scoped atomics, bounded waits and error recording for the
// gpumpi device runtime (gfx950, wave64).
//
// Execution model: one MPI rank is one wavefront.  Rank code runs wave-uniform (all
// 64 lanes execute the program with identical values), so every runtime entry point
// is a wave-collective call: lane 0 performs envelope/flag atomics and the value is
// broadcast; all 64 lanes cooperate on payload copies.  This replaces the reference's
// one-rank-per-thread model, which relied on Volta independent thread scheduling
// (gpu_main/main.cu:150-161) -- lanes of one wavefront cannot spin on each other.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../../include/gpumpi/mpi.h"

namespace gpumpi {

constexpr int kWave = 64;

struct Runtime;

// ---------------------------------------------------------------- lanes / waves
__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int32_t bcast(int32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ uint32_t bcast(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)v);
}
__device__ __forceinline__ uint64_t bcast(uint64_t v) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t bcast(int64_t v) { return (int64_t)bcast((uint64_t)v); }
template <class T>
__device__ __forceinline__ T* bcast(T* p) { return (T*)bcast((uint64_t)(uintptr_t)p); }

__device__ __forceinline__ int32_t readlane(int32_t v, int l) {
    return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ uint32_t readlane(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int32_t)v, l);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int k) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), k);
    return ((uint64_t)hi << 32) | lo;
}

// Wave-wide reductions over 64 lanes (uniform result).
__device__ __forceinline__ int64_t wave_sum(int64_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ int64_t wave_min(int64_t v) {
    for (int o = 32; o > 0; o >>= 1) { int64_t w = __shfl_xor(v, o); v = w < v ? w : v; }
    return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v) {
    for (int o = 32; o > 0; o >>= 1) { int64_t w = __shfl_xor(v, o); v = w > v ? w : v; }
    return v;
}
// Inclusive scan across lanes.
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int l = lane();
    for (int o = 1; o < 64; o <<= 1) {
        int64_t w = __shfl_up(v, o);
        if (l >= o) v += w;
    }
    return v;
}

// ---------------------------------------------------------------- address spaces
__device__ __forceinline__ bool is_private(const void* p) {
    return __builtin_amdgcn_is_private((const __attribute__((address_space(0))) void*)p);
}
__device__ __forceinline__ bool is_shared(const void* p) {
    return __builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void*)p);
}

// ---------------------------------------------------------------- agent-scope atomics
template <class T>
__device__ __forceinline__ T ld_acquire(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T ld_relaxed(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_release(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_relaxed(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T fetch_add(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T fetch_add_relaxed(T* p, T v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T fetch_max(T* p, T v) {
    return __hip_atomic_fetch_max(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ bool cas(T* p, T& expected, T desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_ACQ_REL,
                                                __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
// Orders this wave's payload stores before a following flag store (all lanes call it).
__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
// Waits until every load and store this wave has issued has completed, without writing
// the L2 back.  The workgroup-scope fence keeps the compiler from moving memory operations
// across this point but emits no wait on gfx950 (all waves of a workgroup share one CU),
// so the counter wait is explicit: s_waitcnt with vmcnt, expcnt and lgkmcnt all zero.
// Used before a flag store that either only says "your data has been read" (the reads
// returned, so no later store of another wave can change what they saw) or publishes data
// written with agent-scope atomic stores (sc1: written through to memory, so once they
// have completed they are visible to every XCD and need no L2 write-back).
__device__ __forceinline__ void wait_mem_done() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
}

// ---------------------------------------------------------------- errors
enum ErrorCode : int32_t {
    kErrNone = 0,
    kErrTimeout = 1,        // a bounded wait exceeded the configured limit
    kErrAbort = 2,          // MPI_Abort
    kErrResource = 3,       // request pool / unexpected list / heap exhausted
    kErrArgument = 4,       // invalid rank, type, op, communicator, buffer
    kErrTruncate = 5,       // incoming message larger than the posted receive
    kErrNotSupported = 6,   // API entry the runtime does not provide
    kErrDivergent = 7,      // an MPI call made by a partial wavefront
};

struct ErrorRecord {
    int32_t code;
    int32_t rank;
    int32_t line;
    int32_t detail;
    char msg[112];
};

__device__ Runtime* rt();
__device__ int world_rank();
__device__ void record_error(int32_t code, int line, int32_t detail, const char* msg);
__device__ bool aborted();

#define GPUMPI_FAIL(code, detail, msg) ::gpumpi::record_error((code), __LINE__, (detail), (msg))

// Bounded spin helper: returns false once the runtime was aborted or the wait
// exceeded its limit (the timeout is recorded as an error, so every wait ends).
struct Waiter {
    uint64_t start;
    uint64_t limit;
    uint32_t spins;
    __device__ Waiter();
    __device__ bool keep_going(int line);
};

}  // namespace gpumpi
