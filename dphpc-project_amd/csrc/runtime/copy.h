// Wavefront-cooperative payload copies.
//
// The reference copies a whole message with device memcpy on one thread
// (gpu_libc/cuda_mpi.cu:168, 323; 0.003 GB/s per transfer on V100,
// benchmarks/results_single_gpu_V100.txt:248).  Here all 64 lanes of the rank's
// wavefront move 16-byte vectors: the destination is aligned to 16 bytes, the source
// is read with dword-aligned 16-byte loads (or rebuilt with byte funnel shifts when it
// is not dword aligned), loads are issued U vectors ahead of the stores, and large
// payloads use non-temporal stores so a streaming copy does not evict L2.
#pragma once

#include "common.h"

namespace gpumpi {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 u32x4_a4 __attribute__((aligned(4)));

// Payloads at or above this size bypass L2 on the write side.
constexpr uint64_t kNontemporalBytes = 1u << 20;

__device__ __forceinline__ u32x4 load16_a4(const char* p) {
    return *(const u32x4_a4*)p;
}
__device__ __forceinline__ void store16(char* p, u32x4 v, bool nt) {
    if (nt) __builtin_nontemporal_store(v, (u32x4*)p);
    else *(u32x4*)p = v;
}

// Loads 16 bytes starting at an arbitrary byte address s from dword-aligned words:
// (s & ~3) .. +20.  The extra dword always holds at least one wanted byte, so the
// read never leaves the dwords that the wanted bytes occupy.
__device__ __forceinline__ u32x4 load16_shift(const char* s) {
    const uint32_t* a = (const uint32_t*)((uintptr_t)s & ~(uintptr_t)3);
    const uint32_t sh = (uint32_t)((uintptr_t)s & 3) * 8u;
    u32x4 w = *(const u32x4_a4*)a;
    uint32_t w4 = a[4];
    u32x4 r;
    r.x = __builtin_amdgcn_alignbit(w.y, w.x, sh);
    r.y = __builtin_amdgcn_alignbit(w.z, w.y, sh);
    r.z = __builtin_amdgcn_alignbit(w.w, w.z, sh);
    r.w = __builtin_amdgcn_alignbit(w4, w.w, sh);
    return r;
}

// Copy n bytes from global (or this wave's private) memory to global memory.
// All 64 lanes must call with identical arguments.
template <int U = 4>
__device__ void wave_copy_global(char* d, const char* s, uint64_t n, bool nt) {
    const int l = lane();
    uint64_t head = (uint64_t)((16 - ((uintptr_t)d & 15)) & 15);
    if (head > n) head = n;
    if ((uint64_t)l < head) d[l] = s[l];
    d += head; s += head; n -= head;
    const uint64_t nvec = n >> 4;
    const bool a4 = (((uintptr_t)s) & 3) == 0;
    for (uint64_t base = 0; base < nvec; base += (uint64_t)kWave * U) {
        u32x4 v[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            uint64_t i = base + (uint64_t)k * kWave + l;
            if (i < nvec) v[k] = a4 ? load16_a4(s + 16 * i) : load16_shift(s + 16 * i);
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            uint64_t i = base + (uint64_t)k * kWave + l;
            if (i < nvec) store16(d + 16 * i, v[k], nt);
        }
    }
    const uint64_t tail = n & 15;
    if ((uint64_t)l < tail) d[16 * nvec + l] = s[16 * nvec + l];
}

// Byte-strided copy where lane l moves bytes l, l+64, ...  Used when the source is
// this wave's private memory: every lane reads its own private copy, and wave-uniform
// execution keeps all copies identical.
__device__ __forceinline__ void wave_copy_from_private(char* d, const char* s, uint64_t n) {
    for (uint64_t i = (uint64_t)lane(); i < n; i += kWave) d[i] = s[i];
}

// Every lane writes the whole message into its own private copy of the destination.
__device__ __forceinline__ void lane_copy_to_private(char* d, const char* s, uint64_t n) {
    if ((((uintptr_t)d | (uintptr_t)s | n) & 3) == 0) {
        for (uint64_t i = 0; i < n; i += 4) *(uint32_t*)(d + i) = *(const uint32_t*)(s + i);
    } else {
        for (uint64_t i = 0; i < n; i++) d[i] = s[i];
    }
}

// General copy between any generic pointers.  LDS buffers are not supported (an LDS
// address is only meaningful inside one workgroup).
__device__ __forceinline__ void wave_copy(void* dst, const void* src, uint64_t n) {
    if (n == 0 || dst == src) return;
    char* d = (char*)dst;
    const char* s = (const char*)src;
    if (is_private(d)) {
        lane_copy_to_private(d, s, n);
    } else if (is_private(s)) {
        wave_copy_from_private(d, s, n);
    } else {
        wave_copy_global<4>(d, s, n, n >= kNontemporalBytes);
    }
}

}  // namespace gpumpi
