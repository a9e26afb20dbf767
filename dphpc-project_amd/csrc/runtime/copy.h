// Wavefront-cooperative payload copies.
//
// The reference copies a whole message with device memcpy on one thread
// (gpu_libc/cuda_mpi.cu:168, 323; 0.003 GB/s per transfer on V100,
// benchmarks/results_single_gpu_V100.txt:248).  Here all 64 lanes of the rank's
// wavefront move 16-byte vectors: the destination is aligned to 16 bytes, the source is
// read with unaligned 16-byte global loads, U vectors per lane are loaded before any is
// stored, and large payloads use non-temporal stores so a streaming copy does not evict
// L2.
#pragma once

#include "common.h"

namespace gpumpi {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 u32x4_a4 __attribute__((aligned(4)));
typedef u32x4 u32x4_a1 __attribute__((aligned(1)));

// Global-address-space views: the loads/stores become global_* (not flat_*) instructions.
#define GPUMPI_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ u32x4 gload16(uint64_t a) {
    return *(const GPUMPI_GLOBAL u32x4_a1*)(uintptr_t)a;
}
template <bool NT>
__device__ __forceinline__ void gstore16(uint64_t a, u32x4 v) {
    GPUMPI_GLOBAL u32x4* p = (GPUMPI_GLOBAL u32x4*)(uintptr_t)a;
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Payloads at or above this size bypass L2 on the write side.
constexpr uint64_t kNontemporalBytes = 1u << 20;

// Copy n bytes between global buffers (any byte alignment; the destination side is
// aligned to 16 bytes, the source read with unaligned 16-byte loads, which gfx9 global
// memory supports).  All 64 lanes must call with identical arguments.
template <int U = 4>
__device__ void wave_copy_global(char* d, const char* s, uint64_t n, bool nt) {
    const int l = lane();
    uint64_t head = (uint64_t)((16 - ((uintptr_t)d & 15)) & 15);
    if (head > n) head = n;
    if ((uint64_t)l < head) d[l] = s[l];
    d += head; s += head; n -= head;
    const uint64_t nvec = n >> 4;
    const uint64_t sa = (uint64_t)(uintptr_t)s, dst = (uint64_t)(uintptr_t)d;
    uint64_t base = 0;
    // bulk: whole rounds of kBulk vectors per lane (16 KiB per wave), every load issued
    // before the first store and none behind a bounds check (a lone wavefront's copy rate
    // is set by the bytes it keeps in flight)
    constexpr int kBulk = 8;
    const uint64_t full = nvec - nvec % ((uint64_t)kWave * kBulk);
    for (; base < full; base += (uint64_t)kWave * kBulk) {
        u32x4 v[kBulk];
#pragma unroll
        for (int k = 0; k < kBulk; k++) v[k] = gload16(sa + 16 * (base + (uint64_t)k * kWave + l));
        if (nt) {
#pragma unroll
            for (int k = 0; k < kBulk; k++) gstore16<true>(dst + 16 * (base + (uint64_t)k * kWave + l), v[k]);
        } else {
#pragma unroll
            for (int k = 0; k < kBulk; k++) gstore16<false>(dst + 16 * (base + (uint64_t)k * kWave + l), v[k]);
        }
    }
    for (; base < nvec; base += (uint64_t)kWave * U) {
        u32x4 v[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            uint64_t i = base + (uint64_t)k * kWave + l;
            if (i < nvec) v[k] = gload16(sa + 16 * i);
        }
        if (nt) {
#pragma unroll
            for (int k = 0; k < U; k++) {
                uint64_t i = base + (uint64_t)k * kWave + l;
                if (i < nvec) gstore16<true>(dst + 16 * i, v[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < U; k++) {
                uint64_t i = base + (uint64_t)k * kWave + l;
                if (i < nvec) gstore16<false>(dst + 16 * i, v[k]);
            }
        }
    }
    const uint64_t tail = n & 15;
    if ((uint64_t)l < tail) d[16 * nvec + l] = s[16 * nvec + l];
}

// Byte-strided copy where lane l moves bytes l, l+64, ...  Used when the source is
// this wave's private memory (every lane reads its own private copy, and wave-uniform
// execution keeps all copies identical) or when either side is LDS: generic pointers
// compile to flat loads and stores, which reach LDS as well as global memory.
__device__ __forceinline__ void wave_copy_bytes(char* d, const char* s, uint64_t n) {
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

// LDS staging of large copies.  A lone wave's copy rate is set by the bytes it keeps in
// flight, and holding more of them in VGPRs would raise every rank program's register
// count.  The rank kernel's dynamic LDS therefore holds `s_stage_bytes` per wave (chosen per
// launch, 0 when the launch has none): the wave fills its area with direct global->LDS
// loads (global_load_lds_dwordx4, 16 bytes per lane, no VGPR holds the data in flight),
// waits for them, and moves the area to the destination through ds_read and global
// stores.  Measured with 128 waves copying 64 MiB each at once: 10.8 GB/s per wave through
// registers (8 KiB in flight), 14.4 / 16.7 / 19.1 GB/s through a 16 / 32 / 64 KiB stage.
#define GPUMPI_LDS __attribute__((address_space(3)))
extern __shared__ __attribute__((aligned(16))) char gpumpi_copy_stage[];
__shared__ uint32_t s_stage_bytes;

constexpr uint64_t kStagedMinBytes = 16u << 10;   // smaller copies stay in registers

// Copies the whole stage-sized batches of d[0 .. n) from s (both 16-byte aligned); returns
// the bytes done.  stage_bytes is a multiple of 1 KiB (one 16-byte load per lane).
__device__ __forceinline__ uint64_t wave_copy_staged(char* d, const char* s, uint64_t n, bool nt,
                                                     char GPUMPI_LDS* stage, uint32_t stage_bytes) {
    const int l = lane();
    const int depth = (int)(stage_bytes >> 10);
    const uint64_t full = n - n % stage_bytes;
    for (uint64_t o = 0; o < full; o += stage_bytes) {
#pragma unroll 4
        for (int k = 0; k < depth; k++)
            __builtin_amdgcn_global_load_lds((const GPUMPI_GLOBAL void*)(s + o + 1024 * k + 16 * l),
                                             (GPUMPI_LDS void*)(stage + 1024 * k), 16, 0, 0);
        // the compiler does not order LDS reads after direct loads into LDS: wait here
        // (vmcnt also covers the previous batch's stores, issued before these loads)
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#pragma unroll 4
        for (int k = 0; k < depth; k++) {
            const u32x4 v = *(const GPUMPI_LDS u32x4*)(stage + 1024 * k + 16 * l);
            const uint64_t a = (uint64_t)(uintptr_t)(d + o + 1024 * k + 16 * l);
            if (nt) gstore16<true>(a, v);
            else gstore16<false>(a, v);
        }
    }
    return full;
}

// Copies of more than kWave * 2 bytes.  Out of line on purpose: inlined into the progress
// engine and every collective, the unrolled 16-byte loops raised each rank program's
// register count past 256 per lane (one wave per SIMD, so a 1024-rank grid needed every
// wave slot of the chip); called, they cost one branch per large message.
__device__ __noinline__ void wave_copy_large(char* d, const char* s, uint64_t n) {
    if (is_private(d)) {
        lane_copy_to_private(d, s, n);
    } else if (is_private(s) || is_shared(s) || is_shared(d)) {
        wave_copy_bytes(d, s, n);
    } else {
        const bool nt = n >= kNontemporalBytes;
        const uint32_t sb = s_stage_bytes;
        if (sb != 0 && n >= kStagedMinBytes && (((uintptr_t)d ^ (uintptr_t)s) & 15) == 0) {
            // align both sides to 16 bytes, stage whole batches, the rest through registers
            const uint64_t head = (uint64_t)((16 - ((uintptr_t)d & 15)) & 15);
            if ((uint64_t)lane() < head) d[lane()] = s[lane()];
            char GPUMPI_LDS* stage =
                (char GPUMPI_LDS*)gpumpi_copy_stage + (uint32_t)(threadIdx.x >> 6) * sb;
            const uint64_t done = head + wave_copy_staged(d + head, s + head, n - head, nt, stage, sb);
            d += done;
            s += done;
            n -= done;
        }
        wave_copy_global<4>(d, s, n, nt);
    }
}

// General copy between any generic pointers (global, this wave's private memory, or LDS
// of this workgroup).  All 64 lanes must call with identical arguments.  Up to 128 bytes
// (inline envelopes, small collective slices) every lane moves at most two bytes, both
// loads issued before the stores.
__device__ __forceinline__ void wave_copy(void* dst, const void* src, uint64_t n) {
    if (n == 0 || dst == src) return;
    char* d = (char*)dst;
    const char* s = (const char*)src;
    if (n <= 2 * (uint64_t)kWave && !is_private(d)) {
        const uint64_t l = (uint64_t)lane();
        char a = 0, b = 0;
        if (l < n) a = s[l];
        if (l + kWave < n) b = s[l + kWave];
        if (l < n) d[l] = a;
        if (l + kWave < n) d[l + kWave] = b;
        return;
    }
    wave_copy_large(d, s, n);
}

}  // namespace gpumpi
