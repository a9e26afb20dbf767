// Batched point-to-point step: the matched Send/Recv pairs of one step as two kernels,
// replayable from a hipGraph.
//
// Semantics per pair: those of a matched MPI_Isend / MPI_Irecv pair (gpu_libc/
// cuda_mpi.cu:65-118, gpu_mpi/mpi.cu:617-645) once both have completed -- `bytes` bytes of
// the sender's buffer are in the receiver's.  The list is device memory and is read when
// the step runs, so a kernel earlier on the same stream (or in the same graph) may write
// this step's pairs, sizes included.  Destinations must not overlap each other or any
// source (as MPI's receive buffers must not); zero-byte pairs are allowed.
//
// The reference moves each pair with one thread's memcpy after an envelope handshake.
// Here the step needs no handshake: the pair list is the match.  Work is cut at 8 KiB
// boundaries of each destination address ("chunks"), so every chunk's interior is
// 16-byte aligned on the store side whatever the buffers' alignment:
//   p2p_scan_kernel  one workgroup: exclusive prefix sum of the pairs' chunk counts;
//   p2p_copy_kernel  every wave takes an equal, contiguous range of chunks, finds its first
//                    pair with one 64-ary search of the offsets and walks forward; a chunk is
//                    8 unaligned 16-byte loads per lane, all issued before its stores.
// HBM traffic = 2 x the bytes of the pairs (read once, written once) + 8 B per pair of
// offsets.
#pragma once

#include "common.h"
#include "copy.h"

namespace gpumpi {

struct P2pPair {
    const char* src;
    char* dst;
    uint64_t bytes;
};

constexpr int kP2pChunkShift = 13;                       // 8 KiB chunks of the destination
constexpr uint64_t kP2pChunk = 1ull << kP2pChunkShift;
constexpr int kP2pU = (int)(kP2pChunk / 16 / kWave);     // 16-byte vectors per lane and chunk
constexpr int kP2pScanThreads = 1024;
constexpr int kP2pThreads = 256;

// Branch-free, so that callers can issue the loads of many pairs before using any.
__device__ __forceinline__ uint64_t p2p_chunks(const char* dst, uint64_t bytes) {
    const uint64_t a = (uint64_t)(uintptr_t)dst;
    const uint64_t c = ((a + bytes - 1) >> kP2pChunkShift) - (a >> kP2pChunkShift) + 1;
    return bytes ? c : 0;
}

// chunk_off[i] = chunks of pairs 0 .. i-1, for i = 0 .. n (chunk_off[n] = all chunks).
// Passes of 7168 pairs through LDS (56 KiB): the pairs are read coalesced (pair
// k * 1024 + t by thread t) and their chunk counts parked in LDS; thread t then scans
// counts 7t .. 7t+6, the thread sums are scanned across the workgroup (wave scans, wave
// totals through LDS), and the exclusive offsets go back to LDS and out to chunk_off
// coalesced again.  With one pair
// per thread and pass, 65536 pairs took 64 dependent passes (129 us); with 16 consecutive
// pairs per thread the loads were uncoalesced (99 us).
constexpr int kP2pScanPer = 7;
constexpr int kP2pScanPass = kP2pScanThreads * kP2pScanPer;
static __global__ __launch_bounds__(kP2pScanThreads) void p2p_scan_kernel(const P2pPair* pairs, int64_t n,
                                                                   int64_t* chunk_off) {
    constexpr int kWaves = kP2pScanThreads / kWave;
    __shared__ int64_t s_c[kP2pScanPass];
    __shared__ int64_t s_w[kWaves];
    const int t = threadIdx.x, wv = t >> 6, l = t & 63;
    int64_t carry = 0;
    for (int64_t i0 = 0; i0 < n; i0 += kP2pScanPass) {
        // a clamped index (n >= 1 here) and a masked size: no branch to sink the loads into
#pragma unroll
        for (int k = 0; k < kP2pScanPer; k++) {
            const int64_t j = i0 + (int64_t)k * kP2pScanThreads + t;
            const P2pPair p = pairs[j < n ? j : n - 1];
            s_c[k * kP2pScanThreads + t] = (int64_t)p2p_chunks(p.dst, j < n ? p.bytes : 0);
        }
        __syncthreads();
        int64_t c[kP2pScanPer];
        int64_t sum = 0;
#pragma unroll
        for (int k = 0; k < kP2pScanPer; k++) {
            c[k] = s_c[t * kP2pScanPer + k];
            sum += c[k];
        }
        int64_t incl = sum;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int64_t u = __shfl_up(incl, o);
            if (l >= o) incl += u;
        }
        if (l == kWave - 1) s_w[wv] = incl;
        __syncthreads();
        int64_t before = 0, all = 0;
        for (int u = 0; u < kWaves; u++) {
            before += u < wv ? s_w[u] : 0;
            all += s_w[u];
        }
        int64_t run = carry + before + incl - sum;
#pragma unroll
        for (int k = 0; k < kP2pScanPer; k++) {
            s_c[t * kP2pScanPer + k] = run;
            run += c[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kP2pScanPer; k++) {
            const int64_t j = i0 + (int64_t)k * kP2pScanThreads + t;
            if (j < n) chunk_off[j] = s_c[k * kP2pScanThreads + t];
        }
        carry += all;
        __syncthreads();   // s_c and s_w are rewritten by the next pass
    }
    if (t == 0) chunk_off[n] = carry;
}

// Copies dst bytes [lo, hi) from the matching source bytes (one chunk or less, so at most
// kP2pU vectors per lane).  All 64 lanes call with the same arguments.
__device__ __forceinline__ void p2p_copy_piece(char* dst, const char* src, uint64_t lo, uint64_t hi) {
    const int l = lane();
    const uint64_t d0 = (uint64_t)(uintptr_t)dst;
    const uint64_t a = (lo + 15) & ~(uint64_t)15, b = hi & ~(uint64_t)15;
    if (a >= b) {   // no whole aligned vector: bytes
        for (uint64_t x = lo + (uint64_t)l; x < hi; x += kWave)
            *(char*)(uintptr_t)x = src[x - d0];
        return;
    }
    if ((uint64_t)l < a - lo) *(char*)(uintptr_t)(lo + l) = src[lo + l - d0];
    if ((uint64_t)l < hi - b) *(char*)(uintptr_t)(b + l) = src[b + l - d0];
    const uint64_t nvec = (b - a) >> 4;
    const uint64_t sa = (uint64_t)(uintptr_t)src + (a - d0);
    u32x4 v[kP2pU];
    if (nvec == (uint64_t)kWave * kP2pU) {
#pragma unroll
        for (int u = 0; u < kP2pU; u++) v[u] = gload16(sa + 16 * ((uint64_t)u * kWave + l));
#pragma unroll
        for (int u = 0; u < kP2pU; u++) gstore16<false>(a + 16 * ((uint64_t)u * kWave + l), v[u]);
    } else {
        // every lane loads (vector 0 where it has none: in range, since nvec >= 1), only the
        // stores are conditional -- loads under a condition would leave every vector's
        // register live on both sides of the branch
#pragma unroll
        for (int u = 0; u < kP2pU; u++) {
            const uint64_t k = (uint64_t)u * kWave + l;
            v[u] = gload16(sa + 16 * (k < nvec ? k : 0));
        }
#pragma unroll
        for (int u = 0; u < kP2pU; u++) {
            const uint64_t k = (uint64_t)u * kWave + l;
            if (k < nvec) gstore16<false>(a + 16 * k, v[u]);
        }
    }
}

static __global__ __launch_bounds__(kP2pThreads) void p2p_copy_kernel(const P2pPair* pairs, int64_t n,
                                                               const int64_t* chunk_off) {
    if (n <= 0) return;
    const uint64_t total = (uint64_t)chunk_off[n];
    const uint64_t nw = (uint64_t)gridDim.x * (kP2pThreads / kWave);
    // the wave index through readfirstlane: everything derived from it (the chunk range,
    // the search, the pair fields) is then wave-uniform and kept in SGPRs
    const uint64_t w = (uint64_t)blockIdx.x * (kP2pThreads / kWave) + bcast((uint32_t)(threadIdx.x >> 6));
    const uint64_t c0 = total * w / nw, c1 = total * (w + 1) / nw;
    if (c0 >= c1) return;
    // the pair holding chunk c0: chunk_off[lo] <= c0 < chunk_off[hi] throughout
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const int64_t step = (hi - lo + kWave - 1) / kWave;
        const int64_t probe = lo + (int64_t)lane() * step;
        const bool le = probe < hi && (uint64_t)chunk_off[probe] <= c0;
        const uint64_t m = ballot(le);                 // lane 0 (probe = lo) always holds
        const int k = 63 - __builtin_clzll(m);
        lo += (int64_t)k * step;
        hi = lo + step < hi ? lo + step : hi;
    }
    int64_t i = lo;
    uint64_t next = (uint64_t)chunk_off[i + 1];
    for (uint64_t c = c0; c < c1; c++) {
        while (c >= next) next = (uint64_t)chunk_off[++i + 1];   // skips empty pairs
        const P2pPair p = pairs[i];
        const uint64_t d = (uint64_t)(uintptr_t)p.dst;
        const uint64_t k = c - (uint64_t)chunk_off[i];
        const uint64_t blk = ((d >> kP2pChunkShift) + k) << kP2pChunkShift;
        const uint64_t plo = blk > d ? blk : d;
        const uint64_t phi = blk + kP2pChunk < d + p.bytes ? blk + kP2pChunk : d + p.bytes;
        p2p_copy_piece(p.dst, p.src, plo, phi);
    }
}

}  // namespace gpumpi
