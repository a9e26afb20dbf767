// Element-wise reduction of P rank buffers in the reference's fold order
// (acc = in_0; acc = op(in_r, acc) for r = 1..P-1, gpu_mpi/mpi.cu:209-225, operator bodies
// operators.cu:73-151), in 16-byte columns.
//
// Every element keeps its own left-to-right fold, so results are bit-exact with the
// reference's generic path for every type, FP64 included; the parallelism is over
// elements only.  A lane's columns are loaded from the P sources in rank order with D
// ranks' loads in flight (a ring of 16-byte registers, refilled D ranks ahead), which
// keeps a wave streaming from HBM instead of waiting one load at a time.  Source and destination addresses come from tables of per-rank
// pointers (a stride apart); a wave loads 64 of them at once into its lanes and reads
// each back with readlane.
#pragma once

#include "ops.h"
#include "copy.h"

namespace gpumpi {

// Per-rank pointer tables: entry r at base + r * stride (a uint64_t address).
struct PtrTable {
    const char* base;
    int64_t stride;
    // Lanes past P read entry P - 1 (callers read lanes below P - r0 only).  The load is
    // unconditional, in the global address space, and its value is not touched before
    // the caller's readlane: the compiler counts a load in flight only when it is neither a
    // flat load (whose completion is not ordered with the other counters) nor issued under
    // a lane mask that may skip it, and it waits where the value is first used.  Otherwise
    // each readlane of the result waits for every load and store in flight
    // (s_waitcnt vmcnt(0)) -- which drained the fold's ring of loads at every rank, and a
    // prefetch of the next 64 entries waited as soon as it was issued.
    __device__ __forceinline__ uint64_t lane_chunk(int r0, int P) const {
        const int r = r0 + lane();
        return *(const GPUMPI_GLOBAL uint64_t*)(base + (int64_t)(r < P ? r : P - 1) * stride);
    }
};

// acc = op(x, acc) on the V = 16 / sizeof(T) elements of a 16-byte vector.
template <class T, int OP>
__device__ __forceinline__ u32x4 fold_vec_op(u32x4 x, u32x4 acc) {
    constexpr int V = 16 / (int)sizeof(T);
    T a[V], b[V];
    __builtin_memcpy(a, &x, 16);
    __builtin_memcpy(b, &acc, 16);
#pragma unroll
    for (int i = 0; i < V; i++) b[i] = combine<T, OP>(a[i], b[i]);
    u32x4 out;
    __builtin_memcpy(&out, b, 16);
    return out;
}

// Folds U 16-byte columns per lane -- at byte offsets off + 1024 * u, u < U, of every
// source (ranks 0..P-1 in order) -- with D ranks' loads in flight, into acc[u].  Lanes
// load column u only where bit u of `valid` is set.  A wave thus reads U KiB of a rank's
// buffer per visit.  Wave-uniform P; all lanes must call.
template <class T, int OP, int U, int D>
__device__ __forceinline__ void fold_columns(const PtrTable& src, int P, uint64_t off, uint32_t valid,
                                             u32x4 (&acc)[U]) {
    static_assert(D >= 1 && D <= kWave && kWave % D == 0, "D must divide the wave size");
    u32x4 ring[D][U];
    uint64_t pc = src.lane_chunk(0, P);
    uint64_t pn = P > kWave ? src.lane_chunk(kWave, P) : 0;
#pragma unroll
    for (int k = 0; k < D; k++) {
        const uint64_t b = k < P ? readlane64(pc, k) + off : 0;
#pragma unroll
        for (int u = 0; u < U; u++) {
            ring[k][u] = u32x4{0u, 0u, 0u, 0u};
            if (k < P && (valid >> u & 1u)) ring[k][u] = gload16(b + 1024u * u);
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = u32x4{0u, 0u, 0u, 0u};
    for (int r0 = 0; r0 < P; r0 += kWave) {
        for (int j0 = 0; j0 < kWave; j0 += D) {
            if (r0 + j0 >= P) break;
#pragma unroll
            for (int k = 0; k < D; k++) {
                const int r = r0 + j0 + k;
                if (r < P) {
                    u32x4 x[U];
#pragma unroll
                    for (int u = 0; u < U; u++) x[u] = ring[k][u];
                    if (r + D < P) {
                        const int jn = j0 + k + D;  // chunk-relative index of rank r + D
                        const uint64_t b =
                            (jn < kWave ? readlane64(pc, jn) : readlane64(pn, jn - kWave)) + off;
#pragma unroll
                        for (int u = 0; u < U; u++)
                            if (valid >> u & 1u) ring[k][u] = gload16(b + 1024u * u);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) acc[u] = r == 0 ? x[u] : fold_vec_op<T, OP>(x[u], acc[u]);
                }
            }
        }
        pc = pn;
        pn = r0 + 2 * kWave < P ? src.lane_chunk(r0 + 2 * kWave, P) : 0;
    }
}

// fold_columns for the batched kernels, where registers are plentiful: the same fold with
// no load under a branch.  A column a lane does not fold loads lane 0's first column of the
// same rank instead, ranks past P - 1 load rank P - 1, and the fold takes a value or keeps
// acc by selection.  The compiler can then count the loads in flight and wait for the
// oldest only (s_waitcnt vmcnt(N)), so D ranks' loads stay in flight; in fold_columns the
// masked loads make it wait for all of them (vmcnt(0)) before each rank's readlane.
// Wave-uniform P >= 1; lane 0's column 0 must be valid (the callers' tiles start at a valid
// column).  The rank grid keeps fold_columns: this one's extra registers, through the
// interprocedural allocation, took rank programs past 256 per lane.
template <class T, int OP, int U, int D>
__device__ __forceinline__ void fold_columns_unmasked(const PtrTable& src, int P, uint64_t off,
                                                      uint32_t valid, u32x4 (&acc)[U]) {
    static_assert(D >= 1 && D <= kWave && kWave % D == 0, "D must divide the wave size");
    const uint64_t off0 = bcast(off);   // lane 0's first column
    uint64_t co[U];
#pragma unroll
    for (int u = 0; u < U; u++) co[u] = (valid >> u & 1u) ? off + 1024u * u : off0;
    u32x4 ring[D][U];
    uint64_t pc = src.lane_chunk(0, P);
    uint64_t pn = src.lane_chunk(kWave, P);
#pragma unroll
    for (int k = 0; k < D; k++) {
        const uint64_t b = readlane64(pc, k < P ? k : P - 1);
#pragma unroll
        for (int u = 0; u < U; u++) ring[k][u] = gload16(b + co[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc[u] = u32x4{0u, 0u, 0u, 0u};
    for (int r0 = 0; r0 < P; r0 += kWave) {
        for (int j0 = 0; j0 < kWave; j0 += D) {
            if (r0 + j0 >= P) break;
#pragma unroll
            for (int k = 0; k < D; k++) {
                const int r = r0 + j0 + k;
                u32x4 x[U];
#pragma unroll
                for (int u = 0; u < U; u++) x[u] = ring[k][u];
                // refill with rank r + D (rank P - 1 past the end): its address is in this
                // chunk's pointers or the next chunk's
                const int rn = r + D < P ? r + D : P - 1;
                const int jn = rn - r0;
                const uint64_t bc = readlane64(pc, jn < kWave ? jn : 0);
                const uint64_t bn = readlane64(pn, jn < kWave ? 0 : jn - kWave);
                const uint64_t b = jn < kWave ? bc : bn;
#pragma unroll
                for (int u = 0; u < U; u++) ring[k][u] = gload16(b + co[u]);
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const u32x4 f = r == 0 ? x[u] : fold_vec_op<T, OP>(x[u], acc[u]);
                    acc[u] = r < P ? f : acc[u];
                }
            }
        }
        pc = pn;
        pn = src.lane_chunk(r0 + 2 * kWave, P);
    }
}

// The rank grid's slices: 4 columns per lane (4 KiB of each rank buffer per visit), 4
// ranks in flight.  Measured on 1024 ranks at 2^22 elements: Reduce 5578 / 5790 GB/s
// (MPI_INT / MPI_DOUBLE) against 3619 / 3699 with 2 x 8, which holds as many loads in flight
// but visits every rank buffer twice as often; 8 x 2 takes the rank programs past 256
// registers per lane.
constexpr int kFoldU = 4;
constexpr int kFoldD = 4;
// Slices of at most one tile of 2 columns per lane (Reduce at most 2 KiB per rank and
// slice) fold with fold_columns_unmasked, 2 columns x 4 ranks: latency-bound there, they
// gain from the ring of loads staying in flight.  Same-process A/B on 1024 ranks at 2^16
// elements: Reduce 427 -> 211 us (MPI_INT), 532 -> 248 us (MPI_DOUBLE), Allreduce 602 -> 320
// / 708 -> 358 us; at 2^20 and 2^22 the 4 x 4 fold above was 4-15 % faster.
constexpr int kFoldSmallU = 2;
constexpr int kFoldSmallD = 4;

// Stores a lane's folded column to every destination of the table (ndst of them), or to
// one destination address when `one` is non-zero.
template <bool NT>
__device__ __forceinline__ void fold_store(const PtrTable& dst, int ndst, uint64_t one, uint64_t off,
                                           bool valid, u32x4 v) {
    if (one) {
        if (valid) gstore16<NT>(one + off, v);
        return;
    }
    for (int d0 = 0; d0 < ndst; d0 += kWave) {
        const uint64_t pd = dst.lane_chunk(d0, ndst);
        const int n = ndst - d0 < kWave ? ndst - d0 : kWave;
        for (int j = 0; j < n; j++)
            if (valid) gstore16<NT>(readlane64(pd, j) + off, v);
    }
}

// Stores a lane's U folded columns (offsets off + 1024 * u where bit u of `valid` is set)
// to every destination of the table, or to the one address `one` when non-zero.
template <bool NT, int U>
__device__ __forceinline__ void fold_store_cols(const PtrTable& dst, int ndst, uint64_t one,
                                                uint64_t off, uint32_t valid, const u32x4 (&v)[U]) {
    if (one) {
#pragma unroll
        for (int u = 0; u < U; u++)
            if (valid >> u & 1u) gstore16<NT>(one + off + 1024u * u, v[u]);
        return;
    }
    for (int d0 = 0; d0 < ndst; d0 += kWave) {
        const uint64_t pd = dst.lane_chunk(d0, ndst);
        const int n = ndst - d0 < kWave ? ndst - d0 : kWave;
        for (int j = 0; j < n; j++) {
            const uint64_t b = readlane64(pd, j) + off;
#pragma unroll
            for (int u = 0; u < U; u++)
                if (valid >> u & 1u) gstore16<NT>(b + 1024u * u, v[u]);
        }
    }
}

// Elements [e0, e1) of the P sources, folded one element per lane (the part of a range
// that does not fill a 16-byte column).
template <class T, int OP>
__device__ __forceinline__ void fold_scalar_tail(const PtrTable& src, int P, const PtrTable& dst,
                                                 int ndst, uint64_t one, int64_t e0, int64_t e1) {
    for (int64_t e = e0 + lane(); e < e1; e += kWave) {
        T acc = *(const T*)(*(const uint64_t*)src.base + e * (int64_t)sizeof(T));
        for (int r = 1; r < P; r++) {
            const uint64_t p = *(const uint64_t*)(src.base + (int64_t)r * src.stride);
            acc = combine<T, OP>(*(const T*)(p + e * (int64_t)sizeof(T)), acc);
        }
        if (one) {
            *(T*)(one + e * (int64_t)sizeof(T)) = acc;
        } else {
            for (int d = 0; d < ndst; d++) {
                const uint64_t p = *(const uint64_t*)(dst.base + (int64_t)d * dst.stride);
                *(T*)(p + e * (int64_t)sizeof(T)) = acc;
            }
        }
    }
}

}  // namespace gpumpi
