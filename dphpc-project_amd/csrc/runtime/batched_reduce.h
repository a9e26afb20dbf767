// Batched MPI_Reduce / MPI_Allreduce of one step over P ranks: one kernel, replayed from
// a hipGraph per step.
//
// Semantics: gpu_mpi/mpi.cu:148-235 and 276-282 -- the root (every rank for Allreduce)
// receives, element by element, in_0 folded with in_1 .. in_{P-1} in rank order
// (inout = op(in_r, inout), operators.cu:210-220); the reference's generic path, so
// FP64 results are bit-exact with it.  (Its FP64 SUM on COMM_WORLD used atomicAdd and
// was not deterministic, mpi.cu:169; here every type folds in rank order.)
//
// The reference runs this as P-1 point-to-point messages into a root buffer of P*N
// elements and a serial fold on the root (then a Bcast for Allreduce).  Here the step
// is one grid over the element range: each lane owns U 16-byte columns, folds them over
// the P rank buffers in rank order with D ranks' loads in flight (fold.h), and stores the
// result to the root's buffer or to every rank's.  HBM traffic = the algorithmic bytes:
// P * N * s read, N * s (Reduce) or P * N * s (Allreduce) written.
#pragma once

#include "fold.h"

namespace gpumpi {

constexpr int kReduceThreads = 256;

struct ReduceDev {
    int32_t P;
    int32_t root;            // -1: every rank receives (Allreduce)
    int64_t count;           // elements per rank
    int32_t nt;              // non-temporal stores
    int32_t pad;
    const char* sendbufs;    // [P] device pointers
    const char* recvbufs;    // [P] device pointers
};

// Launch shapes: each wave folds U KiB of every rank buffer per visit (U 16-byte columns
// per lane, 1 KiB apart) with D ranks' loads in flight, U * D columns in all.  Measured on
// MI355X, 1024 ranks (2 x P x N x s bytes per step): at 1 MiB per rank U = 2, D = 8 reached
// 66 % of HBM peak and U = 1 only 40-42 % at any depth (a wave visits every rank buffer
// once per KiB, and the visits, not the bytes in flight, bound it); at 8 MiB per rank
// U = 8, D = 4 reached 73 %, but at 1 MiB it leaves too few waves (128 on 256 CUs).
// The plan takes the wide shape when a rank buffer gives at least four of its tiles per CU.
constexpr int kReduceNarrowU = 2, kReduceNarrowD = 8;
constexpr int kReduceWideU = 8, kReduceWideD = 4;

template <class T, int OP, int U, int D>
__global__ void __launch_bounds__(kReduceThreads) batched_reduce_kernel(ReduceDev p) {
    const uint64_t nvec = (uint64_t)p.count * sizeof(T) / 16;
    const PtrTable src{p.sendbufs, 8}, dst{p.recvbufs, 8};
    const uint64_t one = p.root >= 0 ? *(const uint64_t*)(p.recvbufs + 8 * (int64_t)p.root) : 0;
    constexpr uint64_t kTile = (uint64_t)kWave * U;  // 16-byte columns per wave and visit
    const uint64_t ntiles = (nvec + kTile - 1) / kTile;
    const uint64_t wpb = kReduceThreads / kWave;
    const uint64_t nw = (uint64_t)gridDim.x * wpb;
    for (uint64_t t = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6); t < ntiles; t += nw) {
        const uint64_t v0 = t * kTile + (uint64_t)lane();
        uint32_t valid = 0;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (v0 + (uint64_t)kWave * u < nvec) valid |= 1u << u;
        u32x4 acc[U];
        fold_columns_unmasked<T, OP, U, D>(src, p.P, 16 * v0, valid, acc);
        if (p.nt) fold_store_cols<true, U>(dst, p.P, one, 16 * v0, valid, acc);
        else fold_store_cols<false, U>(dst, p.P, one, 16 * v0, valid, acc);
    }
    // elements past the last whole 16-byte column
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x < kWave)
        fold_scalar_tail<T, OP>(src, p.P, dst, p.P, one, (int64_t)(nvec * 16 / sizeof(T)), p.count);
}

// Integer operators are exact in any order, so below the wide shape their fold splits the
// ranks over the 4 waves of a workgroup: the waves fold ranks [qP/4, (q+1)P/4) of the same
// tile, their partials meet in LDS, and each wave stores the total to a quarter of the
// destinations.  A tile is then read by 4 waves at once (4x the waves of the narrow shape
// at the same 2 KiB per rank visit).  Measured at 1024 ranks x 1 MiB: int32 SUM 70.6 % of
// HBM peak against 64.8 % with the narrow shape on the same box, int64 70.6 % against 56.2 %;
// at 8 MiB per rank the wide shape stays ahead (71.3 % against 58 %).  Floating-point types
// keep the single rank-order fold.  P >= 4.
constexpr int kReduceSplitU = 2, kReduceSplitD = 4;
constexpr int kReduceSplitWaves = kReduceThreads / kWave;

template <class T, int OP, int U, int D>
__global__ void __launch_bounds__(kReduceThreads) batched_reduce_split_kernel(ReduceDev p) {
    constexpr int S = kReduceSplitWaves;
    __shared__ u32x4 part[S][U][kWave];
    const uint64_t nvec = (uint64_t)p.count * sizeof(T) / 16;
    constexpr uint64_t kTile = (uint64_t)kWave * U;
    const uint64_t ntiles = (nvec + kTile - 1) / kTile;
    const int q = (int)bcast((uint32_t)(threadIdx.x >> 6));   // wave-uniform, kept in SGPRs
    const int r0 = (int)((int64_t)p.P * q / S), r1 = (int)((int64_t)p.P * (q + 1) / S);
    const PtrTable src{p.sendbufs + 8 * (int64_t)r0, 8};
    const PtrTable dsub{p.recvbufs + 8 * (int64_t)r0, 8};
    const uint64_t one = p.root >= 0 ? *(const uint64_t*)(p.recvbufs + 8 * (int64_t)p.root) : 0;
    const int l = lane();
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint64_t v0 = t * kTile + (uint64_t)l;
        uint32_t valid = 0;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (v0 + (uint64_t)kWave * u < nvec) valid |= 1u << u;
        u32x4 acc[U];
        fold_columns_unmasked<T, OP, U, D>(src, r1 - r0, 16 * v0, valid, acc);
#pragma unroll
        for (int u = 0; u < U; u++) part[q][u][l] = acc[u];
        __syncthreads();
        u32x4 tot[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            tot[u] = part[0][u][l];
#pragma unroll
            for (int s = 1; s < S; s++) tot[u] = fold_vec_op<T, OP>(part[s][u][l], tot[u]);
        }
        __syncthreads();   // part is rewritten by the next tile
        if (one) {
            if (q == 0) {
                if (p.nt) fold_store_cols<true, U>(dsub, 0, one, 16 * v0, valid, tot);
                else fold_store_cols<false, U>(dsub, 0, one, 16 * v0, valid, tot);
            }
        } else {
            if (p.nt) fold_store_cols<true, U>(dsub, r1 - r0, 0, 16 * v0, valid, tot);
            else fold_store_cols<false, U>(dsub, r1 - r0, 0, 16 * v0, valid, tot);
        }
    }
    // elements past the last whole 16-byte column
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x < kWave) {
        const PtrTable all{p.sendbufs, 8}, dall{p.recvbufs, 8};
        fold_scalar_tail<T, OP>(all, p.P, dall, p.P, one, (int64_t)(nvec * 16 / sizeof(T)), p.count);
    }
}

// Workgroups of the split kernel for `bytes` per rank: one per tile of 64 x kReduceSplitU
// columns.
__host__ inline uint64_t batched_reduce_split_grid(uint64_t bytes) {
    const uint64_t tile = 16ull * kWave * kReduceSplitU;
    const uint64_t g = (bytes / 16 * 16 + tile - 1) / tile;
    return g < 1 ? 1 : g;
}

// Workgroups for `bytes` per rank: one wave per U KiB tile.
__host__ inline uint64_t batched_reduce_grid(uint64_t bytes, int U) {
    const uint64_t tiles = (bytes / 16 + (uint64_t)kWave * U - 1) / ((uint64_t)kWave * U);
    const uint64_t wpb = kReduceThreads / kWave;
    const uint64_t g = (tiles + wpb - 1) / wpb;
    return g < 1 ? 1 : g;
}

// Kernel for a (datatype, built-in operator) pair of the batched path and a shape (0 narrow,
// 1 wide), or nullptr.
template <int U, int D>
__host__ inline const void* batched_reduce_kernel_shape(int dt, int op) {
#define GPUMPI_RED_OPS(T)                                                        \
    switch (op) {                                                                \
    case kOpMax: return (const void*)&batched_reduce_kernel<T, kOpMax, U, D>;    \
    case kOpMin: return (const void*)&batched_reduce_kernel<T, kOpMin, U, D>;    \
    case kOpSum: return (const void*)&batched_reduce_kernel<T, kOpSum, U, D>;    \
    case kOpProd: return (const void*)&batched_reduce_kernel<T, kOpProd, U, D>;  \
    default: return nullptr;                                                     \
    }
    switch (dt) {
    case MPI_INT: case MPI_INT32_T: GPUMPI_RED_OPS(int32_t)
    case MPI_UNSIGNED: case MPI_UINT32_T: GPUMPI_RED_OPS(uint32_t)
    case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG: case MPI_INT64_T: GPUMPI_RED_OPS(int64_t)
    case MPI_UNSIGNED_LONG: case MPI_UNSIGNED_LONG_LONG: case MPI_UINT64_T: GPUMPI_RED_OPS(uint64_t)
    case MPI_FLOAT: GPUMPI_RED_OPS(float)
    case MPI_DOUBLE: case MPI_LONG_DOUBLE: GPUMPI_RED_OPS(double)
    default: return nullptr;
    }
#undef GPUMPI_RED_OPS
}

__host__ inline const void* batched_reduce_kernel_for(int dt, int op, bool wide) {
    return wide ? batched_reduce_kernel_shape<kReduceWideU, kReduceWideD>(dt, op)
                : batched_reduce_kernel_shape<kReduceNarrowU, kReduceNarrowD>(dt, op);
}

// The split kernel for an integer (datatype, built-in operator) pair, or nullptr.
__host__ inline const void* batched_reduce_split_for(int dt, int op) {
#define GPUMPI_RED_SPLIT(T)                                                                          \
    switch (op) {                                                                                    \
    case kOpMax: return (const void*)&batched_reduce_split_kernel<T, kOpMax, kReduceSplitU, kReduceSplitD>;   \
    case kOpMin: return (const void*)&batched_reduce_split_kernel<T, kOpMin, kReduceSplitU, kReduceSplitD>;   \
    case kOpSum: return (const void*)&batched_reduce_split_kernel<T, kOpSum, kReduceSplitU, kReduceSplitD>;   \
    case kOpProd: return (const void*)&batched_reduce_split_kernel<T, kOpProd, kReduceSplitU, kReduceSplitD>; \
    default: return nullptr;                                                                         \
    }
    switch (dt) {
    case MPI_INT: case MPI_INT32_T: GPUMPI_RED_SPLIT(int32_t)
    case MPI_UNSIGNED: case MPI_UINT32_T: GPUMPI_RED_SPLIT(uint32_t)
    case MPI_LONG: case MPI_LONG_LONG_INT: case MPI_LONG_LONG: case MPI_INT64_T: GPUMPI_RED_SPLIT(int64_t)
    case MPI_UNSIGNED_LONG: case MPI_UNSIGNED_LONG_LONG: case MPI_UINT64_T: GPUMPI_RED_SPLIT(uint64_t)
    default: return nullptr;
    }
#undef GPUMPI_RED_SPLIT
}

}  // namespace gpumpi
