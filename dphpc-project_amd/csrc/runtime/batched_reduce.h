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
        fold_columns<T, OP, U, D>(src, p.P, 16 * v0, valid, acc);
        if (p.nt) fold_store_cols<true, U>(dst, p.P, one, 16 * v0, valid, acc);
        else fold_store_cols<false, U>(dst, p.P, one, 16 * v0, valid, acc);
    }
    // elements past the last whole 16-byte column
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x < kWave)
        fold_scalar_tail<T, OP>(src, p.P, dst, p.P, one, (int64_t)(nvec * 16 / sizeof(T)), p.count);
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

}  // namespace gpumpi
