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
// is one grid over the element range: each lane owns a 16-byte column, folds it over
// the P rank buffers in rank order with 16 loads in flight (fold.h), and stores the
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

template <class T, int OP>
__global__ void __launch_bounds__(kReduceThreads) batched_reduce_kernel(ReduceDev p) {
    const uint64_t nvec = (uint64_t)p.count * sizeof(T) / 16;
    const PtrTable src{p.sendbufs, 8}, dst{p.recvbufs, 8};
    const uint64_t one = p.root >= 0 ? *(const uint64_t*)(p.recvbufs + 8 * (int64_t)p.root) : 0;
    const uint64_t stride = (uint64_t)gridDim.x * kReduceThreads;
    for (uint64_t v0 = (uint64_t)blockIdx.x * kReduceThreads + (threadIdx.x & ~(kWave - 1));
         v0 < nvec; v0 += stride) {  // wave-uniform
        const uint64_t v = v0 + lane();
        const bool valid = v < nvec;
        const u32x4 acc = fold_column<T, OP>(src, p.P, 16 * v, valid);
        if (p.nt) fold_store<true>(dst, p.P, one, 16 * v, valid, acc);
        else fold_store<false>(dst, p.P, one, 16 * v, valid, acc);
    }
    // elements past the last whole 16-byte column
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x < kWave)
        fold_scalar_tail<T, OP>(src, p.P, dst, p.P, one, (int64_t)(nvec * 16 / sizeof(T)), p.count);
}

// Kernel for a (datatype, built-in operator) pair of the batched path, or nullptr.
__host__ inline const void* batched_reduce_kernel_for(int dt, int op) {
#define GPUMPI_RED_OPS(T)                                                            \
    switch (op) {                                                                    \
    case kOpMax: return (const void*)&batched_reduce_kernel<T, kOpMax>;              \
    case kOpMin: return (const void*)&batched_reduce_kernel<T, kOpMin>;              \
    case kOpSum: return (const void*)&batched_reduce_kernel<T, kOpSum>;              \
    case kOpProd: return (const void*)&batched_reduce_kernel<T, kOpProd>;            \
    default: return nullptr;                                                         \
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

}  // namespace gpumpi
