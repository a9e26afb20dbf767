// Rank programs that test runtime behaviour the reference's own unit tests do not reach:
//   t_op_table     MPI_Reduce and MPI_Allreduce over a host-supplied list of (datatype,
//                  operator) cases -- every legal pair of the operator table
//                  (gpu_mpi/operators.cu:73-200) and two user operators made with
//                  MPI_Op_create (operators.cu:246-250) that are not commutative, so the
//                  fold order inout = op(in_r, inout), r = 1..P-1 (mpi.cu:209-225) shows;
//   t_lds_buffers  send, receive and collective buffers in LDS (point-to-point through
//                  the bounce and inline paths, collectives through staging).
#pragma once

#include "../runtime/mpi_api.h"
#include "workloads.h"

namespace gpumpi {
namespace progs {

struct OpCase {
    int32_t dtype;  // MPI_Datatype
    int32_t op;     // built-in operator index (MPI_MAX = 0 .. MPI_MINLOC = 11), or -1 / -2:
                    // the user operators below (int64 / double)
};

struct OpTableArgs {
    const OpCase* cases;  // [ncases]
    int32_t ncases;
    int32_t count;        // elements per rank and case
    int32_t root;
    int32_t stride;       // bytes per (case, rank) block of in / out_all, and per case of out_root
    const char* in;       // [ncases][P][stride]
    char* out_all;        // [ncases][P][stride]: MPI_Allreduce results
    char* out_root;       // [ncases][stride]: MPI_Reduce results at root
    int32_t* status;      // [ncases][P]: return codes (Reduce | Allreduce << 8)
};

// User operators: not commutative, so a fold in any other order than the reference's
// gives a different result.  int64: inout = 2 * inout + in (mod 2^64); double:
// inout = 0.5 * inout + in.
__device__ void user_op_int64(void* in, void* inout, int* len, MPI_Datatype* dt) {
    (void)dt;
    const uint64_t* a = (const uint64_t*)in;
    uint64_t* b = (uint64_t*)inout;
    for (int i = 0; i < *len; i++) b[i] = 2u * b[i] + a[i];
}
__device__ void user_op_double(void* in, void* inout, int* len, MPI_Datatype* dt) {
    (void)dt;
    const double* a = (const double*)in;
    double* b = (double*)inout;
    for (int i = 0; i < *len; i++) b[i] = 0.5 * b[i] + a[i];
}

__device__ int t_op_table(int, char**) {
    const OpTableArgs& a = args<OpTableArgs>();
    MPI_Init(nullptr, nullptr);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    MPI_Op uop_i = MPI_OP_NULL, uop_d = MPI_OP_NULL;
    MPI_Op_create(user_op_int64, 0, &uop_i);
    MPI_Op_create(user_op_double, 0, &uop_d);
    for (int c = 0; c < a.ncases; c++) {
        const OpCase oc = a.cases[c];
        const MPI_Op op = oc.op == -1 ? uop_i : oc.op == -2 ? uop_d : &gpumpi_builtin_ops[oc.op];
        const MPI_Datatype dt = (MPI_Datatype)oc.dtype;
        const char* in = a.in + ((int64_t)c * P + rank) * a.stride;
        char* out = a.out_all + ((int64_t)c * P + rank) * a.stride;
        char* root_out = rank == a.root ? a.out_root + (int64_t)c * a.stride : nullptr;
        const int r1 = MPI_Reduce(in, root_out, a.count, dt, op, a.root, MPI_COMM_WORLD);
        const int r2 = MPI_Allreduce(in, out, a.count, dt, op, MPI_COMM_WORLD);
        if (lane() == 0) a.status[(int64_t)c * P + rank] = (r1 & 0xFF) | ((r2 & 0xFF) << 8);
        if (aborted()) break;
    }
    MPI_Op_free(&uop_i);
    MPI_Op_free(&uop_d);
    MPI_Finalize();
    return 0;
}

struct LdsArgs {
    int32_t* ok;  // [P]: 1 when every check of the rank passed
};

__device__ __forceinline__ void lds_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ int t_lds_buffers(int, char**) {
    constexpr int kBig = 300;   // above the inline limit and the two-bytes-per-lane copy
    constexpr int kSmall = 40;  // travels inline in the envelope
    constexpr int kRed = 100;
    __shared__ uint8_t s_send[4][kBig];
    __shared__ uint8_t s_recv[4][kBig];
    __shared__ uint8_t s_small[4][2][kSmall];
    __shared__ int32_t s_red[4][2][kRed];
    const int w = threadIdx.x >> 6;
    const LdsArgs& a = args<LdsArgs>();
    MPI_Init(nullptr, nullptr);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const int next = (rank + 1) % P, prev = (rank + P - 1) % P;
    for (int i = lane(); i < kBig; i += kWave) {
        s_send[w][i] = (uint8_t)(rank * 7 + i);
        s_recv[w][i] = 0;
    }
    for (int i = lane(); i < kSmall; i += kWave) {
        s_small[w][0][i] = (uint8_t)(rank * 3 + 2 * i);
        s_small[w][1][i] = 0;
    }
    for (int i = lane(); i < kRed; i += kWave) {
        s_red[w][0][i] = rank * 1000 + i;
        s_red[w][1][i] = 0;
    }
    lds_wave_sync();
    // ring: LDS -> LDS, one bounced message and one inline message
    MPI_Request rq[2];
    MPI_Irecv(s_recv[w], kBig, MPI_BYTE, prev, 5, MPI_COMM_WORLD, &rq[0]);
    MPI_Irecv(s_small[w][1], kSmall, MPI_BYTE, prev, 6, MPI_COMM_WORLD, &rq[1]);
    MPI_Send(s_send[w], kBig, MPI_BYTE, next, 5, MPI_COMM_WORLD);
    MPI_Send(s_small[w][0], kSmall, MPI_BYTE, next, 6, MPI_COMM_WORLD);
    MPI_Waitall(2, rq, MPI_STATUSES_IGNORE);
    lds_wave_sync();
    bool good = true;
    for (int i = lane(); i < kBig; i += kWave) good = good && s_recv[w][i] == (uint8_t)(prev * 7 + i);
    for (int i = lane(); i < kSmall; i += kWave)
        good = good && s_small[w][1][i] == (uint8_t)(prev * 3 + 2 * i);
    // collectives on LDS buffers: Allreduce (staged input and output), Bcast into LDS
    MPI_Allreduce(s_red[w][0], s_red[w][1], kRed, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
    lds_wave_sync();
    for (int i = lane(); i < kRed; i += kWave) {
        const int32_t exp = 1000 * (P * (P - 1) / 2) + P * i;
        good = good && s_red[w][1][i] == exp;
    }
    MPI_Bcast(s_send[w], kBig, MPI_BYTE, 1 % P, MPI_COMM_WORLD);
    lds_wave_sync();
    for (int i = lane(); i < kBig; i += kWave) good = good && s_send[w][i] == (uint8_t)((1 % P) * 7 + i);
    const bool all_good = ballot(!good) == 0;
    if (lane() == 0) a.ok[rank] = all_good ? 1 : 0;
    MPI_Finalize();
    return 0;
}

}  // namespace progs
}  // namespace gpumpi
