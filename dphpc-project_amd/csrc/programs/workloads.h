// Rank programs for the benchmark workloads (one rank = one wavefront; loops over local
// data are spread over the 64 lanes of the rank).  Each takes a plain-struct argument
// block from the host (rt()->user_args) and writes its results to device buffers named
// there, so the host checks them against the oracle.
#pragma once

#include "../runtime/mpi_api.h"
#include "../runtime/philox.h"

namespace gpumpi {
namespace progs {

template <class T>
__device__ __forceinline__ const T& args() { return *(const T*)rt()->user_args; }

// ---------------------------------------------------------------- Monte Carlo pi
// The reference's toolchain test runs mpich's cpi.c (toolchain_tests/pi); the workload
// here draws n (x, y) pairs per rank from Philox (key {seed_lo, seed_hi ^ rank},
// counter {i, 0, 0, 0}), counts x^2 + y^2 <= 1 in int64 and reduces with
// MPI_Reduce(MPI_INT64_T, MPI_SUM, root 0).
struct McpiArgs {
    uint64_t seed;
    int64_t n;
    int64_t* per_rank;   // [P]
    int64_t* total;      // [1]
    double* seconds;     // [P]
};

__device__ int w_mcpi(int, char**) {
    const McpiArgs& a = args<McpiArgs>();
    MPI_Init(nullptr, nullptr);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32) ^ (uint32_t)rank;
    int64_t hits = 0;
    for (int64_t i = lane(); i < a.n; i += kWave) {
        u32x4v o = philox4x32_10(u32x4v{(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 0u}, k0, k1);
        double x = unit53(o.x, o.y);
        double y = unit53(o.z, o.w);
        double s = x * x + y * y;  // -ffp-contract=off: two products, one sum
        hits += s <= 1.0 ? 1 : 0;
    }
    hits = wave_sum(hits);
    int64_t total = 0;
    MPI_Reduce(&hits, &total, 1, MPI_INT64_T, MPI_SUM, 0, MPI_COMM_WORLD);
    double t1 = MPI_Wtime();
    a.per_rank[rank] = hits;
    a.seconds[rank] = t1 - t0;
    if (rank == 0) *a.total = total;
    MPI_Finalize();
    return 0;
}

// ---------------------------------------------------------------- ping-pong / ring
// Config: P ranks, uint8 payload of `bytes`, generated from Philox (seed, rank, iter 0).
// mode 0: ping-pong pairs (2k, 2k+1), blocking Send/Recv
// mode 1: ping-pong pairs, Isend/Irecv + Waitall
// mode 2: ring shift r -> r+1, blocking (even ranks send first)
// mode 3: ring shift, Isend/Irecv + Waitall
// After the timed loop every rank checks the bytes it holds against the generator and
// reports mismatches.  Timing follows benchmarks/mpi_communication/mpi_communication.c:
// 24-49 (barrier, warmups, timed iterations, per-rank average).
struct P2pArgs {
    uint64_t seed;
    int32_t mode;
    int32_t iters;
    int32_t warmup;
    int32_t pad;
    uint64_t bytes;
    uint8_t* sendbase;    // P * bytes
    uint8_t* recvbase;    // P * bytes
    double* seconds;      // [P] time of the timed loop
    int64_t* mismatches;  // [P]
};

__device__ void fill_payload(uint8_t* p, uint64_t n, uint64_t seed, int rank) {
    // bulk (16-byte aligned buffers): 4 independent Philox blocks per lane and pass, so the
    // multiplies of one block's rounds overlap the others' (one wave: 5.9 -> 7.9 GB/s)
    constexpr int kI = 4;
    uint64_t b0 = 0;
    if (((uintptr_t)p & 15) == 0) {
        const uint64_t nb = n / 16, full = nb - nb % ((uint64_t)kWave * kI);
        for (; b0 < full; b0 += (uint64_t)kWave * kI) {
            u32x4v v[kI];
#pragma unroll
            for (int i = 0; i < kI; i++)
                v[i] = rank_stream(seed, rank, 0u, b0 + (uint64_t)i * kWave + lane(), kStreamBytes);
#pragma unroll
            for (int i = 0; i < kI; i++)
                *(u32x4*)(p + 16 * (b0 + (uint64_t)i * kWave + lane())) = u32x4{v[i].x, v[i].y, v[i].z, v[i].w};
        }
    }
    for (uint64_t b = b0 + (uint64_t)lane(); b * 16 < n; b += kWave) {
        u32x4v v = rank_stream(seed, rank, 0u, b, kStreamBytes);
        uint64_t o = b * 16;
        if (o + 16 <= n && ((uintptr_t)(p + o) & 15) == 0) {
            *(u32x4*)(p + o) = u32x4{v.x, v.y, v.z, v.w};
        } else {
            for (int k = 0; k < 16 && o + k < n; k++) p[o + k] = (uint8_t)(pick(v, k >> 2) >> (8 * (k & 3)));
        }
    }
}

__device__ int64_t check_payload(const uint8_t* p, uint64_t n, uint64_t seed, int rank) {
    int64_t bad = 0;
    for (uint64_t b = (uint64_t)lane(); b * 16 < n; b += kWave) {
        u32x4v v = rank_stream(seed, rank, 0u, b, kStreamBytes);
        uint64_t o = b * 16;
        for (int k = 0; k < 16 && o + k < n; k++)
            bad += p[o + k] != (uint8_t)(pick(v, k >> 2) >> (8 * (k & 3)));
    }
    return wave_sum(bad);
}

__device__ int w_p2p(int, char**) {
    const P2pArgs& a = args<P2pArgs>();
    MPI_Init(nullptr, nullptr);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    uint8_t* sb = a.sendbase + (uint64_t)rank * a.bytes;
    uint8_t* rb = a.recvbase + (uint64_t)rank * a.bytes;
    fill_payload(sb, a.bytes, a.seed, rank);
    const int count = (int)a.bytes;
    const bool pingpong = a.mode < 2;
    const bool nonblocking = (a.mode & 1) != 0;
    int peer_send, peer_recv;
    if (pingpong) {
        peer_send = peer_recv = rank ^ 1;
    } else {
        peer_send = (rank + 1) % P;
        peer_recv = (rank + P - 1) % P;
    }
    const bool active = !pingpong || (rank ^ 1) < P;
    double t0 = 0;
    for (int it = -a.warmup; it < a.iters; it++) {
        if (it == 0) {
            MPI_Barrier(MPI_COMM_WORLD);
            t0 = MPI_Wtime();
        }
        if (!active) continue;
        if (pingpong) {
            // the even rank pings, the odd rank pongs back the bytes it received
            if ((rank & 1) == 0) {
                if (nonblocking) {
                    MPI_Request r[2];
                    MPI_Isend(sb, count, MPI_UINT8_T, peer_send, 1, MPI_COMM_WORLD, &r[0]);
                    MPI_Wait(&r[0], MPI_STATUS_IGNORE);
                    MPI_Irecv(rb, count, MPI_UINT8_T, peer_recv, 2, MPI_COMM_WORLD, &r[1]);
                    MPI_Wait(&r[1], MPI_STATUS_IGNORE);
                } else {
                    MPI_Send(sb, count, MPI_UINT8_T, peer_send, 1, MPI_COMM_WORLD);
                    MPI_Recv(rb, count, MPI_UINT8_T, peer_recv, 2, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                }
            } else {
                if (nonblocking) {
                    MPI_Request r[2];
                    MPI_Irecv(rb, count, MPI_UINT8_T, peer_recv, 1, MPI_COMM_WORLD, &r[0]);
                    MPI_Wait(&r[0], MPI_STATUS_IGNORE);
                    MPI_Isend(rb, count, MPI_UINT8_T, peer_send, 2, MPI_COMM_WORLD, &r[1]);
                    MPI_Wait(&r[1], MPI_STATUS_IGNORE);
                } else {
                    MPI_Recv(rb, count, MPI_UINT8_T, peer_recv, 1, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                    MPI_Send(rb, count, MPI_UINT8_T, peer_send, 2, MPI_COMM_WORLD);
                }
            }
        } else if (nonblocking) {
            MPI_Request r[2];
            MPI_Irecv(rb, count, MPI_UINT8_T, peer_recv, 3, MPI_COMM_WORLD, &r[0]);
            MPI_Isend(sb, count, MPI_UINT8_T, peer_send, 3, MPI_COMM_WORLD, &r[1]);
            MPI_Waitall(2, r, MPI_STATUSES_IGNORE);
        } else if ((rank & 1) == 0) {
            MPI_Send(sb, count, MPI_UINT8_T, peer_send, 3, MPI_COMM_WORLD);
            MPI_Recv(rb, count, MPI_UINT8_T, peer_recv, 3, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        } else {
            MPI_Recv(rb, count, MPI_UINT8_T, peer_recv, 3, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
            MPI_Send(sb, count, MPI_UINT8_T, peer_send, 3, MPI_COMM_WORLD);
        }
    }
    double t1 = MPI_Wtime();
    MPI_Barrier(MPI_COMM_WORLD);
    // even ping-pong ranks hold their own bytes echoed back; odd ones the even peer's;
    // ring ranks hold the left neighbour's bytes
    int origin = pingpong ? ((rank & 1) ? rank ^ 1 : rank) : peer_recv;
    int64_t bad = active && a.iters + a.warmup > 0 ? check_payload(rb, a.bytes, a.seed, origin) : 0;
    a.seconds[rank] = t1 - t0;
    a.mismatches[rank] = bad;
    MPI_Finalize();
    return 0;
}

// ---------------------------------------------------------------- collective sweep
// kind 0 Bcast, 1 Reduce, 2 Allreduce, 3 Alltoall, 4 Allgather, 5 Gather, 6 Scatter.
// Inputs: MPI_INT uniform in [-2^20, 2^20] or MPI_DOUBLE "N(0,1)" from Philox
// (seed, rank, iteration 0).  `count` is the per-rank buffer length in elements
// (Alltoall: per-rank total, count/P per pair).
struct CollArgs {
    uint64_t seed;
    int32_t kind;
    int32_t dtype;       // MPI_INT or MPI_DOUBLE
    int32_t op;          // 0 = MPI_SUM, 1 = MPI_MAX
    int32_t root;
    int64_t count;
    int32_t iters;
    int32_t warmup;
    char* inbase;        // P * in_stride bytes
    char* outbase;       // P * out_stride bytes
    int64_t in_stride;
    int64_t out_stride;
    double* seconds;     // [P]
    int32_t no_fill;     // 1: the inputs are already in place (timing one call per launch)
    int32_t pad;
};

__device__ void fill_input(char* p, int dtype, int64_t n, uint64_t seed, int rank) {
    if (dtype == MPI_DOUBLE) {
        for (int64_t i = lane(); i < n; i += kWave) ((double*)p)[i] = gen_normal(seed, rank, 0u, i);
    } else {
        for (int64_t i = lane(); i < n; i += kWave) ((int32_t*)p)[i] = gen_int(seed, rank, 0u, i);
    }
}

__device__ int w_coll(int, char**) {
    const CollArgs& a = args<CollArgs>();
    MPI_Init(nullptr, nullptr);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    char* in = a.inbase + (uint64_t)rank * a.in_stride;
    char* out = a.outbase + (uint64_t)rank * a.out_stride;
    const MPI_Datatype dt = (MPI_Datatype)a.dtype;
    const MPI_Op op = a.op == 0 ? MPI_SUM : MPI_MAX;
    const int64_t n = a.count;
    int64_t nin = n;
    if (a.kind == 6) nin = rank == a.root ? n * P : 0;
    if (!a.no_fill) fill_input(in, a.dtype, nin, a.seed, rank);
    // Bcast sends the root's `out` (MPI_Bcast leaves the root's buffer as it is, so one
    // copy before the timed calls serves all of them)
    if (a.kind == 0 && rank == a.root) wave_copy(out, in, (uint64_t)n * type_size(dt));
    double t0 = 0;
    for (int it = -a.warmup; it < a.iters; it++) {
        if (it == 0) {
            MPI_Barrier(MPI_COMM_WORLD);
            t0 = MPI_Wtime();
        }
        switch (a.kind) {
        case 0: MPI_Bcast(out, (int)n, dt, a.root, MPI_COMM_WORLD); break;
        case 1: MPI_Reduce(in, out, (int)n, dt, op, a.root, MPI_COMM_WORLD); break;
        case 2: MPI_Allreduce(in, out, (int)n, dt, op, MPI_COMM_WORLD); break;
        case 3: MPI_Alltoall(in, (int)(n / P), dt, out, (int)(n / P), dt, MPI_COMM_WORLD); break;
        case 4: MPI_Allgather(in, (int)n, dt, out, (int)n, dt, MPI_COMM_WORLD); break;
        case 5: MPI_Gather(in, (int)n, dt, out, (int)n, dt, a.root, MPI_COMM_WORLD); break;
        default: MPI_Scatter(in, (int)n, dt, out, (int)n, dt, a.root, MPI_COMM_WORLD); break;
        }
    }
    double t1 = MPI_Wtime();
    a.seconds[rank] = t1 - t0;
    MPI_Finalize();
    return 0;
}

// ---------------------------------------------------------------- heap churn
// Per rank: `ops` operations; with probability 1/2 (or when 64 blocks are live) free a
// random live block, else allocate floor(2^(4 + 12u)) bytes (16 B .. 64 KiB, log-uniform).
// Every block carries a rank/op stamp at both ends that is verified when it is freed.
// Optionally an MPI_Allreduce is interleaved every `coll_every` operations.
struct ChurnArgs {
    uint64_t seed;
    int64_t ops;
    int32_t coll_every;
    int32_t dump;          // write the final live set before freeing it
    int64_t* errors;       // [P]
    int64_t* peak_live;    // [P]
    uint64_t* live_ptrs;   // [P * 64]
    uint64_t* live_sizes;  // [P * 64]
    double* seconds;       // [P]
    int64_t* failed;       // [P] allocations that returned nullptr (arena exhausted); may be null
};

__device__ int w_churn(int, char**) {
    const ChurnArgs& a = args<ChurnArgs>();
    MPI_Init(nullptr, nullptr);
    int rank;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    const int kMaxLive = 64;
    uint64_t* ptrs = (uint64_t*)__gpu_malloc(sizeof(uint64_t) * 2 * kMaxLive);
    uint64_t* sizes = ptrs + kMaxLive;
    int live = 0;
    int64_t errors = 0, failed = 0, peak = 0;
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    for (int64_t op = 0; op < a.ops; op++) {
        u32x4v r = rank_stream(a.seed, rank, 0u, (uint64_t)op, kStreamJobs);
        bool do_free = live == kMaxLive || (live > 0 && (r.x & 1));
        if (do_free) {
            int k = (int)(r.y % (uint32_t)live);
            uint64_t p = ptrs[k], sz = sizes[k];
            uint64_t stamp = ((uint64_t)rank << 40) ^ sz;
            bool ok = *(volatile uint64_t*)(uintptr_t)p == stamp;
            if (sz >= 16) ok = ok && *(volatile uint64_t*)(uintptr_t)(p + sz - 8) == stamp;
            errors += ok ? 0 : 1;
            __gpu_free((void*)(uintptr_t)p);
            ptrs[k] = ptrs[live - 1];
            sizes[k] = sizes[live - 1];
            live--;
        } else {
            double u = unit53(r.z, r.w);
            uint64_t sz = (uint64_t)exp2(4.0 + 12.0 * u);
            if (sz < 8) sz = 8;
            void* p = __gpu_malloc(sz);
            if (!p) {
                failed++;
            } else {
                uint64_t stamp = ((uint64_t)rank << 40) ^ sz;
                *(uint64_t*)p = stamp;
                if (sz >= 16) *(uint64_t*)((char*)p + sz - 8) = stamp;
                ptrs[live] = (uint64_t)(uintptr_t)p;
                sizes[live] = sz;
                live++;
                peak = live > peak ? live : peak;
            }
        }
        if (a.coll_every > 0 && (op + 1) % a.coll_every == 0) {
            int64_t x = live, s = 0;
            MPI_Allreduce(&x, &s, 1, MPI_INT64_T, MPI_SUM, MPI_COMM_WORLD);
        }
    }
    double t1 = MPI_Wtime();
    if (a.dump) {
        for (int k = lane(); k < kMaxLive; k += kWave) {
            a.live_ptrs[rank * kMaxLive + k] = k < live ? ptrs[k] : 0;
            a.live_sizes[rank * kMaxLive + k] = k < live ? sizes[k] : 0;
        }
    }
    MPI_Barrier(MPI_COMM_WORLD);
    for (int k = 0; k < live; k++) __gpu_free((void*)(uintptr_t)ptrs[k]);
    __gpu_free(ptrs);
    // without a `failed` array a failed allocation counts as an error
    a.errors[rank] = a.failed ? errors : errors + failed;
    if (a.failed) a.failed[rank] = failed;
    a.peak_live[rank] = peak;
    a.seconds[rank] = t1 - t0;
    MPI_Finalize();
    return 0;
}

// ---------------------------------------------------------------- Alltoall(v) exchange
// The plain per-call path of the batched exchange: the same buffers as the batched plan
// ([P][P] count/displacement matrices, arrays of per-rank buffer pointers), one MPI call
// per launch (phase 0: Alltoall of counts; phase 1: receive displacements + Alltoallv;
// phase 2: both in one launch).
struct ExchArgs {
    int32_t phase;
    int32_t elem;        // 4: MPI_UNSIGNED, 1: MPI_BYTE, 8: MPI_DOUBLE
    char* const* sendbufs;
    const int32_t* scounts;
    const int32_t* sdispls;
    char* const* recvbufs;
    int32_t* rcounts;
    int32_t* rdispls;
};

__device__ int w_exchange(int, char**) {
    const ExchArgs& a = args<ExchArgs>();
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const MPI_Datatype dt = a.elem == 4 ? MPI_UNSIGNED : a.elem == 8 ? MPI_DOUBLE : MPI_BYTE;
    const int64_t row = (int64_t)rank * P;
    if (a.phase == 0 || a.phase == 2)
        MPI_Alltoall(a.scounts + row, 1, MPI_INT, a.rcounts + row, 1, MPI_INT, MPI_COMM_WORLD);
    if (a.phase == 1 || a.phase == 2) {
        int64_t carry = 0;
        for (int s0 = 0; s0 < P; s0 += kWave) {
            int s = s0 + lane();
            int64_t c = s < P ? a.rcounts[row + s] : 0;
            int64_t inc = wave_incl_scan(c);
            if (s < P) a.rdispls[row + s] = (int32_t)(carry + inc - c);
            carry += __shfl(inc, 63);
        }
        MPI_Alltoallv(a.sendbufs[rank], a.scounts + row, a.sdispls + row, dt, a.recvbufs[rank],
                      a.rcounts + row, a.rdispls + row, dt, MPI_COMM_WORLD);
    }
    return 0;
}

}  // namespace progs
}  // namespace gpumpi
