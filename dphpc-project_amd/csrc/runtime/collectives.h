// Collectives of the rank runtime (the "plain" per-rank path).
//
// Reference algorithms (gpu_mpi/mpi.cu): Bcast is a root loop of P-1 isends, or two grid
// barriers around a shared buffer for COMM_WORLD (mpi.cu:86-140); Reduce receives P-1
// messages into malloc(P*N*s) at the root and folds them in rank order 0..P-1, except
// FP64 SUM on COMM_WORLD, which uses per-element atomicAdd (non-deterministic,
// mpi.cu:148-235); Allreduce = Reduce + Bcast (276-282); Alltoall(v) = P-1 Isend + P-1
// Irecv + self memcpy (333-408); Gather/Scatter/Allgather are linear root loops.
//
// Here every collective runs over the shared address space with all P wavefronts
// working at once:
//   Bcast       small: every rank pulls from the root's buffer; large: rank k stores
//               slice k of the root's buffer into every receiver;
//   Reduce      rank k reduces element slice k over all P inputs in the reference's
//               order (out = in_0; out = op(in_r, out), r = 1..P-1) -- bit-exact with the
//               reference's generic path for every type, FP64 included -- and writes it
//               to the root;
//   Allreduce   same slice reduction, result stored into every rank's receive buffer
//               (2*P*N*s bytes moved, one barrier);
//   Alltoall(v) each rank pulls its P incoming blocks, starting at me+1 so sources are
//               spread;
//   Gather(v)   each rank pushes its block into the root's buffer; Scatter(v) and
//               Allgather(v) pull.
#pragma once

#include "comm.h"
#include "fold.h"
#include "ops.h"

namespace gpumpi {

__device__ __forceinline__ bool check_comm(const CommImpl* c) {
    if (!c) {
        GPUMPI_FAIL(kErrArgument, 0, "invalid communicator (MPI_COMM_NULL)");
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ Bcast
// Two schedules, chosen by the buffer's bytes per rank (the same choice on every rank of a
// correct program, whose ranks all pass the same size):
//   pull  every receiver copies the root's whole buffer -- one wave per receiver, so a
//         receiver's rate is what one wave keeps in flight (small buffers: one wait);
//   push  the root's buffer splits into P slices of whole KiB; rank k loads slice k
//         from the root once and stores it into every receiver's buffer.  Stores do not
//         wait for memory, so a large Bcast runs at the chip's write rate: P*N*s bytes
//         written, N*s read.
// Every rank publishes its buffer and size.  A push checks that all sizes equal the
// root's first; on a mismatch every rank falls back to pull, which copies what fits and
// records the truncation, as the reference's receive does.
// Measured crossovers (MPI_INT, per call): 1024 ranks, pull 183 us / push 160 us at
// 512 KiB and 94 / 134 us at 256 KiB; 64 ranks, 27 / 21 us at 512 KiB and equal at 256 KiB.
constexpr int32_t kBcastPushMinDefault = 512;
constexpr uint64_t kBcastPushMinBytes = 512 << 10;  // and at least this much in all
constexpr int kBcastU = 16;                        // 16-byte columns per lane in a full pass

// Calls f(address of the receiver's buffer) for every receiver (every rank but the root),
// reading the buffer addresses 64 at a time and loading the next 64 while f runs.
template <class F>
__device__ __forceinline__ void for_each_receiver(const PtrTable& dst, int P, int root_in, F&& f) {
    const int root = bcast(root_in);   // wave-uniform: the root test is a scalar branch
    uint64_t pn = dst.lane_chunk(0, P);
    for (int d0 = 0; d0 < P; d0 += kWave) {
        const uint64_t pd = pn;
        pn = dst.lane_chunk(d0 + kWave < P ? d0 + kWave : d0, P);   // unconditional prefetch
        const int n = P - d0 < kWave ? P - d0 : kWave;
        for (int j = 0; j < n; j++)
            if (d0 + j != root) f(readlane64(pd, j));
    }
}

// N KiB of the root's buffer at byte offset `off` of this lane's first column (lane l owns
// columns off + 1024 u, u < N), loaded once and stored into every receiver's buffer, N KiB
// contiguous per receiver.  The last column is only loaded and stored where last_ok.
template <int N>
__device__ __forceinline__ void bcast_push_cols(const PtrTable& dst, int P, int root, uint64_t src,
                                                uint64_t off, bool last_ok) {
    u32x4 x[N];
#pragma unroll
    for (int u = 0; u < N; u++)
        if (u < N - 1 || last_ok) x[u] = gload16(src + off + 1024u * u);
    for_each_receiver(dst, P, root, [&](uint64_t b) {
#pragma unroll
        for (int u = 0; u < N; u++)
            if (u < N - 1 || last_ok) gstore16<false>(b + off + 1024u * u, x[u]);
    });
}

// The last part of a slice, n < kBcastU columns, with N = n instantiated.
template <int N>
__device__ __forceinline__ void bcast_push_rest(int n, const PtrTable& dst, int P, int root, uint64_t src,
                                                uint64_t off, bool last_ok) {
    if constexpr (N > 0) {
        if (n == N) bcast_push_cols<N>(dst, P, root, src, off, last_ok);
        else bcast_push_rest<N - 1>(n, dst, P, root, src, off, last_ok);
    }
}

// Bytes [lo, hi) of the root's buffer stored into every receiver's: full passes of 16 KiB
// (loaded once, then written to one receiver after the other, 16 KiB contiguous each),
// then the rest of the slice in one pass of up to 15 KiB.  Every pass waits for its loads,
// and on gfx9 that wait also drains the stores issued before it (one counter for both), so
// a slice takes as few passes as it can.  Plain stores: on MI355X this store pattern (1024
// waves writing 16 KiB runs into 1024 buffers) measured 5.8 TB/s with them and 5.2 TB/s
// with non-temporal ones.  Out of line: the store loop's registers stay out of the
// collectives that call it.
__device__ __attribute__((noinline)) void bcast_push_slice(const CommImpl* c, int root, uint64_t lo,
                                                           uint64_t hi) {
    const int P = c->size;
    const CollSlot* slots = c->slots;
    const uint64_t src = (uint64_t)(uintptr_t)slots[root].sbuf;
    const PtrTable dst{(const char*)&slots[0].rbuf, (int64_t)sizeof(CollSlot)};
    const uint64_t nv = (hi - lo) / 16;
    constexpr uint64_t kPass = (uint64_t)kWave * kBcastU;
    uint64_t v0 = 0;
    for (; v0 + kPass <= nv; v0 += kPass)
        bcast_push_cols<kBcastU>(dst, P, root, src, lo + 16 * (v0 + (uint64_t)lane()), true);
    if (v0 < nv) {
        const uint64_t rem = nv - v0;  // columns of 16 bytes per wave row, < kPass
        const int n = (int)((rem + kWave - 1) / kWave);
        const bool last_ok = (uint64_t)lane() < rem - (uint64_t)kWave * (n - 1);
        bcast_push_rest<kBcastU - 1>(n, dst, P, root, src, lo + 16 * (v0 + (uint64_t)lane()), last_ok);
    }
    // the bytes past the slice's last whole column (only the last slice has any)
    for (uint64_t i = lo + 16 * nv; i < hi; i++) {
        const char x = ((const char*)src)[i];
        for (int d = lane(); d < P; d += kWave)
            if (d != root) ((char*)slots[d].rbuf)[i] = x;
    }
}

__device__ int coll_bcast(void* buffer, uint64_t bytes, int root, CommImpl* c) {
    const int me = comm_rank_of(c);
    const int P = c->size;
    if (root < 0 || root >= P) {
        GPUMPI_FAIL(kErrArgument, root, "MPI_Bcast: invalid root");
        return MPI_ERR_OTHER;
    }
    if (P == 1 || bytes == 0) return MPI_SUCCESS;
    const int32_t pmin = rt()->bcast_push_min;
    // the default threshold also asks for kBcastPushMinBytes in all: below it a few
    // waves would store to every receiver one after the other
    bool push = pmin >= 0 && bytes / (uint64_t)P >= (uint64_t)pmin &&
                (pmin != kBcastPushMinDefault || bytes >= kBcastPushMinBytes);
    Stager st;
    const void* src = nullptr;
    void* out = buffer;
    if (me == root) src = st.in(buffer, bytes);
    else if (push) out = st.out(buffer, bytes);
    // a null buffer, or one that staging could not move to global memory (the failure is
    // recorded), must not be pushed to or pulled from: its size -1 sends every rank to the
    // pull schedule, and a receiver reads nothing from a root that published -1
    const void* mine = me == root ? src : (const void*)out;
    if (!mine) GPUMPI_FAIL(kErrArgument, me, "MPI_Bcast: null buffer");
    const bool unusable = !mine || (push && needs_staging(mine));
    uint64_t seq = coll_begin(c, me);
    coll_publish(c, me, seq, src, out, nullptr, nullptr, unusable ? -1 : (int64_t)bytes, 0, 0, 0);
    if (push) {
        if (!coll_wait_all(c, seq)) return MPI_ERR_OTHER;
        bool same = true;
        for (int r = lane(); r < P; r += kWave) same &= c->slots[r].a0 == (int64_t)bytes;
        push = ballot(!same) == 0;
    }
    uint64_t got = bytes;  // bytes this rank receives
    if (push) {
        uint64_t per = (bytes + P - 1) / (uint64_t)P;
        per = (per + 1023) / 1024 * 1024;  // whole KiB
        const uint64_t lo = (uint64_t)me * per;
        const uint64_t hi = lo + per < bytes ? lo + per : bytes;
        if (lo < hi) bcast_push_slice(c, root, lo, hi);
    } else if (me != root) {
        if (!coll_wait_slot(c, root, seq)) return MPI_ERR_OTHER;
        const CollSlot* s = &c->slots[root];
        if ((uint64_t)s->a0 != bytes) {
            GPUMPI_FAIL(kErrTruncate, (int32_t)s->a0, "MPI_Bcast: message size mismatch");
        }
        got = s->a0 < 0 ? 0 : (bytes < (uint64_t)s->a0 ? bytes : (uint64_t)s->a0);
        if (got && out) wave_copy(out, s->sbuf, got);
    }
    bool ok = comm_barrier(c);
    if (out != buffer) st.back(buffer, out, got);
    st.release();
    return ok ? MPI_SUCCESS : MPI_ERR_OTHER;
}

// ------------------------------------------------------------------ Reduce / Allreduce
// Slice of the element range owned by comm rank k: multiples of 256 elements so every
// rank's stores stay 16-byte aligned for all element sizes.
__device__ __forceinline__ void reduce_slice(int64_t count, int P, int k, int64_t* lo, int64_t* hi) {
    int64_t per = (count + P - 1) / P;
    per = (per + 255) & ~(int64_t)255;
    *lo = (int64_t)k * per;
    *hi = *lo + per;
    if (*lo > count) *lo = count;
    if (*hi > count) *hi = count;
}

// Value of `v` in lane j (uniform j), for any type of at most 8 bytes.
template <class T>
__device__ __forceinline__ T lane_value(T v, int j) {
    uint64_t b = 0;
    __builtin_memcpy(&b, &v, sizeof(T));
    b = readlane64(b, j);
    T out;
    __builtin_memcpy(&out, &b, sizeof(T));
    return out;
}

constexpr int64_t kReduceSmallSlice = 8;   // elements: at most this, lanes run over ranks

struct ReduceFunctor {
    const CommImpl* c;
    int64_t lo, hi;
    int root;  // -1: all ranks receive
    bool nt;   // non-temporal result stores (large buffers)
    // Out of line, one function per (type, operator): inlined into coll_reduce, the fold's
    // load ring pushed the collective past 256 registers per lane.
    template <class T, int OP>
    __device__ __attribute__((noinline)) void run() {
        const int P = c->size;
        const CollSlot* slots = c->slots;
        if constexpr (sizeof(T) <= 8) {
            if (hi - lo <= kReduceSmallSlice) {
                // A few elements over many ranks: lanes load the inputs of 64 ranks each
                // (8 x 64 loads in flight), and the fold runs in rank order 0..P-1 on values
                // read out of the lanes -- the same order and operations as below.
                constexpr int B = 8;
                for (int64_t i = lo; i < hi; i++) {
                    T acc{};
                    for (int r0 = 0; r0 < P; r0 += B * kWave) {
                        T x[B];
#pragma unroll
                        for (int b = 0; b < B; b++) {
                            const int r = r0 + b * kWave + lane();
                            x[b] = r < P ? ((const T*)slots[r].sbuf)[i] : T{};
                        }
#pragma unroll
                        for (int b = 0; b < B; b++) {
                            for (int j = 0; j < kWave; j++) {
                                const int r = r0 + b * kWave + j;
                                if (r < P) {
                                    const T v = lane_value(x[b], j);
                                    acc = r == 0 ? v : combine<T, OP>(v, acc);
                                }
                            }
                        }
                    }
                    if (root >= 0) {
                        if (lane() == 0) ((T*)slots[root].rbuf)[i] = acc;
                    } else {
                        for (int d = lane(); d < P; d += kWave) ((T*)slots[d].rbuf)[i] = acc;
                    }
                }
                return;
            }
        }
        // The slice in tiles of kFoldU 16-byte columns per lane (kFoldU KiB per rank buffer
        // and visit), each folded over the P inputs in rank order with kFoldD ranks' loads
        // in flight (fold.h); the elements past the last whole column one per lane.  1- and
        // 2-byte types go element by element (a 16-byte column of them unpacks into 8-16
        // values per lane).
        const PtrTable src{(const char*)&slots[0].sbuf, (int64_t)sizeof(CollSlot)};
        const PtrTable dst{(const char*)&slots[0].rbuf, (int64_t)sizeof(CollSlot)};
        const uint64_t one = root >= 0 ? (uint64_t)(uintptr_t)slots[root].rbuf : 0;
        const int64_t b0 = lo * (int64_t)sizeof(T);
        const int64_t nv = sizeof(T) >= 4 ? (hi - lo) * (int64_t)sizeof(T) / 16 : 0;
        constexpr int64_t kTile = (int64_t)kWave * kFoldU;
        if constexpr (sizeof(T) >= 4) {
            if (nv > 0 && nv <= (int64_t)kWave * kFoldSmallU) {
                // one small tile: the unmasked fold
                const int64_t v = lane();
                uint32_t valid = 0;
#pragma unroll
                for (int u = 0; u < kFoldSmallU; u++)
                    if (v + (int64_t)kWave * u < nv) valid |= 1u << u;
                const uint64_t off = (uint64_t)(b0 + 16 * v);
                u32x4 acc[kFoldSmallU];
                fold_columns_unmasked<T, OP, kFoldSmallU, kFoldSmallD>(src, P, off, valid, acc);
                if (nt) fold_store_cols<true, kFoldSmallU>(dst, P, one, off, valid, acc);
                else fold_store_cols<false, kFoldSmallU>(dst, P, one, off, valid, acc);
                fold_scalar_tail<T, OP>(src, P, dst, P, one, lo + nv * 16 / (int64_t)sizeof(T), hi);
                return;
            }
        }
        for (int64_t v0 = 0; v0 < nv; v0 += kTile) {
            const int64_t v = v0 + lane();
            uint32_t valid = 0;
#pragma unroll
            for (int u = 0; u < kFoldU; u++)
                if (v + (int64_t)kWave * u < nv) valid |= 1u << u;
            const uint64_t off = (uint64_t)(b0 + 16 * v);
            if constexpr (sizeof(T) >= 4) {
                u32x4 acc[kFoldU];
                fold_columns<T, OP, kFoldU, kFoldD>(src, P, off, valid, acc);
                if (nt) fold_store_cols<true, kFoldU>(dst, P, one, off, valid, acc);
                else fold_store_cols<false, kFoldU>(dst, P, one, off, valid, acc);
            }
        }
        fold_scalar_tail<T, OP>(src, P, dst, P, one, lo + nv * 16 / (int64_t)sizeof(T), hi);
    }
};

__device__ int coll_reduce(const void* sendbuf, void* recvbuf, int64_t count, int dt,
                           const OpImpl* op, int root, CommImpl* c) {
    const int P = c->size;
    const int me = comm_rank_of(c);
    const int es = type_size(dt);
    if (es <= 0 || !op) {
        GPUMPI_FAIL(kErrArgument, dt, "reduction: invalid datatype or operator");
        return MPI_ERR_OTHER;
    }
    if (root >= P) {
        GPUMPI_FAIL(kErrArgument, root, "MPI_Reduce: invalid root");
        return MPI_ERR_OTHER;
    }
    if (op->builtin >= 0 && !op_allowed(dt, op->builtin)) {
        GPUMPI_FAIL(kErrArgument, dt * 16 + op->builtin,
                    "combination of type and operator is not allowed in reduce operation");
        return MPI_ERR_OTHER;
    }
    if (count <= 0) return MPI_SUCCESS;
    const uint64_t bytes = (uint64_t)count * (uint64_t)es;
    Stager st;
    const void* in = st.in(sendbuf, bytes);
    void* out = nullptr;
    if (root < 0 || me == root) out = st.out(recvbuf, bytes);
    uint64_t seq = coll_begin(c, me);
    coll_publish(c, me, seq, in, out, nullptr, nullptr, 0, 0, 0, 0);
    int64_t lo, hi;
    reduce_slice(count, P, me, &lo, &hi);
    if (lo < hi) {
        // only ranks that reduce a slice read the others' buffers; the rest go straight
        // to the closing barrier, which keeps every buffer alive until the reads are done
        if (!coll_wait_all(c, seq)) return MPI_ERR_OTHER;
        if (op->builtin >= 0) {
            ReduceFunctor f{c, lo, hi, root, bytes >= kNontemporalBytes};
            dispatch_op(dt, op->builtin, f);
        } else {
            // user operator: fold on lane 0 into the first destination, then replicate
            char* dst0 = (char*)c->slots[root >= 0 ? root : 0].rbuf + lo * es;
            uint64_t nb = (uint64_t)(hi - lo) * es;
            wave_copy(dst0, (const char*)c->slots[0].sbuf + lo * es, nb);
            fence_release();
            int len = (int)(hi - lo);
            MPI_Datatype t = (MPI_Datatype)dt;
            for (int r = 1; r < P; r++) {
                if (lane() == 0)
                    op->user_fn((void*)((const char*)c->slots[r].sbuf + lo * es), dst0, &len, &t);
                fence_acquire();
            }
            if (root < 0)
                for (int d = 1; d < P; d++)
                    wave_copy((char*)c->slots[d].rbuf + lo * es, dst0, nb);
        }
    }
    bool ok = comm_barrier(c);
    if (out && out != recvbuf) st.back(recvbuf, out, bytes);
    st.release();
    return ok ? MPI_SUCCESS : MPI_ERR_OTHER;
}

// ------------------------------------------------------------------ Alltoall(v)
constexpr int64_t kAlltoallLaneBytes = 16;   // blocks up to this size: pulled lane by lane

// A uniform Alltoall's (or Allgather's) pull for blocks of whole 16-byte vectors, at most
// VPL per lane (n <= 1 KiB * VPL), from sources (me + k) mod P, k = 1 .. P: source s's
// bytes [soff, soff + n) go to rb + s * n.  The send buffers of 64 sources at a time are
// read into the lanes (one gather) and taken out with readlane; each pass then loads the
// blocks of G = 8 / VPL sources before storing any, so it waits one HBM latency for G
// blocks where a block-by-block pull waits one per block (two, counting the read of the
// source's slot).  Alltoall: a sender's block for rank me starts
// at soff = me * n (a uniform Alltoall's send and receive blocks have the same bytes);
// Allgather: soff = 0.  The destination rb + s * n is 16-byte aligned; sources may be
// unaligned.  Out of line, so its 8 vectors of loads stay out of the rank programs'
// register count.
template <int VPL>
__device__ __noinline__ void coll_pull_uniform(char* rb, const CollSlot* slots, int me, int P,
                                               uint64_t n, uint64_t soff) {
    constexpr int G = 8 / VPL;   // divides 64
    const int l = lane();
    const uint64_t nv = n >> 4;
    for (int k0 = 1; k0 <= P; k0 += kWave) {
        const int nk = P - k0 + 1 < kWave ? P - k0 + 1 : kWave;
        const uint64_t sb = l < nk ? (uint64_t)(uintptr_t)slots[(me + k0 + l) % P].sbuf : 0;
        for (int j0 = 0; j0 < nk; j0 += G) {
            u32x4 v[G][VPL];
#pragma unroll
            for (int b = 0; b < G; b++) {
                if (j0 + b >= nk) break;   // wave-uniform
                const uint64_t sp = readlane64(sb, j0 + b) + soff;
#pragma unroll
                for (int u = 0; u < VPL; u++) {
                    const uint64_t i = (uint64_t)u * kWave + (uint64_t)l;
                    v[b][u] = gload16(sp + 16 * (i < nv ? i : 0));   // clamped: every lane loads
                }
            }
#pragma unroll
            for (int b = 0; b < G; b++) {
                if (j0 + b >= nk) break;
                int s = me + k0 + j0 + b;   // < 2P
                s -= s >= P ? P : 0;
                const uint64_t dp = (uint64_t)(uintptr_t)rb + (uint64_t)s * n;
#pragma unroll
                for (int u = 0; u < VPL; u++) {
                    const uint64_t i = (uint64_t)u * kWave + (uint64_t)l;
                    if (i < nv) gstore16<false>(dp + 16 * i, v[b][u]);
                }
            }
        }
    }
}

// The same pull for blocks of at most 32 vectors (n <= 512 B), which would leave most lanes
// idle: L lanes per source (L = the power of two >= the block's vectors), so a row of 64
// lanes loads R = 64 / L sources, and a pass loads up to 8 rows before storing any.  A
// lane's source address comes from the gathered send buffers by a lane shuffle.
__device__ __noinline__ void coll_pull_spread(char* rb, const CollSlot* slots, int me, int P, uint64_t n,
                                              uint64_t soff, int lg) {
    const int l = lane();
    const int L = 1 << lg, R = kWave >> lg;
    const int sub = l >> lg, vi = l & (L - 1);
    const uint64_t nv = n >> 4;
    const bool active = (uint64_t)vi < nv;
    const uint64_t col = 16 * (uint64_t)(active ? vi : 0);
    for (int k0 = 1; k0 <= P; k0 += kWave) {
        const int nk = P - k0 + 1 < kWave ? P - k0 + 1 : kWave;
        const uint64_t sb = l < nk ? (uint64_t)(uintptr_t)slots[(me + k0 + l) % P].sbuf : 0;
        for (int j0 = 0; j0 < nk; j0 += 8 * R) {
            u32x4 v[8];
#pragma unroll
            for (int b = 0; b < 8; b++) {
                if (j0 + b * R >= nk) break;   // wave-uniform
                const int j = j0 + b * R + sub;
                const uint64_t sp = __shfl(sb, j < nk ? j : 0);
                v[b] = gload16(sp + soff + col);   // every lane loads (source 0 past the end)
            }
#pragma unroll
            for (int b = 0; b < 8; b++) {
                if (j0 + b * R >= nk) break;
                const int j = j0 + b * R + sub;
                int s = me + k0 + j;   // < 2P where j < nk
                s -= s >= P ? P : 0;
                if (j < nk && active) gstore16<false>((uint64_t)(uintptr_t)rb + (uint64_t)s * n + col, v[b]);
            }
        }
    }
}

// coll_pull_uniform for n <= 4 KiB of whole 16-byte vectors (coll_pull_spread up to 512 B).
__device__ __forceinline__ void coll_pull_blocks(char* rb, const CollSlot* slots, int me, int P,
                                                 uint64_t n, uint64_t soff) {
    const uint64_t nv = n / 16;
    if (nv <= kWave / 2) {
        int lg = 0;
        while ((1ull << lg) < nv) lg++;
        coll_pull_spread(rb, slots, me, P, n, soff, lg);
        return;
    }
    const uint64_t vpl = (n / 16 + kWave - 1) / kWave;
    if (vpl == 1) coll_pull_uniform<1>(rb, slots, me, P, n, soff);
    else if (vpl == 2) coll_pull_uniform<2>(rb, slots, me, P, n, soff);
    else coll_pull_uniform<4>(rb, slots, me, P, n, soff);
}

// counts/displs are in elements of the respective datatype (gpu_mpi/mpi.cu:363-408).
__device__ int coll_alltoallv(const void* sendbuf, const int* scounts, const int* sdispls,
                              int sdt, void* recvbuf, const int* rcounts, const int* rdispls,
                              int rdt, CommImpl* c, int uniform_scount, int uniform_rcount) {
    const int P = c->size;
    const int me = comm_rank_of(c);
    const int se = type_size(sdt), re = type_size(rdt);
    if (se <= 0 || re <= 0) {
        GPUMPI_FAIL(kErrArgument, sdt, "alltoall: invalid datatype");
        return MPI_ERR_OTHER;
    }
    const bool uniform = scounts == nullptr;
    Stager st;
    // extents of the local buffers (for staging of private buffers)
    uint64_t sext = 0, rext = 0;
    if (uniform) {
        sext = (uint64_t)uniform_scount * se * P;
        rext = (uint64_t)uniform_rcount * re * P;
    } else if (needs_staging(sendbuf) || needs_staging(recvbuf)) {
        int64_t sm = 0, rm = 0;
        for (int i = lane(); i < P; i += kWave) {
            int64_t a = (int64_t)sdispls[i] + scounts[i];
            int64_t b = (int64_t)rdispls[i] + rcounts[i];
            sm = a > sm ? a : sm;
            rm = b > rm ? b : rm;
        }
        sext = (uint64_t)wave_max(sm) * se;
        rext = (uint64_t)wave_max(rm) * re;
    }
    const void* sb = st.in(sendbuf, sext);
    void* rb = st.out(recvbuf, rext);
    const int* sc = uniform ? nullptr : (const int*)st.in(scounts, sizeof(int) * (uint64_t)P);
    const int* sd = uniform ? nullptr : (const int*)st.in(sdispls, sizeof(int) * (uint64_t)P);
    uint64_t seq = coll_begin(c, me);
    coll_publish(c, me, seq, sb, rb, sc, sd, se, uniform_scount, 0, 0);
    if (!coll_wait_all(c, seq)) return MPI_ERR_OTHER;
    // Blocks of at most kAlltoallLaneBytes into a global receive buffer: lane l pulls the
    // blocks of sources l, l + 64, ... itself (the slot fields are read per lane, after the
    // acquire of coll_wait_all), instead of one wave-wide copy per source -- 16 sources per
    // lane on 1024 ranks instead of 1024 copies one after the other.  The closing barrier's
    // release waits for these stores and its acquire drops this CU's L1 lines, so every
    // lane of the rank then reads every block.  A private receive buffer (staging failed)
    // keeps the wave path: a lane's store would reach only its own copy.
    bool small = !needs_staging(rb);
    if (small) {
        int64_t mx = uniform ? (int64_t)uniform_rcount * re : 0;
        if (!uniform)
            for (int i = lane(); i < P; i += kWave) mx = rcounts[i] * (int64_t)re > mx ? rcounts[i] * (int64_t)re : mx;
        small = wave_max(mx) <= kAlltoallLaneBytes;
    }
    if (small) {
        for (int s = lane(); s < P; s += kWave) {
            const CollSlot* src = &c->slots[s];
            uint64_t n, soff, doff;
            if (uniform) {
                n = (uint64_t)uniform_rcount * re;
                soff = (uint64_t)me * (uint64_t)src->a1 * (uint64_t)src->a0;
                doff = (uint64_t)s * n;
            } else {
                n = (uint64_t)rcounts[s] * re;
                soff = (uint64_t)src->displs[me] * (uint64_t)src->a0;
                doff = (uint64_t)rdispls[s] * re;
            }
            const char* sp = (const char*)src->sbuf + soff;
            char* dp = (char*)rb + doff;
            if (n == 4 && (((uintptr_t)sp | (uintptr_t)dp) & 3) == 0) {
                *(uint32_t*)dp = *(const uint32_t*)sp;
            } else if (n == 16 && (((uintptr_t)sp | (uintptr_t)dp) & 15) == 0) {
                *(u32x4*)dp = *(const u32x4*)sp;
            } else if (n == 8 && (((uintptr_t)sp | (uintptr_t)dp) & 7) == 0) {
                *(uint64_t*)dp = *(const uint64_t*)sp;
            } else {
                for (uint64_t b = 0; b < n; b++) dp[b] = sp[b];
            }
        }
    }
    // Uniform blocks of whole 16-byte vectors up to 4 KiB into an aligned global buffer:
    // several sources per pass (coll_pull_uniform)
    bool batched = false;
    if (!small && uniform && !needs_staging(rb) && ((uintptr_t)rb & 15) == 0) {
        const uint64_t n = (uint64_t)uniform_rcount * re;
        const uint64_t vpl = (n / 16 + kWave - 1) / kWave;
        if (n != 0 && n % 16 == 0 && vpl <= 4) {
            batched = true;
            coll_pull_blocks((char*)rb, c->slots, me, P, n, (uint64_t)me * n);
        }
    }
    for (int k = 1; k <= P && !small && !batched; k++) {
        const int s = (me + k) % P;
        const CollSlot* src = &c->slots[s];
        uint64_t n, soff, doff;
        if (uniform) {
            n = (uint64_t)uniform_rcount * re;
            soff = (uint64_t)me * (uint64_t)src->a1 * (uint64_t)src->a0;
            doff = (uint64_t)s * n;
        } else {
            n = (uint64_t)rcounts[s] * re;
            soff = (uint64_t)src->displs[me] * (uint64_t)src->a0;
            doff = (uint64_t)rdispls[s] * re;
        }
        if (n) wave_copy((char*)rb + doff, (const char*)src->sbuf + soff, n);
    }
    bool ok = comm_barrier(c);
    if (rb != recvbuf) st.back(recvbuf, rb, rext);
    st.release();
    return ok ? MPI_SUCCESS : MPI_ERR_OTHER;
}

// ------------------------------------------------------------------ Gather(v)
// Every rank pushes its block into the root's receive buffer.
__device__ int coll_gatherv(const void* sendbuf, uint64_t sbytes, void* recvbuf,
                            const int* rcounts, const int* displs, int uniform_rcount,
                            int rdt, int root, CommImpl* c, bool all) {
    const int P = c->size;
    const int me = comm_rank_of(c);
    const int re = type_size(rdt);
    if (re <= 0 || root >= P) {
        GPUMPI_FAIL(kErrArgument, rdt, "gather: invalid datatype or root");
        return MPI_ERR_OTHER;
    }
    Stager st;
    const bool uniform = rcounts == nullptr;
    uint64_t rext = 0;
    if (all || me == root) {
        if (uniform) {
            rext = (uint64_t)uniform_rcount * re * P;
        } else {
            int64_t m = 0;
            for (int i = lane(); i < P; i += kWave) {
                int64_t a = (int64_t)displs[i] + rcounts[i];
                m = a > m ? a : m;
            }
            rext = (uint64_t)wave_max(m) * re;
        }
    }
    const void* sb = st.in(sendbuf, sbytes);
    void* rb = st.out(recvbuf, rext);
    const int* rc = (uniform || !(all || me == root)) ? nullptr
                        : (const int*)st.in(rcounts, sizeof(int) * (uint64_t)P);
    const int* rd = (uniform || !(all || me == root)) ? nullptr
                        : (const int*)st.in(displs, sizeof(int) * (uint64_t)P);
    uint64_t seq = coll_begin(c, me);
    coll_publish(c, me, seq, sb, rb, rc, rd, (int64_t)sbytes, uniform_rcount, 0, 0);
    if (all) {
        // Allgather(v): pull every rank's block into my buffer
        if (!coll_wait_all(c, seq)) return MPI_ERR_OTHER;
        bool batched = false;
        if (uniform && !needs_staging(rb) && ((uintptr_t)rb & 15) == 0) {
            // blocks of whole 16-byte vectors up to 4 KiB, every sender's whole (as in a
            // correct program): several sources per pass (coll_pull_uniform)
            const uint64_t n = (uint64_t)uniform_rcount * re;
            const uint64_t vpl = (n / 16 + kWave - 1) / kWave;
            if (n != 0 && n % 16 == 0 && vpl <= 4) {
                bool short_src = false;
                for (int s = lane(); s < P; s += kWave) short_src |= (uint64_t)c->slots[s].a0 < n;
                if (ballot(short_src) == 0) {
                    batched = true;
                    coll_pull_blocks((char*)rb, c->slots, me, P, n, 0);
                }
            }
        }
        for (int k = 0; k < P && !batched; k++) {
            const int s = (me + k) % P;
            uint64_t doff = uniform ? (uint64_t)s * uniform_rcount * re : (uint64_t)rd[s] * re;
            uint64_t n = uniform ? (uint64_t)uniform_rcount * re : (uint64_t)rc[s] * re;
            uint64_t avail = (uint64_t)c->slots[s].a0;
            wave_copy((char*)rb + doff, c->slots[s].sbuf, n < avail ? n : avail);
        }
    } else {
        if (!coll_wait_slot(c, root, seq)) return MPI_ERR_OTHER;
        const CollSlot* rs = &c->slots[root];
        uint64_t doff = uniform ? (uint64_t)me * rs->a1 * re : (uint64_t)rs->displs[me] * re;
        uint64_t n = uniform ? (uint64_t)rs->a1 * re : (uint64_t)rs->counts[me] * re;
        wave_copy((char*)rs->rbuf + doff, sb, n < sbytes ? n : sbytes);
    }
    bool ok = comm_barrier(c);
    if (rb != recvbuf) st.back(recvbuf, rb, rext);
    st.release();
    return ok ? MPI_SUCCESS : MPI_ERR_OTHER;
}

// ------------------------------------------------------------------ Scatter(v)
__device__ int coll_scatterv(const void* sendbuf, const int* scounts, const int* displs,
                             int uniform_scount, int sdt, void* recvbuf, uint64_t rbytes,
                             int root, CommImpl* c) {
    const int P = c->size;
    const int me = comm_rank_of(c);
    const int se = type_size(sdt);
    if (se <= 0 || root >= P) {
        GPUMPI_FAIL(kErrArgument, sdt, "scatter: invalid datatype or root");
        return MPI_ERR_OTHER;
    }
    Stager st;
    const bool uniform = scounts == nullptr;
    const void* sb = nullptr;
    const int* sc = nullptr;
    const int* sd = nullptr;
    if (me == root) {
        uint64_t sext;
        if (uniform) {
            sext = (uint64_t)uniform_scount * se * P;
        } else {
            int64_t m = 0;
            for (int i = lane(); i < P; i += kWave) {
                int64_t a = (int64_t)displs[i] + scounts[i];
                m = a > m ? a : m;
            }
            sext = (uint64_t)wave_max(m) * se;
            sc = (const int*)st.in(scounts, sizeof(int) * (uint64_t)P);
            sd = (const int*)st.in(displs, sizeof(int) * (uint64_t)P);
        }
        sb = st.in(sendbuf, sext);
    }
    uint64_t seq = coll_begin(c, me);
    coll_publish(c, me, seq, sb, nullptr, sc, sd, se, uniform_scount, 0, 0);
    if (!coll_wait_slot(c, root, seq)) return MPI_ERR_OTHER;
    const CollSlot* rs = &c->slots[root];
    uint64_t soff = uniform ? (uint64_t)me * rs->a1 * rs->a0 : (uint64_t)rs->displs[me] * rs->a0;
    uint64_t n = uniform ? (uint64_t)rs->a1 * rs->a0 : (uint64_t)rs->counts[me] * rs->a0;
    wave_copy(recvbuf, (const char*)rs->sbuf + soff, n < rbytes ? n : rbytes);
    bool ok = comm_barrier(c);
    st.release();
    return ok ? MPI_SUCCESS : MPI_ERR_OTHER;
}

}  // namespace gpumpi
