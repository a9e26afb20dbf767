// Served MPI jobs: the job program shared by both execution modes, and the
// device-resident job server.
//
// A job is an independent MPI program on its own communicator of `ranks` ranks
// (the dot-product and Alltoall shapes of the reference's converted benchmarks,
// mpi_malloc_benchmarks/scalar_product.c and alltoall.c):
//   type 0  dot product of two FP64 vectors of `size` elements, block-distributed;
//           x_i, y_i are the seeded "N(0,1)" streams (seed, 0) and (seed, 1) at global
//           index i; local sums run lane-strided then a fixed butterfly over the 64 lanes,
//           then MPI_Allreduce(MPI_DOUBLE, MPI_SUM) in rank order -- the result is the
//           same bits however the job is scheduled;
//   type 1  MPI_Alltoall of `size` bytes per pair (payload = seeded bytes of the sending
//           rank); result = sum over ranks of sum_i w_i * (2i + 1) mod 2^64 over the
//           received 8-byte little-endian words (zero-padded), an order-free checksum.
// Plain mode: one rank-grid launch per job (program "job").  Served mode: a persistent
// rank grid where slot 0 is the scheduler and every other slot a worker.
#pragma once

#include "workloads.h"

namespace gpumpi {
namespace progs {

struct JobDesc {
    uint64_t id;
    uint64_t seed;
    int32_t type;
    int32_t ranks;
    int64_t size;
};

struct JobResult {
    uint64_t id;
    int32_t status;        // 1 done, -1 rejected (does not fit), -2 runtime error
    int32_t ranks;
    double dot;
    uint64_t checksum;
    uint64_t t_admit;      // wall_clock64 ticks
    uint64_t t_done;
    int64_t bytes;         // heap bytes the job needed
};

__host__ __device__ __forceinline__ int64_t pow2_at_least(int64_t v) {
    int64_t p = 16;
    while (p < v) p <<= 1;
    return p;
}
// Heap bytes a job holds: its arena (two buffers per rank in 256-byte granules), which
// the server allocates whole at admission.
__host__ __device__ __forceinline__ int64_t job_bytes(const JobDesc& j) {
    const int64_t per = j.type == 0 ? ((j.size + j.ranks - 1) / j.ranks) * 8
                                    : (int64_t)j.ranks * j.size;
    const int64_t g = ((per > 0 ? per : 1) + 255) & ~(int64_t)255;
    return 2 * g * j.ranks;
}

// sum over i of w_i * (2i + 1) mod 2^64, w_i the 8-byte little-endian words of p[0 .. n)
// (the last one zero-padded), summed over the wave.  A lone wave's read rate is set by the
// bytes it keeps in flight, so the bulk issues 8 unaligned 16-byte loads per lane (8 KiB
// per wave) before summing any.  Vector j = j0 + 64k (j0 = base + lane, k = 0..7) holds
// words 2j and 2j + 1, whose terms add up to 4j*s_k + (w0 + 3*w1), s_k = w0 + w1; over the
// batch that is 4*j0*T + 256*U + R with T = sum s_k, U = sum k*s_k (suffix sums) and
// R = sum (w0 + 3*w1): one variable multiply per batch instead of two per vector.  The
// arithmetic is modular, so the regrouping gives the same bits as the plain sum.
__device__ uint64_t wave_word_checksum(const uint8_t* p, uint64_t n) {
    constexpr int kB = 8;
    const int l = lane();
    const uint64_t a = (uint64_t)(uintptr_t)p;
    const uint64_t nvec = n >> 4;
    const uint64_t full = nvec - nvec % ((uint64_t)kWave * kB);
    uint64_t cs = 0;
    for (uint64_t base = 0; base < full; base += (uint64_t)kWave * kB) {
        u32x4 v[kB];
#pragma unroll
        for (int k = 0; k < kB; k++) v[k] = gload16(a + 16 * (base + (uint64_t)k * kWave + l));
        uint64_t T = 0, U = 0, R = 0;
#pragma unroll
        for (int k = kB - 1; k >= 0; k--) {
            const uint64_t w0 = (uint64_t)v[k].x | (uint64_t)v[k].y << 32;
            const uint64_t w1 = (uint64_t)v[k].z | (uint64_t)v[k].w << 32;
            T += w0 + w1;
            if (k > 0) U += T;   // after the k-th step T = s_k + .. + s_7; summed for k >= 1
            R += w0 + w1 + (w1 << 1);
        }
        cs += ((base + (uint64_t)l) << 2) * T + (U << 8) + R;
    }
    const uint64_t nw = (n + 7) / 8;
    for (uint64_t i = 2 * full + (uint64_t)l; i < nw; i += kWave) {
        uint64_t w = 0;
        if (8 * i + 8 <= n) {
            w = *(const uint64_t*)(p + 8 * i);
        } else {
            for (uint64_t b = 8 * i; b < n; b++) w |= (uint64_t)p[b] << (8 * (b - 8 * i));
        }
        cs += w * (2 * i + 1);
    }
    return (uint64_t)wave_sum((int64_t)cs);
}

// Runs one rank of a job on communicator `comm`; rank 0 fills *res (value fields).
__device__ int job_rank(const JobDesc& j, MPI_Comm comm, JobResult* res) {
    int r, P;
    MPI_Comm_rank(comm, &r);
    MPI_Comm_size(comm, &P);
    int rc = MPI_SUCCESS;
    if (j.type == 0) {
        const int64_t lo = j.size * r / P, hi = j.size * (r + 1) / P, nl = hi - lo;
        double* x = (double*)__gpu_malloc(sizeof(double) * (nl > 0 ? nl : 1));
        double* y = (double*)__gpu_malloc(sizeof(double) * (nl > 0 ? nl : 1));
        if (!x || !y) {
            GPUMPI_FAIL(kErrResource, (int)j.id, "job: heap exhausted");
            return MPI_ERR_OTHER;
        }
        for (int64_t i = lane(); i < nl; i += kWave) {
            x[i] = gen_normal(j.seed, 0, 0u, lo + i);
            y[i] = gen_normal(j.seed, 1, 0u, lo + i);
        }
        double acc = 0.0;
        for (int64_t i = lane(); i < nl; i += kWave) {
            double p = x[i] * y[i];
            acc = acc + p;
        }
        for (int o = 32; o > 0; o >>= 1) acc = acc + __shfl_xor(acc, o);
        double tot = 0.0;
        rc = MPI_Allreduce(&acc, &tot, 1, MPI_DOUBLE, MPI_SUM, comm);
        if (r == 0 && res) {
            res->dot = tot;
            res->checksum = 0;
        }
        __gpu_free(y);
        __gpu_free(x);
    } else {
        const uint64_t m = (uint64_t)j.size, blk = m * (uint64_t)P;
        uint8_t* sb = (uint8_t*)__gpu_malloc(blk ? blk : 1);
        uint8_t* rb = (uint8_t*)__gpu_malloc(blk ? blk : 1);
        if (!sb || !rb) {
            GPUMPI_FAIL(kErrResource, (int)j.id, "job: heap exhausted");
            return MPI_ERR_OTHER;
        }
        fill_payload(sb, blk, j.seed, r);
        rc = MPI_Alltoall(sb, (int)m, MPI_BYTE, rb, (int)m, MPI_BYTE, comm);
        const uint64_t cs = wave_word_checksum(rb, blk);
        uint64_t tot = 0;
        if (rc == MPI_SUCCESS) rc = MPI_Allreduce(&cs, &tot, 1, MPI_UINT64_T, MPI_SUM, comm);
        if (r == 0 && res) {
            res->checksum = tot;
            res->dot = 0.0;
        }
        __gpu_free(rb);
        __gpu_free(sb);
    }
    return rc;
}

// ---------------------------------------------------------------- plain mode
struct JobArgs {
    JobDesc job;
    JobResult* out;
};

__device__ int w_job(int, char**) {
    const JobArgs& a = args<JobArgs>();
    int rc = job_rank(a.job, MPI_COMM_WORLD, a.out);
    if (world_rank() == 0) {
        a.out->id = a.job.id;
        a.out->ranks = a.job.ranks;
        a.out->status = rc == MPI_SUCCESS ? 1 : -2;
        a.out->bytes = job_bytes(a.job);
        a.out->t_done = wall_clock64();
    }
    return rc == MPI_SUCCESS ? 0 : 1;
}

// ---------------------------------------------------------------- served mode
// The server's job heap: a buddy allocator over one region of ntop blocks of 2^top bytes
// (smallest block 64 KiB).  Only the scheduler wave uses it, so it needs no atomics, and
// a freed block merges with its free buddy all the way up -- admission asks it for the
// job's whole arena, so "fits" means "fits now, contiguously".  Every lane applies the
// same updates (each lane reads back its own stores).
constexpr int kBuddyMinShift = 16;
constexpr int kBuddyOrders = 40;

struct Buddy {
    char* base;
    int32_t top;                     // order of the largest blocks (block = 2^(16 + order))
    int32_t ntop;                    // number of largest blocks
    uint64_t* bits;                  // free bitmaps of all orders
    uint64_t off[kBuddyOrders];      // word offset of order o's bitmap in bits
    uint8_t* order_of;               // [ntop << top] order of the allocated block at each 64 KiB
};

__device__ __forceinline__ bool buddy_test(const Buddy& b, int o, uint64_t i) {
    return (b.bits[b.off[o] + (i >> 6)] >> (i & 63)) & 1ull;
}
__device__ __forceinline__ void buddy_set(const Buddy& b, int o, uint64_t i, bool v) {
    uint64_t* w = &b.bits[b.off[o] + (i >> 6)];
    const uint64_t m = 1ull << (i & 63);
    const uint64_t x = *w;
    *w = v ? (x | m) : (x & ~m);
}
// First free block of order o (64 bitmap words per step across the lanes), or -1.
__device__ int64_t buddy_find(const Buddy& b, int o) {
    const uint64_t nblk = (uint64_t)b.ntop << (b.top - o);
    const uint64_t nw = (nblk + 63) / 64;
    for (uint64_t w0 = 0; w0 < nw; w0 += kWave) {
        const uint64_t wi = w0 + lane();
        const uint64_t x = wi < nw ? b.bits[b.off[o] + wi] : 0;
        const uint64_t m = ballot(x != 0);
        if (m) {
            const int L = __builtin_ctzll(m);
            const uint64_t xv = readlane64(x, L);
            return (int64_t)((w0 + L) * 64 + __builtin_ctzll(xv));
        }
    }
    return -1;
}
__device__ char* buddy_alloc(const Buddy& b, uint64_t need) {
    int o = 0;
    while (((uint64_t)1 << (kBuddyMinShift + o)) < need) o++;
    if (o > b.top) return nullptr;
    for (int k = o; k <= b.top; k++) {
        int64_t i = buddy_find(b, k);
        if (i < 0) continue;
        buddy_set(b, k, (uint64_t)i, false);
        for (int j = k - 1; j >= o; j--) {  // split: keep the left half, free the right
            i = 2 * i;
            buddy_set(b, j, (uint64_t)i + 1, true);
        }
        b.order_of[(uint64_t)i << o] = (uint8_t)o;
        return b.base + ((uint64_t)i << (kBuddyMinShift + o));
    }
    return nullptr;
}
__device__ void buddy_free(const Buddy& b, char* p) {
    const uint64_t b0 = (uint64_t)(p - b.base) >> kBuddyMinShift;
    int o = b.order_of[b0];
    uint64_t i = b0 >> o;
    while (o < b.top && buddy_test(b, o, i ^ 1)) {  // merge with a free buddy
        buddy_set(b, o, i ^ 1, false);
        i >>= 1;
        o++;
    }
    buddy_set(b, o, i, true);
}

struct SlotBox {
    uint32_t seq;       // incremented by the scheduler for every assignment
    int32_t job;        // index into the job table
    int32_t rank;       // rank within the job's communicator
    int32_t pad;
};

struct JobSlot {
    JobDesc desc;
    JobResult res;
    CommImpl* comm;
    int32_t remaining;  // ranks still running (atomic)
    int32_t state;      // 0 free, 1 running, 2 finished (set by the last rank)
    int32_t* slots;     // the job's world slots (scheduler-owned list)
    JobArena arena;     // the job's memory, allocated at admission
};

struct ServeShared {
    // host-visible (pinned) memory
    JobDesc* q;                  // [qcap] submission ring written by the host
    JobDesc* dq;                 // [qcap] device copy of the submitted jobs
    uint64_t* q_tail;            // number of jobs submitted (host release)
    JobResult* done;             // [qcap] completion ring written by the device
    uint64_t* done_count;        // number of completions (device release)
    uint64_t* admit_head;        // first job the scheduler has not admitted or rejected yet:
                                 // dq[] entries from here on are still needed (device store)
    int32_t* stop;               // host sets 1: finish admitted jobs and exit
    uint64_t* heartbeat;         // bumped by every host submit/poll call
    // device memory
    SlotBox* boxes;              // [S]
    JobSlot* jobs;               // [jcap] running-job table
    int32_t* free_slots;         // [S] stack of idle worker slots (scheduler-owned)
    int32_t* free_jobs;          // [jcap] stack of free job-table entries (scheduler-owned)
    int32_t* slot_lists;         // [jcap * max_ranks] world slots of each running job
    int32_t* exit_flag;          // scheduler tells workers to leave
    int64_t* stats;              // [4]: admitted, rejected, max concurrent jobs, backfills
    int32_t qcap, jcap, max_ranks, lookahead;
    int64_t heap_budget;         // largest arena a job can ever get (2^top)
    Buddy jobheap;               // the job heap (scheduler-owned)
    uint64_t idle_ticks;         // no heartbeat for this long: behave as if stop was set
};

// Polls of host-written words: coherent loads without the cache invalidate of an acquire;
// the caller fences once it acts on a new value.
__device__ __forceinline__ uint64_t sys_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ int32_t sys_load32(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ void serve_worker(const ServeShared& S) {
    SlotBox* box = &S.boxes[world_rank()];
    uint32_t seen = 0;
    for (;;) {
        uint32_t seq = 0;
        int32_t ex = 0;
        if (lane() == 0) {
            seq = ld_relaxed(&box->seq);
            ex = ld_relaxed(S.exit_flag);
        }
        seq = bcast(seq);
        ex = bcast(ex);
        if (seq != seen) {
            seen = seq;
            fence_acquire();
            const int jb = box->job;
            JobSlot* js = &S.jobs[jb];
            my_state()->arena = &js->arena;
            int rc = job_rank(js->desc, js->comm, &js->res);
            my_state()->arena = nullptr;
            if (rc != MPI_SUCCESS) js->res.status = -2;
            heap_flush_cache(&rt()->heap, &my_state()->cache);
            fence_release();
            if (lane() == 0) {
                int32_t left = __hip_atomic_fetch_add(&js->remaining, -1, __ATOMIC_ACQ_REL,
                                                      __HIP_MEMORY_SCOPE_AGENT) - 1;
                if (left == 0) {
                    js->res.t_done = wall_clock64();
                    st_release(&js->state, 2);
                }
            }
            continue;
        }
        if (ex) break;
        __builtin_amdgcn_s_sleep(4);
    }
}

__device__ void serve_scheduler(const ServeShared& S) {
    Runtime* R = rt();
    const int nslots = R->nranks;
    // free worker slots (slot 0 is this scheduler)
    int nfree = 0;
    for (int s = 1 + lane(); s < nslots; s += kWave) S.free_slots[s - 1] = s;
    nfree = nslots - 1;
    // free job-table entries (a stack, scheduler-owned)
    for (int k = lane(); k < S.jcap; k += kWave) S.free_jobs[k] = S.jcap - 1 - k;
    int nfree_jobs = S.jcap;
    uint64_t head = 0;          // first job not yet admitted or rejected
    uint64_t head_published = 0;
    uint64_t fetched = 0;       // jobs copied from the host ring into dq
    uint64_t completed = 0;
    int running = 0;
    uint64_t beat = 0, beat_time = wall_clock64();
    // admitted-out-of-order bookkeeping: jobs in [head, head+lookahead) may be admitted
    // early (backfill); admitted[] marks them by ring position.
    for (;;) {
        // ---- reclaim finished jobs: 64 job slots per load, then one pass per finished job
        for (int base = 0; base < S.jcap; base += kWave) {
            const int jl = base + lane();
            const int32_t stl = jl < S.jcap ? ld_relaxed(&S.jobs[jl].state) : 0;
            uint64_t fin = ballot(stl == 2);
            if (!fin) continue;
            fence_acquire();
            while (fin) {
                const int jb = base + __builtin_ctzll(fin);
                fin &= fin - 1;
                JobSlot* js = &S.jobs[jb];
                for (int i = lane(); i < js->desc.ranks; i += kWave)
                    S.free_slots[nfree + i] = js->slots[i];
                nfree += js->desc.ranks;
                rt_free(js->comm);
                buddy_free(S.jobheap, js->arena.base);
                // publish the result to the host ring
                const uint64_t pos = completed % (uint64_t)S.qcap;
                JobResult res = js->res;
                res.status = res.status == -2 ? -2 : 1;
                if (lane() == 0) {
                    S.done[pos] = res;
                    __atomic_thread_fence(__ATOMIC_RELEASE);
                    __hip_atomic_store(S.done_count, completed + 1, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                }
                completed++;
                js->state = 0;
                S.free_jobs[nfree_jobs++] = jb;
                running--;
            }
        }
        // ---- admit new jobs (FIFO with bounded backfill)
        uint64_t tail = 0;
        if (lane() == 0) tail = sys_load(S.q_tail);
        tail = bcast(tail);
        // one pass over PCIe per new job: copy the descriptors into device memory
        if (tail != fetched) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the host's ring writes
            for (uint64_t qi = fetched + lane(); qi < tail; qi += kWave)
                S.dq[qi % (uint64_t)S.qcap] = S.q[qi % (uint64_t)S.qcap];
            fetched = tail;
        }
        int32_t stop = 0;
        uint64_t hb = 0;
        if (lane() == 0) {
            stop = sys_load32(S.stop);
            hb = sys_load(S.heartbeat);
        }
        stop = bcast(stop);
        hb = bcast(hb);
        // a host that stopped calling submit/poll is gone: drain and leave, so the grid
        // always ends
        const uint64_t now = wall_clock64();
        if (hb != beat) {
            beat = hb;
            beat_time = now;
        } else if (now - beat_time > S.idle_ticks) {
            stop = 1;
        }
        for (uint64_t qi = head; qi < tail && qi < head + (uint64_t)S.lookahead; qi++) {
            JobDesc d = S.dq[qi % (uint64_t)S.qcap];
            if (d.id == ~0ull) {  // already admitted by backfill
                if (qi == head) head++;
                continue;
            }
            const int64_t need = job_bytes(d);
            if (d.ranks < 1 || d.ranks > S.max_ranks || d.ranks > nslots - 1 ||
                need > S.heap_budget) {
                // can never run: report it as rejected
                JobResult res{};
                res.id = d.id;
                res.status = -1;
                res.ranks = d.ranks;
                res.bytes = need;
                const uint64_t pos = completed % (uint64_t)S.qcap;
                if (lane() == 0) {
                    S.done[pos] = res;
                    __atomic_thread_fence(__ATOMIC_RELEASE);
                    __hip_atomic_store(S.done_count, completed + 1, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_SYSTEM);
                    S.stats[1]++;
                }
                completed++;
                S.dq[qi % (uint64_t)S.qcap].id = ~0ull;
                if (qi == head) head++;
                continue;
            }
            if (d.ranks > nfree) continue;
            if (nfree_jobs == 0) break;
            const int jb = S.free_jobs[nfree_jobs - 1];
            // admission by real headroom: the job's whole arena must be allocatable now
            char* ab = buddy_alloc(S.jobheap, (uint64_t)need);
            if (!ab) {
                if (running == 0) {
                    // nothing will free memory: report it as rejected
                    JobResult res{};
                    res.id = d.id;
                    res.status = -1;
                    res.ranks = d.ranks;
                    res.bytes = need;
                    const uint64_t pos = completed % (uint64_t)S.qcap;
                    if (lane() == 0) {
                        S.done[pos] = res;
                        __atomic_thread_fence(__ATOMIC_RELEASE);
                        __hip_atomic_store(S.done_count, completed + 1, __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_SYSTEM);
                        S.stats[1]++;
                    }
                    completed++;
                    S.dq[qi % (uint64_t)S.qcap].id = ~0ull;
                    if (qi == head) head++;
                }
                continue;
            }
            CommImpl* c = comm_alloc(d.ranks);
            if (!c) {
                buddy_free(S.jobheap, ab);
                break;
            }
            nfree_jobs--;
            JobSlot* js = &S.jobs[jb];
            int32_t* list = S.slot_lists + (int64_t)jb * S.max_ranks;
            for (int i = lane(); i < d.ranks; i += kWave) {
                const int w = S.free_slots[nfree - d.ranks + i];
                list[i] = w;
                c->world_ranks[i] = w;
                c->inv[w] = i;
            }
            nfree -= d.ranks;
            js->desc = d;
            js->comm = c;
            js->slots = list;
            js->remaining = d.ranks;
            js->arena.base = ab;
            js->arena.size = (uint64_t)need;
            js->arena.bump = 0;
            js->res = JobResult{};
            js->res.id = d.id;
            js->res.ranks = d.ranks;
            js->res.bytes = need;
            js->res.t_admit = wall_clock64();
            js->state = 1;
            fence_release();
            for (int i = lane(); i < d.ranks; i += kWave) {
                SlotBox* b = &S.boxes[list[i]];
                b->job = jb;
                b->rank = i;
            }
            fence_release();
            for (int i = lane(); i < d.ranks; i += kWave) {
                SlotBox* b = &S.boxes[list[i]];
                __hip_atomic_fetch_add(&b->seq, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            running++;
            if (lane() == 0) {
                S.stats[0]++;
                if (qi != head) S.stats[3]++;
                if (running > S.stats[2]) S.stats[2] = running;
            }
            S.dq[qi % (uint64_t)S.qcap].id = ~0ull;  // mark admitted
            if (qi == head) head++;
        }
        while (head < tail && S.dq[head % (uint64_t)S.qcap].id == ~0ull) head++;
        if (head != head_published) {
            // the host may reuse ring positions below head; the fence orders this wave's
            // reads of dq[] before the store
            fence_release();
            if (lane() == 0)
                __hip_atomic_store(S.admit_head, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            head_published = head;
        }
        if (stop && head >= tail && running == 0) break;
        __builtin_amdgcn_s_sleep(2);
    }
    fence_release();
    if (lane() == 0) st_release(S.exit_flag, 1);
}

struct ServeArgs {
    ServeShared* shared;
};

__device__ int w_serve(int, char**) {
    const ServeShared& S = *args<ServeArgs>().shared;
    if (world_rank() == 0) serve_scheduler(S);
    else serve_worker(S);
    return 0;
}

}  // namespace progs
}  // namespace gpumpi
