// Communicators, groups, barriers and the slot protocol of the collectives.
//
// Reference: gpu_mpi/communicator.cu:16-221 and gpu_mpi/group.cu:7-121.  There a
// communicator is a per-rank object holding a group, MPI_Comm_rank scans the group
// (O(P), group.cu:110-121), context ids come from an Allreduce(MAX) of per-rank counters
// (communicator.cu:37-46), Barrier only exists for COMM_WORLD (mpi.cu:324-331) and
// Comm_split sorts with an O(P^2) qsort (communicator.cu:164-213).
//
// Here a communicator is ONE object shared by all of its members: world_ranks[] maps
// comm rank -> world rank and inv[] maps back (O(1) Comm_rank), context ids come from a
// global atomic counter taken by the communicator's leader, every communicator has its
// own sense-reversing barrier, and Comm_split ranks by counting (key, rank) pairs with
// 64 lanes (O(P/64) per rank).
//
// Collectives use the shared address space of the single GPU: every member publishes
// its buffer pointers in its CollSlot with seq = the collective's sequence number,
// members pull what they need straight from each other's buffers, and an end barrier
// guarantees the buffers and slots are no longer read when the call returns.
#pragma once

#include "transport.h"

namespace gpumpi {

__device__ __forceinline__ int comm_rank_of(const CommImpl* c) {
    int w = world_rank();
    return c->inv ? c->inv[w] : w;
}
__device__ __forceinline__ int comm_world_rank(const CommImpl* c, int r) {
    return c->world_ranks ? c->world_ranks[r] : r;
}

// Sense-reversing centralized barrier on one communicator.
__device__ bool comm_barrier(CommImpl* c) {
    if (c->size <= 1) return true;
    fence_release();
    uint32_t g = 0, last = 0;
    if (lane() == 0) {
        g = ld_relaxed(&c->gen.value);
        // relaxed arrival: the release fence above and the acquire fence below order
        // the payloads (one L2 write-back and one invalidate per barrier, not per poll)
        uint32_t old = fetch_add_relaxed(&c->arrive.value, 1u);
        if (old == (uint32_t)c->size - 1u) {
            st_relaxed(&c->arrive.value, 0u);
            st_release(&c->gen.value, g + 1u);
            last = 1;
        }
    }
    last = bcast(last);
    g = bcast(g);
    if (!last) {
        Waiter w;
        for (;;) {
            uint32_t cur = 0;
            if (lane() == 0) cur = ld_relaxed(&c->gen.value);
            cur = bcast(cur);
            if (cur != g) break;
            progress();
            if (!w.keep_going(__LINE__)) return false;
        }
    }
    fence_acquire();
    return true;
}

__device__ __forceinline__ uint64_t coll_begin(CommImpl* c, int me) {
    uint64_t s = c->seqs[me] + 1;
    c->seqs[me] = s;
    return s;
}

__device__ void coll_publish(CommImpl* c, int me, uint64_t seq, const void* sbuf, void* rbuf,
                             const int* counts, const int* displs, int64_t a0, int32_t a1,
                             int32_t a2, int64_t a3) {
    CollSlot* s = &c->slots[me];
    s->sbuf = sbuf;
    s->rbuf = rbuf;
    s->counts = counts;
    s->displs = displs;
    s->a0 = a0;
    s->a1 = a1;
    s->a2 = a2;
    s->a3 = a3;
    fence_release();  // with the fence, a relaxed store publishes (one L2 write-back, not two)
    if (lane() == 0) st_relaxed(&s->seq, seq);
}

__device__ bool coll_wait_slot(CommImpl* c, int j, uint64_t seq) {
    Waiter w;
    for (;;) {
        uint64_t s = 0;
        if (lane() == 0) s = ld_relaxed(&c->slots[j].seq);
        s = bcast(s);
        if (s >= seq) break;
        progress();
        if (!w.keep_going(__LINE__)) return false;
    }
    fence_acquire();
    return true;
}

__device__ bool coll_wait_all(CommImpl* c, uint64_t seq) {
    Waiter w;
    for (;;) {
        bool ok = true;
        for (int j = lane(); j < c->size; j += kWave)
            if (ld_relaxed(&c->slots[j].seq) < seq) ok = false;
        if (ballot(!ok) == 0) break;
        progress();
        if (!w.keep_going(__LINE__)) return false;
    }
    fence_acquire();
    return true;
}

// Staging of private-memory (stack) and LDS buffers for the pull-based collectives: other
// ranks cannot address this wavefront's scratch memory or this workgroup's LDS, so inputs
// are copied to the rank's HBM scratch area (or the heap) and outputs are produced there
// and copied back.
__device__ __forceinline__ bool needs_staging(const void* p) { return is_private(p) || is_shared(p); }

struct Stager {
    char* base;
    uint64_t used;
    int nheap;
    void* heap_blocks[6];

    __device__ Stager() : base(my_state()->scratch), used(0), nheap(0) {}

    __device__ void* area(uint64_t bytes) {
        uint64_t b = (bytes + 15) & ~(uint64_t)15;
        if (used + b <= kScratchBytes) {
            void* p = base + used;
            used += b;
            return p;
        }
        if (nheap >= 6) return nullptr;
        void* p = heap_alloc(&rt()->heap, &my_state()->cache, b);
        if (p) heap_blocks[nheap++] = p;
        return p;
    }
    __device__ const void* in(const void* p, uint64_t bytes) {
        if (!p || bytes == 0 || !needs_staging(p)) return p;
        void* a = area(bytes);
        if (!a) {
            GPUMPI_FAIL(kErrResource, (int32_t)bytes, "no staging space for a private buffer");
            return p;
        }
        wave_copy(a, p, bytes);
        return a;
    }
    __device__ void* out(void* p, uint64_t bytes) {
        if (!p || bytes == 0 || !needs_staging(p)) return p;
        void* a = area(bytes);
        if (!a) GPUMPI_FAIL(kErrResource, (int32_t)bytes, "no staging space for a private buffer");
        return a ? a : p;
    }
    __device__ void back(void* priv, const void* staged, uint64_t bytes) {
        if (priv != staged && bytes) wave_copy(priv, staged, bytes);
    }
    __device__ void release() {
        for (int i = 0; i < nheap; i++) heap_free(&rt()->heap, &my_state()->cache, heap_blocks[i]);
        nheap = 0;
    }
};

// ------------------------------------------------------------------ object creation
__device__ void* rt_alloc(uint64_t bytes) {
    void* p = heap_alloc(&rt()->heap, &my_state()->cache, bytes);
    if (!p) GPUMPI_FAIL(kErrResource, (int32_t)bytes, "device heap exhausted");
    return p;
}
__device__ void rt_free(void* p) { heap_free(&rt()->heap, &my_state()->cache, p); }

// Builds an empty communicator of `size` members (called by the leader only).
__device__ CommImpl* comm_alloc(int size) {
    Runtime* R = rt();
    uint64_t bytes = sizeof(CommImpl) + sizeof(CollSlot) * (uint64_t)size +
                     sizeof(uint64_t) * (uint64_t)size + sizeof(int32_t) * (uint64_t)size +
                     sizeof(int32_t) * (uint64_t)R->nranks + sizeof(GroupImpl) + 256;
    char* p = (char*)rt_alloc(bytes);
    if (!p) return nullptr;
    CommImpl* c = (CommImpl*)p;
    char* q = p + ((sizeof(CommImpl) + 63) & ~(uint64_t)63);
    CollSlot* slots = (CollSlot*)q;
    q += sizeof(CollSlot) * (uint64_t)size;
    uint64_t* seqs = (uint64_t*)q;
    q += sizeof(uint64_t) * (uint64_t)size;
    int32_t* wr = (int32_t*)q;
    q += sizeof(int32_t) * (uint64_t)size;
    int32_t* inv = (int32_t*)q;
    q += sizeof(int32_t) * (uint64_t)R->nranks;
    q = (char*)(((uintptr_t)q + 15) & ~(uintptr_t)15);
    GroupImpl* g = (GroupImpl*)q;
    for (int i = lane(); i < size; i += kWave) {
        slots[i].seq = 0;
        seqs[i] = 0;
        wr[i] = -1;
    }
    for (int i = lane(); i < R->nranks; i += kWave) inv[i] = -1;
    uint64_t ctx = 0;
    if (lane() == 0) ctx = fetch_add_relaxed(&R->next_ctx, (uint64_t)1);
    ctx = bcast(ctx);
    c->ctx = (int32_t)ctx;
    c->size = size;
    c->refcount = size;
    c->is_world = 0;
    c->world_ranks = wr;
    c->inv = inv;
    c->slots = slots;
    c->seqs = seqs;
    c->arrive.value = 0;
    c->gen.value = 0;
    g->size = size;
    g->refcount = -1;  // owned by the communicator
    g->ranks = wr;
    c->group = g;
    return c;
}

// Collective over `parent`: members with equal color form a new communicator ordered by
// (key, parent rank); color == MPI_UNDEFINED yields MPI_COMM_NULL.  Restates the
// ordering of gpu_mpi/communicator.cu:154-213.
__device__ int comm_split(CommImpl* parent, int color, int key, CommImpl** out) {
    const int n = parent->size;
    const int me = comm_rank_of(parent);
    uint64_t seq = coll_begin(parent, me);
    coll_publish(parent, me, seq, nullptr, nullptr, nullptr, nullptr, 0, color, key, 0);
    if (!coll_wait_all(parent, seq)) return MPI_ERR_OTHER;
    // rank within color, size of color, leader (smallest (key, rank))
    int64_t nr = 0, ns = 0;
    int64_t best = INT64_MAX;
    for (int j = lane(); j < n; j += kWave) {
        int cj = parent->slots[j].a1;
        int kj = parent->slots[j].a2;
        if (cj != color) continue;
        ns++;
        if (kj < key || (kj == key && j < me)) nr++;
        int64_t order = (int64_t)kj * 4294967296LL + j;
        if (order < best) best = order;
    }
    nr = wave_sum(nr);
    ns = wave_sum(ns);
    best = wave_min(best);
    const int leader = (int)(best - (best >> 32) * 4294967296LL);
    if (!comm_barrier(parent)) return MPI_ERR_OTHER;
    // phase 2: leaders create the shared object
    CommImpl* c = nullptr;
    if (color != MPI_UNDEFINED && nr == 0) c = comm_alloc((int)ns);
    uint64_t seq2 = coll_begin(parent, me);
    coll_publish(parent, me, seq2, nullptr, (void*)c, nullptr, nullptr, 0, color, key, 0);
    if (color != MPI_UNDEFINED) {
        if (!coll_wait_slot(parent, leader, seq2)) return MPI_ERR_OTHER;
        c = (CommImpl*)parent->slots[leader].rbuf;
        if (!c) return MPI_ERR_OTHER;
        const int w = world_rank();
        c->world_ranks[nr] = w;
        c->inv[w] = (int32_t)nr;
    }
    if (!comm_barrier(parent)) return MPI_ERR_OTHER;
    *out = (color == MPI_UNDEFINED) ? nullptr : c;
    return MPI_SUCCESS;
}

}  // namespace gpumpi
