// Point-to-point transport and progress engine.
//
// Reference: CudaMPI::isend/irecv/progress/test/wait (gpu_libc/cuda_mpi.cu:65-471).
// There, each rank locks the receiver's mailbox with a system-scope CAS in managed
// memory, scans 16-entry expected/unexpected queues, and one thread memcpys the payload;
// a standard send completes before the receiver has read the sender's buffer
// (cuda_mpi.cu:197-212) and wildcards do not match arrived messages (cuda_mpi.cu:300-304).
//
// Here:
//   * every rank owns an MPSC envelope ring in HBM.  A sender takes a ticket with one
//     agent-scope fetch_add on the ring tail, waits for the slot's sequence word to
//     say "free for my ticket", writes the 128-byte envelope with 8 lanes and publishes
//     it with a release store of seq = ticket+1.  No locks.
//   * the receiver drains its ring in ticket order (per-sender FIFO, so MPI's
//     non-overtaking rule holds), matching each envelope against its posted receives
//     in post order, else parking it in its unexpected list; a new receive searches the
//     unexpected list in arrival order first.  Wildcards match both ways.
//   * payloads <= 80 bytes travel inline in the envelope (eager, send completes at once);
//     larger ones are pulled by the receiver's whole wavefront straight out of the
//     sender's buffer (rendezvous), after which the receiver release-stores the sender's
//     completion word.  Private-memory sources go through a heap bounce buffer.
//   * every wait calls progress() so a rank blocked anywhere still drains its ring, and
//     every wait is bounded (Waiter) so a kernel always terminates.
#pragma once

#include "copy.h"
#include "state.h"

namespace gpumpi {

__shared__ Runtime* s_rt;

__device__ __forceinline__ Runtime* rt() { return s_rt; }
__device__ __forceinline__ int world_rank() {
    return (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
}
__device__ __forceinline__ RankState* my_state() { return &rt()->ranks[world_rank()]; }
__device__ __forceinline__ bool aborted() { return ld_relaxed(&rt()->abort_flag) != 0; }

// Out of line: inlined with a constant message, the character loop unrolls into every
// caller and the progress engine's register count doubled.
__device__ __noinline__ void record_error(int32_t code, int line, int32_t detail, const char* msg) {
    Runtime* R = rt();
    if (lane() == 0) {
        int32_t expected = 0;
        if (cas(&R->err_lock, expected, 1)) {
            R->err.code = code;
            R->err.rank = world_rank();
            R->err.line = line;
            R->err.detail = detail;
            int i = 0;
            for (; msg && msg[i] && i < (int)sizeof(R->err.msg) - 1; i++) R->err.msg[i] = msg[i];
            R->err.msg[i] = 0;
        }
        st_release(&R->abort_flag, 1);
    }
}

__device__ Waiter::Waiter() : start(wall_clock64()), limit(rt()->timeout_ticks), spins(0) {}

__device__ bool Waiter::keep_going(int line) {
    spins++;
    if (aborted()) return false;
    if ((spins & 63) == 0) {
        if (wall_clock64() - start > limit) {
            record_error(kErrTimeout, line, world_rank(), "bounded wait exceeded its time limit");
            return false;
        }
    }
    if (spins > 16) __builtin_amdgcn_s_sleep(2);
    return true;
}

// ------------------------------------------------------------------ envelopes in lanes
// Lanes 0..7 each hold 16 bytes of a 128-byte envelope.
struct EnvHeader {
    int32_t src, tag, ctx;
    uint32_t flags;
    uint64_t bytes, data, done;
};

__device__ __forceinline__ u32x4 env_load(const Envelope* e) {
    u32x4 v = {0u, 0u, 0u, 0u};
    if (lane() < 8) v = ((const volatile u32x4*)e)[lane()];
    return v;
}
__device__ __forceinline__ void env_store(Envelope* e, u32x4 v) {
    if (lane() < 8) ((u32x4*)e)[lane()] = v;
}
__device__ __forceinline__ EnvHeader env_header(u32x4 v) {
    EnvHeader h;
    h.src = (int32_t)readlane(v.y, 0);
    h.tag = (int32_t)readlane(v.z, 0);
    h.ctx = (int32_t)readlane(v.w, 0);
    h.bytes = ((uint64_t)readlane(v.y, 1) << 32) | readlane(v.x, 1);
    h.data = ((uint64_t)readlane(v.w, 1) << 32) | readlane(v.z, 1);
    h.done = ((uint64_t)readlane(v.y, 2) << 32) | readlane(v.x, 2);
    h.flags = readlane(v.z, 2);
    return h;
}

__device__ __forceinline__ Envelope* ring_of(int r) {
    return rt()->rings + (size_t)r * (size_t)rt()->ring_slots;
}

// ------------------------------------------------------------------ requests
__device__ Request* alloc_request(RankState* rs) {
    int n = rs->n_free;
    if (n <= 0) {
        GPUMPI_FAIL(kErrResource, rt()->max_requests, "request pool exhausted (raise max_requests)");
        return nullptr;
    }
    int idx = rs->free_stack[n - 1];
    rs->n_free = n - 1;
    Request* r = &rs->reqs[idx];
    r->index = idx;
    r->state = kReqActive;
    r->done = 0;
    r->st_error = MPI_SUCCESS;
    r->st_bytes = 0;
    r->st_source = MPI_ANY_SOURCE;
    r->st_tag = MPI_ANY_TAG;
    return r;
}

__device__ void release_request(RankState* rs, Request* r) {
    r->state = kReqFree;
    int n = rs->n_free;
    rs->free_stack[n] = r->index;
    rs->n_free = n + 1;
}

__device__ __forceinline__ bool env_matches(const Request* r, int32_t src, int32_t tag, int32_t ctx) {
    return r->ctx == ctx && (r->peer == MPI_ANY_SOURCE || r->peer == src) &&
           (r->tag == MPI_ANY_TAG || r->tag == tag);
}

// Complete receive r from envelope e (a ring slot or an unexpected-list entry).
__device__ void deliver(Request* r, const EnvHeader& h, const Envelope* e) {
    uint64_t n = h.bytes;
    int32_t err = MPI_SUCCESS;
    if (n > r->bytes) {
        err = MPI_ERR_TRUNCATE;
        GPUMPI_FAIL(kErrTruncate, (int32_t)h.bytes, "message longer than the posted receive");
        n = r->bytes;
    }
    if (h.flags & kEnvInline) {
        wave_copy(r->buf, e->inline_data, n);
    } else {
        wave_copy(r->buf, (const void*)h.data, n);
        fence_release();
        if (h.done && lane() == 0) st_relaxed((uint32_t*)h.done, 1u);
        if (h.flags & kEnvBuffered) heap_free(&rt()->heap, &my_state()->cache, (void*)h.data);
    }
    r->st_source = h.src;
    r->st_tag = h.tag;
    r->st_bytes = h.bytes;
    r->st_error = err;
    r->state = kReqDone;
}

// Index into rs->posted of the first posted receive matching the envelope, or -1.
__device__ int find_posted(RankState* rs, int32_t src, int32_t tag, int32_t ctx) {
    const int n = rs->n_posted;
    for (int base = 0; base < n; base += kWave) {
        int i = base + lane();
        bool m = false;
        if (i < n) m = env_matches(&rs->reqs[rs->posted[i]], src, tag, ctx);
        uint64_t b = ballot(m);
        if (b) return base + (int)__builtin_ctzll(b);
    }
    return -1;
}

__device__ void remove_posted(RankState* rs, int pos) {
    const int n = rs->n_posted;
    for (int base = pos; base < n - 1; base += kWave) {
        int i = base + lane();
        int v = 0;
        if (i < n - 1) v = rs->posted[i + 1];
        __builtin_amdgcn_wave_barrier();
        if (i < n - 1) rs->posted[i] = v;
    }
    rs->n_posted = n - 1;
}

constexpr int32_t kTombstone = (int32_t)0x80000000;

// Unexpected list: arrival order, matched entries become tombstones (ctx field).
__device__ int find_unexp(RankState* rs, const Request* r) {
    const int n = rs->n_unexp;
    for (int base = 0; base < n; base += kWave) {
        int i = base + lane();
        bool m = false;
        if (i < n) {
            const Envelope* e = &rs->unexp[i];
            m = e->ctx != kTombstone && env_matches(r, e->src, e->tag, e->ctx);
        }
        uint64_t b = ballot(m);
        if (b) return base + (int)__builtin_ctzll(b);
    }
    return -1;
}

__device__ void compact_unexp(RankState* rs) {
    const int n = rs->n_unexp;
    int w = 0;
    for (int i = 0; i < n; i++) {
        const Envelope* e = &rs->unexp[i];
        int32_t ctx = e->ctx;
        if (ctx == kTombstone) continue;
        if (w != i) {
            u32x4 v = env_load(e);
            env_store(&rs->unexp[w], v);
        }
        w++;
    }
    rs->n_unexp = w;
}

__device__ void pop_unexp(RankState* rs, int i) {
    rs->unexp[i].ctx = kTombstone;
    int n = rs->n_unexp;
    while (n > 0 && rs->unexp[n - 1].ctx == kTombstone) n--;
    rs->n_unexp = n;
}

// ------------------------------------------------------------------ progress
// Drain the calling rank's envelope ring.  Returns the number of envelopes consumed.
__device__ int progress() {
    Runtime* R = rt();
    RankState* rs = my_state();
    Envelope* ring = ring_of(world_rank());
    const uint32_t N = (uint32_t)R->ring_slots;
    uint64_t h = rs->ring_head;
    int consumed = 0;
    for (uint32_t guard = 0; guard < N; guard++) {
        Envelope* e = &ring[h % N];
        // poll with a plain coherent load; the acquire fence (an L2 invalidate on gfx950)
        // only once an envelope is there
        uint32_t seq = 0;
        if (lane() == 0) seq = ld_relaxed(&e->seq);
        seq = bcast(seq);
        if (seq != (uint32_t)(h + 1)) break;
        fence_acquire();
        u32x4 v = env_load(e);
        EnvHeader hd = env_header(v);
        int p = find_posted(rs, hd.src, hd.tag, hd.ctx);
        if (p >= 0) {
            Request* r = &rs->reqs[rs->posted[p]];
            remove_posted(rs, p);
            deliver(r, hd, e);
        } else {
            if (rs->n_unexp >= R->max_unexpected) compact_unexp(rs);
            if (rs->n_unexp >= R->max_unexpected) break;  // leave it in the ring
            env_store(&rs->unexp[rs->n_unexp], v);
            rs->n_unexp = rs->n_unexp + 1;
        }
        // the slot goes back to its producers once our reads of it are done; what we
        // wrote (the receive buffer, the unexpected queue) is ours and needs no write-back
        wait_mem_done();
        if (lane() == 0) st_relaxed(&e->seq, (uint32_t)(h + N));
        h++;
        rs->ring_head = h;
        consumed++;
    }
    return consumed;
}

// Publish an envelope (lanes 0..7 of v) in dest's ring.  `self_contained`: the envelope
// carries everything the receiver will read (an inline payload), so it is written with
// agent-scope atomic stores, which reach the other XCDs without an L2 write-back, and only
// their completion is awaited; otherwise the receiver will also read memory this wave
// wrote before (the payload it pulls), and a release fence writes the L2 back first.
__device__ bool push_envelope(int dest, u32x4 v, bool self_contained) {
    Runtime* R = rt();
    const uint32_t N = (uint32_t)R->ring_slots;
    uint64_t t = 0;
    if (lane() == 0) t = fetch_add_relaxed(&R->mailbox[dest].tail, (uint64_t)1);
    t = bcast(t);
    Envelope* e = &ring_of(dest)[t % N];
    Waiter w;
    for (;;) {
        uint32_t s = 0;
        if (lane() == 0) s = ld_relaxed(&e->seq);
        s = bcast(s);
        if (s == (uint32_t)t) break;
        progress();  // drain our own ring so two full rings cannot deadlock
        if (!w.keep_going(__LINE__)) return false;
    }
    // no acquire fence: the consumer's reads of this slot completed before it handed the
    // slot back (wait_mem_done in progress), and we only write the slot
    // lane 0 must not overwrite seq with the plain store; it writes words 1..3 only
    if (self_contained) {
        if (lane() >= 1 && lane() < 8) {
            uint64_t* q = (uint64_t*)((u32x4*)e + lane());
            st_relaxed(q, ((uint64_t)v.y << 32) | v.x);
            st_relaxed(q + 1, ((uint64_t)v.w << 32) | v.z);
        }
        if (lane() == 0) {
            st_relaxed((uint32_t*)e + 1, v.y);
            st_relaxed((uint64_t*)e + 1, ((uint64_t)v.w << 32) | v.z);
        }
        wait_mem_done();  // the sc1 stores above have reached memory before seq is written
    } else {
        if (lane() >= 1 && lane() < 8) ((u32x4*)e)[lane()] = v;
        if (lane() == 0) {
            volatile uint32_t* w0 = (volatile uint32_t*)e;
            w0[1] = v.y;
            w0[2] = v.z;
            w0[3] = v.w;
        }
        fence_release();  // with the fence, a relaxed store publishes (one L2 write-back, not two)
    }
    if (lane() == 0) st_relaxed(&e->seq, (uint32_t)(t + 1));
    return true;
}

// ------------------------------------------------------------------ isend / irecv
__device__ Request* isend(const void* buf, uint64_t bytes, int dest_world, int tag, int ctx,
                          const CommImpl* comm) {
    Runtime* R = rt();
    RankState* rs = my_state();
    if (dest_world < 0 || dest_world >= R->nranks) {
        GPUMPI_FAIL(kErrArgument, dest_world, "send to an invalid rank");
        return nullptr;
    }
    Request* r = alloc_request(rs);
    if (!r) return nullptr;
    r->kind = kReqSend;
    r->peer = dest_world;
    r->tag = tag;
    r->ctx = ctx;
    r->buf = (void*)buf;
    r->bytes = bytes;
    r->comm = comm;

    const int l = lane();
    uint32_t flags;
    uint64_t data = 0, done = 0;
    u32x4 v = {0u, 0u, 0u, 0u};
    if (bytes <= (uint64_t)kInlineMax) {
        flags = kEnvInline;
        if (l >= 3 && l < 8) {
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            const uint8_t* b = (const uint8_t*)buf;
            for (int j = 0; j < 16; j++) {
                uint64_t idx = (uint64_t)(16 * (l - 3) + j);
                if (idx < bytes) w[j >> 2] |= (uint32_t)b[idx] << (8 * (j & 3));
            }
            v.x = w[0]; v.y = w[1]; v.z = w[2]; v.w = w[3];
        }
        r->state = kReqDone;
    } else if (is_private(buf) || is_shared(buf)) {
        void* bounce = heap_alloc(&R->heap, &rs->cache, bytes);
        if (!bounce) {
            GPUMPI_FAIL(kErrResource, (int32_t)bytes, "heap exhausted for a send bounce buffer");
            release_request(rs, r);
            return nullptr;
        }
        wave_copy(bounce, buf, bytes);
        flags = kEnvBuffered;
        data = (uint64_t)(uintptr_t)bounce;
        r->state = kReqDone;
    } else {
        flags = 0;
        data = (uint64_t)(uintptr_t)buf;
        done = (uint64_t)(uintptr_t)&r->done;
    }
    if (l == 0) { v.x = 0u; v.y = (uint32_t)world_rank(); v.z = (uint32_t)tag; v.w = (uint32_t)ctx; }
    if (l == 1) { v.x = (uint32_t)bytes; v.y = (uint32_t)(bytes >> 32);
                  v.z = (uint32_t)data; v.w = (uint32_t)(data >> 32); }
    if (l == 2) { v.x = (uint32_t)done; v.y = (uint32_t)(done >> 32); v.z = flags; v.w = 0u; }
    if (!push_envelope(dest_world, v, flags == kEnvInline)) return r;
    return r;
}

__device__ Request* irecv(void* buf, uint64_t bytes, int src_world, int tag, int ctx,
                          const CommImpl* comm) {
    RankState* rs = my_state();
    Request* r = alloc_request(rs);
    if (!r) return nullptr;
    r->kind = kReqRecv;
    r->peer = src_world;
    r->tag = tag;
    r->ctx = ctx;
    r->buf = buf;
    r->bytes = bytes;
    r->comm = comm;
    int u = find_unexp(rs, r);
    if (u >= 0) {
        Envelope* e = &rs->unexp[u];
        EnvHeader hd = env_header(env_load(e));
        deliver(r, hd, e);
        pop_unexp(rs, u);
        return r;
    }
    if (rs->n_posted >= rt()->max_requests) {
        GPUMPI_FAIL(kErrResource, rs->n_posted, "posted receive list full");
        return r;
    }
    rs->posted[rs->n_posted] = r->index;
    rs->n_posted = rs->n_posted + 1;
    progress();
    return r;
}

__device__ bool test_request(Request* r) {
    if (r->state == kReqDone) return true;
    if (r->kind == kReqSend) {
        uint32_t d = 0;
        if (lane() == 0) d = ld_relaxed(&r->done);
        d = bcast(d);
        if (d) {
            fence_acquire();
            r->state = kReqDone;
            return true;
        }
    }
    progress();
    return r->state == kReqDone;
}

// Returns false when the wait was cut short by an error or abort.
__device__ bool wait_request(Request* r) {
    Waiter w;
    while (!test_request(r)) {
        if (!w.keep_going(__LINE__)) return false;
    }
    return true;
}

}  // namespace gpumpi
