// Lock-free size-class device heap (gpu_malloc / gpu_free).
//
// Replaces the reference's coalescing wrappers around the CUDA device heap
// (gpu_malloc/warp_malloc.cu:159-506, dynamic_allocator.cu:9-364,
// no_header_malloc.cu:79-211, dyn_malloc.cu:17-57) and the stdlib entry points that
// dispatch to them (gpu_libc/stdlib.cu:70-269).  Those coalesce the calls of the 32
// ranks of a warp into one CUDA-heap malloc; here a rank is a whole wavefront, so a
// call is already one request, and the design goal is instead no global lock and no
// CUDA-style heap underneath:
//
//   * one arena in HBM, carved into 64 KiB pages from an atomic page bump;
//   * power-of-two classes of 16 B << c; classes below 64 KiB share pages of one class
//     (page_class[] maps a page to its class, so free() needs no block header);
//     classes of 64 KiB and up take a run of whole pages;
//   * per class, nshards Treiber stacks whose 64-bit heads pack a 40-bit block offset
//     (16 B units) with a 24-bit ABA tag; rank r pushes to and pops from shard
//     r % nshards and turns to the other shards only once the page bump is used up (one
//     shared stack per class collapses under contention: with 1024 ranks a CAS round
//     costs hundreds of microseconds).  nshards = one per 64 pages of the arena, at most
//     kHeapShards: every shard holds a partly used page per small class, and on a small
//     arena too many shards would pin the pages that 64 KiB blocks need;
//   * a refill carves a fresh page, links its blocks with all 64 lanes and pushes the
//     chain to the rank's shard with one CAS;
//   * when the page bump is used up, a refill takes a free block of the smallest
//     larger page-run class instead and splits it buddy-style (the upper halves go to
//     the classes in between), so freed large blocks serve later smaller requests;
//   * a small per-rank cache per class absorbs alloc/free churn without atomics;
//   * each stack head sits in its own 128-byte line, and the statistics
//     (bytes_in_use, peak, allocation count) in another: device-scope atomics on one line
//     from every XCD serialise at the memory side;
//   * a rank batches its statistics updates and adds them to the shared counters every
//     kHeapAcctOps operations or kHeapAcctBytes bytes (and when it ends), so a
//     malloc/free served by the cache touches no shared word at all.
#pragma once

#include "common.h"

namespace gpumpi {

constexpr int kHeapClasses = 36;          // 16 B .. 512 GiB
constexpr int kHeapSmallClasses = 12;     // 16 B .. 32 KiB share 64 KiB pages
constexpr int kHeapPageShift = 16;
constexpr uint64_t kHeapPage = 1ull << kHeapPageShift;
constexpr int kHeapCacheDepth = 8;
constexpr uint64_t kHeapOffMask = (1ull << 40) - 1;
constexpr uint64_t kHeapEmpty = kHeapOffMask;  // offset value meaning "empty stack"
constexpr int kHeapHeadStride = 16;            // uint64 words per stack head: 128 bytes
constexpr int kHeapShards = 128;               // stacks per class at most
constexpr uint64_t kHeapPagesPerShard = 64;
constexpr int kHeapAcctOps = 64;               // statistics batch: operations ...
constexpr int64_t kHeapAcctBytes = 1ll << 20;  // ... or net bytes of one rank

struct Heap {
    char* base;
    uint64_t bytes;
    uint64_t npages;
    uint8_t* page_class;             // [npages], 0xFF = untouched
    uint64_t page_bump;              // atomic
    uint64_t* heads;                 // [kHeapClasses][kHeapShards][kHeapHeadStride]
    int32_t nshards;                 // stacks in use per class (heap_shard_count)
    int32_t pad;
    alignas(128) int64_t bytes_in_use;  // atomic: sum of class sizes of live blocks
    int64_t peak_in_use;             // atomic max (sampled when a rank adds its batch)
    int64_t alloc_count;             // atomic, statistics
    int64_t refill_count;            // atomic, statistics
};

__host__ __device__ __forceinline__ int32_t heap_shard_count(uint64_t npages) {
    const uint64_t n = npages / kHeapPagesPerShard;
    return (int32_t)(n < 1 ? 1 : n > (uint64_t)kHeapShards ? (uint64_t)kHeapShards : n);
}

__host__ __device__ __forceinline__ uint64_t* heap_head(Heap* h, int c, int shard) {
    return &h->heads[((size_t)c * kHeapShards + (size_t)shard) * kHeapHeadStride];
}

// Per-rank block cache and statistics batch (live in the rank's private state).
struct HeapCache {
    int8_t n[kHeapSmallClasses];
    int32_t shard;                   // the rank's stack shard: world rank % Heap::nshards
    int32_t acct_ops;                // operations since the batch was added
    int32_t pad;
    int64_t acct_bytes;              // net bytes allocated since then
    int64_t acct_allocs;             // allocations since then
    uint64_t blk[kHeapSmallClasses][kHeapCacheDepth];  // 16-byte-unit offsets
};

// Empties a rank's cache and statistics batch (its shard stays).
__host__ __device__ __forceinline__ void heap_cache_clear(HeapCache* c) {
    for (int k = 0; k < kHeapSmallClasses; k++) c->n[k] = 0;
    c->acct_ops = 0;
    c->acct_bytes = 0;
    c->acct_allocs = 0;
}

__device__ __forceinline__ int heap_class_of(uint64_t n) {
    if (n <= 16) return 0;
    int c = 64 - __clzll(n - 1) - 4;  // ceil(log2(n)) - 4
    return c;
}
__device__ __forceinline__ uint64_t heap_class_bytes(int c) { return 16ull << c; }

__device__ __forceinline__ uint64_t heap_pack(uint64_t off, uint64_t tag) {
    return (off & kHeapOffMask) | (tag << 40);
}

// Rank-private state (the cache) is written by every lane with the same value, so each
// lane reads back its own store; shared words are touched by lane 0 with atomics.

__device__ __forceinline__ int heap_shard_of(const HeapCache* cache) {
    return cache ? cache->shard : 0;
}

// Lane-0 Treiber push of a pre-linked chain first..last (offsets in 16 B units).
__device__ __forceinline__ void heap_push_chain(Heap* h, int c, int shard, uint64_t first,
                                                uint64_t last) {
    uint64_t* head = heap_head(h, c, shard);
    uint64_t old = ld_acquire(head);
    for (;;) {
        *(volatile uint64_t*)(h->base + (last << 4)) = old & kHeapOffMask;
        uint64_t nw = heap_pack(first, (old >> 40) + 1);
        if (cas(head, old, nw)) break;
    }
}

// Lane-0 Treiber pop; returns kHeapEmpty when the stack is empty.
__device__ __forceinline__ uint64_t heap_pop(Heap* h, int c, int shard) {
    uint64_t* head = heap_head(h, c, shard);
    uint64_t old = ld_acquire(head);
    for (;;) {
        uint64_t off = old & kHeapOffMask;
        if (off == kHeapEmpty) return kHeapEmpty;
        uint64_t next = *(volatile uint64_t*)(h->base + (off << 4));
        uint64_t nw = heap_pack(next, (old >> 40) + 1);
        if (cas(head, old, nw)) return off;
    }
}

// Lane-0 pop from the shard `home`, then from every other shard of class c.
__device__ __forceinline__ uint64_t heap_pop_any(Heap* h, int c, int home) {
    const int n = h->nshards;
    for (int k = 0; k < n; k++) {
        const uint64_t off = heap_pop(h, c, (home + k) % n);
        if (off != kHeapEmpty) return off;
    }
    return kHeapEmpty;
}

// A run of 2^(cp - kHeapSmallClasses) fresh pages from the page bump (wave-collective;
// the bump only grows, pages past the end are never handed out), or ~0 when it is used up.
__device__ uint64_t heap_bump_pages(Heap* h, int cp) {
    const uint64_t pages = 1ull << (cp - kHeapSmallClasses);
    uint64_t first = 0;
    if (lane() == 0) first = fetch_add_relaxed(&h->page_bump, pages);
    first = bcast(first);
    return first + pages <= h->npages ? first : ~0ull;
}

// With the bump used up: the first free block of class kstart or larger in any shard,
// split down to page-run class cp -- each split pushes the upper half to the shard
// `home`.  Returns the first page, or ~0 when nothing is left.
__device__ uint64_t heap_split_pages(Heap* h, int cp, int kstart, int home) {
    for (int k = kstart; k < kHeapClasses; k++) {
        uint64_t off = kHeapEmpty;
        if (lane() == 0) off = heap_pop_any(h, k, home);
        off = bcast(off);
        if (off == kHeapEmpty) continue;
        const uint64_t page = (off << 4) >> kHeapPageShift;
        for (int j = k - 1; j >= cp; j--) {
            const uint64_t upper = page + (1ull << (j - kHeapSmallClasses));
            h->page_class[upper] = (uint8_t)j;
            fence_release();
            const uint64_t uoff = (upper << kHeapPageShift) >> 4;
            if (lane() == 0) heap_push_chain(h, j, home, uoff, uoff);
        }
        return page;
    }
    return ~0ull;
}

// Wave-collective allocation.  Returns nullptr when the arena is exhausted.
__device__ void* heap_alloc(Heap* h, HeapCache* cache, uint64_t n);
__device__ void heap_free(Heap* h, HeapCache* cache, void* p);
__device__ uint64_t heap_block_bytes(Heap* h, const void* p);

// Adds a batch of statistics to the shared counters (lane 0).
__device__ __forceinline__ void heap_account_add(Heap* h, int64_t bytes, int64_t allocs) {
    if (lane() == 0) {
        if (bytes != 0) {
            const int64_t now = fetch_add_relaxed(&h->bytes_in_use, bytes) + bytes;
            if (bytes > 0)
                __hip_atomic_fetch_max(&h->peak_in_use, now, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
        if (allocs) fetch_add_relaxed(&h->alloc_count, allocs);
    }
}

// One allocation (delta > 0) or free (delta < 0) of a class-rounded size: into the
// rank's batch, which goes to the shared counters every kHeapAcctOps operations or once
// it holds kHeapAcctBytes net bytes.  Without a cache (job arenas) it goes at once.
__device__ __forceinline__ void heap_account(Heap* h, HeapCache* cache, int64_t delta) {
    const int64_t allocs = delta > 0 ? 1 : 0;
    if (!cache) {
        heap_account_add(h, delta, allocs);
        return;
    }
    // rank-private state: every lane computes and writes the same values
    const int64_t b = cache->acct_bytes + delta;
    const int64_t a = cache->acct_allocs + allocs;
    const int32_t n = cache->acct_ops + 1;
    if (n >= kHeapAcctOps || b >= kHeapAcctBytes || b <= -kHeapAcctBytes) {
        heap_account_add(h, b, a);
        cache->acct_bytes = 0;
        cache->acct_allocs = 0;
        cache->acct_ops = 0;
    } else {
        cache->acct_bytes = b;
        cache->acct_allocs = a;
        cache->acct_ops = n;
    }
}

// Out of line (as heap_free): inlined into the progress engine and the collectives, the
// stack and cache paths added ~70 registers per lane to every rank program.
__device__ __noinline__ void* heap_alloc(Heap* h, HeapCache* cache, uint64_t n) {
    if (n == 0) return nullptr;
    const int c = heap_class_of(n);
    if (c >= kHeapClasses) return nullptr;
    const uint64_t bsz = heap_class_bytes(c);
    // 1. per-rank cache (wave-uniform state, no atomics)
    if (cache && c < kHeapSmallClasses && cache->n[c] > 0) {
        int k = cache->n[c] - 1;
        uint64_t off = cache->blk[c][k];
        cache->n[c] = (int8_t)k;  // rank-private state: every lane writes the same value
        heap_account(h, cache, (int64_t)bsz);
        return h->base + (off << 4);
    }
    // 2. the rank's shard of the class's free stacks
    const int home = heap_shard_of(cache);
    uint64_t off = kHeapEmpty;
    if (lane() == 0) off = heap_pop(h, c, home);
    off = bcast(off);
    if (off != kHeapEmpty) {
        heap_account(h, cache, (int64_t)bsz);
        return h->base + (off << 4);
    }
    // 3. a run of whole pages fresh from the bump; once it is used up, a free block of
    //    the class in another shard, else a larger free block split down
    const int cp = c < kHeapSmallClasses ? kHeapSmallClasses : c;
    uint64_t first = heap_bump_pages(h, cp);
    if (first == ~0ull) {
        if (lane() == 0) off = heap_pop_any(h, c, home);
        off = bcast(off);
        if (off != kHeapEmpty) {
            heap_account(h, cache, (int64_t)bsz);
            return h->base + (off << 4);
        }
        first = heap_split_pages(h, cp, c < kHeapSmallClasses ? cp : cp + 1, home);
    }
    if (first == ~0ull) return nullptr;
    h->page_class[first] = (uint8_t)c;
    if (lane() == 0) fetch_add_relaxed(&h->refill_count, (int64_t)1);
    const uint64_t page_off = (first << kHeapPageShift) >> 4;  // 16 B units
    if (c < kHeapSmallClasses) {
        const uint64_t nb = kHeapPage / bsz;
        const uint64_t step = bsz >> 4;
        if (nb > 1) {
            // block 0 goes to the caller; link blocks 1..nb-1 and push them at once
            for (uint64_t b = 1 + (uint64_t)lane(); b < nb; b += kWave) {
                uint64_t nxt = (b + 1 < nb) ? page_off + (b + 1) * step : kHeapEmpty;
                *(uint64_t*)(h->base + ((page_off + b * step) << 4)) = nxt;
            }
            fence_release();
            if (lane() == 0)
                heap_push_chain(h, c, home, page_off + step, page_off + (nb - 1) * step);
        }
    }
    heap_account(h, cache, (int64_t)bsz);
    return h->base + (page_off << 4);
}

__device__ uint64_t heap_block_bytes(Heap* h, const void* p) {
    uint64_t rel = (uint64_t)((const char*)p - h->base);
    uint64_t page = rel >> kHeapPageShift;
    uint8_t c = h->page_class[page];
    return heap_class_bytes(c);
}

__device__ __noinline__ void heap_free(Heap* h, HeapCache* cache, void* p) {
    if (!p) return;
    const char* cp = (const char*)p;
    if (cp < h->base || cp >= h->base + h->bytes) {
        GPUMPI_FAIL(kErrArgument, 0, "gpu_free: pointer outside the device heap");
        return;
    }
    const uint64_t rel = (uint64_t)(cp - h->base);
    const uint64_t page = rel >> kHeapPageShift;
    const uint8_t cls = h->page_class[page];
    if (cls >= kHeapClasses) {
        GPUMPI_FAIL(kErrArgument, 0, "gpu_free: pointer was not returned by gpu_malloc");
        return;
    }
    const int c = cls;
    const uint64_t bsz = heap_class_bytes(c);
    if ((c < kHeapSmallClasses && ((rel & (kHeapPage - 1)) % bsz) != 0) ||
        (c >= kHeapSmallClasses && (rel & (kHeapPage - 1)) != 0)) {
        GPUMPI_FAIL(kErrArgument, 0, "gpu_free: pointer is not the start of a block");
        return;
    }
    const uint64_t off = rel >> 4;
    heap_account(h, cache, -(int64_t)bsz);
    if (cache && c < kHeapSmallClasses && cache->n[c] < kHeapCacheDepth) {
        int k = cache->n[c];
        cache->blk[c][k] = off;
        cache->n[c] = (int8_t)(k + 1);
        return;
    }
    if (lane() == 0) heap_push_chain(h, c, heap_shard_of(cache), off, off);
}

// Return every cached block of a rank to the shared stacks and add its statistics batch
// to the shared counters (rank exit / job end).
__device__ void heap_flush_cache(Heap* h, HeapCache* cache) {
    if (!cache) return;
    heap_account_add(h, cache->acct_bytes, cache->acct_allocs);
    cache->acct_bytes = 0;
    cache->acct_allocs = 0;
    cache->acct_ops = 0;
    for (int c = 0; c < kHeapSmallClasses; c++) {
        int n = cache->n[c];
        for (int k = 0; k < n; k++) {
            uint64_t off = cache->blk[c][k];
            if (lane() == 0) heap_push_chain(h, c, cache->shard, off, off);
        }
        cache->n[c] = 0;
    }
}

}  // namespace gpumpi
