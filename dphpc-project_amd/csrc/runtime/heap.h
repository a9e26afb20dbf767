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
//   * one Treiber stack per class whose 64-bit head packs a 40-bit block offset (16 B
//     units) with a 24-bit ABA tag;
//   * a refill carves a fresh page, links its blocks with all 64 lanes and pushes the
//     chain with one CAS;
//   * when the page bump is used up, a refill takes a free block of the smallest
//     larger page-run class instead and splits it buddy-style (the upper halves go to
//     the classes in between), so freed large blocks serve later smaller requests;
//   * a small per-rank cache per class absorbs alloc/free churn without atomics;
//   * bytes_in_use is the headroom counter that the serving loop admits against.
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

struct Heap {
    char* base;
    uint64_t bytes;
    uint64_t npages;
    uint8_t* page_class;             // [npages], 0xFF = untouched
    uint64_t page_bump;              // atomic
    uint64_t free_head[kHeapClasses];
    int64_t bytes_in_use;            // atomic: sum of class sizes of live blocks
    int64_t peak_in_use;             // atomic max
    int64_t alloc_count;             // atomic, statistics
    int64_t refill_count;            // atomic, statistics
};

// Per-rank block cache (lives in the rank's private state).
struct HeapCache {
    int8_t n[kHeapSmallClasses];
    uint32_t pad;
    uint64_t blk[kHeapSmallClasses][kHeapCacheDepth];  // 16-byte-unit offsets
};

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

// Lane-0 Treiber push of a pre-linked chain first..last (offsets in 16 B units).
__device__ __forceinline__ void heap_push_chain(Heap* h, int c, uint64_t first, uint64_t last) {
    uint64_t* head = &h->free_head[c];
    uint64_t old = ld_acquire(head);
    for (;;) {
        *(volatile uint64_t*)(h->base + (last << 4)) = old & kHeapOffMask;
        uint64_t nw = heap_pack(first, (old >> 40) + 1);
        if (cas(head, old, nw)) break;
    }
}

// Lane-0 Treiber pop; returns kHeapEmpty when the stack is empty.
__device__ __forceinline__ uint64_t heap_pop(Heap* h, int c) {
    uint64_t* head = &h->free_head[c];
    uint64_t old = ld_acquire(head);
    for (;;) {
        uint64_t off = old & kHeapOffMask;
        if (off == kHeapEmpty) return kHeapEmpty;
        uint64_t next = *(volatile uint64_t*)(h->base + (off << 4));
        uint64_t nw = heap_pack(next, (old >> 40) + 1);
        if (cas(head, old, nw)) return off;
    }
}

// A run of 2^(cp - kHeapSmallClasses) pages (wave-collective): from the page bump while
// it lasts (the bump only grows; pages past the end are never handed out), else the
// first free block of class cp or larger, split down to cp -- each split pushes the
// upper half to its own class.  Returns the first page, or ~0 when nothing is left.
__device__ uint64_t heap_take_pages(Heap* h, int cp) {
    const uint64_t pages = 1ull << (cp - kHeapSmallClasses);
    uint64_t first = 0;
    if (lane() == 0) first = fetch_add_relaxed(&h->page_bump, pages);
    first = bcast(first);
    if (first + pages <= h->npages) return first;
    for (int k = cp; k < kHeapClasses; k++) {
        uint64_t off = kHeapEmpty;
        if (lane() == 0) off = heap_pop(h, k);
        off = bcast(off);
        if (off == kHeapEmpty) continue;
        const uint64_t page = (off << 4) >> kHeapPageShift;
        for (int j = k - 1; j >= cp; j--) {
            const uint64_t upper = page + (1ull << (j - kHeapSmallClasses));
            h->page_class[upper] = (uint8_t)j;
            fence_release();
            const uint64_t uoff = (upper << kHeapPageShift) >> 4;
            if (lane() == 0) heap_push_chain(h, j, uoff, uoff);
        }
        return page;
    }
    return ~0ull;
}

// Wave-collective allocation.  Returns nullptr when the arena is exhausted.
__device__ void* heap_alloc(Heap* h, HeapCache* cache, uint64_t n);
__device__ void heap_free(Heap* h, HeapCache* cache, void* p);
__device__ uint64_t heap_block_bytes(Heap* h, const void* p);

__device__ __forceinline__ void heap_account(Heap* h, int64_t delta) {
    if (lane() == 0) {
        int64_t now = fetch_add_relaxed(&h->bytes_in_use, delta) + delta;
        if (delta > 0) {
            __hip_atomic_fetch_max(&h->peak_in_use, now, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            fetch_add_relaxed(&h->alloc_count, (int64_t)1);
        }
    }
}

__device__ void* heap_alloc(Heap* h, HeapCache* cache, uint64_t n) {
    if (n == 0) return nullptr;
    const int c = heap_class_of(n);
    if (c >= kHeapClasses) return nullptr;
    const uint64_t bsz = heap_class_bytes(c);
    // 1. per-rank cache (wave-uniform state, no atomics)
    if (cache && c < kHeapSmallClasses && cache->n[c] > 0) {
        int k = cache->n[c] - 1;
        uint64_t off = cache->blk[c][k];
        cache->n[c] = (int8_t)k;  // rank-private state: every lane writes the same value
        heap_account(h, (int64_t)bsz);
        return h->base + (off << 4);
    }
    // 2. global free list
    uint64_t off = kHeapEmpty;
    if (lane() == 0) off = heap_pop(h, c);
    off = bcast(off);
    if (off != kHeapEmpty) {
        heap_account(h, (int64_t)bsz);
        return h->base + (off << 4);
    }
    // 3. a run of whole pages: fresh from the bump, else split a larger free block
    const int cp = c < kHeapSmallClasses ? kHeapSmallClasses : c;
    const uint64_t first = heap_take_pages(h, cp);
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
                heap_push_chain(h, c, page_off + step, page_off + (nb - 1) * step);
        }
    }
    heap_account(h, (int64_t)bsz);
    return h->base + (page_off << 4);
}

__device__ uint64_t heap_block_bytes(Heap* h, const void* p) {
    uint64_t rel = (uint64_t)((const char*)p - h->base);
    uint64_t page = rel >> kHeapPageShift;
    uint8_t c = h->page_class[page];
    return heap_class_bytes(c);
}

__device__ void heap_free(Heap* h, HeapCache* cache, void* p) {
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
    heap_account(h, -(int64_t)bsz);
    if (cache && c < kHeapSmallClasses && cache->n[c] < kHeapCacheDepth) {
        int k = cache->n[c];
        cache->blk[c][k] = off;
        cache->n[c] = (int8_t)(k + 1);
        return;
    }
    if (lane() == 0) heap_push_chain(h, c, off, off);
}

// Return every cached block of a rank to the shared stacks (rank exit / job end).
__device__ void heap_flush_cache(Heap* h, HeapCache* cache) {
    if (!cache) return;
    for (int c = 0; c < kHeapSmallClasses; c++) {
        int n = cache->n[c];
        for (int k = 0; k < n; k++) {
            uint64_t off = cache->blk[c][k];
            if (lane() == 0) heap_push_chain(h, c, off, off);
        }
        cache->n[c] = 0;
    }
}

}  // namespace gpumpi
