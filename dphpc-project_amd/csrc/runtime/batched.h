// Batched MPI_Alltoallv step: three kernels replayed as one hipGraph per step.
//
// Semantics: gpu_mpi/mpi.cu:363-408 (block s of rank d's receive buffer <- block d of
// rank s's send buffer, counts/displacements in elements), preceded by the Alltoall of
// the counts and the packed receive displacements that the NPB IS caller computes
// between the two calls (toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c:758-781).
//
// The reference runs this as P-1 Isend + P-1 Irecv per rank, each payload copied by one
// thread.  Here the step is decoupled from ranks and split at the only global dependency
// (the receive displacements need every sender's count):
//
//  plan   one workgroup per destination row d: recv_counts[d][s] = send_counts[s][d],
//         recv_displs = exclusive scan over s, and per segment (d, s) its end offset in
//         the row and its source address.
//  scan   one workgroup: chunks per row (a row's destination lines, counted from the
//         128-byte block holding its first byte, cut into chunks of chunk_lines 16-byte
//         lines) and their exclusive prefix.
//  copy   one workgroup per chunk (grid-stride when the grid is smaller): wave 0 finds
//         the chunk's row and first segment (64-ary searches over L2-resident tables)
//         and stages a window of 64 segments in LDS; each wave then copies a contiguous
//         run of 64*U lines slice by slice, keeping a wave-uniform cursor into the window
//         (it only moves forward; kept in SGPRs) and, when the slice does not lie in one
//         entry, choosing each lane's segment among the cursor's next four with compares.
//         The loop is bound by instruction issue, so the common cases are uniform paths:
//         a slice inside one segment costs one source offset from SGPRs, a slice across
//         one boundary a per-lane choice between two.  A line inside one segment is
//         one unaligned 16-byte load and one aligned 16-byte store; a line that straddles
//         segments is four dword loads (issued with the other lines' loads) and one
//         16-byte store -- or, in
//         deferred mode (a tuning choice that pays when segments are short, as in skewed
//         key distributions), left to one thread per segment run by the grid's first
//         blocks, concurrently with the streaming blocks.  Lines at a row's ends, and
//         dwords that are split or unaligned at the source, go byte by byte; bytes outside
//         the row are never written.
//
// Why not one persistent kernel with a grid barrier: on gfx950 an agent-scope release or
// acquire writes back or invalidates the XCD's L2 (buffer_wbl2 / buffer_inv sc1); a grid
// barrier across 256-2048 workgroups measured 14-150 us, a kernel boundary inside a graph
// a few us.  And a copy grid of many short workgroups streams faster than a persistent
// grid (5.9 vs 5.0-5.2 TB/s for a plain 512 MiB copy on MI355X).
//
// HBM traffic per step = payload read once + payload written once + O(P^2) tables
// (counts and displacements read, 4 + 4 + 8 + 8 bytes per segment written and read).
#pragma once

#include "copy.h"

namespace gpumpi {

constexpr int kA2avWindow = 64;
constexpr int kA2avCopyThreads = 256;

struct A2avDev {
    int32_t P;
    int32_t elem;
    uint32_t chunk_lines;   // destination lines per chunk = kA2avCopyThreads * U
    int32_t straddle_blocks;  // deferred mode: leading copy blocks running the straddle threads
    char* const* sendbufs;
    const int32_t* scounts;
    const int32_t* sdispls;
    char* const* recvbufs;
    int32_t* rcounts;
    int32_t* rdispls;
    uint64_t* seg_end;      // [P*P] end byte offset of segment (d, s) within row d
    uint64_t* seg_src;      // [P*P] address of the first source byte of segment (d, s)
    uint64_t* row_bytes;    // [P]
    uint64_t* chunk_off;    // [P+1] exclusive prefix of chunks per row; [P] = total chunks
    int64_t* total_bytes;   // [1]
    uint32_t* plan_done;    // [1] plan workgroups finished (fused plan; reset by the last one)
};

__device__ __forceinline__ int64_t wave_excl_scan(int64_t v, int64_t* total) {
    int64_t inc = wave_incl_scan(v);
    *total = __shfl(inc, 63);
    return inc - v;
}

// Largest i in [0, n) with a[i] <= x (a non-decreasing, a[0] <= x); wave-uniform,
// 64-ary: each step compares 64 evenly spaced entries.
__device__ __forceinline__ int64_t wave_search_le(const uint64_t* a, int64_t n, uint64_t x) {
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const int64_t step = (hi - lo + 63) / 64;
        const int64_t idx = lo + (int64_t)lane() * step;
        const bool le = idx < hi && a[idx] <= x;
        const uint64_t m = ballot(le);
        const int L = m ? 63 - __builtin_clzll(m) : 0;
        const int64_t nlo = lo + (int64_t)L * step;
        const int64_t nhi = nlo + step;
        lo = nlo;
        hi = nhi < hi ? nhi : hi;
    }
    return lo;
}

// ------------------------------------------------------------------------------ plan
// One workgroup of four wavefronts per destination row d; wave w owns a contiguous
// quarter of the sources.  The first 256 sources of a quarter (all of them up to P = 1024)
// are loaded once -- counts, displacements and send buffers in one round of loads -- and
// kept in registers: the quarters' count totals meet in LDS to give each quarter's
// starting displacement, then each wave scans and writes.  Sources past 256 per quarter
// are summed in a first pass and reloaded in the second.
// COHERENT: the row's byte count is stored with an agent-scope atomic store (written
// through to memory), for a reader in the same kernel on another XCD.
template <bool COHERENT = false>
__device__ __forceinline__ void a2av_plan_row(const A2avDev& p, int64_t d, int64_t* s_tot) {
    const int P = p.P;
    const int l = lane();
    const int w = threadIdx.x >> 6;
    const int Q = (((P + 3) / 4) + kWave - 1) / kWave * kWave;
    const int sb = w * Q;
    const int se = sb + Q < P ? sb + Q : P;
    int64_t c[4], sd[4];
    const char* sbuf[4];
    int64_t part = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int s = sb + q * kWave + l;
        const bool valid = s < se;
        c[q] = valid ? p.scounts[(int64_t)s * P + d] : 0;
        sd[q] = valid ? p.sdispls[(int64_t)s * P + d] : 0;
        sbuf[q] = valid ? p.sendbufs[s] : nullptr;
        part += c[q];
    }
    for (int s0 = sb + 4 * kWave; s0 < se; s0 += kWave) {
        const int s = s0 + l;
        part += s < se ? p.scounts[(int64_t)s * P + d] : 0;
    }
    part = wave_sum(part);
    if (l == 0) s_tot[w] = part;
    __syncthreads();
    int64_t carry = 0, row = 0;
    for (int k = 0; k < 4; k++) {
        carry += k < w ? s_tot[k] : 0;
        row += s_tot[k];
    }
    for (int s0 = sb; s0 < se; s0 += kWave) {
        const int q = (s0 - sb) / kWave;
        const int s = s0 + l;
        int64_t cq, sdq;
        const char* sbq;
        if (q < 4) {
            cq = q == 0 ? c[0] : q == 1 ? c[1] : q == 2 ? c[2] : c[3];
            sdq = q == 0 ? sd[0] : q == 1 ? sd[1] : q == 2 ? sd[2] : sd[3];
            sbq = q == 0 ? sbuf[0] : q == 1 ? sbuf[1] : q == 2 ? sbuf[2] : sbuf[3];
        } else {
            const bool valid = s < se;
            cq = valid ? p.scounts[(int64_t)s * P + d] : 0;
            sdq = valid ? p.sdispls[(int64_t)s * P + d] : 0;
            sbq = valid ? p.sendbufs[s] : nullptr;
        }
        int64_t tot;
        const int64_t rd = wave_excl_scan(cq, &tot) + carry;
        carry += tot;
        if (s < se) {
            const int64_t o = d * P + s;
            p.rcounts[o] = (int32_t)cq;
            p.rdispls[o] = (int32_t)rd;
            p.seg_end[o] = (uint64_t)(rd + cq) * (uint64_t)p.elem;
            p.seg_src[o] = (uint64_t)(uintptr_t)(sbq + sdq * p.elem);
        }
    }
    if (threadIdx.x == 0) {
        const uint64_t rb = (uint64_t)row * (uint64_t)p.elem;
        if constexpr (COHERENT) st_relaxed(&p.row_bytes[d], rb);
        else p.row_bytes[d] = rb;
    }
}

static __global__ void __launch_bounds__(256) a2av_plan_kernel(A2avDev p) {
    __shared__ int64_t s_tot[4];
    a2av_plan_row(p, blockIdx.x, s_tot);
}

// ------------------------------------------------------------------------------ scan
// Run by one workgroup of NT threads (NT a multiple of 64, at most 1024).  COHERENT: the
// row byte counts are read with agent-scope atomic loads (written in the same kernel).
template <int NT, bool COHERENT = false>
__device__ __forceinline__ void a2av_scan_rows(const A2avDev& p, uint64_t* part, int64_t* bpart) {
    constexpr int NW = NT / kWave;
    const int P = p.P;
    const int w = threadIdx.x >> 6;
    const uint64_t C = p.chunk_lines;
    uint64_t carry = 0;
    int64_t bytes = 0;
    for (int base = 0; base < P; base += NT) {
        const int d = base + (int)threadIdx.x;
        uint64_t ch = 0;
        int64_t nb = 0;
        if (d < P) {
            nb = COHERENT ? (int64_t)ld_relaxed(&p.row_bytes[d]) : (int64_t)p.row_bytes[d];
            const uint64_t rb = (uint64_t)(uintptr_t)p.recvbufs[d];
            const uint64_t lines = nb ? ((rb + (uint64_t)nb + 15) >> 4) - ((rb >> 7) << 3) : 0;
            ch = (lines + C - 1) / C;
        }
        const uint64_t inc = (uint64_t)wave_incl_scan((int64_t)ch);
        const int64_t bsum = wave_sum(nb);
        if (lane() == 63) part[w] = inc;
        if (lane() == 0) bpart[w] = bsum;
        __syncthreads();
        uint64_t before = 0, total = 0;
        for (int k = 0; k < NW; k++) {
            if (k < w) before += part[k];
            total += part[k];
            bytes += (threadIdx.x == 0) ? bpart[k] : 0;
        }
        if (d < P) p.chunk_off[d] = carry + before + inc - ch;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        p.chunk_off[P] = carry;
        *p.total_bytes = bytes;
    }
}

static __global__ void __launch_bounds__(1024) a2av_scan_kernel(A2avDev p) {
    __shared__ uint64_t part[16];
    __shared__ int64_t bpart[16];
    a2av_scan_rows<1024>(p, part, bpart);
}

// Plan of the NEXT step run by the first P workgroups of a copy kernel (the step's only
// global dependency -- the receive displacements -- is then resolved while the current
// step streams, instead of in kernels of its own on the critical path).  The tables go to
// the other of two workspaces, so the copy running beside them is not disturbed.
//
// With pn.plan_done set, the workgroup that finishes last also computes the chunk prefix.
// Only the row byte counts cross workgroups inside the kernel (everything else is read by
// the next kernel): each planner stores its count with an agent-scope atomic store, which
// writes through to memory, waits for that store to complete, then counts itself done;
// the last one reads the counts with agent-scope atomic loads, which miss the L2.  No L2
// write-back or invalidate is needed (agent-scope fences in 512 workgroups, writing back
// their XCDs' L2s under the streaming copy, cost 35-60 us per step).
// With pn.plan_done == nullptr the workgroups only plan; a scan kernel follows.
__device__ void a2av_plan_fused(const A2avDev& pn, int64_t d) {
    __shared__ int64_t s_tot[4];
    __shared__ uint64_t s_part[4];
    __shared__ int64_t s_bpart[4];
    __shared__ int32_t s_last;
    if (!pn.plan_done) {  // block-uniform
        a2av_plan_row<false>(pn, d, s_tot);
        return;
    }
    a2av_plan_row<true>(pn, d, s_tot);
    if (threadIdx.x == 0) {
        wait_mem_done();
        const uint32_t old = __hip_atomic_fetch_add(pn.plan_done, 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        s_last = old == (uint32_t)pn.P - 1u;
    }
    __syncthreads();
    if (!s_last) return;
    a2av_scan_rows<kA2avCopyThreads, true>(pn, s_part, s_bpart);
    if (threadIdx.x == 0) *pn.plan_done = 0u;  // next used two kernels later
}

// ------------------------------------------------------------------------------ straddle
// Deferred mode: lines inside a row that hold bytes of two or more segments.  One thread
// per segment of every row; a segment whose end falls strictly inside a 16-byte line, and
// that holds that line's first byte, owns the line: it assembles it per dword from the
// segment holding each dword (per byte where a dword is split or unaligned at its source)
// and stores it once.  Every such line has exactly one owner.  Lines across a row's ends
// stay with the copy loop.  The copy grid's first straddle_blocks blocks run these
// threads: their scattered, latency-bound loads overlap the other blocks' streaming.
__device__ void a2av_straddle_line(const A2avDev& p, uint64_t gid) {
    const int P = p.P;
    if (gid >= (uint64_t)P * (uint64_t)P) return;
    const int64_t d = (int64_t)(gid / (uint64_t)P);
    const int s = (int)(gid - (uint64_t)d * P);
    const uint64_t* se = p.seg_end + d * P;
    // one round of independent loads: this segment, the next one, the row
    const int s1 = s + 1 < P ? s + 1 : s;
    const int64_t prev = (int64_t)se[s > 0 ? s - 1 : 0];
    const int64_t en = (int64_t)se[s];
    const int64_t en1 = (int64_t)se[s1];
    const uint64_t src0 = p.seg_src[d * P + s];
    const uint64_t src1 = p.seg_src[d * P + s1];
    const uint64_t rbase = (uint64_t)(uintptr_t)p.recvbufs[d];
    const int64_t rbytes = (int64_t)p.row_bytes[d];
    const int64_t st = s > 0 ? prev : 0;
    if (en <= st) return;  // empty
    const uint64_t le = rbase + (uint64_t)en;  // one past the segment's last byte
    if ((le & 15) == 0) return;                // ends with its line
    const uint64_t la = (le >> 4) << 4;        // the line holding the segment's last byte
    const int64_t off = (int64_t)la - (int64_t)rbase;
    if (st > off || off < 0 || off + 16 > rbytes) return;
    // the next non-empty segment
    int kb = s + 1;
    int64_t eb = en1;
    uint64_t srcb = src1;
    if (en1 <= en) {
        while ((int64_t)se[kb] <= en) kb++;  // ends before rbytes >= off + 16 > en
        eb = (int64_t)se[kb];
        srcb = p.seg_src[d * P + kb];
    }
    const uint64_t a0 = src0 + (uint64_t)(off - st);  // source of the line's byte 0 in s
    const uint64_t a1 = srcb - (uint64_t)en;          // + row offset, in segment kb
    if (eb >= off + 16) {
        // two segments: four loads in flight
        uint64_t addr[4];
        bool whole[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int64_t q = off + 4 * i;
            const bool first = q < en;
            addr[i] = first ? a0 + 4 * i : a1 + (uint64_t)q;
            whole[i] = (!first || q + 4 <= en) && (addr[i] & 3) == 0;
        }
        uint32_t v[4];
        // a dword split between the segments, or unaligned at its source, is fetched per
        // byte; its aligned-down dword holds a byte of the source and is a safe read
#pragma unroll
        for (int i = 0; i < 4; i++)
            v[i] = *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)(addr[i] & ~(uint64_t)3);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (whole[i]) continue;
            uint32_t x = 0;
            for (int j = 0; j < 4; j++) {
                const int64_t b = off + 4 * i + j;
                const uint64_t ab = b < en ? a0 + 4 * i + j : a1 + (uint64_t)b;
                x |= (uint32_t)*(const GPUMPI_GLOBAL uint8_t*)(uintptr_t)ab << (8 * j);
            }
            v[i] = x;
        }
        *(GPUMPI_GLOBAL u32x4*)(uintptr_t)la = u32x4{v[0], v[1], v[2], v[3]};
        return;
    }
    // three or more segments in the line: walk them
    int k = s;  // holds byte `off`; the walk only moves forward over the line
    uint64_t src[4];
    int whole = 0;
    for (int i = 0; i < 4; i++) {
        const int64_t q = off + 4 * i;
        while (q >= (int64_t)se[k]) k++;
        const int64_t sk = k > 0 ? (int64_t)se[k - 1] : 0;
        src[i] = p.seg_src[d * P + k] + (uint64_t)(q - sk);
        if (q + 4 <= (int64_t)se[k] && (src[i] & 3) == 0) whole |= 1 << i;
    }
    uint32_t v[4];
    k = s;
    for (int i = 0; i < 4; i++) {
        if (whole & (1 << i)) {
            v[i] = *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)src[i];
            continue;
        }
        uint32_t x = 0;
        for (int j = 0; j < 4; j++) {
            const int64_t b = off + 4 * i + j;
            while (b >= (int64_t)se[k]) k++;
            const int64_t sb = k > 0 ? (int64_t)se[k - 1] : 0;
            const uint8_t y = *(const GPUMPI_GLOBAL uint8_t*)(uintptr_t)(p.seg_src[d * P + k] +
                                                                        (uint64_t)(b - sb));
            x |= (uint32_t)y << (8 * j);
        }
        v[i] = x;
    }
    *(GPUMPI_GLOBAL u32x4*)(uintptr_t)la = u32x4{v[0], v[1], v[2], v[3]};
}

// ------------------------------------------------------------------------------ copy
// One segment seen from its row: bytes [st, en) of the row; the source of the row byte at
// offset q is (rbase + q) + D.
struct A2avSegView {
    int64_t st, en;
    uint64_t D;
};

// Segment holding row offset q (0 <= q < row bytes), by binary search over the row's
// segment ends in HBM (the path for lines outside the LDS window).
__device__ __forceinline__ A2avSegView a2av_seg_global(const A2avDev& p, int64_t d, uint64_t rbase,
                                                       int64_t q) {
    const uint64_t* se = p.seg_end + d * p.P;
    int64_t lo = 0, hi = p.P - 1;  // first s with se[s] > q
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)se[mid] > q) hi = mid; else lo = mid + 1;
    }
    A2avSegView v;
    v.st = lo > 0 ? (int64_t)se[lo - 1] : 0;
    v.en = (int64_t)se[lo];
    v.D = p.seg_src[d * p.P + lo] - (rbase + (uint64_t)v.st);
    return v;
}

// Window entry holding row offset q (s_st[0] <= q < coverage): largest k < cnt with
// s_st[k] <= q.  cnt is wave-uniform, so the search takes ceil(log2(cnt)) steps.
__device__ __forceinline__ int a2av_seg_lds(const int64_t* s_st, int cnt, int64_t q) {
    int k = 0;
    for (int b = cnt > 1 ? 1 << (31 - __builtin_clz(cnt - 1)) : 0; b > 0; b >>= 1)
        if (k + b < cnt && s_st[k + b] <= q) k += b;
    return k;
}

struct A2avWin {
    const int64_t* st;
    const int64_t* en;
    const uint64_t* D;
    int cnt;
    int nsearch;        // entries starting before the pass's last line: the search range
    int64_t lo, cover;  // row offsets [lo, cover) are described by the window
};

__device__ __forceinline__ A2avSegView a2av_seg_at(const A2avDev& p, int64_t d, uint64_t rbase,
                                                   const A2avWin& w, int64_t q) {
    if (q >= w.lo && q < w.cover) {
        const int k = a2av_seg_lds(w.st, w.nsearch, q);
        return A2avSegView{w.st[k], w.en[k], w.D[k]};
    }
    return a2av_seg_global(p, d, rbase, q);
}

// A destination line that straddles segments or the row's ends: assembled per dword
// (per byte where a dword is split or its source unaligned); only row bytes are written.
// In deferred mode the lines inside the row are the straddle threads'.
template <bool DEFER>
__device__ void a2av_boundary_line(const A2avDev& p, int64_t d, uint64_t rbase, int64_t rbytes,
                                   const A2avWin& w, uint64_t line) {
    const uint64_t la = line << 4;
    const int64_t off = (int64_t)la - (int64_t)rbase;
    if constexpr (DEFER) {
        if (off >= 0 && off + 16 <= rbytes) return;
    }
    uint32_t v[4] = {0, 0, 0, 0};
    int mask = 0;
    for (int i = 0; i < 4; i++) {
        const int64_t q = off + 4 * i;
        if (q >= 0 && q + 4 <= rbytes) {
            const A2avSegView sg = a2av_seg_at(p, d, rbase, w, q);
            const uint64_t src = la + 4 * i + sg.D;
            if (q + 4 <= sg.en && (src & 3) == 0) {
                v[i] = *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)src;
                mask |= 1 << i;
                continue;
            }
        }
        for (int j = 0; j < 4; j++) {
            const int64_t b = q + j;
            if (b < 0 || b >= rbytes) continue;
            const A2avSegView sg = a2av_seg_at(p, d, rbase, w, b);
            const uint64_t a = la + 4 * i + j;
            *(GPUMPI_GLOBAL uint8_t*)(uintptr_t)a = *(const GPUMPI_GLOBAL uint8_t*)(uintptr_t)(a + sg.D);
        }
    }
    if (mask == 0xF) {
        *(GPUMPI_GLOBAL u32x4*)(uintptr_t)la = u32x4{v[0], v[1], v[2], v[3]};
    } else {
        for (int i = 0; i < 4; i++)
            if (mask & (1 << i)) *(GPUMPI_GLOBAL uint32_t*)(uintptr_t)(la + 4 * i) = v[i];
    }
}

// FUSE: the first pn.P workgroups plan the next step into pn (a2av_plan_fused); the copy
// workgroups follow them.
template <int U, bool NT, bool DEFER, bool FUSE>
__global__ void __launch_bounds__(kA2avCopyThreads) a2av_copy_kernel(A2avDev p, A2avDev pn) {
    __shared__ int64_t s_st[kA2avWindow];
    __shared__ int64_t s_en[kA2avWindow];
    __shared__ uint64_t s_D[kA2avWindow];
    __shared__ int64_t s_meta[10];
    const int P = p.P;
    const int tid = threadIdx.x;
    const int l = lane();
    const uint64_t C = p.chunk_lines;
    uint32_t bid = blockIdx.x, nblk = gridDim.x;
    if constexpr (FUSE) {
        if (bid < (uint32_t)pn.P) {  // block-uniform
            a2av_plan_fused(pn, bid);
            return;
        }
        bid -= (uint32_t)pn.P;
        nblk -= (uint32_t)pn.P;
    }
    uint32_t sb = 0;
    if constexpr (DEFER) {
        sb = (uint32_t)p.straddle_blocks;
        if (bid < sb) {  // block-uniform
            const uint64_t nseg = (uint64_t)P * (uint64_t)P;
            for (uint64_t g = (uint64_t)bid * kA2avCopyThreads + tid; g < nseg;
                 g += (uint64_t)sb * kA2avCopyThreads)
                a2av_straddle_line(p, g);
            return;
        }
    }
    const uint64_t total = p.chunk_off[P];
    for (uint64_t ch = bid - sb; ch < total; ch += nblk - sb) {
        // ---- the chunk's row and line range
        if (tid < kWave) {
            const int64_t d = wave_search_le(p.chunk_off, P, ch);
            const uint64_t c = ch - p.chunk_off[d];
            const uint64_t rbase = (uint64_t)(uintptr_t)p.recvbufs[d];
            const uint64_t rbytes = p.row_bytes[d];
            const uint64_t lend = (rbase + rbytes + 15) >> 4;
            // a row's lines start at the 128-byte block holding its first byte, so every
            // wave's 1 KiB slice covers whole 128-byte blocks (lines before the row's first
            // byte write nothing)
            const uint64_t line0 = ((rbase >> 7) << 3) + c * C;
            const uint64_t line1 = line0 + C < lend ? line0 + C : lend;
            if (l == 0) {
                s_meta[0] = d;
                s_meta[1] = (int64_t)rbase;
                s_meta[2] = (int64_t)rbytes;
                s_meta[3] = (int64_t)line0;
                s_meta[4] = (int64_t)line1;
            }
        }
        __syncthreads();
        // block-uniform values read back from LDS: broadcast so that they live in SGPRs
        // (the copy loop is bound by instruction issue; scalar arithmetic is free beside it)
        const int64_t d = bcast(s_meta[0]);
        const uint64_t rbase = (uint64_t)bcast(s_meta[1]);
        const int64_t rbytes = bcast(s_meta[2]);
        const uint64_t line1 = (uint64_t)bcast(s_meta[4]);
        uint64_t pos = (uint64_t)bcast(s_meta[3]);
        while (pos < line1) {
            // ---- window: 64 segments from the one holding the row offset of line `pos`
            if (tid < kWave) {
                int64_t o = (int64_t)(pos << 4) - (int64_t)rbase;
                if (o < 0) o = 0;
                const uint64_t* se = p.seg_end + d * P;
                int64_t s0 = 0;
                if ((int64_t)se[0] <= o) s0 = wave_search_le(se, P, (uint64_t)o) + 1;
                const int64_t s = s0 + l;
                const bool valid = s < P;
                const int64_t en = valid ? (int64_t)se[s] : 0;
                const int64_t st = valid ? (s > 0 ? (int64_t)se[s - 1] : 0) : 0;
                const uint64_t src = valid ? p.seg_src[d * P + s] : 0;
                const bool ne = valid && en > st;
                const uint64_t m = ballot(ne);
                const int k = __popcll(m & (l == 0 ? 0ull : (~0ull >> (64 - l))));
                if (ne) {
                    s_st[k] = st;
                    s_en[k] = en;
                    s_D[k] = src - (rbase + (uint64_t)st);
                }
                const int64_t cover =
                    s0 + kWave >= P ? rbytes : (int64_t)readlane64((uint64_t)en, kWave - 1);
                // the pass takes the lines that end at or before the coverage, at least one
                uint64_t pend = (rbase + (uint64_t)cover) >> 4;
                if (pend > line1) pend = line1;
                int64_t slow = 0;
                if (pend <= pos) {
                    pend = pos + 1;
                    slow = 1;
                }
                // entries that start before the end of the pass's lines bound the searches
                const int64_t pend_off = (int64_t)(pend << 4) - (int64_t)rbase;
                const uint64_t mb = ballot(ne && st < pend_off);
                if (l == 0) {
                    s_meta[5] = __popcll(m);
                    s_meta[6] = (int64_t)pend;
                    s_meta[7] = slow ? 0 : cover;
                    s_meta[8] = slow ? 0 : o;
                    s_meta[9] = __popcll(mb) > 0 ? __popcll(mb) : 1;
                }
            }
            __syncthreads();
            A2avWin w;
            w.st = s_st;
            w.en = s_en;
            w.D = s_D;
            w.cnt = (int)bcast(s_meta[5]);
            w.nsearch = (int)bcast(s_meta[9]);
            const uint64_t pend = (uint64_t)bcast(s_meta[6]);
            w.cover = bcast(s_meta[7]);
            w.lo = bcast(s_meta[8]);
            // each wave copies a contiguous run of 64*U lines, slice by slice, with a
            // wave-uniform cursor over the window (segments only move forward)
            const int wv = tid >> 6;
            for (uint64_t base = pos; base < pend; base += (uint64_t)kA2avCopyThreads * U) {
                const uint64_t wb = base + (uint64_t)wv * kWave * U;
                if (wb >= pend) continue;
                int64_t off0 = (int64_t)(wb << 4) - (int64_t)rbase;
                int kc = bcast(a2av_seg_lds(s_st, w.nsearch, off0 > w.lo ? off0 : w.lo));
                // entries kc and kc + 1 (ends and source offsets), cached in SGPRs
                int kc1 = kc + 1 < w.cnt ? kc + 1 : kc;
                int64_t Ekc = bcast(s_en[kc]), Ekc1 = bcast(s_en[kc1]);
                uint64_t Dkc = bcast(s_D[kc]), Dkc1 = bcast(s_D[kc1]);
                u32x4 v[U];
                int kind[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const uint64_t sl = wb + (uint64_t)u * kWave;  // slice's first line
                    const uint64_t line = sl + l;
                    kind[u] = line < pend ? 2 : 0;
                    if (sl >= pend) continue;  // wave-uniform
                    const int64_t soff = (int64_t)(sl << 4) - (int64_t)rbase;
                    // move the cursor to the entry holding the slice's first byte: a few
                    // steps forward, or a search when the slice starts further ahead
                    if (soff >= Ekc && kc + 1 < w.nsearch) {
                        if (kc + 3 < w.nsearch && soff >= bcast(s_en[kc + 3]))
                            kc = bcast(a2av_seg_lds(s_st, w.nsearch, soff));
                        else
                            while (kc + 1 < w.nsearch && soff >= bcast(s_en[kc])) kc++;
                        kc1 = kc + 1 < w.cnt ? kc + 1 : kc;
                        Ekc = bcast(s_en[kc]);
                        Dkc = bcast(s_D[kc]);
                        Ekc1 = bcast(s_en[kc1]);
                        Dkc1 = bcast(s_D[kc1]);
                    }
                    const int64_t off = soff + 16 * l;
                    if (soff + 16 * kWave <= Ekc) {
                        // the whole slice lies in entry kc (the common case): one uniform
                        // source offset, no per-lane selection, no LDS access
                        if (line < pend && off >= w.lo) {
                            v[u] = gload16((line << 4) + Dkc);
                            kind[u] = 1;
                        }
                        continue;
                    }
                    if (!DEFER && soff + 16 * kWave <= Ekc1) {
                        // the slice spans entries kc and kc + 1: a lane's line lies in one of
                        // them or straddles their boundary (one lane at most)
                        if (line < pend && off >= w.lo) {
                            const uint64_t la = line << 4;
                            if (off + 16 <= Ekc || off >= Ekc) {
                                v[u] = gload16(la + (off < Ekc ? Dkc : Dkc1));
                                kind[u] = 1;
                            } else if constexpr (!DEFER) {
                                if (((Ekc - off) & 3) == 0 && ((la + Dkc) & 3) == 0 && ((la + Dkc1) & 3) == 0) {
                                    v[u] = u32x4{
                                        *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)(la + (off < Ekc ? Dkc : Dkc1)),
                                        *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)(la + 4 + (off + 4 < Ekc ? Dkc : Dkc1)),
                                        *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)(la + 8 + (off + 8 < Ekc ? Dkc : Dkc1)),
                                        *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)(la + 12 + (off + 12 < Ekc ? Dkc : Dkc1))};
                                    kind[u] = 1;
                                }
                            }
                        }
                        continue;
                    }
                    if (!(line < pend && off >= w.lo && off + 16 <= w.cover)) continue;
                    // this lane's entry: the cursor or one of the next three, else a search
                    int kb = kc;
                    if (kc + 3 < w.cnt && off >= s_en[kc + 3]) kb = a2av_seg_lds(s_st, w.nsearch, off);
                    const int k1 = kb + 1 < w.cnt ? kb + 1 : kb, k2 = kb + 2 < w.cnt ? kb + 2 : k1,
                              k3 = kb + 3 < w.cnt ? kb + 3 : k2;
                    const int64_t e0 = s_en[kb], e1 = s_en[k1], e2 = s_en[k2], e3 = s_en[k3];
                    const uint64_t D0 = s_D[kb], D1 = s_D[k1], D2 = s_D[k2], D3 = s_D[k3];
                    const uint64_t la = line << 4;
                    const int j = (off >= e0) + (off >= e1) + (off >= e2);
                    const int64_t ej = j == 0 ? e0 : j == 1 ? e1 : j == 2 ? e2 : e3;
                    const uint64_t Dj = j == 0 ? D0 : j == 1 ? D1 : j == 2 ? D2 : D3;
                    if (off + 16 <= ej) {
                        v[u] = gload16(la + Dj);
                        kind[u] = 1;
                        continue;
                    }
                    if constexpr (DEFER) continue;
                    // the line straddles segments: one dword from the segment holding it
                    uint64_t a[4];
                    bool ok = true;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int64_t q = off + 4 * i;
                        const int ji = (q >= e0) + (q >= e1) + (q >= e2);
                        const int64_t eq = ji == 0 ? e0 : ji == 1 ? e1 : ji == 2 ? e2 : e3;
                        const uint64_t Dq = ji == 0 ? D0 : ji == 1 ? D1 : ji == 2 ? D2 : D3;
                        a[i] = la + 4 * i + Dq;
                        ok = ok && (a[i] & 3) == 0 && q + 4 <= eq;
                    }
                    if (ok) {
                        v[u] = u32x4{*(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)a[0],
                                     *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)a[1],
                                     *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)a[2],
                                     *(const GPUMPI_GLOBAL uint32_t*)(uintptr_t)a[3]};
                        kind[u] = 1;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (kind[u] == 1) gstore16<NT>((wb + (uint64_t)u * kWave + l) << 4, v[u]);
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (kind[u] == 2)
                        a2av_boundary_line<DEFER>(p, d, rbase, rbytes, w, wb + (uint64_t)u * kWave + l);
            }
            __syncthreads();
            pos = pend;
        }
    }
}

}  // namespace gpumpi
