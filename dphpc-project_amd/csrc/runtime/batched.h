// Batched MPI_Alltoallv step: one persistent kernel per step.
//
// Semantics: gpu_mpi/mpi.cu:363-408 (block s of rank d's receive buffer <- block d of
// rank s's send buffer, counts/displacements in elements), preceded by the Alltoall of
// the counts and the packed receive displacements that the NPB IS caller computes
// between the two calls (toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c:758-781).
//
// The reference runs this as P-1 Isend + P-1 Irecv per rank, each payload copied by one
// thread.  Here the whole step is decoupled from ranks:
//
//  phase 1  one wavefront per destination row d: recv_counts[d][s] = send_counts[s][d],
//           recv_displs = exclusive scan over s, and a segment table entry per (d, s):
//           {src, dst, bytes, first 16-byte line of the row}; lines are counted in
//           absolute 16-byte destination lines, so every full line is one aligned store.
//  barrier  one arrival per workgroup; the last workgroup to arrive scans the per-row
//           line totals into row offsets and releases the others (one grid barrier, no
//           second kernel).
//  phase 2  the global line space is split evenly over ALL wavefronts of the grid, so a
//           Zipf-skewed receiver is copied by many CUs at once.  A wavefront keeps a
//           window of 64 segment records in its lanes (compacted to the non-empty
//           ones), and each lane locates its line's
//           segment: a wave-uniform cursor plus one compare selects between the (at
//           most two) segments of a 64-line slice, with a 6-step shuffle search only
//           for slices crossing several tiny segments.  U lines per lane are loaded
//           before any is stored (16 KiB in flight per wave at U=16).  Full lines are
//           unaligned 16-byte global loads and 16-byte aligned, optionally
//           non-temporal stores; the <=2 partial lines per segment are copied bytewise
//           by the lane that holds the segment.
//
// HBM traffic per step = payload read once + payload written once + the O(P^2) tables.
#pragma once

#include "copy.h"

namespace gpumpi {

struct A2avSeg {
    const char* src;
    char* dst;
    uint64_t bytes;
    uint64_t first;  // first line of this segment, relative to its row
};

struct A2avDev {
    int32_t P;
    int32_t elem;
    int32_t nontemporal;
    int32_t pad;
    char* const* sendbufs;
    const int32_t* scounts;
    const int32_t* sdispls;
    char* const* recvbufs;
    int32_t* rcounts;
    int32_t* rdispls;
    A2avSeg* table;         // [P*P]
    uint64_t* row_lines;    // [P]
    uint64_t* row_off;      // [P+1] exclusive prefix, written by the last phase-1 wave
    int64_t* row_bytes;     // [P]
    int64_t* total_bytes;   // [1]
    uint64_t* arrive;       // grid barrier arrivals, one per workgroup (monotonic across launches)
    uint64_t* release;      // grid barrier release epoch
    int32_t* error;         // set on a timed-out barrier
    uint64_t timeout_ticks;
    uint64_t* stamps;       // optional: [start min, phase-1 end max, barrier exit max, end max]
};

__device__ __forceinline__ void stamp_min(uint64_t* a, uint64_t v) {
    __hip_atomic_fetch_min(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stamp_max(uint64_t* a, uint64_t v) {
    __hip_atomic_fetch_max(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t wave_excl_scan(int64_t v, int64_t* total) {
    int64_t inc = wave_incl_scan(v);
    *total = __shfl(inc, 63);
    return inc - v;
}

// Largest i in [0, n) with a[i] <= x (a non-decreasing, a[0] <= x); wave-uniform.
__device__ __forceinline__ int64_t wave_search_le(const uint64_t* a, int64_t n, uint64_t x,
                                                  int64_t stride_elems = 1, int64_t base = 0) {
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        int64_t span = hi - lo;
        int64_t step = (span + 63) / 64;
        int64_t idx = lo + (int64_t)lane() * step;
        bool le = false;
        if (idx < hi) le = a[base + idx * stride_elems] <= x;
        uint64_t m = ballot(le);
        int L = m ? 63 - __builtin_clzll(m) : 0;
        int64_t nlo = lo + (int64_t)L * step;
        int64_t nhi = nlo + step;
        lo = nlo;
        hi = nhi < hi ? nhi : hi;
    }
    return lo;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int k) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), k);
    return ((uint64_t)hi << 32) | lo;
}
// Forward permute of a 64-bit value: lane i's value lands in lane dst_i.
__device__ __forceinline__ uint64_t push64(uint64_t v, int dst) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_ds_permute(dst << 2, (int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_ds_permute(dst << 2, (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Copies destination lines [line_begin, line_end) of the global line space.
// Each segment k of the current window is described by scalars held in lane k:
//   G  first global line            A  = (dst >> 4) - G   (line j starts at (A + j) * 16)
//   D  = src - dst                  LO = dst, HI = dst + bytes
template <int U, bool NT>
__device__ void a2av_copy_lines(const A2avDev& p, uint64_t line_begin, uint64_t line_end) {
    const int P = p.P;
    const int64_t nseg = (int64_t)P * P;
    const int l = lane();
    if (line_begin >= line_end) return;
    int64_t d = wave_search_le(p.row_off, P, line_begin);
    const uint64_t roff = p.row_off[d];
    int64_t s = wave_search_le((const uint64_t*)&p.table[0].first, P, line_begin - roff,
                               (int64_t)(sizeof(A2avSeg) / sizeof(uint64_t)),
                               (int64_t)d * P * (int64_t)(sizeof(A2avSeg) / sizeof(uint64_t)));
    int64_t seg_base = d * P + s;
    uint64_t line = line_begin;
    const uint64_t lt_mask = l == 0 ? 0ull : (~0ull >> (64 - l));
    while (line < line_end && seg_base < nseg) {
        // ---- load a window of 64 segments, keep the non-empty ones in lanes 0..cnt-1
        const int64_t f = seg_base + l;
        uint64_t G = ~0ull, A = 0, D = 0, LO = 0, HI = 0, nl = 0;
        if (f < nseg) {
            const A2avSeg sg = p.table[f];
            const uint64_t dst = (uint64_t)(uintptr_t)sg.dst;
            G = p.row_off[f / P] + sg.first;
            nl = sg.bytes ? ((dst + sg.bytes + 15) >> 4) - (dst >> 4) : 0;
            A = (dst >> 4) - G;
            D = (uint64_t)(uintptr_t)sg.src - dst;
            LO = dst;
            HI = dst + sg.bytes;
        }
        const bool ne = nl != 0;
        const uint64_t m = ballot(ne);
        const int cnt = __popcll(m);
        seg_base += 64;
        if (cnt == 0) continue;
        const int to = ne ? __popcll(m & lt_mask) : cnt + __popcll(~m & lt_mask);
        G = push64(ne ? G : ~0ull, to);
        A = push64(A, to);
        D = push64(D, to);
        LO = push64(LO, to);
        HI = push64(HI, to);
        const int last = cnt - 1;
        // end line of the window = first line of the last segment + its line count
        const uint64_t g_last = readlane64(G, last);
        const uint64_t hi_last = readlane64(HI, last), lo_last = readlane64(LO, last);
        const uint64_t win_end = g_last + (((hi_last + 15) >> 4) - (lo_last >> 4));
        const uint64_t lim = win_end < line_end ? win_end : line_end;
        // partial first/last lines of each segment: copied bytewise by the segment's lane
        // when the line belongs to [line, lim), the part of the window this wave owns
        if (l < cnt) {
            const uint64_t nlk = ((HI + 15) >> 4) - (LO >> 4);
            for (int e = 0; e < 2; e++) {
                const uint64_t jl = e == 0 ? G : G + nlk - 1;
                if (e == 1 && nlk == 1) break;
                if (jl < line || jl >= lim) continue;
                const uint64_t ls = (A + jl) << 4;
                if (ls >= LO && ls + 16 <= HI) continue;  // full line: main loop
                const uint64_t b0 = ls > LO ? ls : LO;
                const uint64_t b1 = ls + 16 < HI ? ls + 16 : HI;
                char* dp = (char*)(uintptr_t)b0;
                const char* sp = (const char*)(uintptr_t)(b0 + D);
                const uint64_t nb = b1 - b0;
                if (((b0 | (uint64_t)(uintptr_t)sp | nb) & 3) == 0) {
                    for (uint64_t q = 0; q < nb; q += 4)
                        *(uint32_t*)(dp + q) = *(const uint32_t*)(sp + q);
                } else {
                    for (uint64_t q = 0; q < nb; q++) dp[q] = sp[q];
                }
            }
        }
        int kcur = 0;
        while (line < lim) {
            u32x4 v[U];
            uint64_t da[U];
            bool full[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t L0 = line + (uint64_t)u * kWave;
                full[u] = false;
                da[u] = 0;
                if (L0 >= lim) continue;  // wave-uniform
                // advance the current segment to the one holding line L0 (uniform loop)
                while (kcur < last && readlane64(G, kcur + 1) <= L0) kcur++;
                const uint64_t j = L0 + (uint64_t)l;
                const uint64_t g1 = kcur < last ? readlane64(G, kcur + 1) : ~0ull;
                const uint64_t g2 = kcur + 1 < last ? readlane64(G, kcur + 2) : ~0ull;
                uint64_t a, dd, lo, hi;
                if (L0 + (kWave - 1) < g2) {
                    // fast path: the 64-line slice spans at most two segments
                    const bool sel = j >= g1;
                    const int k1 = kcur < last ? kcur + 1 : kcur;
                    a = sel ? readlane64(A, k1) : readlane64(A, kcur);
                    dd = sel ? readlane64(D, k1) : readlane64(D, kcur);
                    lo = sel ? readlane64(LO, k1) : readlane64(LO, kcur);
                    hi = sel ? readlane64(HI, k1) : readlane64(HI, kcur);
                } else {
                    // slow path (many tiny segments): per-lane search over the window
                    const uint64_t jj = j < lim ? j : L0;
                    int k = 0;
#pragma unroll
                    for (int b = 32; b > 0; b >>= 1) {
                        uint64_t g = __shfl(G, k + b);
                        if (k + b <= last && g <= jj) k += b;
                    }
                    a = __shfl(A, k);
                    dd = __shfl(D, k);
                    lo = __shfl(LO, k);
                    hi = __shfl(HI, k);
                }
                const uint64_t ls = (a + j) << 4;  // line start address
                full[u] = j < lim && ls >= lo && ls + 16 <= hi;
                da[u] = ls;
                if (full[u]) v[u] = gload16(ls + dd);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (full[u]) gstore16<NT>(da[u], v[u]);
            line += (uint64_t)U * kWave;
        }
        line = lim;
    }
}

template <int U, bool NT>
__global__ void __launch_bounds__(1024) a2av_step_kernel(A2avDev p) {
    const int P = p.P;
    const int l = lane();
    const uint32_t wpb = blockDim.x >> 6;
    const uint64_t W = (uint64_t)gridDim.x * wpb;
    const uint64_t w = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
    if (p.stamps && threadIdx.x == 0) stamp_min(&p.stamps[0], wall_clock64());

    // ---------------- phase 1: counts transpose, displacements, segment table
    for (int64_t d = (int64_t)w; d < P; d += (int64_t)W) {
        int64_t carry = 0;
        uint64_t carry_lines = 0;
        char* rbase = p.recvbufs[d];
        for (int s0 = 0; s0 < P; s0 += 4 * kWave) {
            // issue the strided column loads of up to 4 chunks before any scan
            int64_t c[4], sd[4];
            const char* sb[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int s = s0 + q * kWave + l;
                const bool valid = s < P;
                c[q] = valid ? p.scounts[(int64_t)s * P + d] : 0;
                sd[q] = valid ? p.sdispls[(int64_t)s * P + d] : 0;
                sb[q] = valid ? p.sendbufs[s] : nullptr;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int s = s0 + q * kWave + l;
                if (s0 + q * kWave >= P) break;
                const bool valid = s < P;
                int64_t tot;
                int64_t rd = wave_excl_scan(c[q], &tot) + carry;
                carry += tot;
                uint64_t nb = (uint64_t)c[q] * (uint64_t)p.elem;
                char* dst = rbase + rd * p.elem;
                uint64_t lines = nb ? (((uint64_t)(uintptr_t)dst + nb + 15) >> 4) -
                                          ((uint64_t)(uintptr_t)dst >> 4)
                                    : 0;
                int64_t ltot;
                uint64_t lfirst = (uint64_t)wave_excl_scan((int64_t)lines, &ltot) + carry_lines;
                carry_lines += (uint64_t)ltot;
                if (valid) {
                    p.rcounts[d * P + s] = (int32_t)c[q];
                    p.rdispls[d * P + s] = (int32_t)rd;
                    A2avSeg sg;
                    sg.src = sb[q] + sd[q] * p.elem;
                    sg.dst = dst;
                    sg.bytes = nb;
                    sg.first = lfirst;
                    p.table[d * P + s] = sg;
                }
            }
        }
        if (l == 0) {
            p.row_lines[d] = carry_lines;
            p.row_bytes[d] = carry * p.elem;
        }
    }

    // ---------------- grid barrier, one arrival per workgroup; the last arriving
    // workgroup's first wavefront builds the row offsets and releases the others
    __shared__ uint32_t s_last;
    __shared__ uint64_t s_epoch;
    fence_release();
    if (p.stamps && l == 0) stamp_max(&p.stamps[1], wall_clock64());
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t B = gridDim.x;
        const uint64_t old = fetch_add(p.arrive, (uint64_t)1);
        s_epoch = old / B + 1;
        s_last = (old % B == B - 1) ? 1u : 0u;
    }
    __syncthreads();
    const uint64_t epoch = s_epoch;
    if (s_last) {
        if (threadIdx.x < 64) {
            fence_acquire();
            uint64_t carry = 0;
            int64_t bytes = 0;
            for (int d0 = 0; d0 < P; d0 += kWave) {
                const int d = d0 + l;
                int64_t v = d < P ? (int64_t)p.row_lines[d] : 0;
                int64_t b = d < P ? p.row_bytes[d] : 0;
                int64_t tot;
                int64_t ex = wave_excl_scan(v, &tot);
                if (d < P) p.row_off[d] = carry + (uint64_t)ex;
                carry += (uint64_t)tot;
                bytes += wave_sum(b);
            }
            if (l == 0) {
                p.row_off[P] = carry;
                *p.total_bytes = bytes;
            }
            fence_release();
            if (l == 0) st_release(p.release, epoch);
        }
    } else if (threadIdx.x == 0) {
        const uint64_t t0 = wall_clock64();
        uint32_t spins = 0;
        for (;;) {
            if (ld_acquire(p.release) >= epoch) break;
            if ((++spins & 255) == 0 && wall_clock64() - t0 > p.timeout_ticks) {
                st_relaxed(p.error, 1);
                s_last = 2u;  // abandon the step
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    if (s_last == 2u) return;
    fence_acquire();
    if (p.stamps && threadIdx.x == 0) stamp_max(&p.stamps[2], wall_clock64());

    // ---------------- phase 2: even split of all destination lines over all waves
    const uint64_t L = p.row_off[P];
    const uint64_t b = L * w / W;          // L < 2^40 lines and W < 2^20 waves
    const uint64_t e = L * (w + 1) / W;
    a2av_copy_lines<U, NT>(p, b, e);
    if (p.stamps && l == 0) stamp_max(&p.stamps[3], wall_clock64());
}

__global__ void a2av_reset_kernel(uint64_t* arrive, uint64_t* release) {
    if (threadIdx.x == 0) {
        *arrive = 0;
        *release = 0;
    }
}

}  // namespace gpumpi
