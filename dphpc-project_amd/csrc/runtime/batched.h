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
//  barrier  the last wavefront to arrive scans the per-row line totals into row offsets
//           and releases the others (one grid barrier, no second kernel).
//  phase 2  the global line space is split evenly over ALL wavefronts of the grid, so a
//           Zipf-skewed receiver is copied by many CUs at once.  A wavefront keeps a
//           window of 64 segment records in its lanes, each lane locates its line's
//           segment with a 6-step shuffle search, and U lines per lane are loaded before
//           any is stored (16 KiB in flight per wave at U=16).  Full lines are
//           dword-aligned 16-byte loads (byte funnel shifts when the source is not
//           dword-aligned) and 16-byte aligned, optionally non-temporal stores; the <=2
//           partial lines per segment are copied bytewise.
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
    uint64_t* arrive;       // grid barrier arrival counter (monotonic across launches)
    uint64_t* release;      // grid barrier release epoch
    int32_t* error;         // set on a timed-out barrier
    uint64_t timeout_ticks;
};

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

template <int U>
__device__ void a2av_copy_lines(const A2avDev& p, uint64_t line_begin, uint64_t line_end) {
    const int P = p.P;
    const int64_t nseg = (int64_t)P * P;
    const int l = lane();
    if (line_begin >= line_end) return;
    // locate the starting segment
    int64_t d = wave_search_le(p.row_off, P, line_begin);
    uint64_t roff = p.row_off[d];
    // segment first-lines of row d live at table[d*P + s].first (stride 4 words)
    int64_t s = wave_search_le((const uint64_t*)&p.table[0].first, P, line_begin - roff,
                               (int64_t)(sizeof(A2avSeg) / sizeof(uint64_t)),
                               (int64_t)d * P * (int64_t)(sizeof(A2avSeg) / sizeof(uint64_t)));
    int64_t seg_base = d * P + s;
    uint64_t line = line_begin;
    const bool nt = p.nontemporal != 0;
    while (line < line_end && seg_base < nseg) {
        // window: lane k holds segment seg_base + k
        int64_t f = seg_base + l;
        uint64_t gfirst = ~0ull, lines = 0;
        const char* src = nullptr;
        char* dst = nullptr;
        uint64_t bytes = 0;
        if (f < nseg) {
            const A2avSeg sg = p.table[f];
            int64_t row = f / P;
            gfirst = p.row_off[row] + sg.first;
            src = sg.src;
            dst = sg.dst;
            bytes = sg.bytes;
            lines = bytes ? (((uint64_t)(uintptr_t)dst + bytes + 15) >> 4) -
                                ((uint64_t)(uintptr_t)dst >> 4)
                          : 0;
        }
        // end of the window = end of its last valid segment
        int last = (int)((nseg - seg_base) < 64 ? (nseg - seg_base) - 1 : 63);
        uint64_t wend = __shfl(gfirst + lines, last);
        uint64_t lim = wend < line_end ? wend : line_end;
        const uint64_t src_lo = (uint64_t)(uintptr_t)src;
        const uint64_t dst_lo = (uint64_t)(uintptr_t)dst;
        while (line < lim) {
            u32x4 v[U];
            char* da[U];
            bool full[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t j0 = line + (uint64_t)u * kWave + l;
                const bool valid = j0 < lim;
                const uint64_t j = valid ? j0 : line;
                // largest k with gfirst_k <= j (6-step search over the lanes; every lane
                // takes part so each shuffle reads active lanes only)
                int k = 0;
#pragma unroll
                for (int b = 32; b > 0; b >>= 1) {
                    uint64_t g = __shfl(gfirst, k + b);
                    if (k + b <= last && g <= j) k += b;
                }
                const uint64_t kd = __shfl(dst_lo, k);
                const uint64_t ks = __shfl(src_lo, k);
                const uint64_t kn = __shfl(bytes, k);
                const uint64_t kg = __shfl(gfirst, k);
                full[u] = false;
                da[u] = nullptr;
                if (valid) {
                    const uint64_t ls = ((kd >> 4) + (j - kg)) << 4;  // line start address
                    const uint64_t lo = ls > kd ? ls : kd;
                    const uint64_t hi = ls + 16 < kd + kn ? ls + 16 : kd + kn;
                    const char* sp = (const char*)(uintptr_t)(ks + (lo - kd));
                    if (lo == ls && hi == ls + 16) {
                        full[u] = true;
                        da[u] = (char*)(uintptr_t)ls;
                        v[u] = ((uintptr_t)sp & 3) == 0 ? load16_a4(sp) : load16_shift(sp);
                    } else if (hi > lo) {
                        char* dp = (char*)(uintptr_t)lo;
                        const uint64_t nb = hi - lo;
                        if (((lo | (uint64_t)(uintptr_t)sp | nb) & 3) == 0) {
                            for (uint64_t q = 0; q < nb; q += 4)
                                *(uint32_t*)(dp + q) = *(const uint32_t*)(sp + q);
                        } else {
                            for (uint64_t q = 0; q < nb; q++) dp[q] = sp[q];
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (full[u]) store16(da[u], v[u], nt);
            line += (uint64_t)U * kWave;
        }
        line = lim;
        seg_base += 64;
    }
}

template <int U>
__global__ void __launch_bounds__(1024) a2av_step_kernel(A2avDev p) {
    const int P = p.P;
    const int l = lane();
    const uint32_t wpb = blockDim.x >> 6;
    const uint64_t W = (uint64_t)gridDim.x * wpb;
    const uint64_t w = (uint64_t)blockIdx.x * wpb + (threadIdx.x >> 6);

    // ---------------- phase 1: counts transpose, displacements, segment table
    for (int64_t d = (int64_t)w; d < P; d += (int64_t)W) {
        int64_t carry = 0;
        uint64_t carry_lines = 0;
        char* rbase = p.recvbufs[d];
        for (int s0 = 0; s0 < P; s0 += kWave) {
            const int s = s0 + l;
            const bool valid = s < P;
            int64_t c = valid ? p.scounts[(int64_t)s * P + d] : 0;
            int64_t sd = valid ? p.sdispls[(int64_t)s * P + d] : 0;
            int64_t tot;
            int64_t rd = wave_excl_scan(c, &tot) + carry;
            carry += tot;
            uint64_t nb = (uint64_t)c * (uint64_t)p.elem;
            char* dst = rbase + rd * p.elem;
            const char* src = valid ? p.sendbufs[s] + sd * p.elem : nullptr;
            uint64_t lines = nb ? (((uint64_t)(uintptr_t)dst + nb + 15) >> 4) -
                                      ((uint64_t)(uintptr_t)dst >> 4)
                                : 0;
            int64_t ltot;
            uint64_t lfirst = (uint64_t)wave_excl_scan((int64_t)lines, &ltot) + carry_lines;
            carry_lines += (uint64_t)ltot;
            if (valid) {
                p.rcounts[d * P + s] = (int32_t)c;
                p.rdispls[d * P + s] = (int32_t)rd;
                A2avSeg sg;
                sg.src = src;
                sg.dst = dst;
                sg.bytes = nb;
                sg.first = lfirst;
                p.table[d * P + s] = sg;
            }
        }
        if (l == 0) {
            p.row_lines[d] = carry_lines;
            p.row_bytes[d] = carry * p.elem;
        }
    }

    // ---------------- grid barrier; the last arriving wave builds the row offsets
    fence_release();
    uint64_t old = 0;
    if (l == 0) old = fetch_add(p.arrive, (uint64_t)1);
    old = bcast(old);
    const uint64_t epoch = old / W + 1;
    if (old % W == W - 1) {
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
    } else {
        const uint64_t t0 = wall_clock64();
        uint32_t spins = 0;
        for (;;) {
            uint64_t e = 0;
            if (l == 0) e = ld_acquire(p.release);
            e = bcast(e);
            if (e >= epoch) break;
            if ((++spins & 255) == 0 && wall_clock64() - t0 > p.timeout_ticks) {
                if (l == 0) st_relaxed(p.error, 1);
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        fence_acquire();
    }

    // ---------------- phase 2: even split of all destination lines over all waves
    const uint64_t L = p.row_off[P];
    const uint64_t b = L * w / W;          // L < 2^40 lines and W < 2^20 waves
    const uint64_t e = L * (w + 1) / W;
    a2av_copy_lines<U>(p, b, e);
}

__global__ void a2av_reset_kernel(uint64_t* arrive, uint64_t* release) {
    if (threadIdx.x == 0) {
        *arrive = 0;
        *release = 0;
    }
}

}  // namespace gpumpi
