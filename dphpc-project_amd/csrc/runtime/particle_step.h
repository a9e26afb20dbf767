// Batched particle step: the particle variant of the uneven exchange as one hipGraph-able
// step of four kernels, with no MPI calls inside.
//
// Workload: P = gx*gy*gz ranks on a periodic grid; rank r owns the unit cell
// (r % gx, (r / gx) % gy, r / (gx * gy)) and a slab of `cap` particle slots; a particle is
// 8 FP64 fields {x, y, z, vx, vy, vz, w, id} (64 bytes).  Each step a particle moves one
// cell along one of the six axis directions with probability kPartMove / 2^32, drawn from
// Philox keyed by (seed, id, step).  The step's result is the one the MPI program gives
// (Alltoall of the counts + MPI_Alltoallv, gpu_mpi/mpi.cu:333-408 semantics): in each
// rank's new slab its stayers first, in order, then the arrivals in ascending source-rank
// order, each source's migrants in their original order, the moved coordinate shifted by
// +-1 and wrapped.
//
// One step, reading (in, cin, dir_in) and writing (out, cout, dir_out):
//   ps_tally_kernel    one workgroup per rank: counts per class of every chunk of kPsChunk
//                      slots (the six directions merged where two reach the same rank, and
//                      stay), their exclusive offsets within the rank, the rank's totals
//   ps_place_kernel    one thread per destination rank: the first slot of every arriving
//                      (source, class) group, the new counts, the overflow flag, the step
//   ps_scatter_kernel  one wavefront per chunk: moves every particle to its final slot,
//                      shifts the migrants' coordinate, and writes the direction byte of the
//                      particle's next step beside its new slot
// dir_in[r*cap + k] is particle k's direction at this step (0..5, or kPsStay).  It is
// written by the previous step's scatter (or by ps_dirs_kernel before the first step), so
// the counting pass reads one byte per particle instead of the particle.
//
// HBM traffic per step: the particles read once and written once (2 x 64 B each) plus
// 2 B of direction bytes; everything else is O(P * chunks).
//
// Exchange mode 1 keeps the MPI program's neighbour mode (Isend/Irecv of each rank's
// migrants toward each neighbour) as messages: the place kernel writes one batched
// point-to-point pair per (source, destination) group -- from the source's send area to the
// group's slots in the destination slab -- the scatter kernel (PACK) writes stayers to their
// slots and migrants, by class, into the sender's send area, and the batched point-to-point
// step (batched_p2p.h) moves the groups.  The result is mode 0's bit for bit; the migrants
// cost 2 x 64 B more each (written to the send area, then copied).
#pragma once

#include "philox.h"
#include "common.h"
#include "copy.h"
#include "batched_p2p.h"

namespace gpumpi {

constexpr uint32_t kStreamParticles = 0x0B17E507u;
constexpr uint32_t kPartMove = 214748365u;   // ~0.05 * 2^32
constexpr int kPartFields = 8;

constexpr int kPsChunk = 256;     // particle slots per wavefront chunk
constexpr int kPsThreads = 256;   // four chunks per workgroup
constexpr int kPsStay = 6;        // direction byte of a particle that stays
constexpr int kPsSkip = 7;        // internal: slot beyond the count, or a dropped migrant
constexpr int kPsClasses = 8;     // per chunk / rank: six directions, stay, padding
constexpr int kPsTallyThreads = 256; // tally workgroup: one per rank
constexpr int kPsTallyBatch = 12;    // 16-byte reads in flight per thread while counting

__host__ __device__ __forceinline__ int grid_neighbour(int r, int d, int gx, int gy, int gz) {
    int x = r % gx, y = (r / gx) % gy, z = r / (gx * gy);
    switch (d) {
    case 0: x = x + 1 == gx ? 0 : x + 1; break;
    case 1: x = x == 0 ? gx - 1 : x - 1; break;
    case 2: y = y + 1 == gy ? 0 : y + 1; break;
    case 3: y = y == 0 ? gy - 1 : y - 1; break;
    case 4: z = z + 1 == gz ? 0 : z + 1; break;
    default: z = z == 0 ? gz - 1 : z - 1; break;
    }
    return x + gx * (y + gy * z);
}

// Direction of particle `id` at `step`: -1 stays, 0..5 = +x -x +y -y +z -z.
__host__ __device__ __forceinline__ int particle_dir(uint64_t seed, uint32_t step, uint64_t id) {
    u32x4v o = philox4x32_10(u32x4v{(uint32_t)id, (uint32_t)(id >> 32), step, kStreamMigrate},
                             (uint32_t)seed, (uint32_t)(seed >> 32));
    if (o.x >= kPartMove) return -1;
    return (int)(((uint64_t)o.y * 6u) >> 32);
}

// Class of each direction of rank r: the smallest direction reaching the same rank (on a
// grid axis of length 1 or 2 several directions reach one rank, possibly r itself), four
// bits per direction.
__host__ __device__ __forceinline__ uint32_t ps_classes(int r, int gx, int gy, int gz) {
    uint32_t packed = 0;
    for (int d = 0; d < 6; d++) {
        const int nb = grid_neighbour(r, d, gx, gy, gz);
        int c = d;
        for (int e = d - 1; e >= 0; e--)
            if (grid_neighbour(r, e, gx, gy, gz) == nb) c = e;
        packed |= (uint32_t)c << (4 * d);
    }
    return packed | ((uint32_t)kPsStay << 24);   // direction byte kPsStay -> class 6
}

struct PartStepDev {
    int32_t gx, gy, gz, P;
    int64_t cap;             // slots per rank slab
    uint64_t seed;
    int32_t nchunks;         // chunks per rank: ceil(cap / kPsChunk)
    int32_t pad;
    const double* in;        // [P][cap][8]
    const int64_t* cin;      // [P]
    const uint8_t* dir_in;   // [P][cap]
    double* out;
    int64_t* cout;
    uint8_t* dir_out;
    const uint32_t* cls_tab; // [P]: ps_classes of every rank
    const int32_t* nbr_tab;  // [P][8]: grid_neighbour of every rank in directions 0..5
    int32_t* chunk_off;      // [P][nchunks][8]: counts, then exclusive offsets in the rank
    int64_t* tot;            // [P][8]: per-rank totals per class
    int64_t* base;           // [P][8]: first slot of (source, class) in its destination slab
    uint32_t* step;          // [0]: the step dir_in belongs to (the place kernel advances it)
    int64_t* stats;          // [0] migrants since prepare, [1] overflow in this step, [2] sticky
    double* send;            // mode 1: [P][cap][8] migrants of each rank, by class
    P2pPair* pairs;          // mode 1: [P][6] messages of the step, by destination
};

__device__ __forceinline__ int64_t ps_count_of(const PartStepDev& a, int r) {
    const int64_t n = a.cin[r];
    return n < 0 ? 0 : n > a.cap ? a.cap : n;
}

__device__ __forceinline__ int ps_class_of(uint32_t packed, int d) {
    d = d > kPsStay ? kPsStay : d;
    return (int)((packed >> (4 * d)) & 0xFu);
}

__device__ __forceinline__ uint32_t ps_lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int32_t ps_wave_scan(int32_t v) {
    const int l = lane();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t w = __shfl_up(v, o);
        if (l >= o) v += w;
    }
    return v;
}

// Particles of the workload's initial state: n0 per rank, positions uniform in the rank's
// cell on a 2^-20 lattice, velocities on a 2^-15 lattice, id = rank * n0 + k.
static __global__ __launch_bounds__(kPsThreads) void ps_init_kernel(int gx, int gy, int gz, int64_t cap,
                                                             int64_t n0, uint64_t seed, double* parts,
                                                             int64_t* counts) {
    const int P = gx * gy * gz;
    const int64_t i = (int64_t)blockIdx.x * kPsThreads + threadIdx.x;
    if (i < P) counts[i] = n0;
    if (i >= (int64_t)P * n0) return;
    const int rank = (int)(i / n0);
    const int64_t k = i - (int64_t)rank * n0;
    const int cx = rank % gx, cy = (rank / gx) % gy, cz = rank / (gx * gy);
    const double q20 = 1.0 / 1048576.0, q15 = 1.0 / 32768.0;
    const u32x4v o = rank_stream(seed, rank, 0u, (uint64_t)k, kStreamParticles);
    double2* q = (double2*)(parts + ((uint64_t)rank * cap + k) * kPartFields);
    q[0] = make_double2((double)cx + (double)(o.x >> 12) * q20, (double)cy + (double)(o.y >> 12) * q20);
    q[1] = make_double2((double)cz + (double)(o.z >> 12) * q20,
                        (double)((int32_t)(o.w & 0xFFFFu) - 32768) * q15);
    q[2] = make_double2((double)((int32_t)(o.w >> 16) - 32768) * q15, (double)(k & 1023) * (1.0 / 1024.0));
    q[3] = make_double2(1.0, (double)((int64_t)rank * n0 + k));
}

// Direction bytes of (in, cin) at `step`; resets the step counter and the statistics.
static __global__ __launch_bounds__(kPsThreads) void ps_dirs_kernel(PartStepDev a, uint32_t step) {
    const int64_t i = (int64_t)blockIdx.x * kPsThreads + threadIdx.x;
    if (i == 0) {
        a.step[0] = step;
        a.stats[0] = 0;
        a.stats[1] = 0;
        a.stats[2] = 0;
    }
    if (i >= (int64_t)a.P * a.cap) return;
    const int r = (int)(i / a.cap);
    const int64_t k = i - (int64_t)r * a.cap;
    if (k >= ps_count_of(a, r)) return;
    const int d = particle_dir(a.seed, step, (uint64_t)a.in[(uint64_t)i * kPartFields + 7]);
    a.dir_out[i] = (uint8_t)(d < 0 ? kPsStay : d);
}

typedef u32x4 ps_u32x4_a1 __attribute__((aligned(1)));   // a slab may start at any byte

// One workgroup per rank: counts and exclusive offsets of its chunks per class, and the
// rank's totals.
//  count: thread t reads 16 direction bytes per round (slots 4096 j + 16 t .. +15, so the
//         16 threads 16 m .. 16 m + 15 cover chunk 16 j + m), kPsTallyBatch rounds in flight,
//         and adds one per slot into 9-bit fields of a 64-bit word (one field per class);
//         the 16 threads of a chunk sum their words and store the chunk's counts;
//  scan:  after a barrier, thread t takes chunk c0 + t of each pass of 256 chunks: wave
//         scans, wave totals through LDS, a running carry between passes.
static __global__ __launch_bounds__(kPsTallyThreads) void ps_tally_kernel(PartStepDev a) {
    const int r = blockIdx.x;
    const int t = threadIdx.x, wv = t >> 6, l = t & 63;
    if (r == 0 && t == 0) a.stats[1] = 0;
    const int64_t n = ps_count_of(a, r);
    const uint32_t cls = a.cls_tab[r];
    const uint8_t* dir = a.dir_in + (uint64_t)r * a.cap;
    int32_t* row = a.chunk_off + (int64_t)r * a.nchunks * kPsClasses;
    constexpr int kChunksPerRound = 16 * kPsTallyThreads / kPsChunk;
    const int rounds = (a.nchunks + kChunksPerRound - 1) / kChunksPerRound;
    for (int j0 = 0; j0 < rounds; j0 += kPsTallyBatch) {
        u32x4 word[kPsTallyBatch];
        int valid[kPsTallyBatch];
#pragma unroll
        for (int i = 0; i < kPsTallyBatch; i++) {
            const int64_t k = (int64_t)(j0 + i) * 16 * kPsTallyThreads + 16 * t;
            valid[i] = k < n ? (n - k < 16 ? (int)(n - k) : 16) : 0;
            word[i] = valid[i] > 0 ? *(const ps_u32x4_a1*)(dir + k) : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int i = 0; i < kPsTallyBatch; i++) {
            uint64_t acc = 0;
#pragma unroll
            for (int b = 0; b < 16; b++) {
                const uint32_t wd = b < 4 ? word[i].x : b < 8 ? word[i].y : b < 12 ? word[i].z : word[i].w;
                const int c = ps_class_of(cls, (int)((wd >> (8 * (b & 3))) & 0xFFu));
                acc += b < valid[i] ? 1ull << (9 * c) : 0ull;
            }
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) acc += __shfl_xor(acc, o);
            const int c = (j0 + i) * kChunksPerRound + (t >> 4);
            const int x = l & 15;
            if (x < kPsClasses && c < a.nchunks)
                row[(int64_t)c * kPsClasses + x] = x <= kPsStay ? (int32_t)((acc >> (9 * x)) & 511u) : 0;
        }
    }
    __syncthreads();   // the chunk counts above are read by other threads of the workgroup

    constexpr int kWaves = kPsTallyThreads / kWave;
    __shared__ int32_t s_w[kWaves][kPsClasses];     // wave totals
    __shared__ int32_t s_pre[kWaves][kPsClasses];   // exclusive prefix over the waves
    __shared__ int32_t s_all[kPsClasses];
    int32_t carry[kPsStay + 1];
#pragma unroll
    for (int x = 0; x <= kPsStay; x++) carry[x] = 0;
    for (int c0 = 0; c0 < a.nchunks; c0 += kPsTallyThreads) {
        const int c = c0 + t;
        int4 v0{0, 0, 0, 0}, v1{0, 0, 0, 0};
        if (c < a.nchunks) {
            v0 = ((const int4*)(row + (int64_t)c * kPsClasses))[0];
            v1 = ((const int4*)(row + (int64_t)c * kPsClasses))[1];
        }
        int32_t v[kPsStay + 1] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z};
        int32_t incl[kPsStay + 1];
#pragma unroll
        for (int x = 0; x <= kPsStay; x++) {
            incl[x] = ps_wave_scan(v[x]);
            if (l == kWave - 1) s_w[wv][x] = incl[x];
        }
        __syncthreads();
        if (t <= kPsStay) {
            int32_t acc = 0;
            for (int u = 0; u < kWaves; u++) {
                s_pre[u][t] = acc;
                acc += s_w[u][t];
            }
            s_all[t] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x <= kPsStay; x++) {
            v[x] = incl[x] - v[x] + s_pre[wv][x] + carry[x];
            carry[x] += s_all[x];
        }
        __syncthreads();
        if (c < a.nchunks) {
            ((int4*)(row + (int64_t)c * kPsClasses))[0] = int4{v[0], v[1], v[2], v[3]};
            ((int4*)(row + (int64_t)c * kPsClasses))[1] = int4{v[4], v[5], v[6], 0};
        }
    }
    int64_t tt = 0;
#pragma unroll
    for (int x = 0; x <= kPsStay; x++) tt = t == x ? carry[x] : tt;
    if (t < kPsClasses) a.tot[(int64_t)r * kPsClasses + t] = tt;
}

// Destination j's slab after the step: its stayers, then for every source s that sends to
// j, in ascending s, s's migrants toward j (one class of s reaches j).  The sources of j
// are its own neighbours: grid_neighbour(s, d) == j exactly when s is j's neighbour in the
// opposite direction.
template <bool P2P>
__global__ __launch_bounds__(kPsThreads) void ps_place_kernel(PartStepDev a) {
    const int j = (int)((int64_t)blockIdx.x * kPsThreads + threadIdx.x);
    if (j == 0) a.step[0] += 1u;
    if (j >= a.P) return;
    // step 1: j's neighbours (its possible sources) and its stayers
    const int4* nt = (const int4*)a.nbr_tab;
    const int4 n0 = nt[2 * j], n1 = nt[2 * j + 1];
    const int64_t stay = a.tot[(int64_t)j * kPsClasses + kPsStay];
    int src[6] = {n0.x, n0.y, n0.z, n0.w, n1.x, n1.y};
    // ascending order (odd-even transposition sort; all indices fixed at compile time)
#pragma unroll
    for (int pass = 0; pass < 6; pass++) {
#pragma unroll
        for (int i = pass & 1; i + 1 < 6; i += 2) {
            const int lo = src[i] < src[i + 1] ? src[i] : src[i + 1];
            const int hi = src[i] < src[i + 1] ? src[i + 1] : src[i];
            src[i] = lo;
            src[i + 1] = hi;
        }
    }
    // step 2: the class of each source toward j (the smallest direction reaching j)
    int cl[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int4 m0 = nt[2 * src[i]], m1 = nt[2 * src[i] + 1];
        const int m[6] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y};
        int c = 0;
#pragma unroll
        for (int d = 5; d >= 0; d--) c = m[d] == j ? d : c;
        cl[i] = c;
    }
    // step 3: the sources' counts toward j
    int64_t cnt[6];
#pragma unroll
    for (int i = 0; i < 6; i++) cnt[i] = a.tot[(int64_t)src[i] * kPsClasses + cl[i]];
    a.base[(int64_t)j * kPsClasses + kPsStay] = 0;
    int64_t acc = stay;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const bool dup = i > 0 && src[i] == src[i - 1];   // a rank reached in two directions
        if constexpr (P2P) {
            // the message of group (src[i], class cl[i]): its place in the source's send
            // area (the classes before it come first) and in j's slab
            P2pPair m{nullptr, nullptr, 0};
            if (!dup && cnt[i] > 0) {
                int64_t sb = 0;
                for (int y = 0; y < cl[i]; y++) sb += a.tot[(int64_t)src[i] * kPsClasses + y];
                m.src = (const char*)(a.send + ((int64_t)src[i] * a.cap + sb) * kPartFields);
                m.dst = (char*)(a.out + ((int64_t)j * a.cap + acc) * kPartFields);
                m.bytes = (uint64_t)cnt[i] * kPartFields * sizeof(double);
            }
            a.pairs[(int64_t)j * 6 + i] = m;
        }
        if (dup) continue;
        a.base[(int64_t)src[i] * kPsClasses + cl[i]] = acc;
        acc += cnt[i];
    }
    a.cout[j] = acc;
    if (acc > a.cap) {
        a.stats[1] = 1;
        a.stats[2] = 1;
        if constexpr (P2P)
            for (int i = 0; i < 6; i++) a.pairs[(int64_t)j * 6 + i].bytes = 0;
    }
    atomicAdd((unsigned long long*)&a.stats[0], (unsigned long long)(acc - stay));
}

// Moves one chunk's particles (PACK: stayers to their slots, migrants into the rank's send
// area by class; their direction bytes still go beside their final slots).  The wavefront
// reads the chunk's 64 particles of a group as
// four fully coalesced 1 KiB rows (lane l holds quarter l % 4 of particle 16 q + l / 4 in
// row q); lane l also decides particle l of the group (its class, its rank within the class
// by ballots, its final slot and its next direction), and the slots reach the lanes
// holding the data by ds_bpermute.  Stores of consecutive particles to consecutive slots
// form 1 KiB rows as well.
template <bool NT, bool PACK>
__global__ __launch_bounds__(kPsThreads) void ps_scatter_kernel(PartStepDev a) {
    constexpr int G = kPsChunk / kWave;
    if constexpr (PACK) {
        // a step that overflows anywhere keeps only the stayers: no message moves
        if (a.stats[1] != 0) {
            const int64_t nt = (int64_t)gridDim.x * kPsThreads;
            for (int64_t i = (int64_t)blockIdx.x * kPsThreads + threadIdx.x; i < (int64_t)a.P * 6; i += nt)
                a.pairs[i].bytes = 0;
        }
    }
    const int64_t w = ((int64_t)blockIdx.x * kPsThreads + threadIdx.x) >> 6;
    if (w >= (int64_t)a.P * a.nchunks) return;
    const int r = (int)(w / a.nchunks);
    const int64_t k0 = (w - (int64_t)r * a.nchunks) * kPsChunk;
    const int64_t n = ps_count_of(a, r);
    if (k0 >= n) return;
    const int l = lane();

    // every load of the chunk first, unconditionally: slots past the count read the
    // chunk's first particle (in bounds) and are dropped below
    const uint64_t slab = (uint64_t)r * a.cap;
    const u32x4* src = (const u32x4*)(a.in + (slab + k0) * kPartFields);
    u32x4 v[G][4];
    uint64_t idbits[G];
    uint32_t dirb[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int p = g * kWave + q * 16 + (l >> 2);
            const int idx = k0 + p < n ? p * 4 + (l & 3) : (l & 3);
            v[g][q] = src[idx];
        }
        const int64_t k = k0 + g * kWave + l < n ? k0 + g * kWave + l : k0;
        idbits[g] = *(const uint64_t*)(a.in + (slab + k) * kPartFields + 7);
        dirb[g] = a.dir_in[slab + k];
    }

    const uint32_t cls = a.cls_tab[r];
    const bool dropped = a.stats[1] != 0;
    const uint32_t next_step = a.step[0];
    // slot cursors per class: destination slab start + group start + chunk offset
    int64_t cur[kPsStay + 1];
    int64_t wcur[kPsStay];   // PACK: migrants' slots in the send area
    {
        const int x = l < kPsClasses ? l : 0;
        const int dest = x < 6 ? a.nbr_tab[(int64_t)r * kPsClasses + x] : r;
        const int32_t off = a.chunk_off[w * kPsClasses + x];
        int64_t cv = (int64_t)dest * a.cap + a.base[(int64_t)r * kPsClasses + x] + off;
#pragma unroll
        for (int c = 0; c <= kPsStay; c++) cur[c] = (int64_t)readlane64((uint64_t)cv, c);
        if constexpr (PACK) {
            // the send area holds the classes in order, each class in chunk order
            const int32_t t6 = l < 6 ? (int32_t)a.tot[(int64_t)r * kPsClasses + l] : 0;
            const int64_t sv = (int64_t)r * a.cap + (ps_wave_scan(t6) - t6) + off;
#pragma unroll
            for (int c = 0; c < kPsStay; c++) wcur[c] = (int64_t)readlane64((uint64_t)sv, c);
        }
    }

    u32x4* dst = (u32x4*)a.out;
#pragma unroll
    for (int g = 0; g < G; g++) {
        // lane l: particle l of the group
        const int d = k0 + g * kWave + l < n ? (dirb[g] > kPsStay ? kPsStay : (int)dirb[g]) : kPsSkip;
        const int c = d == kPsSkip ? kPsSkip : ps_class_of(cls, d);
        int64_t slot = -1, wslot = -1;   // final slot; slot the data is written to
#pragma unroll
        for (int x = 0; x <= kPsStay; x++) {
            const uint64_t m = ballot(c == x);
            const int64_t below = (int64_t)ps_lanes_below(m);
            if (c == x) {
                slot = cur[x] + below;
                wslot = slot;
                if constexpr (PACK)
                    if (x < kPsStay) wslot = wcur[x] + below;
            }
            cur[x] += __popcll(m);
            if constexpr (PACK)
                if (x < kPsStay) wcur[x] += __popcll(m);
        }
        int code = d;
        if (c == kPsSkip || (dropped && c != kPsStay)) code = kPsSkip;
        if (code != kPsSkip) {
            double idd;
            __builtin_memcpy(&idd, &idbits[g], 8);
            const int nd = particle_dir(a.seed, next_step, (uint64_t)idd);
            a.dir_out[slot] = (uint8_t)(nd < 0 ? kPsStay : nd);
        }
        const uint32_t lo = (uint32_t)(uint64_t)wslot;
        const uint32_t hi = ((uint32_t)((uint64_t)wslot >> 32) & 0x00FFFFFFu) | ((uint32_t)code << 24);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int from = (q * 16 + (l >> 2)) << 2;
            const uint32_t plo = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)lo);
            const uint32_t phi = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)hi);
            const int pd = (int)(phi >> 24);
            if (pd == kPsSkip) continue;
            u32x4 val = v[g][q];
            const int part = l & 3;
            const int ax = pd >> 1;
            // the moved coordinate: x, y in quarter 0, z in quarter 1
            if (pd < kPsStay && ((part == 0 && ax < 2) || (part == 1 && ax == 2))) {
                const bool second = part == 0 && ax == 1;
                uint64_t bits = second ? ((uint64_t)val.w << 32 | val.z) : ((uint64_t)val.y << 32 | val.x);
                double cval;
                __builtin_memcpy(&cval, &bits, 8);
                double cn = cval + ((pd & 1) ? -1.0 : 1.0);
                const double dim = (double)(ax == 0 ? a.gx : ax == 1 ? a.gy : a.gz);
                if (cn >= dim) cn -= dim;
                if (cn < 0.0) cn += dim;
                __builtin_memcpy(&bits, &cn, 8);
                if (second) { val.z = (uint32_t)bits; val.w = (uint32_t)(bits >> 32); }
                else { val.x = (uint32_t)bits; val.y = (uint32_t)(bits >> 32); }
            }
            const uint64_t ps = ((uint64_t)(phi & 0x00FFFFFFu) << 32) | plo;
            u32x4* base = dst;
            if constexpr (PACK)
                if (pd < kPsStay) base = (u32x4*)a.send;
            u32x4* p = base + ps * 4 + part;
            if (NT) __builtin_nontemporal_store(val, p);
            else *p = val;
        }
    }
}

}  // namespace gpumpi
