// Device kernels of the bucket-sort caller of the exchange (the NPB IS rank() shape,
// toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c:631-928) for the batched path:
//   gen_keys      seeded keys: uniform [0, 2^key_bits), Zipf(s) by inverse CDF, or the
//                 NPB IS LCG (randlc/find_my_seed/create_seq, is.c:406-537) evaluated in
//                 exact integer arithmetic with per-thread jump-ahead
//   bucket_hist   per-rank bucket histogram (LDS-privatised)
//   bucket_split  NPB splitters (is.c:721-751): rank j takes buckets until the global
//                 running total reaches (j+1)*n; per-(source, destination) send counts
//                 and displacements
//   bucket_scatter stable bucketing (is.c:680-688) -- one wavefront per rank, 64 keys per
//                 step, each lane's rank among equal buckets from `bits` ballots
//   local_rank    counting sort of each rank's received keys over its bucket range
//                 (is.c:785-833, placement as full_verify is.c:545-560)
#pragma once

#include "philox.h"
#include "common.h"

namespace gpumpi {

constexpr uint64_t kNpbA = 1220703125ull;
constexpr uint64_t kNpbSeed = 314159265ull;
constexpr uint64_t kMask46 = (1ull << 46) - 1;

__host__ __device__ __forceinline__ uint64_t lcg_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = (r * a) & kMask46;
        a = (a * a) & kMask46;
        e >>= 1;
    }
    return r;
}

// find_my_seed (is.c:470-510): s * (a^(2^mq))^kn mod 2^46 with mq = floor(log2(nn/np)).
__host__ __device__ __forceinline__ uint64_t npb_rank_seed(int kn, int np, int64_t nn) {
    int64_t nq = nn / np;
    int mq = 0;
    for (; nq > 1; mq++, nq /= 2) {}
    uint64_t an = kNpbA;
    for (int i = 0; i < mq; i++) an = (an * an) & kMask46;
    return (kNpbSeed * lcg_pow(an, (uint64_t)kn)) & kMask46;
}

// NPB IS class parameters and partial-verification tables (is.c:57-104, 264-287).
struct NpbClassInfo {
    int total_keys_log2, max_key_log2, num_buckets_log2;
    int64_t test_index[5], test_rank[5];
};

__host__ __device__ inline bool npb_class_info(char cls, NpbClassInfo& c) {
    int64_t ti[5], tr[5];
    switch (cls) {
    case 'S': {
        c.total_keys_log2 = 16; c.max_key_log2 = 11; c.num_buckets_log2 = 9;
        const int64_t a[5] = {48427, 17148, 23627, 62548, 4431}, b[5] = {0, 18, 346, 64917, 65463};
        for (int i = 0; i < 5; i++) { ti[i] = a[i]; tr[i] = b[i]; }
        break;
    }
    case 'W': {
        c.total_keys_log2 = 20; c.max_key_log2 = 16; c.num_buckets_log2 = 10;
        const int64_t a[5] = {357773, 934767, 875723, 898999, 404505},
                      b[5] = {1249, 11698, 1039987, 1043896, 1048018};
        for (int i = 0; i < 5; i++) { ti[i] = a[i]; tr[i] = b[i]; }
        break;
    }
    case 'A': {
        c.total_keys_log2 = 23; c.max_key_log2 = 19; c.num_buckets_log2 = 10;
        const int64_t a[5] = {2112377, 662041, 5336171, 3642833, 4250760},
                      b[5] = {104, 17523, 123928, 8288932, 8388264};
        for (int i = 0; i < 5; i++) { ti[i] = a[i]; tr[i] = b[i]; }
        break;
    }
    case 'B': {
        c.total_keys_log2 = 25; c.max_key_log2 = 21; c.num_buckets_log2 = 10;
        const int64_t a[5] = {41869, 812306, 5102857, 18232239, 26860214},
                      b[5] = {33422937, 10244, 59149, 33135281, 99};
        for (int i = 0; i < 5; i++) { ti[i] = a[i]; tr[i] = b[i]; }
        break;
    }
    case 'C': {
        c.total_keys_log2 = 27; c.max_key_log2 = 23; c.num_buckets_log2 = 10;
        const int64_t a[5] = {44172927, 72999161, 74326391, 129606274, 21736814},
                      b[5] = {61147, 882988, 266290, 133997595, 133525895};
        for (int i = 0; i < 5; i++) { ti[i] = a[i]; tr[i] = b[i]; }
        break;
    }
    default:
        return false;
    }
    for (int i = 0; i < 5; i++) { c.test_index[i] = ti[i]; c.test_rank[i] = tr[i]; }
    return true;
}

// Expected rank of test key i at `iteration` (is.c:849-905).
__host__ __device__ inline int64_t npb_expected_rank(char cls, int i, int iteration, int64_t r) {
    switch (cls) {
    case 'S': return i <= 2 ? r + iteration : r - iteration;
    case 'W': return i < 2 ? r + iteration - 2 : r - iteration;
    case 'A': return i <= 2 ? r + iteration - 1 : r - (iteration - 1);
    case 'B': return (i == 1 || i == 2 || i == 4) ? r + iteration : r - iteration;
    default: return i <= 2 ? r + iteration : r - iteration;  // 'C'
    }
}

// Key i of rank r for the uniform (dist 0) and Zipf (dist 1) workloads.
__host__ __device__ __forceinline__ uint32_t gen_key(int dist, uint64_t seed, int r, uint32_t step,
                                                     int64_t i, int key_bits, const double* cdf,
                                                     int64_t ncdf) {
    if (dist == 0) {
        u32x4v v = rank_stream(seed, r, step, (uint64_t)(i >> 2), kStreamKeys);
        return pick(v, (int)(i & 3)) >> (32 - key_bits);
    }
    u32x4v v = rank_stream(seed, r, step, (uint64_t)(i >> 1), kStreamZipf);
    const int h = 2 * (int)(i & 1);
    const double u = unit53(pick(v, h), pick(v, h + 1));
    int64_t lo = 0, hi = ncdf - 1;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    return (uint32_t)lo;
}

// NPB IS create_seq key i of rank r: (x1+x2+x3+x4) >> (48 - max_key_log2), where the x's
// are the next four LCG states of the rank's sequence (which starts at the rank seed).
__host__ __device__ __forceinline__ uint32_t npb_key(int r, int P, int64_t total, int64_t i,
                                                     int max_key_log2) {
    uint64_t x = npb_rank_seed(r, P, 4 * total);
    x = (x * lcg_pow(kNpbA, 4ull * (uint64_t)i)) & kMask46;
    uint64_t s = 0;
    for (int j = 0; j < 4; j++) {
        x = (x * kNpbA) & kMask46;
        s += x;
    }
    return (uint32_t)(s >> (48 - max_key_log2));
}

static __global__ void gen_keys_kernel(uint32_t* keys, int P, int64_t n, uint64_t seed, uint32_t step,
                                int dist, int key_bits, const double* cdf, int64_t ncdf) {
    const int64_t total = (int64_t)P * n;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(g / n);
        const int64_t i = g - (int64_t)r * n;
        keys[g] = dist == 2 ? npb_key(r, P, total, i, key_bits)
                            : gen_key(dist, seed, r, step, i, key_bits, cdf, ncdf);
    }
}

// grid: (chunks per rank, P); hist: [P][nb]
static __global__ void bucket_hist_kernel(const uint32_t* keys, int64_t n, int shift, int nb,
                                   uint32_t* hist) {
    extern __shared__ uint32_t lh[];
    const int r = blockIdx.y;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    const uint32_t* k = keys + (int64_t)r * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&lh[k[i] >> shift], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        if (lh[b]) atomicAdd(&hist[(int64_t)r * nb + b], lh[b]);
}

// totals[b] = sum over ranks of hist[r][b]
static __global__ void bucket_totals_kernel(const uint32_t* hist, int P, int nb, int64_t* totals) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        int64_t s = 0;
        for (int r = 0; r < P; r++) s += hist[(int64_t)r * nb + b];
        totals[b] = s;
    }
}

// NPB splitters (is.c:721-751), one thread: bucket range [lo[j], hi[j]] of rank j.
static __global__ void bucket_splitters_kernel(const int64_t* totals, int P, int nb, int64_t n,
                                        int32_t* lo, int32_t* hi) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int64_t acc = 0;
    int j = 0;
    int prev = -1;
    for (int i = 0; i < nb; i++) {
        acc += totals[i];
        if (acc >= (int64_t)(j + 1) * n && j < P) {
            lo[j] = prev + 1;
            hi[j] = i;
            prev = i;
            j++;
        }
    }
    for (; j < P; j++) {
        lo[j] = 1;
        hi[j] = 0;
    }
}

// One wavefront per source rank: send_counts[r][d] = sum of hist[r][lo_d..hi_d],
// send_displs = exclusive prefix over d.
static __global__ void bucket_counts_kernel(const uint32_t* hist, int P, int nb, const int32_t* lo,
                                     const int32_t* hi, int32_t* scounts, int32_t* sdispls) {
    const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (r >= P) return;
    const uint32_t* h = hist + (int64_t)r * nb;
    int64_t carry = 0;
    for (int d0 = 0; d0 < P; d0 += kWave) {
        const int d = d0 + lane();
        int64_t c = 0;
        if (d < P)
            for (int b = lo[d]; b <= hi[d]; b++) c += h[b];
        int64_t inc = wave_incl_scan(c);
        if (d < P) {
            scounts[(int64_t)r * P + d] = (int32_t)c;
            sdispls[(int64_t)r * P + d] = (int32_t)(carry + inc - c);
        }
        carry += __shfl(inc, 63);
    }
}

// Stable bucketing: 4 ranks per 256-thread block, LDS bucket pointers per wave.
static __global__ void __launch_bounds__(256) bucket_scatter_kernel(const uint32_t* keys, int P, int64_t n,
                                                            int shift, int nb, int bits,
                                                            const uint32_t* hist, uint32_t* out) {
    extern __shared__ uint32_t sptr[];
    const int w = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + w;
    uint32_t* ptr = sptr + w * nb;
    if (r < P) {
        // exclusive prefix of this rank's histogram into the wave's LDS pointers
        const uint32_t* h = hist + (int64_t)r * nb;
        uint32_t carry = 0;
        for (int b0 = 0; b0 < nb; b0 += kWave) {
            const int b = b0 + lane();
            uint32_t c = b < nb ? h[b] : 0;
            uint32_t inc = (uint32_t)wave_incl_scan(c);
            if (b < nb) ptr[b] = carry + inc - c;
            carry += (uint32_t)__shfl((int)inc, 63);
        }
    }
    __syncthreads();
    if (r >= P) return;
    const uint32_t* k = keys + (int64_t)r * n;
    uint32_t* o = out + (int64_t)r * n;
    const uint64_t lt = (lane() == 0) ? 0ull : (~0ull >> (64 - lane()));
    for (int64_t base = 0; base < n; base += kWave) {
        const int64_t i = base + lane();
        const bool valid = i < n;
        const uint32_t key = valid ? k[i] : 0u;
        const uint32_t b = valid ? key >> shift : 0u;
        uint64_t match = ballot(valid);
        for (int t = 0; t < bits; t++) {
            const uint64_t m = ballot(valid && ((b >> t) & 1));
            match &= ((b >> t) & 1) ? m : ~m;
        }
        const uint32_t before = (uint32_t)__popcll(match & lt);
        const uint32_t p = valid ? ptr[b] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            o[p + before] = key;
            // the highest lane of each equal-bucket group advances the pointer
            if ((match >> lane()) == 1ull) ptr[b] = p + (uint32_t)__popcll(match);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------ local sort
// Every rank owns a disjoint, ascending key range (the bucket ranges), so the local
// sorts of all ranks together are one counting sort of the concatenated receive buffers:
// rank d's sorted keys are the slice [recv_off_d, recv_off_d + recv_tot_d) of the output.
constexpr int kSortLdsBins = 16384;  // keys below this are counted in LDS first

static __global__ void __launch_bounds__(1024) sort_hist_kernel(const uint32_t* keys, int64_t n,
                                                         uint32_t nbins, uint32_t* hist) {
    __shared__ uint32_t lh[kSortLdsBins];
    const uint32_t lds_bins = nbins < kSortLdsBins ? nbins : kSortLdsBins;
    for (uint32_t b = threadIdx.x; b < lds_bins; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t k = keys[i];
        if (k < lds_bins) atomicAdd(&lh[k], 1u);
        else atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < lds_bins; b += blockDim.x)
        if (lh[b]) atomicAdd(&hist[b], lh[b]);
}

// Exclusive scan of nbins counters in three kernels (4096 counters per 1024-thread block).
static __global__ void __launch_bounds__(1024) scan_block_sums_kernel(const uint32_t* h, uint32_t nbins,
                                                               uint64_t* sums) {
    __shared__ uint64_t red[16];
    const uint64_t base = (uint64_t)blockIdx.x * 4096;
    uint64_t s = 0;
    for (int k = 0; k < 4; k++) {
        const uint64_t i = base + (uint64_t)k * 1024 + threadIdx.x;
        if (i < nbins) s += h[i];
    }
    s = (uint64_t)wave_sum((int64_t)s);
    if (lane() == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < 16; w++) t += red[w];
        sums[blockIdx.x] = t;
    }
}

static __global__ void __launch_bounds__(1024) scan_sums_kernel(uint64_t* sums, int nblocks) {
    // single block, exclusive scan in place (sequential chunks of 1024)
    __shared__ uint64_t part[16];
    uint64_t carry = 0;
    for (int base = 0; base < nblocks; base += 1024) {
        const int i = base + threadIdx.x;
        const uint64_t v = i < nblocks ? sums[i] : 0;
        const uint64_t inc = (uint64_t)wave_incl_scan((int64_t)v);
        if (lane() == 63) part[threadIdx.x >> 6] = inc;
        __syncthreads();
        uint64_t before = 0;
        for (int w = 0; w < (int)(threadIdx.x >> 6); w++) before += part[w];
        uint64_t total = 0;
        for (int w = 0; w < 16; w++) total += part[w];
        if (i < nblocks) sums[i] = carry + before + inc - v;
        carry += total;
        __syncthreads();
    }
}

static __global__ void __launch_bounds__(1024) scan_apply_kernel(const uint32_t* h, uint32_t nbins,
                                                          const uint64_t* sums, uint32_t* prefix) {
    __shared__ uint64_t part[16];
    const uint64_t base = (uint64_t)blockIdx.x * 4096;
    uint64_t carry = sums[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        const uint64_t i = base + (uint64_t)k * 1024 + threadIdx.x;
        const uint64_t v = i < nbins ? h[i] : 0;
        const uint64_t inc = (uint64_t)wave_incl_scan((int64_t)v);
        if (lane() == 63) part[threadIdx.x >> 6] = inc;
        __syncthreads();
        uint64_t before = 0, total = 0;
        for (int w = 0; w < 16; w++) {
            if (w < (int)(threadIdx.x >> 6)) before += part[w];
            total += part[w];
        }
        if (i < nbins) prefix[i] = (uint32_t)(carry + before + inc - v);
        carry += total;
        __syncthreads();
    }
}

// Writes each value's run; runs longer than 256 go to a list filled by whole blocks.
static __global__ void sort_expand_kernel(const uint32_t* hist, const uint32_t* prefix, uint32_t nbins,
                                   uint32_t* out, uint32_t* big, uint32_t* nbig) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nbins;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = hist[v];
        if (c == 0) continue;
        if (c <= 256) {
            const uint32_t o = prefix[v];
            for (uint32_t k = 0; k < c; k++) out[o + k] = (uint32_t)v;
        } else {
            big[atomicAdd(nbig, 1u)] = (uint32_t)v;
        }
    }
}

static __global__ void sort_fill_runs_kernel(const uint32_t* hist, const uint32_t* prefix,
                                      const uint32_t* big, const uint32_t* nbig, uint32_t* out) {
    const uint32_t nb = *nbig;
    for (uint32_t r = blockIdx.x; r < nb; r += gridDim.x) {
        const uint32_t v = big[r];
        const uint32_t c = hist[v], o = prefix[v];
        for (uint32_t k = threadIdx.x; k < c; k += blockDim.x) out[o + k] = v;
    }
}

// Number of keys < bound in keys[0, n) (partial verification of NPB IS).
static __global__ void count_less_kernel(const uint32_t* keys, int64_t n, uint32_t bound,
                                  unsigned long long* out) {
    uint64_t c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        c += keys[i] < bound;
    c = (uint64_t)wave_sum((int64_t)c);
    if (lane() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

}  // namespace gpumpi
