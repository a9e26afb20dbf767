// Distributed bucket sort as a rank program: the NPB IS rank() / full_verify() steps
// (toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c:631-928, 545-621) written against the
// device MPI API, the way the converted benchmark runs on the rank grid.  Every MPI call is
// the reference's: MPI_Allreduce of the bucket sizes, MPI_Alltoall of the send counts,
// MPI_Alltoallv of the keys, MPI_Irecv/MPI_Send of the boundary key, MPI_Reduce of the
// verification count.
//
// Keys: dist 0 uniform in [0, 2^key_bits), dist 1 Zipf by inverse CDF (the exchange
// workload's generators), dist 2 NPB IS create_seq for class `cls` (is.c:520-537) with
// its partial verification (is.c:849-905).  A rank is one wavefront: histograms and
// bucket pointers live in LDS, the ranking counters in HBM.
#pragma once

#include "workloads.h"
#include "particles.h"
#include "../runtime/bucket.h"

namespace gpumpi {
namespace progs {

struct BucketSortArgs {
    int32_t dist;         // 0 uniform, 1 Zipf, 2 NPB IS
    int32_t cls;          // NPB class letter (dist 2)
    int32_t key_bits;     // dist 0/1
    int32_t bucket_bits;  // dist 0/1
    int32_t iterations;   // NPB: MAX_ITERATIONS (10); after one untimed iteration
    int32_t pad;
    int64_t n;            // keys per rank (dist 0/1)
    uint64_t seed;
    const double* cdf;    // dist 1
    int64_t ncdf;
    int64_t kb2_cap;      // per-rank capacity of key_array / key_buff2 (keys)
    int64_t count_cap;    // per-rank capacity of the ranking counters
    uint32_t* key_array;  // [P][kb2_cap]: the keys, after the run the rank's sorted keys
    uint32_t* key_buff1;  // [P][n]
    uint32_t* key_buff2;  // [P][kb2_cap]
    int32_t* counting;    // [P][count_cap]
    int32_t* bsize;       // [P][2^bucket_bits + 5]
    int32_t* btot;        // [P][2^bucket_bits + 5]
    int32_t* cnt;         // [7][P][P]
    int64_t* nlocal;      // [P] keys the rank holds after the last iteration
    int64_t* passed;      // [P] verification passes of the rank
    int64_t* total;       // [1] MPI_Reduce(SUM) of passed at rank 0
    double* seconds;      // [P] timed iterations
};

constexpr int kSortMaxBuckets = 1024;

// Waits for this wave's global stores/atomics and drops stale L1 lines, so a lane reads
// what another lane of the same rank wrote.
__device__ __forceinline__ void rank_mem_sync() {
    fence_release();
    fence_acquire();
}

__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ int w_bucket_sort(int, char**) {
    __shared__ uint32_t s_ptr[4][kSortMaxBuckets];
    __shared__ uint32_t s_cnt[4][kSortMaxBuckets];
    const BucketSortArgs& a = args<BucketSortArgs>();
    MPI_Init(nullptr, nullptr);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    const bool npb = a.dist == 2;
    NpbClassInfo ci{};
    int key_bits = a.key_bits, bucket_bits = a.bucket_bits;
    int64_t n = a.n;
    if (npb) {
        if (!npb_class_info((char)a.cls, ci)) {
            MPI_Abort(MPI_COMM_WORLD, 2);
            return 1;
        }
        key_bits = ci.max_key_log2;
        bucket_bits = ci.num_buckets_log2;
        n = ((int64_t)1 << ci.total_keys_log2) / P;
    }
    const int64_t NB = (int64_t)1 << bucket_bits;
    const int shift = key_bits - bucket_bits;
    const int64_t maxkey = (int64_t)1 << key_bits;
    if (NB > kSortMaxBuckets || shift < 0 || n > a.kb2_cap || (npb && n < 2 * a.iterations + 1)) {
        MPI_Abort(MPI_COMM_WORLD, 2);
        return 1;
    }
    const MPI_Datatype kt = npb ? MPI_INT : MPI_UNSIGNED;
    uint32_t* ka = a.key_array + (int64_t)rank * a.kb2_cap;
    uint32_t* kb1 = a.key_buff1 + (int64_t)rank * n;
    uint32_t* kb2 = a.key_buff2 + (int64_t)rank * a.kb2_cap;
    int32_t* cntg = a.counting + (int64_t)rank * a.count_cap;
    int32_t* bs = a.bsize + (int64_t)rank * (NB + 5);
    int32_t* bt = a.btot + (int64_t)rank * (NB + 5);
    const int64_t PP = (int64_t)P * P;
    int32_t* sc = a.cnt + (int64_t)rank * P;
    int32_t* sd = sc + PP;
    int32_t* rc = sd + PP;
    int32_t* rd = rc + PP;
    int32_t* p1 = rd + PP;
    int32_t* p2 = p1 + PP;
    int32_t* kv = p2 + PP;  // [0] boundary key received, [1] boundary key sent
    uint32_t* lptr = s_ptr[threadIdx.x >> 6];
    uint32_t* lcnt = s_cnt[threadIdx.x >> 6];

    // keys: NPB create_seq with a jump-ahead per lane (lane l makes a contiguous chunk)
    if (npb) {
        const int64_t T = (int64_t)1 << ci.total_keys_log2;
        const int64_t chunk = (n + kWave - 1) / kWave;
        const int64_t i0 = (int64_t)lane() * chunk;
        const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
        uint64_t x = npb_rank_seed(rank, P, 4 * T);
        x = (x * lcg_pow(kNpbA, 4ull * (uint64_t)i0)) & kMask46;
        for (int64_t i = i0; i < i1; i++) {
            uint64_t s = 0;
            for (int j = 0; j < 4; j++) {
                x = (x * kNpbA) & kMask46;
                s += x;
            }
            ka[i] = (uint32_t)(s >> (48 - key_bits));
        }
    } else {
        for (int64_t i = lane(); i < n; i += kWave)
            ka[i] = gen_key(a.dist, a.seed, rank, 0u, i, key_bits, a.cdf, a.ncdf);
    }
    rank_mem_sync();

    int64_t passed = 0, jl = 0, minv = 0, maxv = -1;
    double t0 = 0;
    const int iters = a.iterations;
    for (int step = 0; step <= iters; step++) {
        const int it = step == 0 ? 1 : step;  // step 0: the untimed rank(1) of is.c main
        if (step == 1) {
            MPI_Barrier(MPI_COMM_WORLD);
            t0 = MPI_Wtime();
        }
        // is.c:650-654
        if (npb && rank == 0) {
            ka[it] = (uint32_t)it;
            ka[it + iters] = (uint32_t)(maxkey - it);
            rank_mem_sync();
        }
        // local histogram (is.c:667-675) with the test keys in the extra slots (661-665)
        for (int64_t b = lane(); b < NB; b += kWave) lcnt[b] = 0;
        lds_sync();
        for (int t = 0; t < 5; t++) {
            int32_t v = 0;
            if (npb && ci.test_index[t] / n == rank) v = (int32_t)ka[ci.test_index[t] % n];
            bs[NB + t] = v;
        }
        for (int64_t i = lane(); i < n; i += kWave) atomicAdd(&lcnt[ka[i] >> shift], 1u);
        lds_sync();
        {
            uint32_t carry = 0;
            for (int64_t b0 = 0; b0 < NB; b0 += kWave) {
                const int64_t b = b0 + lane();
                const uint32_t c = b < NB ? lcnt[b] : 0u;
                const uint32_t inc = (uint32_t)wave_incl_scan(c);
                if (b < NB) {
                    bs[b] = (int32_t)c;
                    lptr[b] = carry + inc - c;
                }
                carry += (uint32_t)__shfl((int)inc, 63);
            }
        }
        lds_sync();
        // stable bucketing (is.c:677-688): each lane's place among the lanes with the same
        // bucket comes from one ballot per bucket bit
        const uint64_t lt = lane() == 0 ? 0ull : (~0ull >> (64 - lane()));
        for (int64_t base = 0; base < n; base += kWave) {
            const int64_t i = base + lane();
            const bool valid = i < n;
            const uint32_t key = valid ? ka[i] : 0u;
            const uint32_t b = key >> shift;
            uint64_t match = ballot(valid);
            for (int t = 0; t < bucket_bits; t++) {
                const uint64_t m = ballot(valid && ((b >> t) & 1));
                match &= ((b >> t) & 1) ? m : ~m;
            }
            const uint32_t p = valid ? lptr[b] : 0u;
            lds_sync();
            if (valid) {
                kb1[p + (uint32_t)__popcll(match & lt)] = key;
                if ((match >> lane()) == 1ull) lptr[b] = p + (uint32_t)__popcll(match);
            }
            lds_sync();
        }
        // bucket size totals (is.c:694-701)
        if (MPI_Allreduce(bs, bt, (int)(NB + 5), MPI_INT, MPI_SUM, MPI_COMM_WORLD) != MPI_SUCCESS) break;
        // redistribution (is.c:721-751): a uniform walk over the buckets, 64 at a time
        // loaded into lanes and read back with readlane
        {
            int64_t acc = 0, lacc = 0, j = 0;
            int32_t prev_sc = 0, prev_sd = 0, prev_p2 = 0;
            for (int64_t b0 = 0; b0 < NB; b0 += kWave) {
                const int64_t b = b0 + lane();
                const int32_t vt = b < NB ? bt[b] : 0;
                const int32_t vl = b < NB ? (int32_t)lcnt[b] : 0;
                for (int t = 0; t < kWave && b0 + t < NB; t++) {
                    acc += readlane(vt, t);
                    lacc += readlane(vl, t);
                    if (acc >= (j + 1) * n) {
                        const int32_t i = (int32_t)(b0 + t);
                        const int32_t d = j == 0 ? 0 : prev_sd + prev_sc;
                        const int32_t q1 = j == 0 ? 0 : prev_p2 + 1;
                        sc[j] = (int32_t)lacc;
                        sd[j] = d;
                        p1[j] = q1;
                        p2[j] = i;
                        prev_sc = (int32_t)lacc;
                        prev_sd = d;
                        prev_p2 = i;
                        j++;
                        lacc = 0;
                    }
                }
            }
            const int32_t end = j == 0 ? 0 : prev_sd + prev_sc;
            for (; j < P; j++) {
                sc[j] = 0;
                sd[j] = end;
                p1[j] = 1;
                p2[j] = 0;
            }
        }
        lds_sync();  // lcnt is rewritten by the next iteration
        rank_mem_sync();
        if (MPI_Alltoall(sc, 1, MPI_INT, rc, 1, MPI_INT, MPI_COMM_WORLD) != MPI_SUCCESS) break;
        int64_t nrecv = 0;
        {
            int64_t carry = 0;
            for (int j0 = 0; j0 < P; j0 += kWave) {
                const int j = j0 + lane();
                const int64_t c = j < P ? rc[j] : 0;
                const int64_t inc = wave_incl_scan(c);
                if (j < P) rd[j] = (int32_t)(carry + inc - c);
                carry += __shfl(inc, 63);
            }
            nrecv = carry;
        }
        if (nrecv > a.kb2_cap) {
            MPI_Abort(MPI_COMM_WORLD, 3);
            return 1;
        }
        rank_mem_sync();
        if (MPI_Alltoallv(kb1, sc, sd, kt, kb2, rc, rd, kt, MPI_COMM_WORLD) != MPI_SUCCESS) break;

        // ranking (is.c:785-833)
        minv = (int64_t)p1[rank] << shift;
        maxv = (((int64_t)p2[rank] + 1) << shift) - 1;
        const int64_t range = maxv - minv + 1;
        if (range > a.count_cap) {
            MPI_Abort(MPI_COMM_WORLD, 3);
            return 1;
        }
        for (int64_t i = lane(); i < range; i += kWave) cntg[i] = 0;
        int64_t m = 0, jj = 0;
        for (int k = lane(); k < P; k += kWave) {
            if (k < rank)
                for (int64_t i = p1[k]; i <= p2[k]; i++) m += bt[i];
        }
        for (int64_t i = p1[rank] + lane(); i <= p2[rank]; i += kWave) jj += bt[i];
        m = wave_sum(m);
        jl = wave_sum(jj);
        rank_mem_sync();
        for (int64_t i = lane(); i < jl; i += kWave) atomicAdd(&cntg[(int64_t)kb2[i] - minv], 1);
        rank_mem_sync();
        {
            int64_t carry = 0;
            for (int64_t b0 = 0; b0 < range; b0 += kWave) {
                const int64_t b = b0 + lane();
                const int64_t c = b < range ? cntg[b] : 0;
                const int64_t inc = wave_incl_scan(c);
                if (b < range) cntg[b] = (int32_t)(carry + inc);
                carry += __shfl(inc, 63);
            }
        }
        rank_mem_sync();
        // partial verification (is.c:838-905)
        if (npb && step > 0) {
            for (int t = 0; t < 5; t++) {
                const int64_t k = bt[NB + t];
                if (minv <= k && k <= maxv) {
                    const int64_t key_rank = (k - 1 - minv >= 0 ? cntg[k - 1 - minv] : 0) + m;
                    if (key_rank == npb_expected_rank((char)a.cls, t, it, ci.test_rank[t])) passed++;
                }
            }
        }
    }
    const double t1 = MPI_Wtime();

    // full_verify (is.c:545-621): place the keys, pass the largest key to the next rank,
    // count keys out of order
    for (int64_t i = lane(); i < jl; i += kWave) {
        const uint32_t k = kb2[i];
        const int32_t pos = atomicSub(&cntg[(int64_t)k - minv], 1) - 1;
        ka[pos] = k;
    }
    rank_mem_sync();
    const int64_t last = jl < 1 ? 0 : jl - 1;
    kv[1] = (int32_t)ka[last];
    rank_mem_sync();
    MPI_Request req = MPI_REQUEST_NULL;
    if (rank > 0) MPI_Irecv(&kv[0], 1, kt, rank - 1, 1000, MPI_COMM_WORLD, &req);
    if (rank < P - 1) MPI_Send(&kv[1], 1, kt, rank + 1, 1000, MPI_COMM_WORLD);
    if (rank > 0) MPI_Wait(&req, MPI_STATUS_IGNORE);
    int64_t bad = 0;
    if (rank > 0 && jl > 0 && (uint32_t)kv[0] > ka[0]) bad++;
    int64_t lb = 0;
    for (int64_t i = 1 + lane(); i < jl; i += kWave) lb += ka[i - 1] > ka[i] ? 1 : 0;
    bad += wave_sum(lb);
    if (bad == 0) passed++;
    int64_t total = 0;
    MPI_Reduce(&passed, &total, 1, MPI_INT64_T, MPI_SUM, 0, MPI_COMM_WORLD);
    a.nlocal[rank] = jl;
    a.passed[rank] = passed;
    a.seconds[rank] = t1 - t0;
    if (rank == 0) *a.total = total;
    MPI_Finalize();
    return 0;
}

}  // namespace progs
}  // namespace gpumpi
