/*
 * Host-MPI version of the bucket-sort exchange workload (TEST INFRASTRUCTURE / CPU baseline).
 *
 * Each rank draws n uint32 keys (uniform in [0, 2^key_bits) or Zipf(1.1)) from the same
 * Philox streams as the GPU runtime, buckets them NPB-style (NPB IS rank(),
 * toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c:673-781: local histogram, MPI_Allreduce
 * of bucket sizes, splitters, MPI_Alltoall of counts, MPI_Alltoallv of keys) and sorts
 * its received keys locally.  Timing follows benchmarks/mpi_communication/
 * mpi_communication.c:24-49: barrier, warm-ups, timed repetitions, per-rank average,
 * MPI_Reduce(MAX) to rank 0.  Only the exchange (Alltoall + Alltoallv) is timed.
 *
 * usage: a2av_bucket n dist(0 uniform, 1 zipf) seed iters warmup [dump_prefix]
 * prints: "a2av ranks=P n=N dist=D payload_bytes=B seconds=T sorted_ok=K checksum=C"
 */
#include <math.h>
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gpumpi_oracle.h"

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int me, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &me);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (argc < 6) {
        if (me == 0) fprintf(stderr, "usage: %s n dist seed iters warmup [dump_prefix]\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int64_t n = atoll(argv[1]);
    const int dist = atoi(argv[2]);
    const uint64_t seed = strtoull(argv[3], NULL, 0);
    const int iters = atoi(argv[4]);
    const int warmup = atoi(argv[5]);
    const int key_bits = 23, bucket_bits = 10;
    const int NB = 1 << bucket_bits, shift = key_bits - bucket_bits;

    uint32_t *keys = malloc(sizeof(uint32_t) * n);
    if (dist == 0) {
        orc_fill_keys_uniform(seed, me, 0, keys, n, key_bits);
    } else {
        int64_t nk = (int64_t)1 << key_bits;
        double *cdf = malloc(sizeof(double) * nk);
        double acc = 0;
        for (int64_t k = 0; k < nk; k++) { acc += pow((double)(k + 1), -1.1); cdf[k] = acc; }
        for (int64_t k = 0; k < nk; k++) cdf[k] /= acc;
        orc_fill_keys_zipf(seed, me, 0, cdf, nk, keys, n);
        free(cdf);
    }
    /* local histogram + stable bucketing (is.c:673-688) */
    int *bsize = calloc(NB, sizeof(int)), *btot = calloc(NB, sizeof(int));
    for (int64_t i = 0; i < n; i++) bsize[keys[i] >> shift]++;
    int *bptr = malloc(sizeof(int) * NB);
    bptr[0] = 0;
    for (int b = 1; b < NB; b++) bptr[b] = bptr[b - 1] + bsize[b - 1];
    uint32_t *kb1 = malloc(sizeof(uint32_t) * n);
    for (int64_t i = 0; i < n; i++) kb1[bptr[keys[i] >> shift]++] = keys[i];
    MPI_Allreduce(bsize, btot, NB, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
    /* splitters (is.c:721-751) */
    int *sc = calloc(P, sizeof(int)), *sd = calloc(P, sizeof(int));
    int *rc = calloc(P, sizeof(int)), *rd = calloc(P, sizeof(int));
    int64_t acc = 0, lacc = 0;
    int j = 0;
    for (int b = 0; b < NB; b++) {
        acc += btot[b];
        lacc += bsize[b];
        if (acc >= (int64_t)(j + 1) * n && j < P) { sc[j++] = (int)lacc; lacc = 0; }
    }
    for (int d = 1; d < P; d++) sd[d] = sd[d - 1] + sc[d - 1];

    uint32_t *kb2 = NULL;
    int64_t nrecv = 0;
    double tsum = 0;
    for (int it = -warmup; it < iters; it++) {
        MPI_Barrier(MPI_COMM_WORLD);
        double t0 = MPI_Wtime();
        MPI_Alltoall(sc, 1, MPI_INT, rc, 1, MPI_INT, MPI_COMM_WORLD);
        rd[0] = 0;
        for (int d = 1; d < P; d++) rd[d] = rd[d - 1] + rc[d - 1];
        nrecv = rd[P - 1] + rc[P - 1];
        if (!kb2) kb2 = malloc(sizeof(uint32_t) * (nrecv ? nrecv : 1));
        MPI_Alltoallv(kb1, sc, sd, MPI_UNSIGNED, kb2, rc, rd, MPI_UNSIGNED, MPI_COMM_WORLD);
        double t1 = MPI_Wtime();
        if (it >= 0) tsum += t1 - t0;
    }
    double tavg = iters > 0 ? tsum / iters : 0, tmax = 0;
    MPI_Reduce(&tavg, &tmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    int64_t bytes = nrecv * 4, total_bytes = 0;
    MPI_Reduce(&bytes, &total_bytes, 1, MPI_INT64_T, MPI_SUM, 0, MPI_COMM_WORLD);

    /* local sort + global order check across neighbours (is.c:545-621) */
    qsort(kb2, (size_t)nrecv, sizeof(uint32_t), cmp_u32);
    uint64_t cs = 0;
    for (int64_t i = 0; i < nrecv; i++) cs = cs * 1000003u + kb2[i];
    uint32_t last = nrecv ? kb2[nrecv - 1] : 0, prev_last = 0;
    int have = nrecv > 0, prev_have = 0;
    MPI_Status st;
    if (me > 0) {
        MPI_Recv(&prev_last, 1, MPI_UNSIGNED, me - 1, 7, MPI_COMM_WORLD, &st);
        MPI_Recv(&prev_have, 1, MPI_INT, me - 1, 8, MPI_COMM_WORLD, &st);
    }
    if (me < P - 1) {
        uint32_t send_last = have ? last : prev_last;
        int send_have = have || prev_have;
        MPI_Send(&send_last, 1, MPI_UNSIGNED, me + 1, 7, MPI_COMM_WORLD);
        MPI_Send(&send_have, 1, MPI_INT, me + 1, 8, MPI_COMM_WORLD);
    }
    int ok = !(prev_have && have && prev_last > kb2[0]), all_ok = 0;
    MPI_Reduce(&ok, &all_ok, 1, MPI_INT, MPI_SUM, 0, MPI_COMM_WORLD);
    uint64_t css = 0;
    MPI_Reduce(&cs, &css, 1, MPI_UINT64_T, MPI_BXOR, 0, MPI_COMM_WORLD);
    if (argc > 6) {
        char path[512];
        snprintf(path, sizeof(path), "%s.%d.bin", argv[6], me);
        FILE *f = fopen(path, "wb");
        if (f) { fwrite(kb2, sizeof(uint32_t), (size_t)nrecv, f); fclose(f); }
    }
    if (me == 0)
        printf("a2av ranks=%d n=%lld dist=%d payload_bytes=%lld seconds=%.9f sorted_ok=%d checksum=%llu\n",
               P, (long long)n, dist, (long long)total_bytes, tmax, all_ok, (unsigned long long)css);
    free(keys); free(kb1); free(kb2); free(bsize); free(btot); free(bptr);
    free(sc); free(sd); free(rc); free(rd);
    MPI_Finalize();
    return 0;
}
