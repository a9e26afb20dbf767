/*
 * Host-MPI check of the oracle's Alltoallv restatement (TEST INFRASTRUCTURE ONLY).
 *
 * Every rank draws uneven per-pair counts (0..maxc elements, a fifth of the pairs empty),
 * gapped send displacements and Philox payloads of elem-byte elements, and runs MPICH's
 * MPI_Alltoallv.  Rank 0 then gathers every rank's send buffer, counts and displacements,
 * runs orc_alltoallv (the restatement of gpu_mpi/mpi.cu:363-408) on them and compares
 * every rank's receive buffer with MPICH's, byte for byte.
 *
 * usage: a2av_check seed maxc elem        prints: "a2av_check ranks=P bytes=B mismatches=M"
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gpumpi_oracle.h"

static uint32_t rnd(uint64_t seed, int rank, uint64_t i, uint32_t w)
{
    uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 0x0A2AC4EC, w};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32) ^ (uint32_t)rank};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return o[0];
}

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int me, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &me);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (argc < 4) {
        if (me == 0) fprintf(stderr, "usage: %s seed maxc elem\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const uint64_t seed = strtoull(argv[1], NULL, 0);
    const int maxc = atoi(argv[2]);
    const int elem = atoi(argv[3]);
    MPI_Datatype t = elem == 1 ? MPI_BYTE : elem == 2 ? MPI_SHORT : elem == 4 ? MPI_INT : MPI_DOUBLE;
    int *sc = malloc(sizeof(int) * P), *sd = malloc(sizeof(int) * P);
    int off = (int)(rnd(seed, me, 0, 1) % 5);
    for (int d = 0; d < P; d++) {
        uint32_t r = rnd(seed, me, (uint64_t)d, 2);
        sc[d] = (r % 5 == 0) ? 0 : (int)(r % (uint32_t)(maxc + 1));
        sd[d] = off;
        off += sc[d] + (int)(rnd(seed, me, (uint64_t)d, 3) % 3);
    }
    const int sext = off;
    uint8_t *sb = malloc((size_t)(sext > 0 ? sext : 1) * elem);
    orc_fill_bytes(seed, me, 7, sb, (int64_t)sext * elem);
    int *rc = malloc(sizeof(int) * P), *rd = malloc(sizeof(int) * P);
    MPI_Alltoall(sc, 1, MPI_INT, rc, 1, MPI_INT, MPI_COMM_WORLD);
    int rext = 0;
    for (int s = 0; s < P; s++) { rd[s] = rext; rext += rc[s]; }
    uint8_t *rb = malloc((size_t)(rext > 0 ? rext : 1) * elem);
    MPI_Alltoallv(sb, sc, sd, t, rb, rc, rd, t, MPI_COMM_WORLD);

    /* gather everything to rank 0 */
    int *all_sc = NULL, *all_sd = NULL, *sext_all = NULL, *rext_all = NULL;
    if (me == 0) {
        all_sc = malloc(sizeof(int) * P * P);
        all_sd = malloc(sizeof(int) * P * P);
        sext_all = malloc(sizeof(int) * P);
        rext_all = malloc(sizeof(int) * P);
    }
    MPI_Gather(sc, P, MPI_INT, all_sc, P, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Gather(sd, P, MPI_INT, all_sd, P, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Gather(&sext, 1, MPI_INT, sext_all, 1, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Gather(&rext, 1, MPI_INT, rext_all, 1, MPI_INT, 0, MPI_COMM_WORLD);
    long long mismatches = 0, bytes = 0;
    if (me == 0) {
        void **sbufs = malloc(sizeof(void *) * P), **rbufs = malloc(sizeof(void *) * P);
        void **mpich = malloc(sizeof(void *) * P);
        int *all_rc = malloc(sizeof(int) * P * P), *all_rd = malloc(sizeof(int) * P * P);
        for (int r = 0; r < P; r++) {
            sbufs[r] = malloc((size_t)(sext_all[r] > 0 ? sext_all[r] : 1) * elem);
            rbufs[r] = calloc((size_t)(rext_all[r] > 0 ? rext_all[r] : 1), (size_t)elem);
            mpich[r] = malloc((size_t)(rext_all[r] > 0 ? rext_all[r] : 1) * elem);
            if (r == 0) {
                memcpy(sbufs[0], sb, (size_t)sext * elem);
                memcpy(mpich[0], rb, (size_t)rext * elem);
            } else {
                MPI_Recv(sbufs[r], sext_all[r] * elem, MPI_BYTE, r, 1, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
                MPI_Recv(mpich[r], rext_all[r] * elem, MPI_BYTE, r, 2, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
            }
        }
        for (int d = 0; d < P; d++) {
            int acc = 0;
            for (int s = 0; s < P; s++) {
                all_rc[d * P + s] = all_sc[s * P + d];
                all_rd[d * P + s] = acc;
                acc += all_sc[s * P + d];
            }
        }
        orc_alltoallv(P, elem, elem, (const void *const *)sbufs, all_sc, all_sd, rbufs, all_rc, all_rd);
        for (int r = 0; r < P; r++) {
            bytes += (long long)rext_all[r] * elem;
            const uint8_t *a = rbufs[r], *b = mpich[r];
            for (long long i = 0; i < (long long)rext_all[r] * elem; i++) mismatches += a[i] != b[i];
        }
        printf("a2av_check ranks=%d bytes=%lld mismatches=%lld\n", P, bytes, mismatches);
    } else {
        MPI_Send(sb, sext * elem, MPI_BYTE, 0, 1, MPI_COMM_WORLD);
        MPI_Send(rb, rext * elem, MPI_BYTE, 0, 2, MPI_COMM_WORLD);
    }
    MPI_Finalize();
    return 0;
}
