/*
 * Host-MPI check of the oracle's reduction restatement (TEST INFRASTRUCTURE ONLY).
 *
 * Every rank fills n elements from the workload generators (orc_fill_int: uniform
 * integers in [-2^20, 2^20]; orc_fill_normal: the FP64 inputs) and calls MPICH's
 * MPI_Reduce and MPI_Allreduce with MPI_SUM and MPI_MAX.  Rank 0 gathers the inputs and
 * compares MPICH's results with orc_reduce (the reference's rank-order fold,
 * gpu_mpi/mpi.cu:209-225): integers and FP64 MAX must be equal; FP64 SUM, folded by
 * MPICH in its own order, within 1e-12 * sum |x_r| per element.
 *
 * usage: reduce_check seed n
 * prints: "reduce_check ranks=P n=N int_mismatch=A dmax_mismatch=B dsum_worst=C allreduce_mismatch=D"
 *         (dsum_worst: the largest |MPICH - oracle| / (1e-12 * sum |x|); <= 1 passes)
 */
#include <math.h>
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gpumpi_oracle.h"

enum { ORC_MPI_INT = 2, ORC_MPI_DOUBLE = 13, ORC_OP_MAX = 0, ORC_OP_SUM = 2 };

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int me, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &me);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (argc < 3) {
        if (me == 0) fprintf(stderr, "usage: %s seed n\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const uint64_t seed = strtoull(argv[1], NULL, 0);
    const int n = atoi(argv[2]);
    int32_t *xi = malloc(sizeof(int32_t) * n), *ri = malloc(sizeof(int32_t) * n), *mi = malloc(sizeof(int32_t) * n);
    double *xd = malloc(sizeof(double) * n), *rd = malloc(sizeof(double) * n), *md = malloc(sizeof(double) * n);
    double *ad = malloc(sizeof(double) * n);
    orc_fill_int(seed, me, 0, xi, n);
    orc_fill_normal(seed, me, 0, xd, n);
    MPI_Reduce(xi, ri, n, MPI_INT, MPI_SUM, 0, MPI_COMM_WORLD);
    MPI_Reduce(xi, mi, n, MPI_INT, MPI_MAX, 0, MPI_COMM_WORLD);
    MPI_Reduce(xd, rd, n, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    MPI_Reduce(xd, md, n, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    MPI_Allreduce(xd, ad, n, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    int32_t *gi = NULL;
    double *gd = NULL, *ga = NULL;
    if (me == 0) {
        gi = malloc(sizeof(int32_t) * (size_t)n * P);
        gd = malloc(sizeof(double) * (size_t)n * P);
        ga = malloc(sizeof(double) * (size_t)n * P);
    }
    MPI_Gather(xi, n, MPI_INT, gi, n, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Gather(xd, n, MPI_DOUBLE, gd, n, MPI_DOUBLE, 0, MPI_COMM_WORLD);
    MPI_Gather(ad, n, MPI_DOUBLE, ga, n, MPI_DOUBLE, 0, MPI_COMM_WORLD);
    if (me == 0) {
        const void **ins_i = malloc(sizeof(void *) * P), **ins_d = malloc(sizeof(void *) * P);
        for (int r = 0; r < P; r++) {
            ins_i[r] = gi + (size_t)r * n;
            ins_d[r] = gd + (size_t)r * n;
        }
        int32_t *oi = malloc(sizeof(int32_t) * n), *omi = malloc(sizeof(int32_t) * n);
        double *od = malloc(sizeof(double) * n), *omd = malloc(sizeof(double) * n);
        orc_reduce(P, ORC_MPI_INT, ORC_OP_SUM, (const void *const *)ins_i, n, oi);
        orc_reduce(P, ORC_MPI_INT, ORC_OP_MAX, (const void *const *)ins_i, n, omi);
        orc_reduce(P, ORC_MPI_DOUBLE, ORC_OP_SUM, (const void *const *)ins_d, n, od);
        orc_reduce(P, ORC_MPI_DOUBLE, ORC_OP_MAX, (const void *const *)ins_d, n, omd);
        long long imis = 0, dmax = 0, amis = 0;
        double worst = 0.0;
        for (int i = 0; i < n; i++) {
            imis += (oi[i] != ri[i]) + (omi[i] != mi[i]);
            dmax += memcmp(&omd[i], &md[i], sizeof(double)) != 0;
            double norm = 0.0;
            for (int r = 0; r < P; r++) norm += fabs(gd[(size_t)r * n + i]);
            const double e = fabs(rd[i] - od[i]) / (1e-12 * (norm > 0 ? norm : 1.0));
            if (e > worst) worst = e;
            for (int r = 0; r < P; r++) amis += memcmp(&ga[(size_t)r * n + i], &rd[i], sizeof(double)) != 0 &&
                                               fabs(ga[(size_t)r * n + i] - od[i]) > 1e-12 * norm;
        }
        printf("reduce_check ranks=%d n=%d int_mismatch=%lld dmax_mismatch=%lld dsum_worst=%.6f "
               "allreduce_mismatch=%lld\n", P, n, imis, dmax, worst, amis);
    }
    MPI_Finalize();
    return 0;
}
