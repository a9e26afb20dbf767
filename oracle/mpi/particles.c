/*
 * Host-MPI version of the particle-exchange workload (TEST INFRASTRUCTURE / CPU baseline).
 *
 * gx*gy*gz ranks on a periodic grid, n0 particles of 8 FP64 fields per rank.  Each step
 * every particle moves with probability ~0.05 one cell along one of the six axis
 * directions (decision from Philox keyed by particle id and step); each rank packs its
 * migrants by destination rank, exchanges counts with MPI_Alltoall and the particles with
 * MPI_Alltoallv (gpu_mpi/mpi.cu:363-408 semantics: receive block of source s at rdispls[s],
 * after the particles that stayed).  Timing as benchmarks/mpi_communication/
 * mpi_communication.c:24-49 (barrier, timed steps, MPI_Reduce(MAX) to rank 0).
 *
 * usage: particles gx gy gz n0 seed steps [dump_prefix]
 * prints: "particles ranks=P n0=N steps=S moved=M bytes=B seconds=T"
 * dump:   <prefix>.<rank> = int64 count, then count*8 doubles (final state)
 */
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gpumpi_oracle.h"

int main(int argc, char **argv)
{
    MPI_Init(&argc, &argv);
    int me, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &me);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (argc < 7) {
        if (me == 0) fprintf(stderr, "usage: %s gx gy gz n0 seed steps [dump_prefix]\n", argv[0]);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int gx = atoi(argv[1]), gy = atoi(argv[2]), gz = atoi(argv[3]);
    const int64_t n0 = atoll(argv[4]);
    const uint64_t seed = strtoull(argv[5], NULL, 0);
    const int steps = atoi(argv[6]);
    if (gx * gy * gz != P) {
        if (me == 0) fprintf(stderr, "grid %dx%dx%d needs %d ranks, got %d\n", gx, gy, gz, gx * gy * gz, P);
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int F = ORC_PART_FIELDS;
    const int dims[3] = {gx, gy, gz};
    const int64_t cap = 2 * n0 + 1024;
    double *cur = malloc(sizeof(double) * cap * F), *nxt = malloc(sizeof(double) * cap * F);
    double *send = malloc(sizeof(double) * cap * F);
    int *sc = malloc(sizeof(int) * P), *sd = malloc(sizeof(int) * P);
    int *rc = malloc(sizeof(int) * P), *rd = malloc(sizeof(int) * P), *pos = malloc(sizeof(int) * P);
    orc_particles_init_rank(gx, gy, gz, me, n0, seed, cur);
    int64_t n = n0, moved = 0, bytes = 0;
    MPI_Barrier(MPI_COMM_WORLD);
    double t0 = MPI_Wtime();
    for (int s = 0; s < steps; s++) {
        memset(sc, 0, sizeof(int) * P);
        int64_t ns = 0;
        for (int64_t k = 0; k < n; k++) {
            int d = orc_particle_dir(seed, (uint32_t)s, (uint64_t)cur[k * F + 7]);
            if (d < 0) ns++;
            else sc[orc_grid_neighbour(gx, gy, gz, me, d)] += F;
        }
        int acc = 0;
        for (int j = 0; j < P; j++) { sd[j] = acc; pos[j] = acc; acc += sc[j]; }
        int64_t st = 0;
        for (int64_t k = 0; k < n; k++) {
            const double *q = cur + k * F;
            int d = orc_particle_dir(seed, (uint32_t)s, (uint64_t)q[7]);
            if (d < 0) { memcpy(nxt + st * F, q, sizeof(double) * F); st++; continue; }
            int j = orc_grid_neighbour(gx, gy, gz, me, d);
            double *m = send + pos[j];
            memcpy(m, q, sizeof(double) * F);
            int ax = d >> 1;
            double c = m[ax] + ((d & 1) ? -1.0 : 1.0);
            if (c >= (double)dims[ax]) c -= (double)dims[ax];
            if (c < 0.0) c += (double)dims[ax];
            m[ax] = c;
            pos[j] += F;
        }
        moved += acc / F;
        MPI_Alltoall(sc, 1, MPI_INT, rc, 1, MPI_INT, MPI_COMM_WORLD);
        int64_t r_acc = ns * F;
        for (int j = 0; j < P; j++) { rd[j] = (int)r_acc; r_acc += rc[j]; }
        if (r_acc > cap * F) {
            fprintf(stderr, "rank %d: %lld particles exceed capacity\n", me, (long long)(r_acc / F));
            MPI_Abort(MPI_COMM_WORLD, 2);
        }
        MPI_Alltoallv(send, sc, sd, MPI_DOUBLE, nxt, rc, rd, MPI_DOUBLE, MPI_COMM_WORLD);
        bytes += 2 * (int64_t)acc * 8;
        n = r_acc / F;
        double *t = cur; cur = nxt; nxt = t;
    }
    double dt = MPI_Wtime() - t0, tmax = 0;
    int64_t moved_all = 0, bytes_all = 0;
    MPI_Reduce(&dt, &tmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    MPI_Reduce(&moved, &moved_all, 1, MPI_INT64_T, MPI_SUM, 0, MPI_COMM_WORLD);
    MPI_Reduce(&bytes, &bytes_all, 1, MPI_INT64_T, MPI_SUM, 0, MPI_COMM_WORLD);
    if (argc > 7) {
        char path[4096];
        snprintf(path, sizeof path, "%s.%d", argv[7], me);
        FILE *f = fopen(path, "wb");
        if (!f || fwrite(&n, sizeof n, 1, f) != 1 ||
            fwrite(cur, sizeof(double) * F, (size_t)n, f) != (size_t)n) {
            fprintf(stderr, "rank %d: cannot write %s\n", me, path);
            MPI_Abort(MPI_COMM_WORLD, 3);
        }
        fclose(f);
    }
    if (me == 0)
        printf("particles ranks=%d n0=%lld steps=%d moved=%lld bytes=%lld seconds=%.9f\n", P,
               (long long)n0, steps, (long long)moved_all, (long long)bytes_all, tmax);
    free(cur); free(nxt); free(send); free(sc); free(sd); free(rc); free(rd); free(pos);
    MPI_Finalize();
    return 0;
}
