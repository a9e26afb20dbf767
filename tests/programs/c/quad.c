/* Composite Simpson quadrature of 4 / (1 + x^2) over [0, 1], one contiguous block of the
 * interval per rank.
 *
 * Rank 0 picks the number of panels per block (argv[1], rounded up to even) and a
 * tolerance, and broadcasts both as one array of doubles.  Every rank integrates its block
 * [r / P, (r + 1) / P].  MPI_Allreduce sums the blocks on every rank, and MPI_Gather brings
 * the block values to rank 0.  Rank 0 checks two things:
 *   - the gathered values, added in rank order, agree with the reduced total to rounding
 *     (P ulps of the total; a runtime that folds in rank order gives the same bits, and
 *     both values are printed so a test can check that);
 *   - the total is pi within the tolerance.
 * The result is broadcast, and every rank exits with status 0 only if both hold. */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include "mpi.h"

static double integrand(double x) { return 4.0 / (1.0 + x * x); }

/* Simpson's rule on [lo, hi] with m panels (m even): weights 1, 4, 2, 4, ..., 4, 1. */
static double simpson(double lo, double hi, long m) {
    const double h = (hi - lo) / (double)m;
    double s = integrand(lo) + integrand(hi);
    for (long k = 1; k < m; k++)
        s += (k % 2 ? 4.0 : 2.0) * integrand(lo + h * (double)k);
    return s * h / 3.0;
}

int main(int argc, char **argv) {
    const double pi = 3.14159265358979323846;
    int rank, size;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);

    double setup[2] = {0.0, 0.0};   /* panels per block, tolerance */
    if (rank == 0) {
        long m = argc > 1 ? strtol(argv[1], NULL, 10) : 2000;
        if (m < 2) m = 2;
        setup[0] = (double)(m + m % 2);
        setup[1] = 1e-10;
    }
    MPI_Bcast(setup, 2, MPI_DOUBLE, 0, MPI_COMM_WORLD);

    const double block = simpson((double)rank / size, (double)(rank + 1) / size, (long)setup[0]);
    double total = 0.0;
    MPI_Allreduce(&block, &total, 1, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);

    double *blocks = NULL;
    if (rank == 0) blocks = (double *)malloc(sizeof(double) * (size_t)size);
    MPI_Gather(&block, 1, MPI_DOUBLE, blocks, 1, MPI_DOUBLE, 0, MPI_COMM_WORLD);

    int passed = 1;
    if (rank == 0) {
        double sum = blocks[0];
        for (int r = 1; r < size; r++) sum = blocks[r] + sum;
        const double err = fabs(total - pi);
        passed = fabs(sum - total) <= size * 2.3e-16 * total && err < setup[1];
        printf("quad: %d blocks x %ld panels, integral %.15f, |error| %.3e, %s\n", size,
               (long)setup[0], total, err, passed ? "ok" : "FAILED");
        printf("quad: reduced %.17g gathered %.17g\n", total, sum);
        free(blocks);
    }
    MPI_Bcast(&passed, 1, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Finalize();
    return passed ? 0 : 1;
}
