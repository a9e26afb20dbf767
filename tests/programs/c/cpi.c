/* Midpoint-rule pi over MPI ranks, in plain C: the shape of mpich's cpi example. */
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include "mpi.h"

static double f(double a) { return 4.0 / (1.0 + a * a); }

int main(int argc, char *argv[]) {
    int n = 0, myid, numprocs, i;
    double PI25DT = 3.141592653589793238462643;
    double mypi, pi = 0.0, h, sum, x;

    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &numprocs);
    MPI_Comm_rank(MPI_COMM_WORLD, &myid);
    if (myid == 0)
        n = argc > 1 ? atoi(argv[1]) : 100000;
    MPI_Bcast(&n, 1, MPI_INT, 0, MPI_COMM_WORLD);
    h = 1.0 / (double)n;
    sum = 0.0;
    for (i = myid + 1; i <= n; i += numprocs) {
        x = h * ((double)i - 0.5);
        sum += f(x);
    }
    mypi = h * sum;
    MPI_Reduce(&mypi, &pi, 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    int ok = 1;
    if (myid == 0) {
        printf("pi is approximately %.16f, error is %.16f\n", pi, fabs(pi - PI25DT));
        ok = fabs(pi - PI25DT) < 1e-8;
    }
    MPI_Bcast(&ok, 1, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Finalize();
    return ok ? 0 : 1;
}
