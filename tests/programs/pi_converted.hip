// A converted MPI program in the shape the reference's converter emits
// (source_converter/converter.cpp:64-177): main renamed __gpu_main, every function
// __device__, the global `int visits;` turned into a per-rank array used through
// __gpu_global, and __gpu_max_threads() giving the array size.  The computation is the
// midpoint-rule pi of mpich's cpi example (the reference's toolchain_tests/pi).
#include <gpumpi/global_vars.h>
#include <gpumpi/mpi.h>
#include <gpumpi/gpu_malloc.h>

__device__ int __gpu_max_threads() { return 512; }
__device__ int visits[512];

__device__ double f(double a) { return 4.0 / (1.0 + a * a); }

__device__ int __gpu_main(int argc, char** argv) {
    int n = 0, myid, numprocs;
    double mypi, pi = 0.0, h, sum, x;
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &numprocs);
    MPI_Comm_rank(MPI_COMM_WORLD, &myid);
    __gpu_global(visits) = 0;
    if (myid == 0) n = argc > 1 ? 1000 * (argv[1][0] - '0') : 100000;
    MPI_Bcast(&n, 1, MPI_INT, 0, MPI_COMM_WORLD);
    h = 1.0 / (double)n;
    sum = 0.0;
    for (int i = myid + 1; i <= n; i += numprocs) {
        x = h * ((double)i - 0.5);
        sum += f(x);
        __gpu_global(visits)++;
    }
    mypi = h * sum;
    MPI_Reduce(&mypi, &pi, 1, MPI_DOUBLE, MPI_SUM, 0, MPI_COMM_WORLD);
    int total = 0;
    MPI_Allreduce(&__gpu_global(visits), &total, 1, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
    int ok = total == n;
    if (myid == 0) {
        ok = ok && fabs(pi - 3.141592653589793) < 1e-8;
        gpumpi_printf("pi is approximately %.16f (%d intervals, %d ranks)\n", pi, n, numprocs);
    }
    MPI_Bcast(&ok, 1, MPI_INT, 0, MPI_COMM_WORLD);
    MPI_Finalize();
    return ok ? 0 : 1;
}

#include <gpumpi/gpumpi_main.h>
