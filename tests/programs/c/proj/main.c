/* Two-file program: every rank passes a token around the ring `size` times; the token
 * comes back to its owner, and each rank counted `size` hops in the global of util.c. */
#include <stdio.h>
#include "mpi.h"
#include "util.h"

int main(int argc, char **argv) {
    int rank, size, i;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    int token = rank * 3 + 1;
    for (i = 0; i < size; i++) token = ring_pass(token, rank, size);
    int ok = token == rank * 3 + 1 && hops == size;
    int all = 0;
    MPI_Allreduce(&ok, &all, 1, MPI_INT, MPI_MIN, MPI_COMM_WORLD);
    if (rank == 0) printf("ring: %s after %d hops\n", all ? "ok" : "FAILED", hops);
    MPI_Finalize();
    return all ? 0 : 1;
}
