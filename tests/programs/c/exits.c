/* exit() ends the calling rank; a non-zero status makes the program's status 1. */
#include <stdio.h>
#include <stdlib.h>
#include "mpi.h"

int main(int argc, char **argv) {
    int rank;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Barrier(MPI_COMM_WORLD);
    if (rank == 0) printf("before exit\n");
    exit(rank == 1 ? 3 : 0);
    printf("not reached\n");
    return 0;
}
