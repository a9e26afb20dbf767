/* malloc and MPI_Alltoall in a call tree: with --malloc-version the converter threads a
 * `__coalesced` flag through the functions that allocate or exchange outside branches and
 * loops, `true` from main.  The exchange result is checked on every rank. */
#include <stdlib.h>
#include "mpi.h"

static int *make_block(int n) {
    int *p = malloc(n * sizeof(int));
    return p;
}

static int exchange(int *send, int n, int rank, int size) {
    int *recv = malloc(n * size * sizeof(int));
    MPI_Alltoall(send, n, MPI_INT, recv, n, MPI_INT, MPI_COMM_WORLD);
    int ok = 1, s, k;
    for (s = 0; s < size; s++)
        for (k = 0; k < n; k++)
            if (recv[s * n + k] != s * 1000 + rank * 10 + k) ok = 0;
    free(recv);
    return ok;
}

int main(int argc, char **argv) {
    int rank, size, d, k, n = 4;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    int *send = make_block(n * size);
    for (d = 0; d < size; d++)
        for (k = 0; k < n; k++) send[d * n + k] = rank * 1000 + d * 10 + k;
    int ok = exchange(send, n, rank, size);
    if (argc > 5) {
        ok = exchange(send, n, rank, size) && ok;
    }
    free(send);
    int all = 0;
    MPI_Allreduce(&ok, &all, 1, MPI_INT, MPI_MIN, MPI_COMM_WORLD);
    return all ? 0 : 1;
}
