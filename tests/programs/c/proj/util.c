#include "util.h"
#include "mpi.h"

int hops = 0;

int ring_pass(int value, int rank, int size) {
    int got = -1;
    MPI_Status st;
    int next = (rank + 1) % size, prev = (rank + size - 1) % size;
    MPI_Sendrecv(&value, 1, MPI_INT, next, RING_TAG, &got, 1, MPI_INT, prev, RING_TAG,
                 MPI_COMM_WORLD, &st);
    hops++;
    return got;
}
