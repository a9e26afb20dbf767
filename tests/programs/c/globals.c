/* Per-rank global state of a converted program: every rank owns a copy of each global
 * and static variable, initialized from the C initializer, including through
 * redeclarations, macros that name a global, multi-dimensional and pointer arrays and
 * partially initialized static locals.  Also exercises the C-to-C++ fixes (implicit
 * void* conversions, an identifier named `class`, an old-style definition) and the
 * device C library (string functions, sprintf, atoi, qsort, rand).  Exit status 0 when
 * every check holds on every rank. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mpi.h"

extern int counter;
extern int counter;
int counter = -7;
double weights[3] = {0.5, 1.5, 2.5};
int grid[2][3] = {{1, 2, 3}, {4, 5, 6}};
int *slots[2] = {NULL, NULL};
const int kTable[4] = {10, 20, 30, 40};
int class = 3;
#define ALIAS_WEIGHTS weights

static int bad;
#define CHECK(c) do { if (!(c)) { printf("rank check failed: %s (line %d)\n", #c, __LINE__); bad = 1; } } while (0)

static int bump(void) {
    static int calls[4] = {100};
    calls[0] += 1;
    return calls[0];
}

int oldstyle(a, b) int a; double b; { return a + (int)b; }

static int cmp_int(const void *x, const void *y) {
    int a = *(const int *)x, b = *(const int *)y;
    return a < b ? -1 : a > b;
}

int main(int argc, char **argv) {
    int rank, size, i;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);

    CHECK(counter == -7);
    CHECK(weights[2] == 2.5 && ALIAS_WEIGHTS[0] == 0.5);
    CHECK(grid[1][2] == 6 && slots[1] == NULL && kTable[3] == 40 && class == 3);
    counter = rank;
    ALIAS_WEIGHTS[1] = rank * 2.0;
    grid[0][0] = rank + 100;
    class = rank;
    CHECK(bump() == 101 && bump() == 102);

    int *buf = malloc(8 * sizeof(int));
    slots[0] = buf;
    for (i = 0; i < 8; i++) buf[i] = (i * 37 + rank) % 11;
    qsort(buf, 8, sizeof(int), cmp_int);
    for (i = 1; i < 8; i++) CHECK(buf[i - 1] <= buf[i]);

    char text[64];
    sprintf(text, "r%d-%s-%.2f", rank, "x", 1.25);
    char expect[64];
    strcpy(expect, "r");
    CHECK(strncmp(text, expect, 1) == 0 && strlen(text) > 5 && strstr(text, "-x-1.25") != NULL);
    CHECK(atoi(text + 1) == rank);
    CHECK(oldstyle(2, 3.5) == 5);
    srand(rank + 1);
    int r1 = rand(), r2 = rand();
    CHECK(r1 >= 0 && r2 >= 0 && r1 != r2);

    MPI_Barrier(MPI_COMM_WORLD);
    /* other ranks' writes did not touch this rank's copies */
    CHECK(counter == rank && weights[1] == rank * 2.0 && grid[0][0] == rank + 100 && class == rank);
    CHECK(weights[0] == 0.5 && grid[1][0] == 4 && slots[0] == buf);
    int total = 0;
    MPI_Allreduce(&counter, &total, 1, MPI_INT, MPI_SUM, MPI_COMM_WORLD);
    CHECK(total == size * (size - 1) / 2);
    free(buf);
    int anybad = 0;
    MPI_Allreduce(&bad, &anybad, 1, MPI_INT, MPI_MAX, MPI_COMM_WORLD);
    if (rank == 0) printf("globals: %s on %d ranks\n", anybad ? "FAILED" : "ok", size);
    MPI_Finalize();
    return anybad;
}
