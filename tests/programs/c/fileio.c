/* File I/O from ranks through the host channel: every rank writes its own file with
 * fprintf, fputs, fputc and fwrite, reads it back with fscanf, fgetc, fgets, fread and
 * feof; opening a missing file fails; rank 0 reads one number from stdin when asked.
 * Exit status 0 when every rank's checks hold. */
#include <stdio.h>
#include <string.h>
#include <math.h>
#include "mpi.h"

int main(int argc, char **argv) {
    int rank, size, i, bad = 0;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    char name[256];
    sprintf(name, "%s/rank%d.txt", argc > 1 ? argv[1] : ".", rank);
    FILE *f = fopen(name, "w");
    int data[1500], back[1500];
    for (i = 0; i < 1500; i++) data[i] = rank * 100000 + i * 7;
    if (!f) {
        printf("rank %d: cannot open %s for writing\n", rank, name);
        bad = 1;
    } else {
        fprintf(f, "rank %d of %d pi=%.6f\n", rank, size, 3.14159265);
        fputs("line two", f);
        fputc('\n', f);
        if (fwrite(data, sizeof(int), 1500, f) != 1500) bad = 1;  /* 6000 bytes: two requests */
        fclose(f);
        f = fopen(name, "r");
        int r = -1, s = -1;
        double pi = 0;
        char line[128];
        if (!f || fscanf(f, "rank %d of %d pi=%lf", &r, &s, &pi) != 3) bad = 1;
        if (r != rank || s != size || fabs(pi - 3.141593) > 1e-9) bad = 1;
        if (fgetc(f) != '\n') bad = 1;
        if (!fgets(line, sizeof line, f) || strcmp(line, "line two\n") != 0) bad = 1;
        if (fread(back, sizeof(int), 1500, f) != 1500 || memcmp(back, data, sizeof data) != 0) bad = 1;
        if (fgetc(f) != EOF || !feof(f)) bad = 1;
        fclose(f);
    }
    if (fopen("/nonexistent-dir/none.txt", "r") != NULL) bad = 1;
    if (argc > 2 && rank == 0) {
        int v = 0;
        if (scanf("%d", &v) != 1 || v != 4242) bad = 1;
    }
    if (bad) printf("rank %d: file checks failed\n", rank);
    int anybad = 0;
    MPI_Allreduce(&bad, &anybad, 1, MPI_INT, MPI_MAX, MPI_COMM_WORLD);
    if (rank == 0) printf("fileio: %s on %d ranks\n", anybad ? "FAILED" : "ok", size);
    MPI_Finalize();
    return anybad;
}
