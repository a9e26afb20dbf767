// Device stdio for converted programs (the subset of gpu_libs/gpu_libc/stdio.cuh:1-31 that
// needs no host delegation).  A rank is a wavefront running the program wave-uniformly,
// so output is printed by lane 0 only; stdout and stderr both go to the device printf
// stream, and streams are not evaluated (the host FILE objects do not exist on the GPU).
// Files need device-to-host delegation, which this subset does not provide: fopen finds
// no file (returns NULL) and reads report end-of-file, so a program that probes for an
// optional input file (NPB's timer.flag) runs as it does when the file is absent.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdarg.h>
#include <stdint.h>
#include <math.h>
#include "../mpi.h"
#include "format.h"

#define GPUMPI_LIBC_STDIO 1
#define printf(...) (__lane_id() == 0 ? printf(__VA_ARGS__) : 0)
#define fprintf(stream, ...) printf(__VA_ARGS__)
#define puts(s) printf("%s\n", (s))
#define putchar(c) printf("%c", (c))
#define fputs(s, stream) printf("%s", (s))
#define fflush(stream) 0
#define sprintf(buf, ...) ::gpumpi_libc::format((buf), (size_t)-1, __VA_ARGS__)
#define snprintf(buf, cap, ...) ::gpumpi_libc::format((buf), (cap), __VA_ARGS__)
#define vsprintf(buf, fmt, ap) ::gpumpi_libc::vformat((buf), (size_t)-1, (fmt), (ap))
#define vsnprintf(buf, cap, fmt, ap) ::gpumpi_libc::vformat((buf), (cap), (fmt), (ap))
#define fopen(path, mode) ((FILE*)nullptr)
#define fclose(stream) 0
#define fgets(s, n, stream) ((char*)nullptr)
#define fgetc(stream) (-1)
#define getc(stream) (-1)
#define feof(stream) 1
#define fscanf(stream, ...) (-1)
