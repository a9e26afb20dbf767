// Device stdio for converted programs (gpu_libs/gpu_libc/stdio.cuh:1-31).  A rank is a
// wavefront running the program wave-uniformly, so console output is printed by lane 0
// only, through the device printf stream.  Files go through the device-to-host I/O
// channel (gpumpi_io: one pinned request slot per rank, served by the launcher's host
// thread): fopen/fclose/fread/fwrite/fgets/fgetc/ungetc/fputs/fprintf/fscanf/feof/
// fflush/fseek/ftell/rewind on host files, and scanf/getchar on the host's stdin.  A
// FILE* on the device is a handle, never dereferenced; stdout and stderr are fixed ones.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdarg.h>
#include <stdint.h>
#include <math.h>
#include "../mpi.h"
#include "../gpu_malloc.h"
#include "../io_ops.h"
#include "format.h"

namespace gpumpi_libc {
__device__ inline int64_t handle(FILE* f) { return (int64_t)(intptr_t)f; }
__device__ inline bool console(FILE* f) {
    return handle(f) == ::gpumpi::kIoStdout || handle(f) == ::gpumpi::kIoStderr;
}
__device__ inline int c_len(const char* s) {
    int n = 0;
    while (s[n]) n++;
    return n;
}
__device__ inline FILE* f_open(const char* path, const char* mode) {
    const int64_t h = gpumpi_io(::gpumpi::kIoOpen, 0, 0, 0, path, c_len(path) + 1, mode,
                                c_len(mode) + 1, nullptr, 0);
    return h > 0 ? (FILE*)(intptr_t)h : nullptr;
}
__device__ inline int f_close(FILE* f) {
    if (!f) return -1;
    if (handle(f) < 0) return 0;
    return (int)gpumpi_io(::gpumpi::kIoClose, handle(f), 0, 0, nullptr, 0, nullptr, 0, nullptr, 0);
}
__device__ inline size_t f_write(const void* ptr, size_t size, size_t n, FILE* f) {
    if (!f || size == 0) return 0;
    const char* p = (const char*)ptr;
    const size_t total = size * n;
    size_t done = 0;
    while (done < total) {
        const int len = (int)(total - done < (size_t)::gpumpi::kIoData ? total - done : (size_t)::gpumpi::kIoData);
        const int64_t w = gpumpi_io(::gpumpi::kIoWrite, handle(f), 0, 0, p + done, len, nullptr, 0, nullptr, 0);
        if (w <= 0) break;
        done += (size_t)w;
        if (w < len) break;
    }
    return done / size;
}
__device__ inline size_t f_read(void* ptr, size_t size, size_t n, FILE* f) {
    if (!f || size == 0) return 0;
    char* p = (char*)ptr;
    const size_t total = size * n;
    size_t done = 0;
    while (done < total) {
        const int want = (int)(total - done < (size_t)::gpumpi::kIoData ? total - done : (size_t)::gpumpi::kIoData);
        const int64_t r = gpumpi_io(::gpumpi::kIoRead, handle(f), want, 0, nullptr, 0, nullptr, 0, p + done, want);
        if (r <= 0) break;
        done += (size_t)r;
        if (r < want) break;
    }
    return done / size;
}
__device__ inline char* f_gets(char* s, int n, FILE* f) {
    if (!f || n <= 0) return nullptr;
    const int64_t r = gpumpi_io(::gpumpi::kIoGets, handle(f), n, 0, nullptr, 0, nullptr, 0, s, n);
    return r > 0 ? s : nullptr;
}
__device__ inline int f_getc(FILE* f) {
    return f ? (int)gpumpi_io(::gpumpi::kIoGetc, handle(f), 0, 0, nullptr, 0, nullptr, 0, nullptr, 0) : -1;
}
__device__ inline int f_ungetc(int c, FILE* f) {
    return f ? (int)gpumpi_io(::gpumpi::kIoUngetc, handle(f), c, 0, nullptr, 0, nullptr, 0, nullptr, 0) : -1;
}
__device__ inline int f_simple(int op, FILE* f, int64_t a1 = 0, int64_t a2 = 0) {
    return f ? (int)gpumpi_io(op, handle(f), a1, a2, nullptr, 0, nullptr, 0, nullptr, 0) : -1;
}
__device__ inline int f_puts(const char* s, FILE* f) {
    if (console(f)) return __lane_id() == 0 ? printf("%s", s) : 0;
    return f_write(s, 1, (size_t)c_len(s), f) == (size_t)c_len(s) ? 1 : -1;
}
__device__ inline int f_putc(int c, FILE* f) {
    const char ch = (char)c;
    if (console(f)) return __lane_id() == 0 ? printf("%c", ch) : ch;
    return f_write(&ch, 1, 1, f) == 1 ? (unsigned char)ch : -1;
}
// fprintf to a file: formatted into a heap buffer, then written.
__device__ inline int f_vprintf(FILE* f, const char* fmt, va_list ap) {
    va_list ap2;
    va_copy(ap2, ap);
    const int n = vformat(nullptr, 0, fmt, ap2);
    va_end(ap2);
    char* buf = (char*)__gpu_malloc((size_t)n + 1);
    if (!buf) return -1;
    vformat(buf, (size_t)n + 1, fmt, ap);
    const size_t w = f_write(buf, 1, (size_t)n, f);
    __gpu_free(buf);
    return w == (size_t)n ? n : -1;
}
__device__ inline int f_printf(FILE* f, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const int n = f_vprintf(f, fmt, ap);
    va_end(ap);
    return n;
}
struct FileSource {
    FILE* f;
    __device__ int get() { return f_getc(f); }
    __device__ void unget(int c) { f_ungetc(c, f); }
};
__device__ inline int f_scanf(FILE* f, const char* fmt, ...) {
    FileSource src{f};
    va_list ap;
    va_start(ap, fmt);
    const int n = vscan(src, fmt, ap);
    va_end(ap);
    return n;
}
}  // namespace gpumpi_libc

#define GPUMPI_LIBC_STDIO 1
#undef stdin
#undef stdout
#undef stderr
#define stdin ((FILE*)(intptr_t)::gpumpi::kIoStdin)
#define stdout ((FILE*)(intptr_t)::gpumpi::kIoStdout)
#define stderr ((FILE*)(intptr_t)::gpumpi::kIoStderr)
#define printf(...) (__lane_id() == 0 ? printf(__VA_ARGS__) : 0)
#define fprintf(stream, ...) \
    (::gpumpi_libc::console(stream) ? printf(__VA_ARGS__) : ::gpumpi_libc::f_printf((stream), __VA_ARGS__))
#define puts(s) printf("%s\n", (s))
#define putchar(c) printf("%c", (c))
#define fputs(s, stream) ::gpumpi_libc::f_puts((s), (stream))
#define fputc(c, stream) ::gpumpi_libc::f_putc((c), (stream))
#define putc(c, stream) ::gpumpi_libc::f_putc((c), (stream))
#define fflush(stream) ((stream) ? ::gpumpi_libc::f_simple(::gpumpi::kIoFlush, (stream)) : 0)
#define sprintf(buf, ...) ::gpumpi_libc::format((buf), (size_t)-1, __VA_ARGS__)
#define snprintf(buf, cap, ...) ::gpumpi_libc::format((buf), (cap), __VA_ARGS__)
#define vsprintf(buf, fmt, ap) ::gpumpi_libc::vformat((buf), (size_t)-1, (fmt), (ap))
#define vsnprintf(buf, cap, fmt, ap) ::gpumpi_libc::vformat((buf), (cap), (fmt), (ap))
#define vfprintf(stream, fmt, ap) ::gpumpi_libc::f_vprintf((stream), (fmt), (ap))
#define fopen(path, mode) ::gpumpi_libc::f_open((path), (mode))
#define fclose(stream) ::gpumpi_libc::f_close(stream)
#define fwrite(ptr, size, n, stream) ::gpumpi_libc::f_write((ptr), (size), (n), (stream))
#define fread(ptr, size, n, stream) ::gpumpi_libc::f_read((ptr), (size), (n), (stream))
#define fgets(s, n, stream) ::gpumpi_libc::f_gets((s), (n), (stream))
#define fgetc(stream) ::gpumpi_libc::f_getc(stream)
#define getc(stream) ::gpumpi_libc::f_getc(stream)
#define getchar() ::gpumpi_libc::f_getc(stdin)
#define ungetc(c, stream) ::gpumpi_libc::f_ungetc((c), (stream))
#define feof(stream) ::gpumpi_libc::f_simple(::gpumpi::kIoEof, (stream))
#define fseek(stream, off, whence) ::gpumpi_libc::f_simple(::gpumpi::kIoSeek, (stream), (off), (whence))
#define ftell(stream) ((long)::gpumpi_libc::f_simple(::gpumpi::kIoTell, (stream)))
#define rewind(stream) ((void)::gpumpi_libc::f_simple(::gpumpi::kIoSeek, (stream), 0, 0))
#define fscanf(stream, ...) ::gpumpi_libc::f_scanf((stream), __VA_ARGS__)
#define scanf(...) ::gpumpi_libc::f_scanf(stdin, __VA_ARGS__)
#define sscanf(str, ...) ::gpumpi_libc::scan_string((str), __VA_ARGS__)
