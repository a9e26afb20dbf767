// Number parsing behind the device atoi/strtol/atof/strtod of converted programs.  Plain
// C++ with every function marked GPUMPI_LIBC_FN (device code by default), so that the
// same source also builds for the host and is checked there against the C library.
#pragma once
#include <math.h>
#include <stddef.h>

#ifndef GPUMPI_LIBC_FN
#define GPUMPI_LIBC_FN __device__ inline
#endif

namespace gpumpi_libc {
GPUMPI_LIBC_FN bool is_space(int c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

GPUMPI_LIBC_FN long long parse_int(const char* s, char** end, int base) {
    const char* p = s;
    while (is_space(*p)) p++;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    // "0x" starts a hex number only when a hex digit follows (p[2] exists: p[1] is 'x')
    const bool prefix = p[0] == '0' && (p[1] == 'x' || p[1] == 'X') &&
                        ((p[2] >= '0' && p[2] <= '9') || (p[2] >= 'a' && p[2] <= 'f') ||
                         (p[2] >= 'A' && p[2] <= 'F'));
    if ((base == 0 || base == 16) && prefix) {
        p += 2;
        base = 16;
    } else if (base == 0) {
        base = p[0] == '0' ? 8 : 10;
    }
    unsigned long long v = 0;
    const char* digits = p;
    for (;; p++) {
        int d;
        if (*p >= '0' && *p <= '9') d = *p - '0';
        else if (*p >= 'a' && *p <= 'z') d = *p - 'a' + 10;
        else if (*p >= 'A' && *p <= 'Z') d = *p - 'A' + 10;
        else break;
        if (d >= base) break;
        v = v * (unsigned)base + (unsigned)d;
    }
    if (end) *end = (char*)(p == digits ? s : p);
    return neg ? -(long long)v : (long long)v;
}

// Decimal/exponent parsing to the nearest double for up to 19 significant digits
// (exact for the integers and short decimals command lines carry).
GPUMPI_LIBC_FN double parse_double(const char* s, char** end) {
    const char* p = s;
    while (is_space(*p)) p++;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    unsigned long long mant = 0;
    int digits = 0, exp10 = 0;
    const char* start = p;
    for (; *p >= '0' && *p <= '9'; p++) {
        if (digits < 19) { mant = mant * 10 + (unsigned)(*p - '0'); digits += mant != 0; }
        else exp10++;
    }
    if (*p == '.') {
        p++;
        for (; *p >= '0' && *p <= '9'; p++)
            if (digits < 19) { mant = mant * 10 + (unsigned)(*p - '0'); digits += mant != 0; exp10--; }
    }
    if (p == start || (p == start + 1 && *start == '.')) {
        if (end) *end = (char*)s;
        return 0.0;
    }
    if (*p == 'e' || *p == 'E') {
        char* q;
        const long long e = parse_int(p + 1, &q, 10);
        if (q != p + 1) { exp10 += (int)e; p = q; }
    }
    double v = (double)mant;
    if (exp10 > 0) v *= pow(10.0, (double)exp10);
    else if (exp10 < 0) v /= pow(10.0, (double)-exp10);
    if (end) *end = (char*)p;
    return neg ? -v : v;
}

}  // namespace gpumpi_libc
