// The formatter behind the device sprintf/snprintf of converted programs.  Plain C++
// with every function marked GPUMPI_LIBC_FN (device code by default), so that the same
// source also builds for the host and is checked there against the C library.
#pragma once
#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>
#include "parse.h"

#ifndef GPUMPI_LIBC_FN
#define GPUMPI_LIBC_FN __device__ inline
#endif

namespace gpumpi_libc {
// snprintf's conversions: d i u o x X c s p f F e E g G %, flags - + space 0 #, width and
// precision (also '*'), length hh h l ll z j t.  Returns the length the full output
// would have; writes at most cap - 1 characters and a terminating zero.
struct Sink {
    char* out;
    size_t cap, n;
    GPUMPI_LIBC_FN void put(char c) {
        if (n + 1 < cap) out[n] = c;
        n++;
    }
    GPUMPI_LIBC_FN void pad(int count, char c) {
        for (int i = 0; i < count; i++) put(c);
    }
};

GPUMPI_LIBC_FN int utoa(unsigned long long v, unsigned base, bool upper, char* buf) {
    const char* dig = upper ? "0123456789ABCDEF" : "0123456789abcdef";
    int n = 0;
    do {
        buf[n++] = dig[v % base];
        v /= base;
    } while (v);
    return n;  // reversed
}

GPUMPI_LIBC_FN void emit_number(Sink& s, const char* sign, const char* digits_rev, int nd,
                                   int prec, int width, bool left, bool zero) {
    const int sl = sign[0] ? (sign[1] ? 2 : 1) : 0;
    const int body = (prec > nd ? prec : nd) + sl;
    if (!left && !zero) s.pad(width - body, ' ');
    for (int i = 0; i < sl; i++) s.put(sign[i]);
    if (!left && zero) s.pad(width - body, '0');
    s.pad(prec - nd, '0');
    for (int i = nd - 1; i >= 0; i--) s.put(digits_rev[i]);
    if (left) s.pad(width - body, ' ');
}

// Rounds x * p (p a power of ten, exact) to an integer as the exact product would round:
// the residual of the floating-point product (fma) settles products that land on .5,
// and exact ties go to the even integer, as the C library does.
GPUMPI_LIBC_FN double round_scaled(double x, double p) {
    const double t = x * p;
    const double err = fma(x, p, -t);
    const double f = floor(t);
    const double d = t - f;
    if (d > 0.5) return f + 1;
    if (d < 0.5) return f;
    if (err > 0) return f + 1;
    if (err < 0) return f;
    return fmod(f, 2.0) == 0 ? f : f + 1;
}

// round(v * 10^k) as the exact value would round, for |k| <= 22 (10^|k| exact): the
// residual of the product or quotient (fma) settles values that land on .5.
GPUMPI_LIBC_FN double round_pow10(double v, int k) {
    if (k >= 0) return round_scaled(v, pow(10.0, (double)k));
    const double p = pow(10.0, (double)-k);
    const double q = v / p;
    const double r = fma(-q, p, v);  // v - q * p, exactly
    const double f = floor(q);
    const double d = q - f;
    if (d > 0.5) return f + 1;
    if (d < 0.5) return f;
    if (r > 0) return f + 1;
    if (r < 0) return f;
    return fmod(f, 2.0) == 0 ? f : f + 1;
}

// Fixed and exponent forms of a double, digit by digit from its decimal scaling;
// correct to the last printed digit for precisions up to 15.
GPUMPI_LIBC_FN void emit_double(Sink& s, double v, char conv, int prec, int width, bool left,
                                   bool zero, bool plus, bool space, bool alt) {
    char buf[400];
    int n = 0;
    const bool neg = v < 0 || (v == 0 && 1.0 / v < 0);
    if (neg) v = -v;
    const char* sign = neg ? "-" : plus ? "+" : space ? " " : "";
    if (v != v || v > 1.7976931348623157e308) {
        const char* t = v != v ? "nan" : "inf";
        const int l = (int)(sign[0] != 0) + 3;
        if (!left) s.pad(width - l, ' ');
        if (sign[0]) s.put(sign[0]);
        for (int i = 0; i < 3; i++) s.put(conv >= 'a' ? t[i] : (char)(t[i] - 32));
        if (left) s.pad(width - l, ' ');
        return;
    }
    if (prec < 0) prec = 6;
    bool gstyle = conv == 'g' || conv == 'G';
    int exp10 = 0;
    if (v != 0) {
        exp10 = (int)floor(log10(v));
        if (v / pow(10.0, (double)exp10) >= 10.0) exp10++;
        if (v / pow(10.0, (double)exp10) < 1.0) exp10--;
    }
    bool efmt = conv == 'e' || conv == 'E';
    if (gstyle) {
        const int P = prec == 0 ? 1 : prec;
        // the exponent after rounding to P significant digits
        double r = v == 0 ? 0 : round_pow10(v, P - 1 - exp10);
        int X = exp10;
        if (r >= pow(10.0, (double)P)) X++;
        if (P > X && X >= -4) { efmt = false; prec = P - 1 - X; }
        else { efmt = true; prec = P - 1; }
    }
    if (efmt) {
        double scaled = v == 0 ? 0 : round_pow10(v, prec - exp10);
        if (scaled >= pow(10.0, (double)(prec + 1))) { exp10++; scaled = round_pow10(v, prec - exp10); }
        char d[32];
        int nd = 0;
        unsigned long long u = (unsigned long long)scaled;
        do { d[nd++] = (char)('0' + u % 10); u /= 10; } while (u);
        while (nd < prec + 1) d[nd++] = '0';
        buf[n++] = d[nd - 1];
        if (prec > 0 || alt) buf[n++] = '.';
        for (int i = nd - 2; i >= 0; i--) buf[n++] = d[i];
        if (gstyle && !alt) {
            while (n > 1 && buf[n - 1] == '0') n--;
            if (buf[n - 1] == '.') n--;
        }
        buf[n++] = (conv == 'E' || conv == 'G') ? 'E' : 'e';
        int e = exp10;
        buf[n++] = e < 0 ? '-' : '+';
        if (e < 0) e = -e;
        if (e >= 100) buf[n++] = (char)('0' + e / 100);
        buf[n++] = (char)('0' + (e / 10) % 10);
        buf[n++] = (char)('0' + e % 10);
    } else {
        double ip = floor(v);
        double fp = v - ip;
        double scaled;
        if (prec == 0) {
            // the tie at .5 goes to the even integer, as the C library rounds
            scaled = 0;
            ip = round_scaled(v, 1.0);
        } else {
            scaled = round_scaled(fp, pow(10.0, (double)prec));
            if (scaled >= pow(10.0, (double)prec)) { ip += 1; scaled -= pow(10.0, (double)prec); }
        }
        char d[330];
        int nd = 0;
        if (ip < 1e19) {
            unsigned long long u = (unsigned long long)ip;
            do { d[nd++] = (char)('0' + u % 10); u /= 10; } while (u);
        } else {
            double t = ip;
            while (t >= 1 && nd < 320) { double q = floor(t / 10); d[nd++] = (char)('0' + (int)(t - q * 10)); t = q; }
        }
        for (int i = nd - 1; i >= 0; i--) buf[n++] = d[i];
        if (prec > 0 || alt) buf[n++] = '.';
        char f[32];
        unsigned long long u = (unsigned long long)scaled;
        for (int i = 0; i < prec && i < 32; i++) { f[i] = (char)('0' + u % 10); u /= 10; }
        for (int i = (prec < 32 ? prec : 32) - 1; i >= 0; i--) buf[n++] = f[i];
        if (gstyle && !alt && prec > 0) {
            while (buf[n - 1] == '0') n--;
            if (buf[n - 1] == '.') n--;
        }
    }
    const int sl = sign[0] ? 1 : 0;
    if (!left && !zero) s.pad(width - n - sl, ' ');
    if (sl) s.put(sign[0]);
    if (!left && zero) s.pad(width - n - sl, '0');
    for (int i = 0; i < n; i++) s.put(buf[i]);
    if (left) s.pad(width - n - sl, ' ');
}

GPUMPI_LIBC_FN int vformat(char* out, size_t cap, const char* fmt, va_list ap) {
    Sink s{out, cap, 0};
    for (const char* p = fmt; *p; p++) {
        if (*p != '%') { s.put(*p); continue; }
        p++;
        bool left = false, plus = false, space = false, zero = false, alt = false;
        for (;; p++) {
            if (*p == '-') left = true;
            else if (*p == '+') plus = true;
            else if (*p == ' ') space = true;
            else if (*p == '0') zero = true;
            else if (*p == '#') alt = true;
            else break;
        }
        int width = 0, prec = -1;
        if (*p == '*') { width = va_arg(ap, int); if (width < 0) { left = true; width = -width; } p++; }
        else while (*p >= '0' && *p <= '9') width = width * 10 + (*p++ - '0');
        if (*p == '.') {
            p++;
            prec = 0;
            if (*p == '*') { prec = va_arg(ap, int); p++; }
            else while (*p >= '0' && *p <= '9') prec = prec * 10 + (*p++ - '0');
        }
        int len = 0;  // 0 int, 1 long, 2 long long, -1 short, -2 char
        for (;; p++) {
            if (*p == 'l') len = len == 1 ? 2 : 1;
            else if (*p == 'h') len = len == -1 ? -2 : -1;
            else if (*p == 'z' || *p == 'j' || *p == 't') len = 2;
            else break;
        }
        const char c = *p;
        if (!c) break;
        char digits[64];
        switch (c) {
        case 'd': case 'i': {
            long long v = len == 2 ? va_arg(ap, long long) : len == 1 ? va_arg(ap, long) : va_arg(ap, int);
            if (len == -1) v = (short)v;
            if (len == -2) v = (signed char)v;
            const bool neg = v < 0;
            const unsigned long long u = neg ? 0ull - (unsigned long long)v : (unsigned long long)v;
            int nd = (prec == 0 && u == 0) ? 0 : utoa(u, 10, false, digits);
            emit_number(s, neg ? "-" : plus ? "+" : space ? " " : "", digits, nd, prec, width, left,
                        zero && prec < 0);
            break;
        }
        case 'u': case 'x': case 'X': case 'o': {
            unsigned long long u = len == 2 ? va_arg(ap, unsigned long long)
                                 : len == 1 ? va_arg(ap, unsigned long) : va_arg(ap, unsigned);
            if (len == -1) u = (unsigned short)u;
            if (len == -2) u = (unsigned char)u;
            const unsigned base = c == 'u' ? 10 : c == 'o' ? 8 : 16;
            int nd = (prec == 0 && u == 0) ? 0 : utoa(u, base, c == 'X', digits);
            const char* pre = (alt && u && base == 16) ? (c == 'X' ? "0X" : "0x") : (alt && base == 8) ? "0" : "";
            emit_number(s, pre, digits, nd, prec, width, left, zero && prec < 0);
            break;
        }
        case 'p': {
            const unsigned long long u = (unsigned long long)(uintptr_t)va_arg(ap, void*);
            emit_number(s, "0x", digits, utoa(u, 16, false, digits), -1, width, left, false);
            break;
        }
        case 'c': {
            const char ch = (char)va_arg(ap, int);
            if (!left) s.pad(width - 1, ' ');
            s.put(ch);
            if (left) s.pad(width - 1, ' ');
            break;
        }
        case 's': {
            const char* str = va_arg(ap, const char*);
            if (!str) str = "(null)";
            int l = 0;
            while (str[l] && (prec < 0 || l < prec)) l++;
            if (!left) s.pad(width - l, ' ');
            for (int i = 0; i < l; i++) s.put(str[i]);
            if (left) s.pad(width - l, ' ');
            break;
        }
        case 'f': case 'F': case 'e': case 'E': case 'g': case 'G':
            emit_double(s, va_arg(ap, double), c, prec, width, left, zero, plus, space, alt);
            break;
        case '%':
            s.put('%');
            break;
        default:
            s.put('%');
            s.put(c);
        }
    }
    if (cap > 0) out[s.n < cap ? s.n : cap - 1] = 0;
    return (int)s.n;
}

GPUMPI_LIBC_FN int format(char* out, size_t cap, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    const int n = vformat(out, cap, fmt, ap);
    va_end(ap);
    return n;
}

// ---------------------------------------------------------------------------- scanning
// The conversions of C99 scanf: d i u o x X f e g E G a s c n % and scansets [...],
// '*' suppression, field widths and the lengths hh h l ll z j t L.  Src provides
// get() (next byte, -1 at the end) and unget(byte).  Returns the number of assignments,
// or -1 when the input ends before the first conversion.
template <class Src>
GPUMPI_LIBC_FN int vscan(Src& src, const char* fmt, va_list ap) {
    int assigned = 0, count = 0;  // count: bytes consumed (%n)
    bool any_conversion = false;
    auto get = [&]() -> int { const int c = src.get(); if (c >= 0) count++; return c; };
    auto unget = [&](int c) { if (c >= 0) { src.unget(c); count--; } };
    auto space = [](int c) { return c == ' ' || (c >= '\t' && c <= '\r'); };
    for (const char* p = fmt; *p; p++) {
        if (space((unsigned char)*p)) {
            int c;
            while (space(c = get())) {}
            unget(c);
            continue;
        }
        if (*p != '%' || p[1] == '%') {
            if (*p == '%') p++;
            int c = get();
            if (*p == '%') { while (space(c)) c = get(); }
            if (c != (unsigned char)*p) {
                unget(c);
                if (c < 0 && !any_conversion) return -1;
                return assigned;
            }
            continue;
        }
        p++;
        const bool suppress = *p == '*';
        if (suppress) p++;
        int width = 0;
        while (*p >= '0' && *p <= '9') width = width * 10 + (*p++ - '0');
        int len = 0;  // -2 hh, -1 h, 0, 1 l, 2 ll, 3 L
        for (;; p++) {
            if (*p == 'h') len = len == -1 ? -2 : -1;
            else if (*p == 'l') len = len == 1 ? 2 : 1;
            else if (*p == 'L') len = 3;
            else if (*p == 'z' || *p == 'j' || *p == 't') len = 2;
            else break;
        }
        const char conv = *p;
        if (!conv) break;
        if (conv == 'n') {
            if (!suppress) {
                if (len == 2) *va_arg(ap, long long*) = count;
                else if (len == 1) *va_arg(ap, long*) = count;
                else *va_arg(ap, int*) = count;
            }
            continue;
        }
        int c;
        if (conv != 'c' && conv != '[') {
            while (space(c = get())) {}
            unget(c);
        }
        if (width <= 0) width = conv == 'c' ? 1 : 1 << 30;
        char buf[400];
        int n = 0;
        if (conv == 'c' || conv == 's' || conv == '[') {
            bool set[256] = {};
            bool neg = false;
            if (conv == '[') {
                p++;
                if (*p == '^') { neg = true; p++; }
                if (*p == ']') { set[(unsigned char)']'] = true; p++; }
                for (; *p && *p != ']'; p++) {
                    if (p[1] == '-' && p[2] && p[2] != ']') {
                        for (int x = (unsigned char)p[0]; x <= (unsigned char)p[2]; x++) set[x] = true;
                        p += 2;
                    } else {
                        set[(unsigned char)*p] = true;
                    }
                }
                if (!*p) break;
            }
            char* dst = suppress ? nullptr : va_arg(ap, char*);
            while (n < width) {
                c = get();
                if (c < 0) break;
                const bool ok = conv == 'c' ? true : conv == 's' ? !space(c) : (set[c] != neg);
                if (!ok) { unget(c); break; }
                if (dst) dst[n] = (char)c;
                n++;
            }
            if (n == 0 || (conv == 'c' && n < width)) {  // input or matching failure
                if (c < 0 && !any_conversion) return -1;
                return assigned;
            }
            if (dst && conv != 'c') dst[n] = 0;
            any_conversion = true;
            if (dst) assigned++;
            continue;
        }
        // numbers: collect the longest valid prefix into buf
        const bool flt = conv == 'f' || conv == 'e' || conv == 'g' || conv == 'E' ||
                         conv == 'G' || conv == 'a' || conv == 'F';
        int base = conv == 'o' ? 8 : (conv == 'x' || conv == 'X' || conv == 'p') ? 16 : conv == 'i' ? 0 : 10;
        auto take = [&](int ch) { if (n < 399 && n < width) { buf[n++] = (char)ch; return true; } return false; };
        c = get();
        if ((c == '+' || c == '-') && take(c)) c = get();
        if (!flt) {
            if ((base == 0 || base == 16) && c == '0' && n < width) {
                take(c);
                c = get();
                if ((c == 'x' || c == 'X') && n < width) {
                    const int c2 = get();
                    const bool hex = (c2 >= '0' && c2 <= '9') || (c2 >= 'a' && c2 <= 'f') || (c2 >= 'A' && c2 <= 'F');
                    if (hex) { take(c); base = 16; c = c2; }
                    else { unget(c2); }
                } else if (base == 0) {
                    base = 8;
                }
            }
            if (base == 0) base = 10;
            for (;; c = get()) {
                int d = c >= '0' && c <= '9' ? c - '0' : c >= 'a' && c <= 'z' ? c - 'a' + 10 :
                        c >= 'A' && c <= 'Z' ? c - 'A' + 10 : 99;
                if (d >= base || n >= width || !take(c)) break;
            }
        } else {
            bool digits = false;
            while (c >= '0' && c <= '9' && take(c)) { digits = true; c = get(); }
            if (c == '.' && take(c)) {
                c = get();
                while (c >= '0' && c <= '9' && take(c)) { digits = true; c = get(); }
            }
            if (digits && (c == 'e' || c == 'E') && n < width) {
                const int save = n;
                take(c);
                int c2 = get();
                if ((c2 == '+' || c2 == '-') && take(c2)) c2 = get();
                bool ed = false;
                while (c2 >= '0' && c2 <= '9' && take(c2)) { ed = true; c2 = get(); }
                if (!ed) n = save;  // no exponent digits: not part of the number
                c = c2;
            }
        }
        unget(c);
        buf[n] = 0;
        const bool sign_only = n == 1 && (buf[0] == '+' || buf[0] == '-');
        if (n == 0 || sign_only || (n == 2 && buf[1] == '.' && !(buf[0] >= '0' && buf[0] <= '9')) ||
            (n == 1 && buf[0] == '.')) {
            if (n == 0 && c < 0 && !any_conversion) return -1;
            return assigned;
        }
        any_conversion = true;
        if (suppress) continue;
        if (flt) {
            const double v = parse_double(buf, nullptr);
            if (len >= 1) *va_arg(ap, double*) = v;
            else *va_arg(ap, float*) = (float)v;
        } else if (conv == 'p') {
            *va_arg(ap, void**) = (void*)(uintptr_t)parse_int(buf, nullptr, 16);
        } else {
            const long long v = parse_int(buf, nullptr, base);
            if (len == 2) *va_arg(ap, long long*) = v;
            else if (len == 1) *va_arg(ap, long*) = (long)v;
            else if (len == -1) *va_arg(ap, short*) = (short)v;
            else if (len == -2) *va_arg(ap, signed char*) = (signed char)v;
            else *va_arg(ap, int*) = (int)v;
        }
        assigned++;
    }
    return assigned;
}

// A string as a scanning source (sscanf).
struct StringSource {
    const char* s;
    size_t i;
    GPUMPI_LIBC_FN int get() { return s[i] ? (unsigned char)s[i++] : -1; }
    GPUMPI_LIBC_FN void unget(int) { i--; }
};

GPUMPI_LIBC_FN int scan_string(const char* str, const char* fmt, ...) {
    StringSource src{str, 0};
    va_list ap;
    va_start(ap, fmt);
    const int n = vscan(src, fmt, ap);
    va_end(ap);
    return n;
}
}  // namespace gpumpi_libc
