// Host build of the device C library's formatter and number parsers
// (include/gpumpi/libc/format.h, parse.h), compared with the C library on fixed and
// random inputs.  Prints one line per mismatch and "fails=N"; exit status 1 on any.
#define GPUMPI_LIBC_FN inline
#include "gpumpi/libc/format.h"
#include "gpumpi/libc/parse.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

static int fails = 0;

template <class... A>
static void chk(const char* f, A... a) {
    char x[512], y[512];
    const int n1 = gpumpi_libc::format(x, sizeof x, f, a...);
    const int n2 = snprintf(y, sizeof y, f, a...);
    if (n1 != n2 || std::strcmp(x, y)) {
        std::printf("format %s: ours [%s] %d, libc [%s] %d\n", f, x, n1, y, n2);
        fails++;
    }
}

int main() {
    chk("[%d|%5d|%-5d|%05d|%+d|% d|%.3d|%.0d]", -42, 7, 7, -7, 3, 9, 5, 0);
    chk("[%x|%#X|%o|%#o|%lld|%u|%lu|%hhd|%hd]", 255, 255, 8, 8, -123456789012LL, 4000000000u,
        123ul, 300, 70000);
    chk("[%f|%.3f|%e|%.2e|%E]", 3.14159265358979, -2.5, 12345.678, 0.000123, 1e-300);
    chk("[%g|%g|%g|%g|%g|%G|%#g]", 0.0001, 123456789.0, 100.0, 1e-5, 0.5, 1e20, 1.0);
    chk("[%10.4f|%-10.2f|%010.3f|%+.1f|%+.1f]", 2.0 / 3.0, 1.5, -3.25, 0.05, 0.25);
    chk("[%s|%10s|%-6s|%.2s|%c|%5c|%%]", "str", "right", "left", "trunc", 'z', 'q');
    chk("[%.0f|%.0e|%#.0f|%g|%g|%.0f|%.0f]", 2.5, 15.0, 3.0, 0.0, -0.0, 0.5, 1.5);
    chk("%.12f %.15e %.10g %.3g", 1.0 / 7.0, 1.0 / 3.0, 123.456789012345, 0.0009995);
    chk("%5.2f%%  %d/%d %s", 99.5, 3, 4, (const char*)nullptr);
    chk("%*d|%-*d|%.*f", 6, 42, 4, 7, 3, 1.0 / 3.0);
    chk("%f %e %g", INFINITY, -INFINITY, NAN);
    // random magnitudes and precisions against the C library
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> mant(1.0, 10.0);
    std::uniform_int_distribution<int> ex(-12, 15), pr(0, 9);
    char fmt[32];
    for (int i = 0; i < 4000; i++) {
        const double v = (rng() & 1 ? -1 : 1) * mant(rng) * std::pow(10.0, ex(rng));
        const char convs[] = "feg";
        std::snprintf(fmt, sizeof fmt, "%%.%d%c", pr(rng), convs[i % 3]);
        chk(fmt, v);
    }
    for (int i = 0; i < 2000; i++) {
        const long long v = (long long)(rng() >> (rng() % 63)) * (rng() & 1 ? -1 : 1);
        chk("%lld %llx %+12lld", v, (unsigned long long)v, v);
    }
    const char* ints[] = {"42", "  -17xyz", "+0x1F", "0755", "abc", "9223372036854775807", "-0", "0x", "12e3"};
    for (const char* s : ints) {
        char *e1, *e2;
        const long long a = gpumpi_libc::parse_int(s, &e1, 0), b = std::strtoll(s, &e2, 0);
        if (a != b || e1 != e2) {
            std::printf("parse_int %s: ours %lld libc %lld\n", s, a, b);
            fails++;
        }
    }
    const char* dbls[] = {"3.25", "-1e-3", "  6.02e23", "0.1", "12345678.875", "1.5x", ".5", "x", "7.", "1e", "-0.0"};
    for (const char* s : dbls) {
        char *e1, *e2;
        const double a = gpumpi_libc::parse_double(s, &e1), b = std::strtod(s, &e2);
        if ((a != b && std::fabs(a - b) > 1e-15 * std::fabs(b)) || e1 != e2) {
            std::printf("parse_double %s: ours %.17g libc %.17g\n", s, a, b);
            fails++;
        }
    }
    // scanning: the same inputs through vscan and the C library's sscanf
    struct ScanCase { const char* in; const char* fmt; };
    const ScanCase scans[] = {
        {"42 -17 0x1f 0755", "%d %d %i %i"}, {"  12abc", "%d%s"}, {"3.5e2 -0.25 7", "%lf %f %lf"},
        {"a,b,c", "%c,%c,%c"}, {"key=value;rest", "%[^=]=%[^;];%s"}, {"12345", "%3d%2d"},
        {"x 5", "%d"}, {"", "%d"}, {"   ", "%d"}, {"10 20", "%*d %d"}, {"abc 99", "abc %d"},
        {"1.5x", "%lf%c"}, {"-", "%d"}, {"ff 17", "%x %o"}, {"9 %", "%d %%"}, {"1e", "%lf%s"},
        {"+7 -0 .5 5.", "%d %d %lf %lf"}, {"abc", "%2s%s"}, {"Class = S", " Class = %c"},
        {"  2.5  3 end", "%lf %d %n"},
    };
    for (const ScanCase& t : scans) {
        int a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
        double da[4] = {0, 0, 0, 0}, db[4] = {0, 0, 0, 0};
        float fa = 0, fb = 0;
        char sa[3][64] = {{0}}, sb[3][64] = {{0}};
        char ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
        int r1, r2;
        const std::string f = t.fmt;
        if (f == "%lf %f %lf") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &da[0], &fa, &da[1]);
            r2 = std::sscanf(t.in, t.fmt, &db[0], &fb, &db[1]);
        } else if (f == "%c,%c,%c") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &ca[0], &ca[1], &ca[2]);
            r2 = std::sscanf(t.in, t.fmt, &cb[0], &cb[1], &cb[2]);
        } else if (f == "%[^=]=%[^;];%s" || f == "%2s%s") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, sa[0], sa[1], sa[2]);
            r2 = std::sscanf(t.in, t.fmt, sb[0], sb[1], sb[2]);
        } else if (f == "%d%s" || f == "%lf%s") {
            if (f == "%d%s") {
                r1 = gpumpi_libc::scan_string(t.in, t.fmt, &a[0], sa[0]);
                r2 = std::sscanf(t.in, t.fmt, &b[0], sb[0]);
            } else {
                r1 = gpumpi_libc::scan_string(t.in, t.fmt, &da[0], sa[0]);
                r2 = std::sscanf(t.in, t.fmt, &db[0], sb[0]);
            }
        } else if (f == "%lf%c") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &da[0], &ca[0]);
            r2 = std::sscanf(t.in, t.fmt, &db[0], &cb[0]);
        } else if (f == "%d %d %lf %lf") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &a[0], &a[1], &da[0], &da[1]);
            r2 = std::sscanf(t.in, t.fmt, &b[0], &b[1], &db[0], &db[1]);
        } else if (f == " Class = %c") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &ca[0]);
            r2 = std::sscanf(t.in, t.fmt, &cb[0]);
        } else if (f == "%lf %d %n") {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &da[0], &a[0], &a[1]);
            r2 = std::sscanf(t.in, t.fmt, &db[0], &b[0], &b[1]);
        } else {
            r1 = gpumpi_libc::scan_string(t.in, t.fmt, &a[0], &a[1], &a[2], &a[3]);
            r2 = std::sscanf(t.in, t.fmt, &b[0], &b[1], &b[2], &b[3]);
        }
        const bool same = r1 == r2 && std::memcmp(a, b, sizeof a) == 0 && std::memcmp(da, db, sizeof da) == 0 &&
                          fa == fb && std::memcmp(sa, sb, sizeof sa) == 0 && std::memcmp(ca, cb, sizeof ca) == 0;
        if (!same) {
            std::printf("scan \"%s\" with \"%s\": ours %d [%d %d %d %d | %g %g %g | %s %s %s | %c%c%c], "
                        "libc %d [%d %d %d %d | %g %g %g | %s %s %s | %c%c%c]\n", t.in, t.fmt,
                        r1, a[0], a[1], a[2], a[3], da[0], da[1], (double)fa, sa[0], sa[1], sa[2], ca[0], ca[1], ca[2],
                        r2, b[0], b[1], b[2], b[3], db[0], db[1], (double)fb, sb[0], sb[1], sb[2], cb[0], cb[1], cb[2]);
            fails++;
        }
    }
    std::printf("fails=%d\n", fails);
    return fails != 0;
}
