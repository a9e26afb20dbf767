/*
 * gpumpi oracle -- CPU restatement (TEST INFRASTRUCTURE ONLY; see gpumpi_oracle.h).
 * Built with -ffp-contract=off so every FP64 operation rounds as written.
 */
#include "gpumpi_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel Random Numbers: As Easy */
/* as 1, 2, 3", SC'11).  Multipliers and Weyl key increments are the published */
/* constants of the algorithm.                                                 */
/* ------------------------------------------------------------------------- */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void rank_key(uint64_t seed, int rank, uint32_t key[2])
{
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32) ^ (uint32_t)rank;
}

static inline void philox_at(uint64_t seed, int rank, uint32_t w0, uint64_t block,
                             uint32_t stream, uint32_t out[4])
{
    uint32_t key[2], ctr[4];
    rank_key(seed, rank, key);
    ctr[0] = w0;
    ctr[1] = (uint32_t)block;
    ctr[2] = (uint32_t)(block >> 32);
    ctr[3] = stream;
    orc_philox4x32_10(ctr, key, out);
}

static inline double u53(uint32_t hi, uint32_t lo)
{
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)v * (1.0 / 9007199254740992.0); /* 2^-53, exact */
}

/* Config-1 workload: rank r draws n (x, y) pairs from Philox with key
 * (seed_lo, seed_hi ^ r) and counter (i, 0, 0, 0). */
int64_t orc_mcpi_hits(uint64_t seed, int rank, int64_t n)
{
    uint32_t key[2], ctr[4], o[4];
    rank_key(seed, rank, key);
    int64_t hits = 0;
    for (int64_t i = 0; i < n; i++) {
        ctr[0] = (uint32_t)i; ctr[1] = (uint32_t)((uint64_t)i >> 32); ctr[2] = 0; ctr[3] = 0;
        orc_philox4x32_10(ctr, key, o);
        double x = u53(o[0], o[1]);
        double y = u53(o[2], o[3]);
        double xx = x * x;
        double yy = y * y;
        double s = xx + yy;
        hits += (s <= 1.0) ? 1 : 0;
    }
    return hits;
}

void orc_fill_bytes(uint64_t seed, int rank, uint32_t iter, uint8_t *buf, int64_t nbytes)
{
    uint32_t o[4];
    for (int64_t b = 0; b * 16 < nbytes; b++) {
        philox_at(seed, rank, iter, (uint64_t)b, ORC_STREAM_BYTES, o);
        for (int k = 0; k < 16 && b * 16 + k < nbytes; k++)
            buf[b * 16 + k] = (uint8_t)(o[k >> 2] >> (8 * (k & 3)));
    }
}

void orc_fill_int(uint64_t seed, int rank, uint32_t iter, int32_t *out, int64_t n)
{
    uint32_t o[4];
    for (int64_t i = 0; i < n; i++) {
        if ((i & 3) == 0) philox_at(seed, rank, iter, (uint64_t)(i >> 2), ORC_STREAM_INT, o);
        out[i] = (int32_t)(o[i & 3] % 2097153u) - 1048576;
    }
}

/* ln, sqrt and cos(2 pi u) from +, -, *, / and exact bit manipulation only, in the
 * device generator's order (dphpc-project_amd/csrc/runtime/philox.h), so the oracle draws
 * the device's bits: this file is built with -ffp-contract=off like the device code, and
 * IEEE double +, -, *, / round the same way on both.  Pinned against libm in
 * tests/test_oracle.py (orc_det_log / orc_det_sqrt / orc_det_cos_2pi). */
static const double DET_LN2_HI = 0.6931471805599453;
static const double DET_LN2_LO = 2.3190468138462996e-17;
static const double DET_HALF_PI = 1.5707963267948966;
static const double DET_SQRT2 = 1.4142135623730951;

static double det_from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t det_to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m-1)/(m+1), to s^23. */
double orc_det_log(double x)
{
    const uint64_t b = det_to_bits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = det_from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > DET_SQRT2) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    static const int odd[11] = {21, 19, 17, 15, 13, 11, 9, 7, 5, 3, 1};
    double p = 1.0 / 23.0;
    for (int k = 0; k < 11; k++) p = p * s2 + 1.0 / (double)odd[k];
    const double lnm = (2.0 * s) * p;
    const double de = (double)e;
    return de * DET_LN2_HI + (de * DET_LN2_LO + lnm);
}

/* Newton from 2^floor(e/2), seven steps. */
double orc_det_sqrt(double x)
{
    if (x == 0.0) return 0.0;
    const int e = (int)((det_to_bits(x) >> 52) & 0x7FF) - 1023;
    const int h = e >= 0 ? e / 2 : -((1 - e) / 2);
    double y = det_from_bits((uint64_t)(h + 1023) << 52);
    for (int k = 0; k < 7; k++) y = 0.5 * (y + x / y);
    return y;
}

/* quadrant of 4u (exact), then the Taylor series of cos and sin of (4u - q) pi/2 in
 * Horner form: cos t = 1 - t^2/(1*2) (1 - t^2/(3*4) (... (1 - t^2/(21*22)))),
 * sin t = t (1 - t^2/(2*3) (... (1 - t^2/(22*23)))). */
double orc_det_cos_2pi(double u)
{
    const double f = 4.0 * u;
    const int q = (int)f;
    const double t = (f - (double)q) * DET_HALF_PI;
    const double t2 = t * t;
    double c = 1.0, sn = 1.0;
    for (int k = 21; k >= 1; k -= 2) c = 1.0 - t2 * (1.0 / (double)(k * (k + 1))) * c;
    for (int k = 22; k >= 2; k -= 2) sn = 1.0 - t2 * (1.0 / (double)(k * (k + 1))) * sn;
    sn = t * sn;
    return q == 0 ? c : q == 1 ? -sn : q == 2 ? -c : sn;
}

/* N(0,1) by Box-Muller on Philox block i: sqrt(-2 ln u1) cos(2 pi u2), u1 in (0, 1] from
 * words 0-1, u2 in [0, 1) from words 2-3, 53 bits each. */
static void orc_fill_normal_one(uint64_t seed, int rank, uint32_t iter, int64_t i, double *out)
{
    uint32_t o[4];
    philox_at(seed, rank, iter, (uint64_t)i, ORC_STREAM_NORMAL, o);
    const double u1 = (double)(((((uint64_t)o[0] << 32) | o[1]) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (double)((((uint64_t)o[2] << 32) | o[3]) >> 11) * (1.0 / 9007199254740992.0);
    *out = orc_det_sqrt(-2.0 * orc_det_log(u1)) * orc_det_cos_2pi(u2);
}

void orc_fill_normal(uint64_t seed, int rank, uint32_t iter, double *out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) orc_fill_normal_one(seed, rank, iter, i, &out[i]);
}

/* Elements idx[0..n) of the orc_fill_int / orc_fill_normal streams: element i depends
 * only on its own Philox counter(s), so samples of a long input are cheap to draw. */
void orc_fill_int_at(uint64_t seed, int rank, uint32_t iter, const int64_t *idx, int64_t n,
                     int32_t *out)
{
    uint32_t o[4];
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = idx[k];
        philox_at(seed, rank, iter, (uint64_t)(i >> 2), ORC_STREAM_INT, o);
        out[k] = (int32_t)(o[i & 3] % 2097153u) - 1048576;
    }
}

void orc_fill_normal_at(uint64_t seed, int rank, uint32_t iter, const int64_t *idx, int64_t n,
                        double *out)
{
    for (int64_t k = 0; k < n; k++) orc_fill_normal_one(seed, rank, iter, idx[k], &out[k]);
}

void orc_fill_keys_uniform(uint64_t seed, int rank, uint32_t step, uint32_t *out,
                           int64_t n, int key_bits)
{
    uint32_t o[4];
    for (int64_t i = 0; i < n; i++) {
        if ((i & 3) == 0) philox_at(seed, rank, step, (uint64_t)(i >> 2), ORC_STREAM_KEYS, o);
        out[i] = o[i & 3] >> (32 - key_bits);
    }
}

void orc_fill_keys_zipf(uint64_t seed, int rank, uint32_t step, const double *cdf,
                        int64_t ncdf, uint32_t *out, int64_t n)
{
    uint32_t o[4];
    for (int64_t i = 0; i < n; i++) {
        if ((i & 1) == 0) philox_at(seed, rank, step, (uint64_t)(i >> 1), ORC_STREAM_ZIPF, o);
        double u = u53(o[2 * (i & 1)], o[2 * (i & 1) + 1]);
        /* first k with cdf[k] > u */
        int64_t lo = 0, hi = ncdf - 1;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] > u) hi = mid; else lo = mid + 1;
        }
        out[i] = (uint32_t)lo;
    }
}

/* ------------------------------------------------------------------------- */
/* Datatypes (gpu_mpi/mpi_types_list.cuh:4-33) and operators                   */
/* ------------------------------------------------------------------------- */
enum {
    T_CHAR, T_SHORT, T_INT, T_LONG, T_LONG_LONG_INT, T_LONG_LONG, T_SIGNED_CHAR,
    T_UNSIGNED_CHAR, T_UNSIGNED_SHORT, T_UNSIGNED, T_UNSIGNED_LONG, T_UNSIGNED_LONG_LONG,
    T_FLOAT, T_DOUBLE, T_LONG_DOUBLE, T_WCHAR, T_C_BOOL, T_INT8, T_INT16, T_INT32, T_INT64,
    T_UINT8, T_UINT16, T_UINT32, T_UINT64, T_C_COMPLEX, T_C_FLOAT_COMPLEX, T_C_DOUBLE_COMPLEX,
    T_C_LONG_DOUBLE_COMPLEX, T_BYTE, T_COUNT
};
enum { O_MAX, O_MIN, O_SUM, O_PROD, O_LAND, O_BAND, O_LOR, O_BOR, O_LXOR, O_BXOR,
       O_MAXLOC, O_MINLOC, O_COUNT };

static const int type_sizes[T_COUNT] = {
    1, 2, 4, 8, 8, 8, 1, 1, 2, 4, 8, 8, 4, 8, 8, 4, 1, 1, 2, 4, 8, 1, 2, 4, 8, 8, 8, 16, 16, 1
};

int orc_type_size(int t) { return (t >= 0 && t < T_COUNT) ? type_sizes[t] : -1; }

static int is_integer(int t)
{
    switch (t) {
    case T_SHORT: case T_INT: case T_LONG: case T_LONG_LONG_INT: case T_LONG_LONG:
    case T_SIGNED_CHAR: case T_UNSIGNED_CHAR: case T_UNSIGNED_SHORT: case T_UNSIGNED:
    case T_UNSIGNED_LONG: case T_UNSIGNED_LONG_LONG: case T_INT8: case T_INT16: case T_INT32:
    case T_INT64: case T_UINT8: case T_UINT16: case T_UINT32: case T_UINT64: return 1;
    default: return 0;
    }
}
static int is_floating(int t) { return t == T_FLOAT || t == T_DOUBLE || t == T_LONG_DOUBLE; }
static int is_logical(int t) { return t == T_C_BOOL; }
static int is_complex(int t)
{
    return t == T_C_COMPLEX || t == T_C_FLOAT_COMPLEX || t == T_C_DOUBLE_COMPLEX ||
           t == T_C_LONG_DOUBLE_COMPLEX;
}
static int is_byte(int t) { return t == T_BYTE; }

/* gpu_mpi/operators.cu:153-200 */
int orc_op_allowed(int t, int op)
{
    if (t < 0 || t >= T_COUNT) return 0;
    switch (op) {
    case O_MAX: case O_MIN: return is_integer(t) || is_floating(t);
    case O_SUM: case O_PROD: return is_integer(t) || is_floating(t) || is_complex(t);
    case O_LAND: case O_LOR: case O_LXOR: return is_integer(t) || is_logical(t);
    case O_BAND: case O_BOR: case O_BXOR: return is_integer(t) || is_byte(t);
    default: return 0; /* MAXLOC/MINLOC have no legal type in the type list */
    }
}

/* inout[i] = op(in[i], inout[i]) -- operators.cu:73-151, 210-220 */
#define APPLY_SCALAR(T)                                                               \
    do {                                                                              \
        const T *a = (const T *)in; T *b = (T *)inout;                                \
        for (int64_t i = 0; i < n; i++) {                                             \
            T x = a[i], y = b[i];                                                     \
            switch (op) {                                                             \
            case O_MAX: b[i] = x > y ? x : y; break;                                  \
            case O_MIN: b[i] = x > y ? y : x; break;                                  \
            case O_SUM: b[i] = (T)(x + y); break;                                     \
            case O_PROD: b[i] = (T)(x * y); break;                                    \
            case O_LAND: b[i] = (T)(x && y); break;                                   \
            case O_LOR: b[i] = (T)(x || y); break;                                    \
            case O_LXOR: b[i] = (T)(x != y); break;                                   \
            default: break;                                                           \
            }                                                                         \
        }                                                                             \
    } while (0)
#define APPLY_BITS(T)                                                                 \
    do {                                                                              \
        const T *a = (const T *)in; T *b = (T *)inout;                                \
        for (int64_t i = 0; i < n; i++) {                                             \
            T x = a[i], y = b[i];                                                     \
            switch (op) {                                                             \
            case O_BAND: b[i] = (T)(x & y); break;                                    \
            case O_BOR: b[i] = (T)(x | y); break;                                     \
            case O_BXOR: b[i] = (T)(x ^ y); break;                                    \
            default: break;                                                           \
            }                                                                         \
        }                                                                             \
    } while (0)
#define APPLY_COMPLEX(T)                                                              \
    do {                                                                              \
        const T *a = (const T *)in; T *b = (T *)inout;                                \
        for (int64_t i = 0; i < n; i++) {                                             \
            T xr = a[2 * i], xi = a[2 * i + 1], yr = b[2 * i], yi = b[2 * i + 1];     \
            if (op == O_SUM) { b[2 * i] = xr + yr; b[2 * i + 1] = xi + yi; }          \
            else {                                                                    \
                T rr = xr * yr; T ii = xi * yi; T ri = xr * yi; T ir = xi * yr;       \
                b[2 * i] = rr - ii; b[2 * i + 1] = ri + ir;                           \
            }                                                                         \
        }                                                                             \
    } while (0)

static void apply_op(int t, int op, const void *in, void *inout, int64_t n)
{
    int bits = (op == O_BAND || op == O_BOR || op == O_BXOR);
    switch (t) {
#define CASE_INT(TT, CT) case TT: if (bits) APPLY_BITS(CT); else APPLY_SCALAR(CT); break;
    CASE_INT(T_SHORT, short)
    CASE_INT(T_INT, int)
    CASE_INT(T_LONG, long)
    CASE_INT(T_LONG_LONG_INT, long long)
    CASE_INT(T_LONG_LONG, long long)
    CASE_INT(T_SIGNED_CHAR, signed char)
    CASE_INT(T_UNSIGNED_CHAR, unsigned char)
    CASE_INT(T_UNSIGNED_SHORT, unsigned short)
    CASE_INT(T_UNSIGNED, unsigned)
    CASE_INT(T_UNSIGNED_LONG, unsigned long)
    CASE_INT(T_UNSIGNED_LONG_LONG, unsigned long long)
    CASE_INT(T_INT8, int8_t)
    CASE_INT(T_INT16, int16_t)
    CASE_INT(T_INT32, int32_t)
    CASE_INT(T_INT64, int64_t)
    CASE_INT(T_UINT8, uint8_t)
    CASE_INT(T_UINT16, uint16_t)
    CASE_INT(T_UINT32, uint32_t)
    CASE_INT(T_UINT64, uint64_t)
    CASE_INT(T_BYTE, unsigned char)
    CASE_INT(T_C_BOOL, unsigned char)
#undef CASE_INT
    case T_FLOAT: APPLY_SCALAR(float); break;
    case T_DOUBLE: case T_LONG_DOUBLE: APPLY_SCALAR(double); break;
    case T_C_COMPLEX: case T_C_FLOAT_COMPLEX: APPLY_COMPLEX(float); break;
    case T_C_DOUBLE_COMPLEX: case T_C_LONG_DOUBLE_COMPLEX: APPLY_COMPLEX(double); break;
    default: break;
    }
}

/* gpu_mpi/mpi.cu:209-225: rank 0's data seeds the result, then
 * recv = op(buf_r, recv) for r = 1 .. P-1. */
int orc_reduce(int P, int t, int op, const void *const *inputs, int64_t count, void *out)
{
    if (!orc_op_allowed(t, op)) return -1;
    size_t bytes = (size_t)count * (size_t)type_sizes[t];
    memcpy(out, inputs[0], bytes);
    for (int r = 1; r < P; r++) apply_op(t, op, inputs[r], out, count);
    return 0;
}

/* gpu_mpi/mpi.cu:363-408: block i of rank me's recvbuf <- rank i's send block me */
void orc_alltoallv(int P, int send_elem, int recv_elem,
                   const void *const *sendbufs, const int *sendcounts, const int *sdispls,
                   void *const *recvbufs, const int *recvcounts, const int *rdispls)
{
    for (int d = 0; d < P; d++) {
        for (int s = 0; s < P; s++) {
            size_t nb = (size_t)recvcounts[d * P + s] * (size_t)recv_elem;
            const char *src = (const char *)sendbufs[s] + (size_t)sdispls[s * P + d] * send_elem;
            char *dst = (char *)recvbufs[d] + (size_t)rdispls[d * P + s] * recv_elem;
            if (nb) memcpy(dst, src, nb);
        }
    }
}

/* gpu_mpi/communicator.cu:154-213: members of one color ordered by (key, old rank) */
void orc_comm_split(int P, const int *color, const int *key, int *new_rank, int *new_size)
{
    for (int r = 0; r < P; r++) {
        int nr = 0, ns = 0;
        for (int j = 0; j < P; j++) {
            if (color[j] != color[r]) continue;
            ns++;
            if (key[j] < key[r] || (key[j] == key[r] && j < r)) nr++;
        }
        new_rank[r] = nr;
        new_size[r] = ns;
    }
}

/* ------------------------------------------------------------------------- */
/* NPB IS 3.4.1 (is.c).  randlc (406-452) is the linear congruential generator  */
/* x <- a*x mod 2^46 returning x*2^-46; here it is evaluated in exact integer  */
/* arithmetic, which the double-precision original reproduces exactly on any  */
/* machine with >= 48 mantissa bits (is.c:371-386).                            */
/* ------------------------------------------------------------------------- */
#define NPB_MASK46 ((1ULL << 46) - 1)
#define NPB_A 1220703125ULL
#define NPB_SEED 314159265ULL
#define NPB_MAX_IT 10
#define NPB_TEST 5

static inline uint64_t lcg_mul(uint64_t a, uint64_t x) { return (a * x) & NPB_MASK46; }

int orc_npb_class_params(char cls, orc_npb_class *o)
{
    o->cls = cls;
    switch (cls) { /* is.c:57-104 */
    case 'S': o->total_keys_log2 = 16; o->max_key_log2 = 11; o->num_buckets_log2 = 9; return 0;
    case 'W': o->total_keys_log2 = 20; o->max_key_log2 = 16; o->num_buckets_log2 = 10; return 0;
    case 'A': o->total_keys_log2 = 23; o->max_key_log2 = 19; o->num_buckets_log2 = 10; return 0;
    case 'B': o->total_keys_log2 = 25; o->max_key_log2 = 21; o->num_buckets_log2 = 10; return 0;
    case 'C': o->total_keys_log2 = 27; o->max_key_log2 = 23; o->num_buckets_log2 = 10; return 0;
    default: return -1;
    }
}

/* is.c:264-279 */
static void npb_test_arrays(char cls, int64_t idx[5], int64_t rnk[5])
{
    static const int64_t S_i[5] = {48427, 17148, 23627, 62548, 4431};
    static const int64_t S_r[5] = {0, 18, 346, 64917, 65463};
    static const int64_t W_i[5] = {357773, 934767, 875723, 898999, 404505};
    static const int64_t W_r[5] = {1249, 11698, 1039987, 1043896, 1048018};
    static const int64_t A_i[5] = {2112377, 662041, 5336171, 3642833, 4250760};
    static const int64_t A_r[5] = {104, 17523, 123928, 8288932, 8388264};
    static const int64_t B_i[5] = {41869, 812306, 5102857, 18232239, 26860214};
    static const int64_t B_r[5] = {33422937, 10244, 59149, 33135281, 99};
    static const int64_t C_i[5] = {44172927, 72999161, 74326391, 129606274, 21736814};
    static const int64_t C_r[5] = {61147, 882988, 266290, 133997595, 133525895};
    const int64_t *pi, *pr;
    switch (cls) {
    case 'S': pi = S_i; pr = S_r; break;
    case 'W': pi = W_i; pr = W_r; break;
    case 'A': pi = A_i; pr = A_r; break;
    case 'B': pi = B_i; pr = B_r; break;
    default: pi = C_i; pr = C_r; break;
    }
    memcpy(idx, pi, sizeof(int64_t) * 5);
    memcpy(rnk, pr, sizeof(int64_t) * 5);
}

/* is.c:849-905: per-class shift of the expected test ranks */
static int64_t npb_adjust(char cls, int i, int it, int64_t r)
{
    switch (cls) {
    case 'S': return (i <= 2) ? r + it : r - it;
    case 'W': return (i < 2) ? r + it - 2 : r - it;
    case 'A': return (i <= 2) ? r + it - 1 : r - (it - 1);
    case 'B': return (i == 1 || i == 2 || i == 4) ? r + it : r - it;
    default:  return (i <= 2) ? r + it : r - it; /* 'C' */
    }
}

/* find_my_seed (is.c:470-510): s * (a^(2^mq))^kn mod 2^46, mq = floor(log2(nn/np)) */
static uint64_t npb_my_seed(int kn, int np, int64_t nn)
{
    int64_t nq = nn / np;
    int mq = 0;
    for (; nq > 1; mq++, nq /= 2) {}
    uint64_t an = NPB_A;
    for (int i = 1; i <= mq; i++) an = lcg_mul(an, an);
    uint64_t t1 = NPB_SEED, t2 = an;
    int64_t kk = kn;
    for (int i = 1; i <= 100; i++) {
        int64_t ik = kk / 2;
        if (2 * ik != kk) t1 = lcg_mul(t2, t1);
        if (ik == 0) break;
        t2 = lcg_mul(t2, t2);
        kk = ik;
    }
    return t1;
}

/* create_seq (is.c:520-537): key = floor((MAX_KEY/4) * (r1+r2+r3+r4)), which in
 * integers is (x1+x2+x3+x4) >> (48 - max_key_log2). */
void orc_npb_create_keys(const orc_npb_class *c, int P, int rank, int32_t *keys)
{
    int64_t total = 1LL << c->total_keys_log2;
    int64_t nk = total / P;
    uint64_t x = npb_my_seed(rank, P, 4 * total);
    int sh = 48 - c->max_key_log2;
    for (int64_t i = 0; i < nk; i++) {
        uint64_t s = 0;
        for (int j = 0; j < 4; j++) { x = lcg_mul(NPB_A, x); s += x; }
        keys[i] = (int32_t)(s >> sh);
    }
}

typedef struct npb_state {
    const orc_npb_class *c;
    int P;
    int64_t nk, nb, maxkey;
    int shift;
    int64_t tidx[5], trank[5];
    int32_t *keys;      /* P * nk, key_array per rank (rank-major) */
    int32_t *kb1;       /* per-rank work array, sized maxkey+1 */
    int32_t **kb2;      /* received keys per rank */
    int64_t *kb2n;
    int32_t **kbp_save; /* saved key_buff_ptr arrays for full verify */
    int64_t *min_save;
} npb_state;

/* One rank() call on all P ranks (is.c:631-928).  Returns partial-verify passes. */
static int64_t npb_rank(npb_state *st, int it, int keep)
{
    const orc_npb_class *c = st->c;
    int P = st->P;
    int64_t nk = st->nk, NB = st->nb;
    int shift = st->shift;
    int64_t passed = 0;

    /* is.c:650-654: rank 0 alters two keys each iteration */
    st->keys[it] = it;
    st->keys[it + NPB_MAX_IT] = (int32_t)(st->maxkey - it);

    int64_t *bsize = calloc((size_t)P * (NB + 5), sizeof(int64_t));
    int64_t *tot = calloc((size_t)(NB + 5), sizeof(int64_t));
    int32_t *kb1 = malloc(sizeof(int32_t) * (size_t)P * nk);
    int *scount = calloc((size_t)P * P, sizeof(int));
    int *sdispl = calloc((size_t)P * P, sizeof(int));
    int64_t *pbd1 = calloc((size_t)P, sizeof(int64_t));
    int64_t *pbd2 = calloc((size_t)P, sizeof(int64_t));

    for (int r = 0; r < P; r++) {
        int64_t *bs = bsize + (size_t)r * (NB + 5);
        int32_t *ka = st->keys + (size_t)r * nk;
        for (int i = 0; i < NPB_TEST; i++)
            if (st->tidx[i] / nk == r) bs[NB + i] = ka[st->tidx[i] % nk];
        for (int64_t i = 0; i < nk; i++) bs[ka[i] >> shift]++;
        int64_t *ptr = malloc(sizeof(int64_t) * NB);
        ptr[0] = 0;
        for (int64_t i = 1; i < NB; i++) ptr[i] = ptr[i - 1] + bs[i - 1];
        for (int64_t i = 0; i < nk; i++) {
            int32_t k = ka[i];
            kb1[(size_t)r * nk + ptr[k >> shift]++] = k;
        }
        free(ptr);
    }
    /* Allreduce SUM */
    for (int r = 0; r < P; r++)
        for (int64_t i = 0; i < NB + 5; i++) tot[i] += bsize[(size_t)r * (NB + 5) + i];

    /* Splitters (is.c:721-751); send counts use each rank's local bucket sizes. */
    for (int r = 0; r < P; r++) {
        int64_t *bs = bsize + (size_t)r * (NB + 5);
        int64_t acc = 0, lacc = 0;
        int64_t j = 0;
        int *sc = scount + (size_t)r * P, *sd = sdispl + (size_t)r * P;
        sd[0] = 0;
        if (r == 0) pbd1[0] = 0;
        for (int64_t i = 0; i < NB; i++) {
            acc += tot[i];
            lacc += bs[i];
            if (acc >= (j + 1) * nk) {
                sc[j] = (int)lacc;
                if (j != 0) {
                    sd[j] = sd[j - 1] + sc[j - 1];
                    if (r == 0) pbd1[j] = pbd2[j - 1] + 1;
                }
                if (r == 0) pbd2[j] = i;
                j++;
                lacc = 0;
            }
        }
        while (j < P) {
            sc[j] = 0;
            if (r == 0) pbd1[j] = 1;
            j++;
        }
    }
    /* Alltoall(counts) + Alltoallv(keys) */
    for (int d = 0; d < P; d++) {
        int64_t n = 0;
        for (int s = 0; s < P; s++) n += scount[(size_t)s * P + d];
        free(st->kb2[d]);
        st->kb2[d] = malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
        st->kb2n[d] = n;
        int64_t off = 0;
        for (int s = 0; s < P; s++) {
            int cnt = scount[(size_t)s * P + d];
            memcpy(st->kb2[d] + off, kb1 + (size_t)s * nk + sdispl[(size_t)s * P + d],
                   sizeof(int32_t) * (size_t)cnt);
            off += cnt;
        }
    }
    /* Local ranking and partial verification (is.c:785-905) */
    for (int r = 0; r < P; r++) {
        int64_t minv = pbd1[r] << shift;
        int64_t maxv = ((pbd2[r] + 1) << shift) - 1;
        int64_t range = maxv - minv + 1;
        int32_t *cnt = calloc((size_t)(range > 0 ? range : 1), sizeof(int32_t));
        int64_t m = 0;
        for (int k = 0; k < r; k++)
            for (int64_t i = pbd1[k]; i <= pbd2[k]; i++) m += tot[i];
        int64_t jl = 0;
        for (int64_t i = pbd1[r]; i <= pbd2[r]; i++) jl += tot[i];
        for (int64_t i = 0; i < jl; i++) cnt[st->kb2[r][i] - minv]++;
        for (int64_t i = minv; i < maxv; i++) cnt[i + 1 - minv] += cnt[i - minv];
        for (int i = 0; i < NPB_TEST; i++) {
            int64_t k = tot[NB + i];
            if (minv <= k && k <= maxv) {
                int64_t kr = (k - 1 - minv >= 0 ? cnt[k - 1 - minv] : 0) + m;
                int64_t tr = npb_adjust(c->cls, i, it, st->trank[i]);
                if (kr == tr) passed++;
            }
        }
        if (keep) {
            free(st->kbp_save[r]);
            st->kbp_save[r] = cnt;
            st->min_save[r] = minv;
            st->kb2n[r] = jl;
        } else {
            free(cnt);
        }
    }
    free(bsize); free(tot); free(kb1); free(scount); free(sdispl); free(pbd1); free(pbd2);
    return passed;
}

static void npb_init(npb_state *st, const orc_npb_class *c, int P, int32_t *keys)
{
    memset(st, 0, sizeof(*st));
    st->c = c; st->P = P;
    st->nk = (1LL << c->total_keys_log2) / P;
    st->nb = 1LL << c->num_buckets_log2;
    st->maxkey = 1LL << c->max_key_log2;
    st->shift = c->max_key_log2 - c->num_buckets_log2;
    npb_test_arrays(c->cls, st->tidx, st->trank);
    st->keys = keys;
    st->kb2 = calloc((size_t)P, sizeof(int32_t *));
    st->kb2n = calloc((size_t)P, sizeof(int64_t));
    st->kbp_save = calloc((size_t)P, sizeof(int32_t *));
    st->min_save = calloc((size_t)P, sizeof(int64_t));
}

static void npb_free(npb_state *st)
{
    for (int r = 0; r < st->P; r++) { free(st->kb2[r]); free(st->kbp_save[r]); }
    free(st->kb2); free(st->kb2n); free(st->kbp_save); free(st->min_save);
}

int64_t orc_npb_is_iteration(const orc_npb_class *c, int P, int iteration, int32_t *keys)
{
    npb_state st;
    npb_init(&st, c, P, keys);
    int64_t p = npb_rank(&st, iteration, 0);
    npb_free(&st);
    return p;
}

int64_t orc_npb_is_run(const orc_npb_class *c, int P)
{
    int64_t nk = (1LL << c->total_keys_log2) / P;
    int32_t *keys = malloc(sizeof(int32_t) * (size_t)P * nk);
    for (int r = 0; r < P; r++) orc_npb_create_keys(c, P, r, keys + (size_t)r * nk);
    npb_state st;
    npb_init(&st, c, P, keys);
    npb_rank(&st, 1, 0); /* untimed warm-up, is.c:1069 */
    int64_t passed = 0;
    for (int it = 1; it <= NPB_MAX_IT; it++) passed += npb_rank(&st, it, it == NPB_MAX_IT);
    /* full_verify (is.c:545-621): place keys, then check global order */
    int32_t prev_last = 0;
    int have_prev = 0;
    for (int r = 0; r < P; r++) {
        int64_t n = st.kb2n[r];
        int32_t *out = malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
        int32_t *cnt = st.kbp_save[r];
        int64_t minv = st.min_save[r];
        for (int64_t i = 0; i < n; i++) {
            int32_t k = st.kb2[r][i];
            out[--cnt[k - minv]] = k;
        }
        int64_t bad = 0;
        if (r > 0 && n > 0 && have_prev && prev_last > out[0]) bad++;
        for (int64_t i = 1; i < n; i++)
            if (out[i - 1] > out[i]) bad++;
        if (bad == 0) passed++;
        if (n > 0) { prev_last = out[n - 1]; have_prev = 1; }
        free(out);
    }
    npb_free(&st);
    free(keys);
    return passed;
}

/* ------------------------------------------------------------------------- */
/* NPB-style bucket exchange for the workload keys                             */
/* ------------------------------------------------------------------------- */
void orc_bucket_plan(int P, int64_t n, int key_bits, int bucket_bits, const uint32_t *keys,
                     uint32_t *key_buff1, int *send_count, int *send_displ,
                     int *bucket_lo, int *bucket_hi)
{
    int64_t NB = 1LL << bucket_bits;
    int shift = key_bits - bucket_bits;
    int64_t *bs = calloc((size_t)P * NB, sizeof(int64_t));
    int64_t *tot = calloc((size_t)NB, sizeof(int64_t));
    for (int r = 0; r < P; r++) {
        const uint32_t *ka = keys + (size_t)r * n;
        int64_t *b = bs + (size_t)r * NB;
        for (int64_t i = 0; i < n; i++) b[ka[i] >> shift]++;
        int64_t *ptr = malloc(sizeof(int64_t) * NB);
        ptr[0] = 0;
        for (int64_t i = 1; i < NB; i++) ptr[i] = ptr[i - 1] + b[i - 1];
        for (int64_t i = 0; i < n; i++) key_buff1[(size_t)r * n + ptr[ka[i] >> shift]++] = ka[i];
        free(ptr);
        for (int64_t i = 0; i < NB; i++) tot[i] += b[i];
    }
    for (int r = 0; r < P; r++) {
        int64_t *b = bs + (size_t)r * NB;
        int *sc = send_count + (size_t)r * P, *sd = send_displ + (size_t)r * P;
        int64_t acc = 0, lacc = 0, j = 0;
        int64_t prev_hi = -1;
        for (int64_t i = 0; i < NB; i++) {
            acc += tot[i];
            lacc += b[i];
            if (acc >= (j + 1) * n) {
                sc[j] = (int)lacc;
                if (r == 0) { bucket_lo[j] = (int)(prev_hi + 1); bucket_hi[j] = (int)i; }
                prev_hi = i;
                j++;
                lacc = 0;
            }
        }
        while (j < P) {
            sc[j] = 0;
            if (r == 0) { bucket_lo[j] = 1; bucket_hi[j] = 0; }
            j++;
        }
        sd[0] = 0;
        for (int d = 1; d < P; d++) sd[d] = sd[d - 1] + sc[d - 1];
    }
    free(bs); free(tot);
}

void orc_counts_transpose(int P, const int *send_count, int *recv_count, int *recv_displ)
{
    for (int d = 0; d < P; d++) {
        int acc = 0;
        for (int s = 0; s < P; s++) {
            int c = send_count[(size_t)s * P + d];
            recv_count[(size_t)d * P + s] = c;
            recv_displ[(size_t)d * P + s] = acc;
            acc += c;
        }
    }
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

void orc_local_sort(int64_t nkeys, const uint32_t *keys, uint32_t *out)
{
    memcpy(out, keys, sizeof(uint32_t) * (size_t)nkeys);
    qsort(out, (size_t)nkeys, sizeof(uint32_t), cmp_u32);
}

/* ------------------------------------------------------------------------- */
/* Particle exchange.  The workload is the particle variant of the batched uneven
 * exchange; its one communication step is an MPI_Alltoallv whose semantics are
 * those of gpu_mpi/mpi.cu:363-408 (restated by orc_alltoallv). */
int orc_particle_dir(uint64_t seed, uint32_t step, uint64_t id)
{
    const uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), step, ORC_STREAM_MIGR};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    if (o[0] >= ORC_PART_MOVE) return -1;
    return (int)(((uint64_t)o[1] * 6u) >> 32);
}

int orc_grid_neighbour(int gx, int gy, int gz, int rank, int dir)
{
    int x = rank % gx, y = (rank / gx) % gy, z = rank / (gx * gy);
    switch (dir) {
    case 0: x = (x + 1) % gx; break;
    case 1: x = (x + gx - 1) % gx; break;
    case 2: y = (y + 1) % gy; break;
    case 3: y = (y + gy - 1) % gy; break;
    case 4: z = (z + 1) % gz; break;
    default: z = (z + gz - 1) % gz; break;
    }
    return x + gx * (y + gy * z);
}

void orc_particles_init_rank(int gx, int gy, int gz, int rank, int64_t n0, uint64_t seed,
                             double *parts)
{
    const double q20 = 1.0 / 1048576.0, q15 = 1.0 / 32768.0;
    const int cx = rank % gx, cy = (rank / gx) % gy, cz = rank / (gx * gy);
    (void)gz;
    for (int64_t k = 0; k < n0; k++) {
        uint32_t o[4];
        philox_at(seed, rank, 0u, (uint64_t)k, ORC_STREAM_PART, o);
        double *q = parts + k * ORC_PART_FIELDS;
        q[0] = (double)cx + (double)(o[0] >> 12) * q20;
        q[1] = (double)cy + (double)(o[1] >> 12) * q20;
        q[2] = (double)cz + (double)(o[2] >> 12) * q20;
        q[3] = (double)((int32_t)(o[3] & 0xFFFFu) - 32768) * q15;
        q[4] = (double)((int32_t)(o[3] >> 16) - 32768) * q15;
        q[5] = (double)(k & 1023) * (1.0 / 1024.0);
        q[6] = 1.0;
        q[7] = (double)((int64_t)rank * n0 + k);
    }
}

void orc_particles_init(int gx, int gy, int gz, int64_t n0, uint64_t seed, int64_t cap,
                        double *parts, int64_t *counts)
{
    const int P = gx * gy * gz;
    for (int r = 0; r < P; r++) {
        orc_particles_init_rank(gx, gy, gz, r, n0, seed, parts + (size_t)r * cap * ORC_PART_FIELDS);
        counts[r] = n0;
    }
}

int64_t orc_particles_step(int gx, int gy, int gz, uint64_t seed, uint32_t step, int64_t cap,
                           const double *in, const int64_t *cin, double *out, int64_t *cout)
{
    const int P = gx * gy * gz;
    const int dims[3] = {gx, gy, gz};
    const size_t F = ORC_PART_FIELDS;
    int *sc = calloc((size_t)P * P, sizeof(int)), *sd = calloc((size_t)P * P, sizeof(int));
    int *rc = calloc((size_t)P * P, sizeof(int)), *rd = calloc((size_t)P * P, sizeof(int));
    double *send = malloc(sizeof(double) * (size_t)P * cap * F);
    const void **sbufs = malloc(sizeof(void *) * P);
    void **rbufs = malloc(sizeof(void *) * P);
    int64_t *stay = malloc(sizeof(int64_t) * P);
    int64_t moved = 0, status = 0;
    for (int r = 0; r < P; r++) {
        const double *p = in + (size_t)r * cap * F;
        double *o = out + (size_t)r * cap * F;
        double *sb = send + (size_t)r * cap * F;
        int64_t n = cin[r], ns = 0;
        /* pass 1: migrants per destination rank; pass 2: pack them in ascending
         * destination order, each group in the original particle order */
        for (int64_t k = 0; k < n; k++) {
            int d = orc_particle_dir(seed, step, (uint64_t)p[k * F + 7]);
            if (d < 0) ns++;
            else sc[(size_t)r * P + orc_grid_neighbour(gx, gy, gz, r, d)] += (int)F;
        }
        int acc = 0;
        for (int j = 0; j < P; j++) { sd[(size_t)r * P + j] = acc; acc += sc[(size_t)r * P + j]; }
        int *cur = malloc(sizeof(int) * P);
        memcpy(cur, sd + (size_t)r * P, sizeof(int) * P);
        int64_t st = 0;
        for (int64_t k = 0; k < n; k++) {
            const double *q = p + k * F;
            int d = orc_particle_dir(seed, step, (uint64_t)q[7]);
            if (d < 0) { memcpy(o + st * F, q, sizeof(double) * F); st++; continue; }
            int j = orc_grid_neighbour(gx, gy, gz, r, d);
            double *m = sb + cur[j];
            memcpy(m, q, sizeof(double) * F);
            const int ax = d >> 1;
            double c = m[ax] + ((d & 1) ? -1.0 : 1.0);
            if (c >= (double)dims[ax]) c -= (double)dims[ax];
            if (c < 0.0) c += (double)dims[ax];
            m[ax] = c;
            cur[j] += (int)F;
            moved++;
        }
        free(cur);
        stay[r] = ns;
        sbufs[r] = sb;
        rbufs[r] = o;
    }
    /* Alltoall of the counts, then receive displacements after the stayers */
    for (int r = 0; r < P; r++) {
        int64_t acc = stay[r] * (int64_t)F;
        for (int s = 0; s < P; s++) {
            rc[(size_t)r * P + s] = sc[(size_t)s * P + r];
            rd[(size_t)r * P + s] = (int)acc;
            acc += rc[(size_t)r * P + s];
        }
        if (acc > cap * (int64_t)F) status = -1;
        cout[r] = acc / (int64_t)F;
    }
    if (status == 0)
        orc_alltoallv(P, 8, 8, sbufs, sc, sd, rbufs, rc, rd);
    free(sc); free(sd); free(rc); free(rd); free(send); free(sbufs); free(rbufs); free(stay);
    return status ? status : moved;
}

/* ------------------------------------------------------------------------- */
typedef struct { uint64_t p, s; int64_t i; } blk_t;
static int cmp_blk(const void *a, const void *b)
{
    const blk_t *x = a, *y = b;
    return (x->p > y->p) - (x->p < y->p);
}

int64_t orc_heap_check(const uint64_t *ptrs, const uint64_t *sizes, int64_t n,
                       uint64_t base, uint64_t arena)
{
    blk_t *v = malloc(sizeof(blk_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) { v[i].p = ptrs[i]; v[i].s = sizes[i]; v[i].i = i; }
    qsort(v, (size_t)n, sizeof(blk_t), cmp_blk);
    int64_t bad = 0;
    for (int64_t i = 0; i < n && !bad; i++) {
        if (v[i].p < base || v[i].p + v[i].s > base + arena) bad = i + 1;
        else if (v[i].s >= 16 && (v[i].p & 15)) bad = i + 1;
        else if (i > 0 && v[i - 1].p + v[i - 1].s > v[i].p) bad = i + 1;
    }
    free(v);
    return bad;
}
