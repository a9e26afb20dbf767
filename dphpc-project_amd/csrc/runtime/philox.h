// Counter-based random streams for the workload programs (device side).
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2,
// 3", SC'11): 10 rounds of two 32x32->64 multiplies with the published multipliers, key
// bumped by the published Weyl constants between rounds.  Stream layout shared with the
// host generators: key = {seed_lo, seed_hi ^ rank}, counter = {iteration, block_lo,
// block_hi, stream id}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpumpi {

enum : uint32_t {
    kStreamBytes = 0x0B17E500u,
    kStreamInt = 0x0B17E501u,
    kStreamNormal = 0x0B17E502u,
    kStreamKeys = 0x0B17E503u,
    kStreamZipf = 0x0B17E504u,
    kStreamMigrate = 0x0B17E505u,
    kStreamJobs = 0x0B17E506u,
};

struct u32x4v { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ u32x4v philox4x32_10(u32x4v c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        u32x4v n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
    }
    return c;
}

__host__ __device__ __forceinline__ u32x4v rank_stream(uint64_t seed, int rank, uint32_t w0,
                                                       uint64_t block, uint32_t stream) {
    u32x4v c{w0, (uint32_t)block, (uint32_t)(block >> 32), stream};
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32) ^ (uint32_t)rank);
}

__host__ __device__ __forceinline__ double unit53(uint32_t hi, uint32_t lo) {
    return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

__host__ __device__ __forceinline__ uint32_t pick(const u32x4v& v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// Element i of the seeded int32 stream: uniform in [-2^20, 2^20].
__host__ __device__ __forceinline__ int32_t gen_int(uint64_t seed, int rank, uint32_t it, int64_t i) {
    u32x4v v = rank_stream(seed, rank, it, (uint64_t)(i >> 2), kStreamInt);
    return (int32_t)(pick(v, (int)(i & 3)) % 2097153u) - 1048576;
}

// ln, sqrt and cos(2 pi u) from +, -, *, / and exact bit manipulation only, evaluated in
// the same order as the oracle's C restatement, so device and host draw the same bits
// (both are built with -ffp-contract=off; IEEE double +, -, *, / round the same way on
// gfx950 and x86-64).  Accuracy is a few ulps, which the N(0,1) inputs need; matching bits
// are what the bit-exact reduction tests need.
constexpr double kDetLn2Hi = 0.6931471805599453;        // ln 2 rounded to double
constexpr double kDetLn2Lo = 2.3190468138462996e-17;    // ln 2 - kDetLn2Hi
constexpr double kDetHalfPi = 1.5707963267948966;
constexpr double kDetSqrt2 = 1.4142135623730951;

__host__ __device__ __forceinline__ double det_bits_to_double(uint64_t b) {
    double d;
    __builtin_memcpy(&d, &b, 8);
    return d;
}
__host__ __device__ __forceinline__ uint64_t det_double_to_bits(double d) {
    uint64_t b;
    __builtin_memcpy(&b, &d, 8);
    return b;
}

// ln x for a normal x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)) (exact), and
// ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^23/23), s = (m - 1) / (m + 1), |s| <= 0.1716.
__host__ __device__ __forceinline__ double det_log(double x) {
    uint64_t b = det_double_to_bits(x);
    int e = (int)((b >> 52) & 0x7FF) - 1023;
    double m = det_bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > kDetSqrt2) {
        m = m * 0.5;
        e = e + 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    const double lnm = (2.0 * s) * p;
    const double de = (double)e;
    return de * kDetLn2Hi + (de * kDetLn2Lo + lnm);
}

// sqrt x for x >= 0 (finite): Newton's iteration y <- (y + x / y) / 2 from 2^floor(e/2),
// which is within a factor of 2 of the root; seven steps reach full precision.
__host__ __device__ __forceinline__ double det_sqrt(double x) {
    if (x == 0.0) return 0.0;
    const int e = (int)((det_double_to_bits(x) >> 52) & 0x7FF) - 1023;
    const int h = e >= 0 ? e / 2 : -((1 - e) / 2);
    double y = det_bits_to_double((uint64_t)(h + 1023) << 52);
    for (int k = 0; k < 7; k++) y = 0.5 * (y + x / y);
    return y;
}

// cos(2 pi u) for u in [0, 1): quadrant q = floor(4u) and r = 4u - q are exact; with
// t = r pi/2 in [0, pi/2), cos and sin of t by their Taylor series to t^22 and t^23
// (Horner in t^2; the last terms are below 3e-17), then the quadrant's sign and swap.
__host__ __device__ __forceinline__ double det_cos_2pi(double u) {
    const double f = 4.0 * u;
    const int q = (int)f;
    const double t = (f - (double)q) * kDetHalfPi;
    const double t2 = t * t;
    double c = 1.0;
    c = 1.0 - t2 * (1.0 / 462.0) * c;   // 21 * 22
    c = 1.0 - t2 * (1.0 / 380.0) * c;   // 19 * 20
    c = 1.0 - t2 * (1.0 / 306.0) * c;
    c = 1.0 - t2 * (1.0 / 240.0) * c;
    c = 1.0 - t2 * (1.0 / 182.0) * c;
    c = 1.0 - t2 * (1.0 / 132.0) * c;
    c = 1.0 - t2 * (1.0 / 90.0) * c;
    c = 1.0 - t2 * (1.0 / 56.0) * c;
    c = 1.0 - t2 * (1.0 / 30.0) * c;
    c = 1.0 - t2 * (1.0 / 12.0) * c;
    c = 1.0 - t2 * (1.0 / 2.0) * c;     // cos t
    double sn = 1.0;
    sn = 1.0 - t2 * (1.0 / 506.0) * sn;  // 22 * 23
    sn = 1.0 - t2 * (1.0 / 420.0) * sn;
    sn = 1.0 - t2 * (1.0 / 342.0) * sn;
    sn = 1.0 - t2 * (1.0 / 272.0) * sn;
    sn = 1.0 - t2 * (1.0 / 210.0) * sn;
    sn = 1.0 - t2 * (1.0 / 156.0) * sn;
    sn = 1.0 - t2 * (1.0 / 110.0) * sn;
    sn = 1.0 - t2 * (1.0 / 72.0) * sn;
    sn = 1.0 - t2 * (1.0 / 42.0) * sn;
    sn = 1.0 - t2 * (1.0 / 20.0) * sn;
    sn = 1.0 - t2 * (1.0 / 6.0) * sn;
    sn = t * sn;                          // sin t
    return q == 0 ? c : q == 1 ? -sn : q == 2 ? -c : sn;
}

// Element i of the seeded N(0,1) stream: Box-Muller on Philox block i,
// z = sqrt(-2 ln u1) cos(2 pi u2) with u1 in (0, 1] from words (x, y) and u2 in [0, 1) from
// words (z, w), 53 bits each.  |z| <= sqrt(-2 ln 2^-53) = 8.57.
__host__ __device__ __forceinline__ double gen_normal(uint64_t seed, int rank, uint32_t it, int64_t i) {
    const u32x4v v = rank_stream(seed, rank, it, (uint64_t)i, kStreamNormal);
    const double u1 = (double)(((((uint64_t)v.x << 32) | v.y) >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = (double)((((uint64_t)v.z << 32) | v.w) >> 11) * (1.0 / 9007199254740992.0);
    return det_sqrt(-2.0 * det_log(u1)) * det_cos_2pi(u2);
}

// Byte o of the seeded payload stream.
__host__ __device__ __forceinline__ uint8_t gen_byte(uint64_t seed, int rank, uint32_t it, uint64_t o) {
    u32x4v v = rank_stream(seed, rank, it, o >> 4, kStreamBytes);
    return (uint8_t)(pick(v, (int)((o >> 2) & 3)) >> (8 * (o & 3)));
}

}  // namespace gpumpi
