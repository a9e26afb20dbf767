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

// Element i of the seeded "N(0,1)" stream: Irwin-Hall(12) - 6 in exact integer arithmetic.
__host__ __device__ __forceinline__ double gen_normal(uint64_t seed, int rank, uint32_t it, int64_t i) {
    uint64_t s = 0;
    for (int j = 0; j < 3; j++) {
        u32x4v v = rank_stream(seed, rank, it, (uint64_t)(3 * i + j), kStreamNormal);
        s += (uint64_t)v.x + v.y + v.z + v.w;
    }
    return (double)((int64_t)s - (int64_t)6 * 4294967296LL) * (1.0 / 4294967296.0);
}

// Byte o of the seeded payload stream.
__host__ __device__ __forceinline__ uint8_t gen_byte(uint64_t seed, int rank, uint32_t it, uint64_t o) {
    u32x4v v = rank_stream(seed, rank, it, o >> 4, kStreamBytes);
    return (uint8_t)(pick(v, (int)((o >> 2) & 3)) >> (8 * (o & 3)));
}

}  // namespace gpumpi
