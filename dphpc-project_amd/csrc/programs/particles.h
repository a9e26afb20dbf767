// Particle exchange on a periodic gx*gy*gz rank grid: the particle variant of the batched
// uneven exchange workload.  A particle is 8 FP64 fields {x, y, z, vx, vy, vz, w, id}
// (64 bytes); rank r owns the unit cell (r % gx, (r / gx) % gy, r / (gx * gy)).  Each step
// a particle moves one cell along one of the six axis directions with probability
// kPartMove / 2^32 (~0.05), drawn from Philox keyed by (seed, id, step), so the decision
// does not depend on where the particle sits in memory.  Positions are multiples of 2^-20,
// so the +-1 move and the periodic wrap are exact in FP64.
//
// Per step each rank splits its particles stably: stayers go to the front of its next
// buffer, migrants are packed by destination rank (ascending, original order within a
// destination) into its send area.  The exchange is then either
//   mode 0: MPI_Alltoall of the counts + MPI_Alltoallv of the particles (MPI_DOUBLE), or
//   mode 1: Isend/Irecv of counts and particles with the (at most six) neighbours,
// and arrivals land after the stayers in ascending source-rank order -- the receive
// layout of MPI_Alltoallv (gpu_mpi/mpi.cu:363-408), so both modes give the same state.
#pragma once

#include "workloads.h"
#include "../runtime/particle_step.h"

namespace gpumpi {
namespace progs {

struct ParticleArgs {
    uint64_t seed;
    int32_t gx, gy, gz;
    int32_t mode;        // 0 collective, 1 neighbour p2p
    int32_t steps;
    int32_t first_step;  // global index of the first step of this launch
    int64_t n0;          // > 0: initialise buf[0] with n0 particles per rank
    int64_t cap;         // particles per rank region of buf[0], buf[1] and sendbuf
    double* buf[2];      // [P][cap][8]; the launch reads buf[0] and ends in buf[steps & 1]
    double* sendbuf;     // [P][cap][8]
    int64_t* counts;     // [2][P] particles per rank in buf[0] / buf[1]
    int32_t* cnt;        // [4][P][P] send counts, send displs, recv counts, recv displs
    int64_t* moved;      // [P] migrants sent by the rank in this launch
    double* seconds;     // [P]
};

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// One particle in registers: four 16-byte vectors, read as eight doubles.
union PartReg {
    u32x4 v[4];
    double f[kPartFields];
};

__device__ __forceinline__ void particle_load(const double* p, PartReg& r) {
#pragma unroll
    for (int i = 0; i < 4; i++) r.v[i] = ((const u32x4*)p)[i];
}
__device__ __forceinline__ void particle_store(double* p, const PartReg& r) {
#pragma unroll
    for (int i = 0; i < 4; i++) ((u32x4*)p)[i] = r.v[i];
}

__device__ void particles_init_rank(const ParticleArgs& a, int rank, double* p) {
    const int cx = rank % a.gx, cy = (rank / a.gx) % a.gy, cz = rank / (a.gx * a.gy);
    const double q20 = 1.0 / 1048576.0, q15 = 1.0 / 32768.0;
    for (int64_t k = lane(); k < a.n0; k += kWave) {
        u32x4v o = rank_stream(a.seed, rank, 0u, (uint64_t)k, kStreamParticles);
        double* q = p + k * kPartFields;
        q[0] = (double)cx + (double)(o.x >> 12) * q20;
        q[1] = (double)cy + (double)(o.y >> 12) * q20;
        q[2] = (double)cz + (double)(o.z >> 12) * q20;
        q[3] = (double)((int32_t)(o.w & 0xFFFFu) - 32768) * q15;
        q[4] = (double)((int32_t)(o.w >> 16) - 32768) * q15;
        q[5] = (double)(k & 1023) * (1.0 / 1024.0);
        q[6] = 1.0;
        q[7] = (double)((int64_t)rank * a.n0 + k);
    }
}

__device__ int w_particles(int, char**) {
    const ParticleArgs& a = args<ParticleArgs>();
    MPI_Init(nullptr, nullptr);
    int rank, P;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    if (a.gx * a.gy * a.gz != P || a.n0 > a.cap || a.cap * kPartFields > 0x7FFFFFFF) {
        MPI_Abort(MPI_COMM_WORLD, 2);  // grid does not match the rank count, or capacity
        return 1;
    }
    const int64_t cap8 = a.cap * kPartFields;

    // distinct neighbours in ascending rank order (a grid axis of size 1 or 2 maps several
    // directions onto one rank, possibly this one)
    int U[6];
    int nU = 0;
    for (int d = 0; d < 6; d++) {
        int nb = grid_neighbour(rank, d, a.gx, a.gy, a.gz);
        bool seen = false;
        for (int c = 0; c < nU; c++) seen = seen || U[c] == nb;
        if (seen) continue;
        int c = nU++;
        while (c > 0 && U[c - 1] > nb) { U[c] = U[c - 1]; c--; }
        U[c] = nb;
    }
    int nbr[6];
    for (int d = 0; d < 6; d++) nbr[d] = grid_neighbour(rank, d, a.gx, a.gy, a.gz);

    if (a.n0 > 0) {
        particles_init_rank(a, rank, a.buf[0] + (uint64_t)rank * cap8);
        a.counts[rank] = a.n0;
    }
    int32_t* sc = a.cnt + (int64_t)rank * P;
    int32_t* sd = sc + (int64_t)P * P;
    int32_t* rc = sd + (int64_t)P * P;
    int32_t* rd = rc + (int64_t)P * P;
    int64_t moved = 0;
    MPI_Barrier(MPI_COMM_WORLD);
    const double t0 = MPI_Wtime();
    for (int s = 0; s < a.steps; s++) {
        const int cur = s & 1;
        const uint32_t step = (uint32_t)(a.first_step + s);
        const double* in = a.buf[cur] + (uint64_t)rank * cap8;
        double* out = a.buf[cur ^ 1] + (uint64_t)rank * cap8;
        double* sb = a.sendbuf + (uint64_t)rank * cap8;
        const int64_t n = a.counts[(int64_t)cur * P + rank];

        // pass 1: particles per destination group
        int64_t gcount[6] = {0, 0, 0, 0, 0, 0};
        int64_t nstay = 0;
        for (int64_t base = 0; base < n; base += kWave) {
            const int64_t k = base + lane();
            const int d = k < n ? particle_dir(a.seed, step, (uint64_t)in[k * kPartFields + 7]) : -2;
            int dst = -1;
#pragma unroll
            for (int dd = 0; dd < 6; dd++) dst = d == dd ? nbr[dd] : dst;
            nstay += __popcll(ballot(d == -1));
#pragma unroll
            for (int c = 0; c < 6; c++)
                if (c < nU) gcount[c] += __popcll(ballot(d >= 0 && dst == U[c]));
        }
        int64_t goff[6], gcur[6], acc = 0;
#pragma unroll
        for (int c = 0; c < 6; c++) { goff[c] = acc; gcur[c] = acc; acc += c < nU ? gcount[c] : 0; }
        const int64_t nmove = acc;

        // pass 2: stable split, migrants moved one cell with the periodic wrap
        int64_t st = 0;
        for (int64_t base = 0; base < n; base += kWave) {
            const int64_t k = base + lane();
            const bool valid = k < n;
            PartReg v;
            if (valid) particle_load(in + k * kPartFields, v);
            const int d = valid ? particle_dir(a.seed, step, (uint64_t)v.f[7]) : -2;
            int dst = -1;
#pragma unroll
            for (int dd = 0; dd < 6; dd++) dst = d == dd ? nbr[dd] : dst;
            const uint64_t ms = ballot(d == -1);
            int64_t pos = st + lanes_below(ms);
            st += __popcll(ms);
            double* target = out;
#pragma unroll
            for (int c = 0; c < 6; c++) {
                if (c < nU) {
                    const bool mine = d >= 0 && dst == U[c];
                    const uint64_t m = ballot(mine);
                    if (mine) { pos = gcur[c] + lanes_below(m); target = sb; }
                    gcur[c] += __popcll(m);
                }
            }
            if (d >= 0) {
                const int ax = d >> 1;
                const double g = (double)(ax == 0 ? a.gx : ax == 1 ? a.gy : a.gz);
                double c = (ax == 0 ? v.f[0] : ax == 1 ? v.f[1] : v.f[2]) + ((d & 1) ? -1.0 : 1.0);
                if (c >= g) c -= g;
                if (c < 0.0) c += g;
                if (ax == 0) v.f[0] = c;
                else if (ax == 1) v.f[1] = c;
                else v.f[2] = c;
            }
            if (valid) particle_store(target + pos * kPartFields, v);
        }
        moved += nmove;

        // exchange: counts in doubles (8 per particle), arrivals after the stayers
        int64_t total = nstay;
        if (a.mode == 0) {
            for (int j = lane(); j < P; j += kWave) {
                int32_t c_j = 0, d_j = 0;
#pragma unroll
                for (int c = 0; c < 6; c++)
                    if (c < nU && U[c] == j) { c_j = (int32_t)(gcount[c] * kPartFields); d_j = (int32_t)(goff[c] * kPartFields); }
                sc[j] = c_j;
                sd[j] = d_j;
            }
            if (MPI_Alltoall(sc, 1, MPI_INT, rc, 1, MPI_INT, MPI_COMM_WORLD) != MPI_SUCCESS) break;
            int64_t carry = nstay * kPartFields;
            for (int j0 = 0; j0 < P; j0 += kWave) {
                const int j = j0 + lane();
                const int64_t c = j < P ? rc[j] : 0;
                const int64_t inc = wave_incl_scan(c);
                if (j < P) rd[j] = (int32_t)(carry + inc - c);
                carry += __shfl(inc, 63);
            }
            total = carry / kPartFields;
            if (total > a.cap) {
                MPI_Abort(MPI_COMM_WORLD, 3);
                return 1;
            }
            if (MPI_Alltoallv(sb, sc, sd, MPI_DOUBLE, out, rc, rd, MPI_DOUBLE, MPI_COMM_WORLD) != MPI_SUCCESS)
                break;
        } else {
            MPI_Request req[12];
            for (int c = 0; c < nU; c++) {
                sc[U[c]] = (int32_t)gcount[c];
                MPI_Irecv(&rc[U[c]], 1, MPI_INT, U[c], 7, MPI_COMM_WORLD, &req[2 * c]);
                MPI_Isend(&sc[U[c]], 1, MPI_INT, U[c], 7, MPI_COMM_WORLD, &req[2 * c + 1]);
            }
            if (MPI_Waitall(2 * nU, req, MPI_STATUSES_IGNORE) != MPI_SUCCESS) break;
            int64_t rpos[6];
            for (int c = 0; c < nU; c++) {
                rpos[c] = total;
                total += *(volatile int32_t*)&rc[U[c]];
            }
            if (total > a.cap) {
                MPI_Abort(MPI_COMM_WORLD, 3);
                return 1;
            }
            for (int c = 0; c < nU; c++) {
                const int rn = *(volatile int32_t*)&rc[U[c]];
                MPI_Irecv(out + rpos[c] * kPartFields, rn * kPartFields, MPI_DOUBLE, U[c], 8,
                          MPI_COMM_WORLD, &req[2 * c]);
                MPI_Isend(sb + goff[c] * kPartFields, (int)(gcount[c] * kPartFields), MPI_DOUBLE,
                          U[c], 8, MPI_COMM_WORLD, &req[2 * c + 1]);
            }
            if (MPI_Waitall(2 * nU, req, MPI_STATUSES_IGNORE) != MPI_SUCCESS) break;
        }
        a.counts[(int64_t)(cur ^ 1) * P + rank] = total;
    }
    const double t1 = MPI_Wtime();
    MPI_Barrier(MPI_COMM_WORLD);
    a.moved[rank] = moved;
    a.seconds[rank] = t1 - t0;
    MPI_Finalize();
    return 0;
}

}  // namespace progs
}  // namespace gpumpi
