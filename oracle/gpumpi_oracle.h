/*
 * gpumpi oracle -- CPU restatement of the reference's device-MPI data path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  It is the checker,
 * never the product: the product path lives in dphpc-project_amd/csrc and
 * fails loudly when its HIP library is missing.
 *
 * What each function restates (reference file:line) and what pins it:
 *
 *   orc_philox4x32_10   Philox4x32-10 (Salmon et al., SC'11).  Pinned by the
 *                       Random123 known-answer vectors and by the MIT-licensed
 *                       rocRAND header in /opt/rocm (tests/test_oracle.py).
 *   orc_fill_*          Seeded input generators for the five workloads.  They are
 *                       definitions (the device generators must match them bit
 *                       for bit); pinned through the Philox KATs plus the
 *                       closed-form ranges/moments checked in the tests
 *                       (tests/test_oracle.py).  orc_fill_*_at draw the same
 *                       streams at given indices, equal to the full streams there.
 *   orc_mcpi_hits       Monte-Carlo pi hit count per rank.  Pinned by brute
 *                       force on tiny n and by the pi estimate's error bound.
 *   orc_alltoallv       MPI_Alltoallv semantics, gpu_mpi/mpi.cu:363-408 (counts
 *                       and displacements in elements).  Pinned by an independent
 *                       Python block permutation (tests/test_oracle.py) and by host
 *                       MPICH's MPI_Alltoallv byte for byte (oracle/mpi/a2av_check.c,
 *                       tests/test_oracle_pins.py).
 *   orc_counts_transpose  The Alltoall of the send counts and the packed receive
 *                       displacements (is.c:758-781).  Pinned by a transpose and an
 *                       exclusive prefix sum (tests/test_oracle_pins.py).
 *   orc_reduce          Reference fold order of MPI_Reduce, gpu_mpi/mpi.cu:209-225
 *                       (recv = buf_0; recv = op(buf_r, recv) for r = 1..P-1),
 *                       operator bodies gpu_mpi/operators.cu:73-151 and the
 *                       type/op legality table operators.cu:153-200.  Pinned by
 *                       the reference test known answers (tests/collectives_test.cu
 *                       :6-98), a cancellation case that fixes the left-to-right
 *                       fold, numpy for integer and bitwise operators, and host
 *                       MPICH's MPI_Reduce/MPI_Allreduce: integers and FP64 MAX
 *                       equal, FP64 SUM within 1e-12 * sum |x| per element
 *                       (oracle/mpi/reduce_check.c, tests/test_oracle_pins.py).
 *   orc_comm_split      MPI_Comm_split ordering by (key, old rank),
 *                       gpu_mpi/communicator.cu:154-213.  Pinned by the
 *                       reference table tests/communicator_test.cu:93-152.
 *   orc_npb_*           NPB IS 3.4.1 (toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/
 *                       IS/is.c): randlc 406-452 restated as the exact integer
 *                       LCG x <- a*x mod 2^46, find_my_seed 470-510,
 *                       create_seq 520-537, rank() 631-928, full_verify
 *                       545-621.  Pinned by NPB's own partial-verification
 *                       known answers (is.c:264-287, adjustments 849-905):
 *                       passed == 5*MAX_ITERATIONS + P for classes S, W, A, and
 *                       orc_npb_is_iteration passing all five test keys in each of
 *                       iterations 1..10 run in order on one key array.
 *   orc_bucket_*        The NPB-style bucket exchange applied to the workload's
 *                       uniform / Zipf keys (same rank() steps).  Pinned by the
 *                       global order equalling a plain sort of all keys.
 *   orc_local_sort      The counting rank and placement of rank() / full_verify
 *                       (is.c:807-833, 545-560) as an ascending sort.  Pinned by
 *                       numpy.sort (tests/test_oracle_pins.py).
 *   orc_particles_*     Particle exchange on a periodic gx*gy*gz rank grid: every
 *                       rank packs its migrants by destination rank and the step
 *                       is one MPI_Alltoallv (orc_alltoallv above), stayers first.
 *                       Pinned by conservation (ids are a permutation), every
 *                       particle inside its owner's cell, a migrant moves exactly
 *                       one cell along one axis, the migration rate against its
 *                       binomial bounds, and host MPICH running the same program
 *                       (oracle/mpi/particles.c) bit for bit.
 */
#ifndef GPUMPI_ORACLE_H
#define GPUMPI_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stream identifiers folded into Philox counter word 3 so that the
 * generators of different workloads never share a counter. */
enum {
    ORC_STREAM_BYTES  = 0x0B17E500u,
    ORC_STREAM_INT    = 0x0B17E501u,
    ORC_STREAM_NORMAL = 0x0B17E502u,
    ORC_STREAM_KEYS   = 0x0B17E503u,
    ORC_STREAM_ZIPF   = 0x0B17E504u,
    ORC_STREAM_MIGR   = 0x0B17E505u,
    ORC_STREAM_JOBS   = 0x0B17E506u,
    ORC_STREAM_PART   = 0x0B17E507u
};

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- generators (rank key = {seed_lo, seed_hi ^ rank}) ---- */
int64_t orc_mcpi_hits(uint64_t seed, int rank, int64_t n);
void orc_fill_bytes(uint64_t seed, int rank, uint32_t iter, uint8_t *buf, int64_t nbytes);
void orc_fill_int(uint64_t seed, int rank, uint32_t iter, int32_t *out, int64_t n);
void orc_fill_normal(uint64_t seed, int rank, uint32_t iter, double *out, int64_t n);
/* The generator's ln, sqrt and cos(2 pi u), exported for their pins. */
double orc_det_log(double x);
double orc_det_sqrt(double x);
double orc_det_cos_2pi(double u);
/* The same streams at the element indices idx[0..n). */
void orc_fill_int_at(uint64_t seed, int rank, uint32_t iter, const int64_t *idx, int64_t n,
                     int32_t *out);
void orc_fill_normal_at(uint64_t seed, int rank, uint32_t iter, const int64_t *idx, int64_t n,
                        double *out);
void orc_fill_keys_uniform(uint64_t seed, int rank, uint32_t step, uint32_t *out,
                           int64_t n, int key_bits);
void orc_fill_keys_zipf(uint64_t seed, int rank, uint32_t step, const double *cdf,
                        int64_t ncdf, uint32_t *out, int64_t n);

/* ---- MPI semantics ---- */
int  orc_type_size(int datatype);
int  orc_op_allowed(int datatype, int op);
/* Alltoallv over P ranks; per-rank arrays are row-major [P][P]. */
void orc_alltoallv(int P, int send_elem, int recv_elem,
                   const void *const *sendbufs, const int *sendcounts, const int *sdispls,
                   void *const *recvbufs, const int *recvcounts, const int *rdispls);
/* Reduction in reference fold order; returns 0, or -1 for an illegal type/op. */
int  orc_reduce(int P, int datatype, int op, const void *const *inputs, int64_t count,
                void *out);
void orc_comm_split(int P, const int *color, const int *key, int *new_rank, int *new_size);

/* ---- NPB IS ---- */
typedef struct orc_npb_class {
    int total_keys_log2, max_key_log2, num_buckets_log2;
    char cls;
} orc_npb_class;
int  orc_npb_class_params(char cls, orc_npb_class *out);
/* Keys of one rank after create_seq (num_keys = 2^total / P entries). */
void orc_npb_create_keys(const orc_npb_class *c, int P, int rank, int32_t *keys);
/* Run the whole benchmark (rank(1) warm-up, 10 iterations, full verify) on
 * P simulated ranks; returns passed_verification as NPB reports it. */
int64_t orc_npb_is_run(const orc_npb_class *c, int P);
/* Partial-verification count of one rank() iteration of the NPB IS caller given
 * the bucket sizes totals and the ranked key buffers -- used to check the
 * device bucket exchange against the class known answers.  keys is P*num_keys. */
int64_t orc_npb_is_iteration(const orc_npb_class *c, int P, int iteration, int32_t *keys);

/* ---- NPB-style bucket exchange of the workload keys ---- */
/* Given P ranks x n keys (rank-major), compute per rank the stable bucketed send
 * buffer, send counts [P][P] and displacements, NPB splitters (is.c:721-751).
 * Buffers: key_buff1 (P*n), send_count/send_displ (P*P). */
void orc_bucket_plan(int P, int64_t n, int key_bits, int bucket_bits, const uint32_t *keys,
                     uint32_t *key_buff1, int *send_count, int *send_displ,
                     int *bucket_lo, int *bucket_hi);
/* recv_count/recv_displ [P][P] from send_count (the Alltoall of counts). */
void orc_counts_transpose(int P, const int *send_count, int *recv_count, int *recv_displ);
/* Counting sort of each rank's received keys (rank() 807-833 then full_verify's
 * placement); out gets the keys in ascending order per rank. */
void orc_local_sort(int64_t nkeys, const uint32_t *keys, uint32_t *out);

/* ---- particle exchange (8 FP64 fields per particle: x y z vx vy vz w id) ---- */
#define ORC_PART_FIELDS 8
/* Probability of a move per particle and step is ORC_PART_MOVE / 2^32 (~0.05). */
#define ORC_PART_MOVE 214748365u
/* Destination direction of particle `id` at `step`: -1 stays, 0..5 = +x -x +y -y +z -z. */
int  orc_particle_dir(uint64_t seed, uint32_t step, uint64_t id);
int  orc_grid_neighbour(int gx, int gy, int gz, int rank, int dir);
/* Initial particles of one rank: n0 particles into parts[0 .. n0*8). */
void orc_particles_init_rank(int gx, int gy, int gz, int rank, int64_t n0, uint64_t seed,
                             double *parts);
/* Initial state: n0 particles per rank in buffer [P][cap][8]; counts[P] = n0. */
void orc_particles_init(int gx, int gy, int gz, int64_t n0, uint64_t seed, int64_t cap,
                        double *parts, int64_t *counts);
/* One exchange step from (in, cin) to (out, cout).  Returns the number of migrants,
 * or -1 if a rank would exceed cap. */
int64_t orc_particles_step(int gx, int gy, int gz, uint64_t seed, uint32_t step, int64_t cap,
                           const double *in, const int64_t *cin, double *out, int64_t *cout);

/* ---- heap invariant checker ---- */
/* Returns 0 if the live blocks [ptr, ptr+size) are pairwise disjoint, lie in
 * [base, base+arena) and are 16-byte aligned for size >= 16; else the index+1 of
 * the first offending block (after sorting by address). */
int64_t orc_heap_check(const uint64_t *ptrs, const uint64_t *sizes, int64_t n,
                       uint64_t base, uint64_t arena);

#ifdef __cplusplus
}
#endif
#endif
