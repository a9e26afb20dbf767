/*
 * gpumpi host C ABI (libgpumpi.so) -- launches rank grids on one MI355X.
 *
 * Plain C types only (pointers, sizes, ints); device buffers are passed as raw device
 * addresses and streams as hipStream_t cast to void*.  Every entry point returns 0 on
 * success, a positive program exit status, or a negative gpumpi_status code.
 *
 * Reference interfaces replaced:
 *   gpumpi_main ............ gpu_libs/gpu_main/main.cu:134-322 (the launcher's main():
 *                            "<prog args> ---gpumpi -n -g -b -s -p -e", exit code of the
 *                            rank program, main.cu:49-98, 127-130, 321)
 *   gpumpi_config .......... the ---gpumpi options (main.cu:71-79) plus the hard-coded
 *                            recvListSize (main.cu:219) and pendingBufferSize (-e)
 *   gpumpi_runtime_create /
 *   gpumpi_run ............. SharedState::Holder + the single cudaLaunchKernel of
 *                            __gpu_main_caller (main.cu:217-289), split so one runtime
 *                            serves many launches (the plain per-call path)
 *   gpumpi_last_error ...... CudaMPI::recordError / printLastError (gpu_libc/assert.cu:24-71)
 *   gpumpi_a2av_*  ......... no reference counterpart: the batched persistent-kernel
 *                            MPI_Alltoallv step (gpu_mpi/mpi.cu:363-408 semantics)
 *   gpumpi_p2p_* ........... the matched MPI_Isend/MPI_Irecv pairs of one step
 *                            (gpu_libc/cuda_mpi.cu:65-118, gpu_mpi/mpi.cu:617-645) as one
 *                            batched, graph-replayable step
 *   gpumpi_particles_* ..... no reference counterpart: the batched particle step (the
 *                            particle variant of the uneven exchange)
 *   gpumpi_serve_*, gpumpi_run_job ... no reference counterpart: the device-resident job
 *                            server and the one-launch-per-job plain path
 *
 * A reference-side maintainer binds these with ctypes/cffi (dphpc-project_amd/__init__.py
 * is such a binding) or links them directly (-I include -lgpumpi).
 */
#ifndef GPUMPI_GPUMPI_H
#define GPUMPI_GPUMPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gpumpi_status {
    GPUMPI_OK = 0,
    GPUMPI_E_ARG = -1,          /* invalid argument or configuration */
    GPUMPI_E_HIP = -2,          /* a HIP runtime call failed */
    GPUMPI_E_RESIDENCY = -3,    /* the rank grid cannot be co-resident on the GPU */
    GPUMPI_E_PROGRAM = -4,      /* unknown rank program */
    GPUMPI_E_DEVICE = -5,       /* the device runtime recorded an error (see last_error) */
    GPUMPI_E_NOMEM = -6         /* device memory allocation failed */
};

typedef struct gpumpi_config {
    int32_t num_ranks;          /* -n: MPI ranks (one wavefront each) */
    int32_t ranks_per_block;    /* wavefronts (ranks) per workgroup, 1..4 (-b / 64); 0 = 1 */
    int32_t ring_slots;         /* envelope ring slots per rank; 0 = 256 */
    int32_t max_requests;       /* -e: request pool per rank; 0 = 1024 */
    int32_t max_unexpected;     /* unexpected-message list per rank; 0 = 1024 */
    int32_t bcast_push_min;     /* MPI_Bcast of at least this many bytes per rank (buffer
                                   bytes / ranks) runs as a push: every rank stores one
                                   slice of the root's buffer into all receivers; smaller
                                   ones are pulled by each receiver.  0 = 512 with at
                                   least 512 KiB in all (the measured crossover on 64 and
                                   1024 ranks), < 0 never */
    int32_t copy_stage_kib;     /* LDS staging area per rank (wavefront) for its copies of
                                   16 KiB and more: 0 = the largest of 64, 32 and 16 KiB
                                   with which the grid keeps its residency headroom and the
                                   workgroup fits the device's LDS limit (none if no size
                                   does), 1..64 = this many KiB if it fits, < 0 never.  The
                                   launch gives it as dynamic LDS. */
    int32_t reserved1;          /* must be 0 */
    uint64_t heap_bytes;        /* -p: device heap arena; 0 = 1 GiB */
    uint64_t stack_bytes;       /* -s: per-lane stack limit; 0 = runtime default */
    double timeout_s;           /* bound on every device-side wait; 0 = 20 s */
} gpumpi_config;

typedef struct gpumpi_runtime gpumpi_runtime;

void gpumpi_config_default(gpumpi_config* cfg);

/* Create / destroy a rank-grid runtime (mailboxes, request pools, heap, COMM_WORLD). */
int gpumpi_runtime_create(const gpumpi_config* cfg, gpumpi_runtime** out);
void gpumpi_runtime_destroy(gpumpi_runtime* rt);

/* Launch the built-in rank program `program` on every rank and wait for it.
 * args/args_bytes is copied to device memory and visible to ranks as the program's
 * argument block.  Returns 0, 1 if any rank returned non-zero (main.cu:127-130), or a
 * negative gpumpi_status.  *elapsed_ms (optional) receives the kernel's wall time. */
int gpumpi_run(gpumpi_runtime* rt, const char* program, int argc, const char* const* argv,
               const void* args, size_t args_bytes, void* stream, double* elapsed_ms);

/* Launch without waiting (the caller synchronises the stream). */
int gpumpi_run_async(gpumpi_runtime* rt, const char* program, const void* args,
                     size_t args_bytes, void* stream);

/* Bytes of LDS staging per wavefront that the last launch on `rt` gave its ranks
 * (gpumpi_config.copy_stage_kib), 0 if none or nothing was launched yet. */
int gpumpi_runtime_copy_stage(const gpumpi_runtime* rt);

/* Launch a rank-grid kernel compiled into the caller's own module (a converted program,
 * see gpumpi/gpumpi_main.h) with the same semantics as gpumpi_run. */
int gpumpi_run_kernel(gpumpi_runtime* rt, const void* kernel, int argc, const char* const* argv,
                      const void* args, size_t args_bytes, void* stream, double* elapsed_ms);
/* CLI of a compiled converted program: argv = "<exe> <args> ---gpumpi -n N ...". */
int gpumpi_main_kernel(const void* kernel, int argc, const char* const* argv);
/* Layout stamp of the runtime structures this library was built with (no reference
 * counterpart): gpumpi_main.h compares it with the stamp of the headers a program was
 * compiled against and refuses to launch on a mismatch (a stale build). */
uint64_t gpumpi_layout_stamp(void);

/* One-shot: create, run, destroy (what main.cu does for one converted program). */
int gpumpi_launch(const char* program, int argc, const char* const* argv, const void* args,
                  size_t args_bytes, const gpumpi_config* cfg);

/* CLI form: argv = "<program> <program args> ---gpumpi -n N [-g G] [-b B] [-s S] [-p P]
 * [-e E] [-t seconds]".  Returns the program's exit code as main.cu:321 does. */
int gpumpi_main(int argc, const char* const* argv);

/* Text of the last error (device error record or host failure); "" when none. */
const char* gpumpi_last_error(gpumpi_runtime* rt);
/* Device error record of the last run: code, rank (negative values when none). */
int gpumpi_last_device_error(gpumpi_runtime* rt, int32_t* code, int32_t* rank);

/* Number of built-in rank programs and their names. */
int gpumpi_program_count(void);
const char* gpumpi_program_name(int i);

/* Heap statistics of a runtime: bytes of live blocks, peak, capacity. */
int gpumpi_heap_stats(gpumpi_runtime* rt, int64_t* in_use, int64_t* peak, int64_t* capacity);

/* Device properties used by the launcher (CU count, wall-clock rate in kHz). */
int gpumpi_device_info(int32_t* num_cus, int32_t* wallclock_khz, int32_t* max_waves_per_cu);

/* ------------------------------------------------------------------------------------
 * Batched Alltoallv step (plan + scan + copy kernels, hipGraph replay, autotuned shape).
 *
 * Describes one MPI_Alltoallv over P ranks whose buffers are already resident in HBM:
 *   sendbufs[P], recvbufs[P]        device arrays of device pointers
 *   send_counts[P*P]                device int32, [s][d] = elements rank s sends to d
 *   send_displs[P*P]                device int32, [s][d] = element displacement in s's sendbuf
 * The step performs the Alltoall of the counts (recv_counts[d][s] = send_counts[s][d]),
 * the receive displacements (exclusive prefix over s, written to recv_displs[d][s]) and
 * the payload exchange: three dependent kernels (plan, scan, copy) that graph_replay
 * captures once and replays as one hipGraph launch.  recv_counts/recv_displs are device
 * int32 [P*P] outputs.  elem_bytes is the datatype size.
 * ------------------------------------------------------------------------------------ */
typedef struct gpumpi_a2av_plan gpumpi_a2av_plan;

typedef struct gpumpi_a2av_tune {
    int32_t block_threads;      /* copy workgroup size: 256 (0 = default) */
    int32_t blocks_per_cu;      /* copy workgroups per CU at most; 0 = one per chunk,
                                   -1 = default */
    int32_t lines_per_lane;     /* 16-byte lines per thread and round: 2, 4, 8; 0 = default */
    int32_t nontemporal;        /* 1 = non-temporal stores, 0 = cached, -1 = default */
    int32_t chunk_rounds;       /* rounds of block_threads * lines per chunk: 1, 2, 4, 8;
                                   0 = default */
    int32_t defer_straddle;     /* 1 = lines holding several segments are assembled by one
                                   thread per segment in leading copy blocks (pays for short
                                   segments), 0 = by the copy loop, -1 = default (0) */
    int32_t pipeline;           /* graph replays (two workspaces, the next step planned
                                   while the current one copies): 0 = plan, scan, copy
                                   kernels per step; 1 = each copy kernel also plans the next
                                   step in its first P workgroups, a scan kernel follows;
                                   2 = as 1 with the scan in the plan's last workgroup
                                   (one kernel per step);
                                   3 = plan and scan on a concurrent graph branch;
                                   -1 = default (2) */
} gpumpi_a2av_tune;

int gpumpi_a2av_plan_create(int32_t P, int32_t elem_bytes, void* const* sendbufs,
                            const int32_t* send_counts, const int32_t* send_displs,
                            void* const* recvbufs, int32_t* recv_counts,
                            int32_t* recv_displs, const gpumpi_a2av_tune* tune,
                            gpumpi_a2av_plan** out);
void gpumpi_a2av_plan_destroy(gpumpi_a2av_plan* plan);
/* Run one step on `stream` (asynchronous). */
int gpumpi_a2av_execute(gpumpi_a2av_plan* plan, void* stream);
/* Capture `steps_per_graph` steps into a hipGraph once, then replay it. */
int gpumpi_a2av_graph_replay(gpumpi_a2av_plan* plan, int32_t steps_per_graph, void* stream);
/* One step like gpumpi_a2av_plan_create + gpumpi_a2av_execute + gpumpi_a2av_plan_destroy,
 * allocating nothing: no hipMalloc or hipFree, and no call that allocates.  No counterpart
 * in the reference; for callers that cannot allocate -- code under hipGraph capture, tight
 * serving loops, harnesses that forbid allocation.  The plan tables live in a static
 * device array of about 4 MiB per GPU, held until the process ends, so P is at most 512
 * (GPUMPI_E_ARG beyond it).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (concurrent streams, threads) race and give undefined results.  Arguments
 * as for gpumpi_a2av_plan_create; asynchronous on `stream`. */
int gpumpi_a2av_execute_noalloc(int32_t P, int32_t elem_bytes, void* const* sendbufs,
                                const int32_t* send_counts, const int32_t* send_displs,
                                void* const* recvbufs, int32_t* recv_counts,
                                int32_t* recv_displs, void* stream);
/* Sweep launch shapes on the live buffers (single steps), then the pipeline modes (10-step
 * graph replays), keep the fastest; returns the chosen shape and best_us = its time per
 * step in graph replays. */
int gpumpi_a2av_autotune(gpumpi_a2av_plan* plan, void* stream, gpumpi_a2av_tune* chosen,
                         double* best_us);
/* Kernel durations of `steps` (>= 2) steps of pipeline mode 1 or 2 launched directly, timed
 * with HIP events on `stream`: us[0] = the first step's plan + scan, us[1] = the average
 * copy kernel (each but the last also planning the next step; in mode 1 with the scan
 * kernel after it), us[2] = the sequence's time per step. */
int gpumpi_a2av_pipeline_times(gpumpi_a2av_plan* plan, int32_t steps, void* stream, double* us);
/* Per-kernel durations of a step (profiling): enable = 1 makes gpumpi_a2av_execute record
 * HIP events around its kernels, 0 stops it, -1 reads us[3] = {plan, scan, copy} in
 * microseconds for the last executed step (graph replays are not timed). */
int gpumpi_a2av_phase_times(gpumpi_a2av_plan* plan, int32_t enable, double* us);
/* Launch shape in use and the number of workgroups of the copy grid. */
int gpumpi_a2av_get_tune(gpumpi_a2av_plan* plan, gpumpi_a2av_tune* tune, int32_t* grid);
/* Total payload bytes moved by the last completed step (reads them back from the device). */
int gpumpi_a2av_payload_bytes(gpumpi_a2av_plan* plan, int64_t* bytes);

/* ------------------------------------------------------------------------------------
 * Batched MPI_Reduce / MPI_Allreduce: the reduction of one step over P ranks as one
 * kernel (replayable from a hipGraph), for callers that hold every rank's buffer on the
 * GPU.  Replaces the call MPI_Reduce(sendbuf, recvbuf, count, datatype, op, root, comm)
 * (gpu_libs/gpu_mpi/mpi.cu:148-235; root = -1: MPI_Allreduce, mpi.cu:276-282) made by
 * all P ranks of a step.  Every element is folded in the reference's rank order
 * (inout = op(in_r, inout), r = 1..P-1, mpi.cu:209-225), so FP64 results are the rank
 * grid's bit for bit.  sendbufs / recvbufs: device arrays of P device pointers (recvbufs
 * entries other than the root's are not written by MPI_Reduce); count elements per rank.
 * datatype is an MPI_Datatype, op a built-in operator index (MPI_MAX = 0, MPI_MIN = 1,
 * MPI_SUM = 2, MPI_PROD = 3).  The batched path takes int32, uint32, int64, uint64, float
 * and double with those four; any other pair returns GPUMPI_E_ARG (the rank grid's
 * MPI_Reduce takes every legal pair).
 * ------------------------------------------------------------------------------------ */
typedef struct gpumpi_reduce_plan gpumpi_reduce_plan;
int gpumpi_reduce_plan_create(int32_t P, int64_t count, int32_t datatype, int32_t op, int32_t root,
                              void* const* sendbufs, void* const* recvbufs,
                              gpumpi_reduce_plan** out);
void gpumpi_reduce_plan_destroy(gpumpi_reduce_plan* plan);
/* One step on `stream` (asynchronous). */
int gpumpi_reduce_execute(gpumpi_reduce_plan* plan, void* stream);
/* Capture `steps` steps into a hipGraph once per step count, then replay it. */
int gpumpi_reduce_graph_replay(gpumpi_reduce_plan* plan, int32_t steps, void* stream);
/* Average duration of the kernel over `reps` launches, HIP events on `stream` (us). */
int gpumpi_reduce_kernel_time(gpumpi_reduce_plan* plan, int32_t reps, void* stream, double* us);

/* ------------------------------------------------------------------------------------
 * Batched particle step: the particle variant of the exchange (tally, place and scatter
 * kernels per step, hipGraph replay), with the result of the MPI program that
 * exchanges the migrants by MPI_Alltoall of the counts + MPI_Alltoallv
 * (gpu_mpi/mpi.cu:333-408 semantics).  No reference counterpart.
 *
 * P = gx*gy*gz ranks on a periodic grid; a state is parts[P][cap][8] FP64 {x, y, z, vx,
 * vy, vz, w, id} and counts[P] (int64), device buffers.  Per step a particle moves to a
 * face neighbour with probability ~0.05 by Philox(seed, id, step); each rank's new slab
 * holds its stayers in order, then the arrivals in ascending source rank.  If any rank
 * would receive more than cap particles, the step writes the stayers and the new counts
 * only and sets the overflow flag (the MPI program's error path).
 * The plan keeps a direction byte per slot for the state it will read next:
 * gpumpi_particles_prepare computes them for a state at a given step, every step writes
 * them for its output, so steps must follow each other on the same chain of buffers.
 * ------------------------------------------------------------------------------------ */
typedef struct gpumpi_particle_plan gpumpi_particle_plan;

/* nontemporal: 1 = non-temporal stores in the scatter kernel. */
int gpumpi_particles_plan_create(int32_t gx, int32_t gy, int32_t gz, int64_t cap, uint64_t seed,
                                 int32_t nontemporal, gpumpi_particle_plan** out);
void gpumpi_particles_plan_destroy(gpumpi_particle_plan* plan);
/* The workload's initial state: n0 particles per rank (asynchronous on `stream`). */
int gpumpi_particles_init(gpumpi_particle_plan* plan, double* parts, int64_t* counts, int64_t n0,
                          void* stream);
/* Direction bytes of (parts, counts) at `step`; resets the plan's statistics. */
int gpumpi_particles_prepare(gpumpi_particle_plan* plan, const double* parts, const int64_t* counts,
                             uint32_t step, void* stream);
/* One step in -> out (asynchronous; in != out). */
int gpumpi_particles_step(gpumpi_particle_plan* plan, const double* in, const int64_t* cin,
                          double* out, int64_t* cout, void* stream);
/* `steps` (even) steps a -> b -> a ..., captured into a hipGraph once, then replayed; the
 * state ends in (a, ca).  The step index carries on across replays. */
int gpumpi_particles_graph_replay(gpumpi_particle_plan* plan, double* a, int64_t* ca, double* b,
                                  int64_t* cb, int32_t steps, void* stream);
/* Synchronises the device; migrants moved since prepare, sticky overflow flag, and the
 * step index of the next step. */
int gpumpi_particles_status(gpumpi_particle_plan* plan, int64_t* moved, int32_t* overflow,
                            uint32_t* next_step);
/* enable = 1/0 arms/disarms HIP events around the kernels of gpumpi_particles_step;
 * -1 reads us[4] = {tally, place, scatter, batched p2p step (0 in exchange mode 0)} of the
 * last timed step. */
int gpumpi_particles_phase_times(gpumpi_particle_plan* plan, int32_t enable, double* us);
/* How a step moves its migrants.  0 (default): the scatter kernel writes every particle to
 * its final slot.  1: the MPI program's neighbour mode (Isend/Irecv of each rank's migrants
 * toward each neighbour) -- the scatter packs each rank's migrants by destination into a
 * send area, and one batched point-to-point step (gpumpi_p2p_*, one pair per source and
 * destination) moves them into the destination slabs.  Both give the same state bit for
 * bit.  The first switch to 1 allocates the send area (P * cap * 64 bytes). */
int gpumpi_particles_exchange_mode(gpumpi_particle_plan* plan, int32_t mode);

/* ------------------------------------------------------------------------------------
 * Batched point-to-point step: the matched MPI_Isend/MPI_Irecv pairs of one step
 * (gpu_libc/cuda_mpi.cu:65-118, gpu_mpi/mpi.cu:617-645) in two kernels, without envelopes
 * -- the pair list is the match.  Each pair copies `bytes` bytes from src to dst, as the
 * matched receive would deliver them.  The list is device memory, read when the step runs,
 * so device code earlier on the stream (or in the same graph) may write this step's pairs.
 * Destinations must not overlap each other or a source; zero-byte pairs are allowed.
 * ------------------------------------------------------------------------------------ */
typedef struct gpumpi_p2p_pair {
    const void* src;
    void* dst;
    uint64_t bytes;
} gpumpi_p2p_pair;
typedef struct gpumpi_p2p_plan gpumpi_p2p_plan;

/* A plan for lists of up to max_pairs pairs (allocates 8 * (max_pairs + 1) bytes). */
int gpumpi_p2p_plan_create(int64_t max_pairs, gpumpi_p2p_plan** out);
void gpumpi_p2p_plan_destroy(gpumpi_p2p_plan* plan);
/* One step over pairs[0 .. npairs) on `stream`: asynchronous, allocates nothing, may be
 * captured into a caller's graph. */
int gpumpi_p2p_execute(gpumpi_p2p_plan* plan, const gpumpi_p2p_pair* pairs, int64_t npairs, void* stream);
/* `steps` steps over the same list, captured into a hipGraph once (per list, count and
 * steps) and replayed. */
int gpumpi_p2p_graph_replay(gpumpi_p2p_plan* plan, const gpumpi_p2p_pair* pairs, int64_t npairs,
                            int32_t steps, void* stream);
/* reps timed steps (synchronous); us[2] = average {scan, copy} kernel times. */
int gpumpi_p2p_kernel_times(gpumpi_p2p_plan* plan, const gpumpi_p2p_pair* pairs, int64_t npairs,
                            int32_t reps, void* stream, double* us);

/* ------------------------------------------------------------------------------------
 * Bucket-sort caller of the exchange (the NPB IS rank() shape,
 * toolchain_tests/npb/NPB3.4.1/NPB3.4-MPI/IS/is.c:631-781), on device buffers.
 * ------------------------------------------------------------------------------------ */
/* keys[P][n]: dist 0 = uniform [0, 2^key_bits) from Philox (seed, rank, step);
 * 1 = Zipf by inverse CDF over cdf[ncdf] (device doubles); 2 = NPB IS create_seq keys
 * of P ranks with max_key_log2 = key_bits (is.c:520-537, seeds is.c:470-510). */
int gpumpi_gen_keys(uint32_t* keys, int32_t P, int64_t n, uint64_t seed, uint32_t step,
                    int32_t dist, int32_t key_bits, const double* cdf, int64_t ncdf, void* stream);
/* NPB-style plan: per-rank bucket histograms, global totals, splitters (is.c:721-751),
 * send_counts/send_displs [P][P] and the stably bucketed send buffers key_buff1[P][n];
 * bucket_lo/bucket_hi[P] are each destination's bucket range (empty when lo > hi).
 * Synchronous on `stream`. */
int gpumpi_bucket_plan(const uint32_t* keys, int32_t P, int64_t n, int32_t key_bits,
                       int32_t bucket_bits, uint32_t* key_buff1, int32_t* send_counts,
                       int32_t* send_displs, int32_t* bucket_lo, int32_t* bucket_hi,
                       void* stream);

/* Counting sort of n keys < 2^key_bits into out (device buffers, synchronous).  Applied
 * to the concatenated receive buffers of a bucket exchange it is every rank's local sort
 * at once (ranks own disjoint ascending key ranges; is.c:785-833 ranks by counting). */
int gpumpi_local_sort(const uint32_t* keys, int64_t n, int32_t key_bits, uint32_t* out,
                      void* stream);
/* NPB IS class cls ('S','W','A','B','C') on P ranks (power of two) through the batched
 * exchange: returns NPB's verification count in *passed (correct run: 5 * 10 + P,
 * is.c:1137-1138) and the summed exchange time of the 10 timed iterations. */
int gpumpi_npb_is_run(char cls, int32_t P, int64_t* passed, double* exchange_ms);

/* ------------------------------------------------------------------------------------
 * Served MPI jobs.  A job is an independent MPI program on its own communicator:
 *   type 0: dot product of two FP64 vectors of `size` elements over `ranks` ranks
 *           (MPI_Allreduce SUM of the local dots), result in .dot;
 *   type 1: MPI_Alltoall of `size` bytes per pair, result = checksum of the received
 *           bytes in .checksum.
 * gpumpi_serve_*: a persistent rank grid of cfg->num_ranks wavefront slots; slot 0 admits
 * jobs in arrival order (bounded backfill) when enough slots are idle and the job's
 * heap need fits in heap_fraction of the device heap; the others run jobs.  The grid
 * holds the GPU until gpumpi_serve_stop, or until cfg->timeout_s passes without a
 * gpumpi_serve_submit/poll call (a host that went away): then it finishes the admitted
 * jobs and exits by itself.
 * gpumpi_run_job: the plain path -- one rank-grid launch per job.
 * ------------------------------------------------------------------------------------ */
typedef struct gpumpi_job {
    uint64_t id;
    uint64_t seed;
    int32_t type;
    int32_t ranks;
    int64_t size;
} gpumpi_job;

typedef struct gpumpi_job_result {
    uint64_t id;
    int32_t status;             /* 1 done, -1 rejected (never fits), -2 device error */
    int32_t ranks;
    double dot;
    uint64_t checksum;
    uint64_t t_admit_ticks;     /* device wall clock (100 MHz) at admission */
    uint64_t t_done_ticks;      /* device wall clock when the last rank finished */
    int64_t bytes;              /* heap bytes reserved for the job */
} gpumpi_job_result;

typedef struct gpumpi_server gpumpi_server;

int gpumpi_serve_start(const gpumpi_config* cfg, int32_t queue_capacity, int32_t max_job_ranks,
                       double heap_fraction, gpumpi_server** out);
/* Non-blocking; GPUMPI_E_NOMEM while the ring is full: queue_capacity jobs submitted since
 * the oldest job the server has not yet admitted, or since the oldest result this host
 * has not yet polled.  Poll, then submit again. */
int gpumpi_serve_submit(gpumpi_server* server, const gpumpi_job* job);
/* Copies up to max new completions; returns how many. */
int gpumpi_serve_poll(gpumpi_server* server, gpumpi_job_result* out, int32_t max);
/* Finishes every submitted job, stops the server; stats[4] (optional) = admitted,
 * rejected, max concurrently running jobs, jobs admitted ahead of an older one. */
int gpumpi_serve_stop(gpumpi_server* server, int64_t* stats);
/* Plain path: launch job->ranks ranks of `rt` for one job on `stream` (asynchronous);
 * the result lands in dev_result (device memory). */
int gpumpi_run_job(gpumpi_runtime* rt, const gpumpi_job* job, gpumpi_job_result* dev_result,
                   void* stream);
/* Heap bytes a job reserves (the admission rule of the server). */
int64_t gpumpi_job_heap_bytes(const gpumpi_job* job);

#ifdef __cplusplus
}
#endif
#endif
