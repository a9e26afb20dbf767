// Device-resident runtime state: one Runtime per rank grid, one RankState per rank.
//
// Layout in HBM (all hipMalloc, coarse-grained device memory; cross-XCD visibility is
// by agent-scope release/acquire):
//   Mailbox[P]            128 B each; tail ticket counter on its own cache line
//   Envelope[P][ring]     128 B slots, per-slot sequence numbers (bounded MPSC ring)
//   RankState[P]          rank-private bookkeeping (written only by the owning wave)
//   Request[P][maxreq]    request pool per rank
//   Envelope[P][maxunexp] unexpected-message list per rank
//   scratch[P][64 KiB]    staging area for private-memory buffers in collectives
//   Heap arena            lock-free size-class heap (heap.h)
// The reference keeps the equivalent structures in CUDA managed memory guarded by a
// system-scope CAS lock per receiver (gpu_libc/cuda_mpi.cuh:86-113, 372-390, 459-533).
#pragma once

#include <stddef.h>

#include "common.h"
#include "heap.h"
#include "../../../include/gpumpi/io_ops.h"

namespace gpumpi {

// Message envelope (reference: CudaMPI::MessageDescriptor, gpu_libc/cuda_mpi.cuh:372-379).
constexpr int kInlineMax = 80;
enum EnvFlags : uint32_t { kEnvInline = 1, kEnvBuffered = 2 };

struct Envelope {
    uint32_t seq;     // ring protocol: == ticket+1 when full, ticket+ring when free
    int32_t src;      // sender world rank
    int32_t tag;
    int32_t ctx;      // communicator context id
    uint64_t bytes;
    uint64_t data;    // sender buffer (rendezvous) or bounce buffer (buffered)
    uint64_t done;    // address of the sender's completion word (rendezvous)
    uint32_t flags;
    uint32_t pad;
    uint8_t inline_data[kInlineMax];
};
static_assert(sizeof(Envelope) == 128, "envelope is two 64-byte lines");

struct alignas(128) Mailbox {
    uint64_t tail;    // next ticket (atomic, senders)
    uint64_t pad[15];
};

enum ReqState : int32_t { kReqFree = 0, kReqActive = 1, kReqDone = 2 };
enum ReqKind : int32_t { kReqSend = 1, kReqRecv = 2 };

struct CommImpl;

// Reference: CudaMPI::PendingOperation (gpu_libc/cuda_mpi.cuh:193-230) + MPI_Request_impl
// (gpu_mpi/mpi.cu:20-29).
struct Request {
    int32_t state;
    int32_t kind;
    int32_t peer;     // send: dest world rank; recv: source world rank or ANY (-1)
    int32_t tag;
    int32_t ctx;
    int32_t index;    // slot index in the rank's pool
    void* buf;
    uint64_t bytes;
    uint32_t done;    // rendezvous: set to 1 by the receiver after the copy
    int32_t st_source;
    int32_t st_tag;
    int32_t st_error;
    uint64_t st_bytes;
    const CommImpl* comm;
};

constexpr uint64_t kScratchBytes = 64 * 1024;

// Bump region of one served job: the job's ranks allocate from it, the scheduler frees
// it in one piece when the job ends.
struct JobArena {
    char* base;
    uint64_t size;
    uint64_t bump;  // atomic, bytes handed out
};

struct RankState {
    int32_t world_rank;
    int32_t n_posted;
    int32_t n_unexp;
    int32_t n_free;
    uint64_t ring_head;
    int32_t* posted;      // [max_requests] request indices, post order
    Envelope* unexp;      // [max_unexpected], arrival order
    Request* reqs;        // [max_requests]
    int32_t* free_stack;  // [max_requests]
    char* scratch;        // [kScratchBytes]
    int32_t initialized;
    int32_t finalized;
    HeapCache cache;
    JobArena* arena;      // non-null while the rank runs a served job
};

struct GroupImpl {
    int32_t size;
    int32_t refcount;     // < 0: static (never freed)
    int32_t* ranks;       // world ranks
};

// Per-member slot of the shared-address-space collectives: each rank publishes its
// buffer pointers, the others pull.  seq == the collective sequence number.
struct alignas(64) CollSlot {
    uint64_t seq;
    const void* sbuf;
    void* rbuf;
    const int* counts;
    const int* displs;
    int64_t a0;
    int32_t a1;
    int32_t a2;
    int64_t a3;
};

struct alignas(128) BarrierWord {
    uint32_t value;
    uint32_t pad[31];
};

// A communicator object shared by all of its members (MPI_Comm is a pointer to it).
// Reference: MPI_Comm_impl (gpu_mpi/communicator.cu:16-30), one object per rank with an
// O(P) group scan per MPI_Comm_rank (gpu_mpi/group.cu:110-121); here rank lookup is O(1)
// through inv[].
struct CommImpl {
    int32_t ctx;
    int32_t size;
    int32_t refcount;     // atomic; < 0 for MPI_COMM_WORLD (static)
    int32_t is_world;
    int32_t* world_ranks; // [size] comm rank -> world rank
    int32_t* inv;         // [nworld] world rank -> comm rank or -1 (nullptr: identity)
    CollSlot* slots;      // [size]
    uint64_t* seqs;       // [size] next collective sequence number of each member
    BarrierWord arrive;
    BarrierWord gen;
    GroupImpl* group;     // group of this communicator (shared, read-only)
};

// ---------------------------------------------------------------- host I/O delegation
// Each rank owns one request slot in pinned host memory: the rank writes the arguments
// and up to kIoData bytes, then publishes an odd seq with a system-scope release; the
// launcher's host thread performs the C library call and answers with seq + 1.  The
// request codes (IoOp) are in include/gpumpi/io_ops.h.
struct IoSlot {
    uint32_t seq;
    int32_t op;
    int64_t a0, a1, a2;
    int64_t result;
    int32_t len;
    int32_t pad;
    char data[kIoData];
};

struct Runtime {
    int32_t nranks;
    int32_t ranks_per_block;
    int32_t ring_slots;
    int32_t max_requests;
    int32_t max_unexpected;
    int32_t abort_flag;   // atomic
    int32_t exit_code;    // atomic or
    int32_t err_lock;     // atomic
    int32_t bcast_push_min;  // Bcast pushes slices from this many bytes per rank; < 0 never
    int32_t copy_stage_bytes;  // LDS staging area per wave of the running launch (0 = none)
    uint64_t timeout_ticks;
    double wtick;         // seconds per wall_clock64 tick
    uint64_t next_ctx;    // atomic context-id allocator
    Mailbox* mailbox;
    Envelope* rings;
    RankState* ranks;
    CommImpl* world;
    CommImpl* null_comm;
    GroupImpl* world_group;
    GroupImpl* empty_group;
    const void* user_args;
    Heap heap;
    ErrorRecord err;
    IoSlot* io;           // nranks request slots in pinned host memory
};

// Layout of the structures shared by the library's host code and init kernels and by the
// device code of a program compiled separately against these headers: a program built
// against other headers than the library would read them at other offsets.  The launcher
// main() compares its own value with the library's before it launches anything.
constexpr uint64_t layout_mix(uint64_t h, uint64_t v) { return (h ^ v) * 0x100000001B3ull; }
constexpr uint64_t kLayoutStamp =
    layout_mix(layout_mix(layout_mix(layout_mix(layout_mix(layout_mix(layout_mix(
        0xCBF29CE484222325ull, sizeof(Runtime)), sizeof(RankState)), sizeof(Heap)),
        sizeof(HeapCache)), offsetof(Runtime, heap)), offsetof(RankState, cache)),
        offsetof(Heap, bytes_in_use));

}  // namespace gpumpi
