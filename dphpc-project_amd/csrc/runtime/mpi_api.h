// Definitions of the device MPI API declared in include/gpumpi/mpi.h.
// Each function cites the reference entry point whose behaviour it provides
// (gpu_libs/gpu_mpi/mpi.cu, communicator.cu, group.cu, operators.cu).
#pragma once

#include "collectives.h"
#include "../../../include/gpumpi/gpu_malloc.h"

// Weak: each translation unit that includes the runtime is its own code object and needs
// its own copy; the host-side shadow symbols of the copies are merged at link time.
__device__ __attribute__((weak)) gpumpi::OpImpl gpumpi_builtin_ops[12] = {
    {0, nullptr, 1}, {1, nullptr, 1}, {2, nullptr, 1}, {3, nullptr, 1},
    {4, nullptr, 1}, {5, nullptr, 1}, {6, nullptr, 1}, {7, nullptr, 1},
    {8, nullptr, 1}, {9, nullptr, 1}, {10, nullptr, 1}, {11, nullptr, 1}};

namespace gpumpi {
__device__ MPI_Comm world_comm() { return rt()->world; }
__device__ MPI_Comm null_comm() { return nullptr; }
__device__ MPI_Group world_group() { return rt()->world_group; }
__device__ MPI_Group empty_group() { return rt()->empty_group; }

__device__ __forceinline__ bool wave_converged() {
    return __builtin_amdgcn_read_exec() == ~0ull;
}
// All MPI entry points require the whole wavefront (the rank) to call them.
#define GPUMPI_ENTER()                                                                   \
    do {                                                                                 \
        if (!::gpumpi::wave_converged()) {                                               \
            ::gpumpi::record_error(::gpumpi::kErrDivergent, __LINE__, 0,                 \
                                   "MPI call from a partial wavefront");                 \
            return MPI_ERR_OTHER;                                                        \
        }                                                                                \
        if (::gpumpi::aborted()) return MPI_ERR_OTHER;                                   \
    } while (0)

__device__ __forceinline__ bool bytes_of(int count, MPI_Datatype t, uint64_t* out) {
    int s = type_size(t);
    if (s <= 0 || count < 0) {
        GPUMPI_FAIL(kErrArgument, (int)t, "invalid datatype or count");
        return false;
    }
    *out = (uint64_t)count * (uint64_t)s;
    return true;
}

__device__ __forceinline__ void fill_status(MPI_Status* st, const Request* r) {
    if (!st) return;
    if (r->kind == kReqRecv) {
        int src = r->st_source;
        st->MPI_SOURCE = (src >= 0 && r->comm) ? (r->comm->inv ? r->comm->inv[src] : src) : src;
        st->MPI_TAG = r->st_tag;
        st->gpumpi_bytes = (long long)r->st_bytes;
    } else {
        st->MPI_SOURCE = MPI_ANY_SOURCE;
        st->MPI_TAG = MPI_ANY_TAG;
        st->gpumpi_bytes = 0;
    }
    st->MPI_ERROR = r->st_error;
}
}  // namespace gpumpi

using namespace gpumpi;

// ---------------------------------------------------------------- environment
// mpi.cu:44-75
__device__ int MPI_Init(int*, char***) {
    GPUMPI_ENTER();
    my_state()->initialized = 1;
    return MPI_SUCCESS;
}
__device__ int MPI_Init_thread(int* argc, char*** argv, int, int* provided) {
    if (provided) *provided = MPI_THREAD_SINGLE;
    return MPI_Init(argc, argv);
}
__device__ int MPI_Initialized(int* flag) {
    *flag = my_state()->initialized;
    return MPI_SUCCESS;
}
__device__ int MPI_Finalize(void) {
    GPUMPI_ENTER();
    // Drain what other ranks may still push to this rank before it leaves.
    progress();
    my_state()->finalized = 1;
    return MPI_SUCCESS;
}
__device__ int MPI_Abort(MPI_Comm, int errorcode) {
    record_error(kErrAbort, __LINE__, errorcode, "MPI_Abort called");
    return MPI_SUCCESS;
}
__device__ void gpumpi_exit(int status) {
    Runtime* R = rt();
    if (status != 0 && lane() == 0)
        __hip_atomic_fetch_or(&R->exit_code, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    heap_flush_cache(&R->heap, &my_state()->cache);
    __builtin_amdgcn_endpgm();
}
// One request of this rank to the launcher's host I/O thread (see IoSlot).  Called
// wave-uniformly; in1/in2 are copied back to back into the request, up to outcap bytes of
// the answer into out (private, LDS-free memory handled like any message buffer).
__device__ int64_t gpumpi_io(int op, int64_t a0, int64_t a1, int64_t a2, const void* in1,
                             int len1, const void* in2, int len2, void* out, int outcap) {
    Runtime* R = rt();
    if (!R->io || len1 < 0 || len2 < 0 || len1 + len2 > kIoData) {
        record_error(kErrArgument, __LINE__, len1 + len2, "host I/O request too large or no I/O channel");
        return -1;
    }
    IoSlot* slot = R->io + world_rank();
    uint32_t seq = 0;
    if (lane() == 0) seq = __hip_atomic_load(&slot->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    seq = bcast(seq);
    if (len1) wave_copy(slot->data, in1, (uint64_t)len1);
    if (len2) wave_copy(slot->data + len1, in2, (uint64_t)len2);
    if (lane() == 0) {
        slot->op = op;
        slot->a0 = a0;
        slot->a1 = a1;
        slot->a2 = a2;
        slot->len = len1 + len2;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // the request before its seq, to the host
    if (lane() == 0) __hip_atomic_store(&slot->seq, seq + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    Waiter w;
    bool answered = false;
    for (;;) {
        uint32_t s = 0;
        if (lane() == 0) s = __hip_atomic_load(&slot->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (bcast(s) == seq + 2) {
            answered = true;
            break;
        }
        if (!w.keep_going(__LINE__)) break;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    if (!answered) return -1;
    const int64_t res = bcast((uint64_t)slot->result);
    if (out && outcap > 0 && (op == kIoRead || op == kIoGets) && res > 0) {
        const int64_t n = res < outcap ? res : outcap;
        wave_copy(out, slot->data, (uint64_t)n);
    }
    return res;
}

__device__ int MPI_Get_processor_name(char* name, int* resultlen) {
    const char n[] = "MI355X wavefront";
    for (int i = 0; i < (int)sizeof(n); i++) name[i] = n[i];
    *resultlen = (int)sizeof(n) - 1;
    return MPI_SUCCESS;
}
// mpi.cu:142-146, 271-274: a device clock; here the constant 100 MHz wall clock.
__device__ double MPI_Wtick(void) { return rt()->wtick; }
__device__ double MPI_Wtime(void) { return (double)wall_clock64() * rt()->wtick; }

// ---------------------------------------------------------------- point to point
// mpi.cu:617-645
__device__ int MPI_Isend(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                         MPI_Comm comm, MPI_Request* request) {
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t bytes;
    if (!bytes_of(count, datatype, &bytes)) return MPI_ERR_OTHER;
    if (dest < 0 || dest >= comm->size) {
        GPUMPI_FAIL(kErrArgument, dest, "MPI_Isend: invalid destination rank");
        return MPI_ERR_OTHER;
    }
    Request* r = isend(buf, bytes, comm_world_rank(comm, dest), tag, comm->ctx, comm);
    *request = r;
    return r ? MPI_SUCCESS : MPI_ERR_OTHER;
}
__device__ int MPI_Irecv(void* buf, int count, MPI_Datatype datatype, int source, int tag,
                         MPI_Comm comm, MPI_Request* request) {
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t bytes;
    if (!bytes_of(count, datatype, &bytes)) return MPI_ERR_OTHER;
    if (source != MPI_ANY_SOURCE && (source < 0 || source >= comm->size)) {
        GPUMPI_FAIL(kErrArgument, source, "MPI_Irecv: invalid source rank");
        return MPI_ERR_OTHER;
    }
    int src = source == MPI_ANY_SOURCE ? MPI_ANY_SOURCE : comm_world_rank(comm, source);
    Request* r = irecv(buf, bytes, src, tag, comm->ctx, comm);
    if (request) *request = r;
    return r ? MPI_SUCCESS : MPI_ERR_OTHER;
}
// mpi.cu:672-691 (Wait on MPI_REQUEST_NULL is legal here)
__device__ int MPI_Wait(MPI_Request* request, MPI_Status* status) {
    GPUMPI_ENTER();
    if (!request || *request == MPI_REQUEST_NULL) {
        if (status) *status = MPI_Status();
        return MPI_SUCCESS;
    }
    Request* r = *request;
    bool ok = wait_request(r);
    fill_status(status, r);
    int err = r->st_error;
    release_request(my_state(), r);
    *request = MPI_REQUEST_NULL;
    return ok ? err : MPI_ERR_OTHER;
}
__device__ int MPI_Test(MPI_Request* request, int* flag, MPI_Status* status) {
    GPUMPI_ENTER();
    if (!request || *request == MPI_REQUEST_NULL) {
        *flag = 1;
        if (status) *status = MPI_Status();
        return MPI_SUCCESS;
    }
    Request* r = *request;
    if (test_request(r)) {
        *flag = 1;
        fill_status(status, r);
        release_request(my_state(), r);
        *request = MPI_REQUEST_NULL;
    } else {
        *flag = 0;
    }
    return MPI_SUCCESS;
}
// mpi.cu:653-659 (MPI_STATUSES_IGNORE is legal here)
__device__ int MPI_Waitall(int count, MPI_Request reqs[], MPI_Status* statuses) {
    int err = MPI_SUCCESS;
    for (int i = 0; i < count; i++) {
        int e = MPI_Wait(&reqs[i], statuses ? &statuses[i] : nullptr);
        if (e != MPI_SUCCESS) err = e;
    }
    return err;
}
__device__ int MPI_Testall(int count, MPI_Request reqs[], int* flag, MPI_Status statuses[]) {
    GPUMPI_ENTER();
    bool all = true;
    for (int i = 0; i < count; i++)
        if (reqs[i] != MPI_REQUEST_NULL && !test_request(reqs[i])) all = false;
    if (!all) {
        *flag = 0;
        return MPI_SUCCESS;
    }
    *flag = 1;
    return MPI_Waitall(count, reqs, statuses);
}
__device__ int MPI_Waitsome(int incount, MPI_Request reqs[], int* outcount, int indices[],
                            MPI_Status statuses[]) {
    GPUMPI_ENTER();
    bool any_active = false;
    for (int i = 0; i < incount; i++) any_active |= reqs[i] != MPI_REQUEST_NULL;
    if (!any_active) {
        *outcount = MPI_UNDEFINED;
        return MPI_SUCCESS;
    }
    Waiter w;
    for (;;) {
        int n = 0;
        for (int i = 0; i < incount; i++) {
            if (reqs[i] != MPI_REQUEST_NULL && test_request(reqs[i])) {
                fill_status(statuses ? &statuses[n] : nullptr, reqs[i]);
                release_request(my_state(), reqs[i]);
                reqs[i] = MPI_REQUEST_NULL;
                indices[n++] = i;
            }
        }
        if (n > 0) {
            *outcount = n;
            return MPI_SUCCESS;
        }
        if (!w.keep_going(__LINE__)) {
            *outcount = 0;
            return MPI_ERR_OTHER;
        }
    }
}
__device__ int MPI_Request_free(MPI_Request* request) {
    GPUMPI_ENTER();
    if (!request || *request == MPI_REQUEST_NULL) return MPI_SUCCESS;
    // MPI semantics: the operation still completes; the handle is released once it has.
    wait_request(*request);
    release_request(my_state(), *request);
    *request = MPI_REQUEST_NULL;
    return MPI_SUCCESS;
}
__device__ int MPI_Get_count(const MPI_Status* status, MPI_Datatype datatype, int* count) {
    int s = type_size(datatype);
    if (!status || s <= 0) return MPI_ERR_OTHER;
    *count = (int)(status->gpumpi_bytes / s);
    return MPI_SUCCESS;
}
// mpi.cu:247-269
__device__ int MPI_Send(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                        MPI_Comm comm) {
    MPI_Request r = MPI_REQUEST_NULL;
    int e = MPI_Isend(buf, count, datatype, dest, tag, comm, &r);
    if (e != MPI_SUCCESS) return e;
    return MPI_Wait(&r, MPI_STATUS_IGNORE);
}
__device__ int MPI_Recv(void* buf, int count, MPI_Datatype datatype, int source, int tag,
                        MPI_Comm comm, MPI_Status* status) {
    MPI_Request r = MPI_REQUEST_NULL;
    int e = MPI_Irecv(buf, count, datatype, source, tag, comm, &r);
    if (e != MPI_SUCCESS) return e;
    return MPI_Wait(&r, status);
}
// mpi.cu:255-260 is a stub in the reference; here a real exchange.
__device__ int MPI_Sendrecv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, int dest,
                            int sendtag, void* recvbuf, int recvcount, MPI_Datatype recvtype,
                            int source, int recvtag, MPI_Comm comm, MPI_Status* status) {
    MPI_Request rr = MPI_REQUEST_NULL, sr = MPI_REQUEST_NULL;
    int e = MPI_Irecv(recvbuf, recvcount, recvtype, source, recvtag, comm, &rr);
    if (e != MPI_SUCCESS) return e;
    e = MPI_Isend(sendbuf, sendcount, sendtype, dest, sendtag, comm, &sr);
    if (e != MPI_SUCCESS) return e;
    e = MPI_Wait(&sr, MPI_STATUS_IGNORE);
    int e2 = MPI_Wait(&rr, status);
    return e != MPI_SUCCESS ? e : e2;
}

// ---------------------------------------------------------------- collectives
__device__ int MPI_Barrier(MPI_Comm comm) {  // mpi.cu:324-331
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    return comm_barrier(comm) ? MPI_SUCCESS : MPI_ERR_OTHER;
}
__device__ int MPI_Bcast(void* buffer, int count, MPI_Datatype datatype, int root,
                         MPI_Comm comm) {  // mpi.cu:101-140
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t bytes;
    if (!bytes_of(count, datatype, &bytes)) return MPI_ERR_OTHER;
    return coll_bcast(buffer, bytes, root, comm);
}
__device__ int MPI_Reduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype datatype,
                          MPI_Op op, int root, MPI_Comm comm) {  // mpi.cu:181-235
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    if (root < 0) {
        GPUMPI_FAIL(kErrArgument, root, "MPI_Reduce: invalid root");
        return MPI_ERR_OTHER;
    }
    return coll_reduce(sendbuf, recvbuf, count, datatype, op, root, comm);
}
__device__ int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count,
                             MPI_Datatype datatype, MPI_Op op, MPI_Comm comm) {  // mpi.cu:276-282
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    return coll_reduce(sendbuf, recvbuf, count, datatype, op, -1, comm);
}
__device__ int MPI_Gather(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                          void* recvbuf, int recvcount, MPI_Datatype recvtype, int root,
                          MPI_Comm comm) {  // mpi.cu:291-322
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t sb;
    if (!bytes_of(sendcount, sendtype, &sb)) return MPI_ERR_OTHER;
    return coll_gatherv(sendbuf, sb, recvbuf, nullptr, nullptr, recvcount, recvtype, root, comm,
                        false);
}
__device__ int MPI_Gatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                           void* recvbuf, const int recvcounts[], const int displs[],
                           MPI_Datatype recvtype, int root, MPI_Comm comm) {  // mpi.cu:503-531
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t sb;
    if (!bytes_of(sendcount, sendtype, &sb)) return MPI_ERR_OTHER;
    return coll_gatherv(sendbuf, sb, recvbuf, recvcounts, displs, 0, recvtype, root, comm,
                        false);
}
__device__ int MPI_Allgather(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                             void* recvbuf, int recvcount, MPI_Datatype recvtype,
                             MPI_Comm comm) {  // mpi.cu:484-493
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t sb;
    if (!bytes_of(sendcount, sendtype, &sb)) return MPI_ERR_OTHER;
    return coll_gatherv(sendbuf, sb, recvbuf, nullptr, nullptr, recvcount, recvtype, 0, comm,
                        true);
}
__device__ int MPI_Allgatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                              void* recvbuf, const int recvcounts[], const int displs[],
                              MPI_Datatype recvtype, MPI_Comm comm) {  // mpi.cu:495-501 (stub)
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t sb;
    if (!bytes_of(sendcount, sendtype, &sb)) return MPI_ERR_OTHER;
    return coll_gatherv(sendbuf, sb, recvbuf, recvcounts, displs, 0, recvtype, 0, comm, true);
}
__device__ int MPI_Scatter(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                           void* recvbuf, int recvcount, MPI_Datatype recvtype, int root,
                           MPI_Comm comm) {  // mpi.cu:532-563
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t rb;
    if (!bytes_of(recvcount, recvtype, &rb)) return MPI_ERR_OTHER;
    return coll_scatterv(sendbuf, nullptr, nullptr, sendcount, sendtype, recvbuf, rb, root, comm);
}
__device__ int MPI_Scatterv(const void* sendbuf, const int sendcounts[], const int displs[],
                            MPI_Datatype sendtype, void* recvbuf, int recvcount,
                            MPI_Datatype recvtype, int root, MPI_Comm comm) {  // mpi.cu:564-591
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    uint64_t rb;
    if (!bytes_of(recvcount, recvtype, &rb)) return MPI_ERR_OTHER;
    return coll_scatterv(sendbuf, sendcounts, displs, 0, sendtype, recvbuf, rb, root, comm);
}
__device__ int MPI_Alltoall(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                            void* recvbuf, int recvcount, MPI_Datatype recvtype,
                            MPI_Comm comm) {  // mpi.cu:333-362
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    return coll_alltoallv(sendbuf, nullptr, nullptr, sendtype, recvbuf, nullptr, nullptr,
                          recvtype, comm, sendcount, recvcount);
}
__device__ int MPI_Alltoallv(const void* sendbuf, const int sendcounts[], const int sdispls[],
                             MPI_Datatype sendtype, void* recvbuf, const int recvcounts[],
                             const int rdispls[], MPI_Datatype recvtype,
                             MPI_Comm comm) {  // mpi.cu:363-408
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    return coll_alltoallv(sendbuf, sendcounts, sdispls, sendtype, recvbuf, recvcounts, rdispls,
                          recvtype, comm, 0, 0);
}

// mpi.cu:410-482: the coalescing variants differ only in which allocator served the
// request arrays; here there are no request arrays, so they are the plain collectives.
#define GPUMPI_ATA_DEFINE(suffix)                                                          \
    __device__ int MPI_Alltoall_##suffix(const void* sendbuf, int sendcount,               \
                                         MPI_Datatype sendtype, void* recvbuf,             \
                                         int recvcount, MPI_Datatype recvtype,             \
                                         MPI_Comm comm, bool) {                            \
        return MPI_Alltoall(sendbuf, sendcount, sendtype, recvbuf, recvcount, recvtype,    \
                            comm);                                                         \
    }                                                                                      \
    __device__ int MPI_Alltoallv_##suffix(const void* sendbuf, const int sendcounts[],     \
                                          const int sdispls[], MPI_Datatype sendtype,      \
                                          void* recvbuf, const int recvcounts[],           \
                                          const int rdispls[], MPI_Datatype recvtype,      \
                                          MPI_Comm comm, bool) {                           \
        return MPI_Alltoallv(sendbuf, sendcounts, sdispls, sendtype, recvbuf, recvcounts,  \
                             rdispls, recvtype, comm);                                     \
    }
GPUMPI_ATA_DEFINE(coalesce)
GPUMPI_ATA_DEFINE(v1)
GPUMPI_ATA_DEFINE(v2)
GPUMPI_ATA_DEFINE(v3)
GPUMPI_ATA_DEFINE(v4)
GPUMPI_ATA_DEFINE(v5)
GPUMPI_ATA_DEFINE(v6)
#undef GPUMPI_ATA_DEFINE

// ---------------------------------------------------------------- operators, datatypes
__device__ int MPI_Op_create(MPI_User_function* user_fn, int commute, MPI_Op* op) {  // operators.cu:246-250
    GPUMPI_ENTER();
    OpImpl* o = (OpImpl*)rt_alloc(sizeof(OpImpl));
    if (!o) return MPI_ERR_OTHER;
    o->builtin = -1;
    o->user_fn = user_fn;
    o->commute = commute;
    *op = o;
    return MPI_SUCCESS;
}
__device__ int MPI_Op_free(MPI_Op* op) {
    GPUMPI_ENTER();
    if (op && *op && (*op)->builtin < 0) rt_free(*op);
    if (op) *op = MPI_OP_NULL;
    return MPI_SUCCESS;
}
__device__ int MPI_Type_size(MPI_Datatype datatype, int* size) {
    *size = type_size(datatype);
    return *size > 0 ? MPI_SUCCESS : MPI_ERR_OTHER;
}
// Derived datatypes cannot be expressed by the reference's enum MPI_Datatype
// (datatypes.cuh:12-14); like the reference (mpi.cu:237-245) they are not provided.
__device__ int MPI_Type_contiguous(int, MPI_Datatype, MPI_Datatype*) {
    record_error(kErrNotSupported, __LINE__, 0, "MPI_Type_contiguous is not supported");
    return MPI_ERR_OTHER;
}
__device__ int MPI_Type_commit(MPI_Datatype*) {
    record_error(kErrNotSupported, __LINE__, 0, "MPI_Type_commit is not supported");
    return MPI_ERR_OTHER;
}

// ---------------------------------------------------------------- communicators
__device__ int MPI_Comm_size(MPI_Comm comm, int* size) {  // communicator.cu:215-217
    if (!comm) return MPI_ERR_OTHER;
    *size = comm->size;
    return MPI_SUCCESS;
}
__device__ int MPI_Comm_rank(MPI_Comm comm, int* rank) {  // communicator.cu:219-221
    if (!comm) return MPI_ERR_OTHER;
    *rank = comm_rank_of(comm);
    return MPI_SUCCESS;
}
__device__ int MPI_Comm_group(MPI_Comm comm, MPI_Group* group) {  // communicator.cu:96-100
    if (!comm) return MPI_ERR_OTHER;
    *group = comm->group;
    return MPI_SUCCESS;
}
__device__ int MPI_Comm_split(MPI_Comm comm, int color, int key, MPI_Comm* newcomm) {  // 164-213
    GPUMPI_ENTER();
    if (!check_comm(comm)) return MPI_ERR_OTHER;
    return comm_split(comm, color, key, newcomm);
}
__device__ int MPI_Comm_dup(MPI_Comm comm, MPI_Comm* newcomm) {  // communicator.cu:118-128
    GPUMPI_ENTER();
    if (!comm) return MPI_FAILURE;
    return comm_split(comm, 0, comm_rank_of(comm), newcomm);
}
__device__ int MPI_Comm_create(MPI_Comm comm, MPI_Group group, MPI_Comm* newcomm) {  // 102-116
    GPUMPI_ENTER();
    if (!check_comm(comm) || !group) return MPI_ERR_OTHER;
    int pos = MPI_UNDEFINED;
    MPI_Group_rank(group, &pos);
    return comm_split(comm, pos == MPI_UNDEFINED ? MPI_UNDEFINED : 0, pos, newcomm);
}
__device__ int MPI_Comm_free(MPI_Comm* comm) {  // communicator.cu:80-94
    GPUMPI_ENTER();
    if (!comm || !*comm) return MPI_SUCCESS;
    CommImpl* c = *comm;
    *comm = MPI_COMM_NULL;
    if (c->is_world) return MPI_SUCCESS;
    int32_t left = 0;
    if (lane() == 0) left = __hip_atomic_fetch_add(&c->refcount, -1, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_AGENT) - 1;
    left = bcast(left);
    if (left == 0) rt_free(c);
    return MPI_SUCCESS;
}
// Cartesian topologies: not provided by the reference either (communicator.cu:138-146).
__device__ int MPI_Cart_create(MPI_Comm, int, const int[], const int[], int, MPI_Comm*) {
    record_error(kErrNotSupported, __LINE__, 0, "MPI_Cart_create is not supported");
    return MPI_ERR_OTHER;
}
__device__ int MPI_Cart_sub(MPI_Comm, const int[], MPI_Comm*) {
    record_error(kErrNotSupported, __LINE__, 0, "MPI_Cart_sub is not supported");
    return MPI_ERR_OTHER;
}
// Balanced factorisation of nnodes over the zero entries of dims (MPI standard 7.5.2):
// the prime factors of the free part are dealt largest-first to the currently smallest
// free dimension, then the free dimensions are sorted non-increasing.
__device__ int MPI_Dims_create(int nnodes, int ndims, int dims[]) {
    int fixed = 1, nfree = 0;
    for (int i = 0; i < ndims; i++) {
        if (dims[i] < 0) return MPI_ERR_OTHER;
        if (dims[i] > 0) fixed *= dims[i]; else nfree++;
    }
    if (fixed <= 0 || nnodes % fixed != 0) return MPI_ERR_OTHER;
    int rem = nnodes / fixed;
    if (nfree == 0) return rem == 1 ? MPI_SUCCESS : MPI_ERR_OTHER;
    int fac[32], nf = 0;
    for (int p = 2; rem > 1 && nf < 32; ) {
        if (rem % p == 0) { fac[nf++] = p; rem /= p; }
        else p = (p == 2) ? 3 : p + 2;
        if ((long long)p * p > rem && rem > 1) { fac[nf++] = rem; rem = 1; }
    }
    int vals[16];
    for (int i = 0; i < nfree && i < 16; i++) vals[i] = 1;
    for (int f = nf - 1; f >= 0; f--) {
        int m = 0;
        for (int i = 1; i < nfree; i++) if (vals[i] < vals[m]) m = i;
        vals[m] *= fac[f];
    }
    for (int i = 0; i < nfree; i++)
        for (int j = i + 1; j < nfree; j++)
            if (vals[j] > vals[i]) { int t = vals[i]; vals[i] = vals[j]; vals[j] = t; }
    for (int i = 0, k = 0; i < ndims; i++) if (dims[i] == 0) dims[i] = vals[k++];
    return MPI_SUCCESS;
}

// ---------------------------------------------------------------- groups (group.cu)
__device__ int MPI_Group_size(MPI_Group group, int* size) {
    if (!group) return MPI_ERR_OTHER;
    *size = group->size;
    return MPI_SUCCESS;
}
__device__ int MPI_Group_rank(MPI_Group group, int* rank) {  // group.cu:110-121
    if (!group) return MPI_ERR_OTHER;
    const int w = world_rank();
    int64_t pos = INT64_MAX;
    for (int i = lane(); i < group->size; i += kWave) {
        int gr = group->ranks ? group->ranks[i] : i;
        if (gr == w && i < pos) pos = i;
    }
    pos = wave_min(pos);
    *rank = pos == INT64_MAX ? MPI_UNDEFINED : (int)pos;
    return MPI_SUCCESS;
}
__device__ int MPI_Group_incl(MPI_Group group, int n, const int ranks[], MPI_Group* newgroup) {  // 57-68
    GPUMPI_ENTER();
    if (!group || n < 0) return MPI_ERR_OTHER;
    GroupImpl* g = (GroupImpl*)rt_alloc(sizeof(GroupImpl) + sizeof(int32_t) * (uint64_t)(n + 1));
    if (!g) return MPI_ERR_OTHER;
    int32_t* rk = (int32_t*)(g + 1);
    for (int i = 0; i < n; i++) rk[i] = group->ranks ? group->ranks[ranks[i]] : ranks[i];
    g->size = n;
    g->refcount = 1;
    g->ranks = rk;
    *newgroup = g;
    return MPI_SUCCESS;
}
__device__ int MPI_Group_range_incl(MPI_Group group, int n, int ranges[][3], MPI_Group* newgroup) {  // 70-93
    GPUMPI_ENTER();
    if (!group) return MPI_ERR_OTHER;
    int total = 0;
    for (int i = 0; i < n; i++) {
        int first = ranges[i][0], last = ranges[i][1], stride = ranges[i][2];
        if (stride == 0) return MPI_ERR_OTHER;
        int cnt = (last - first) / stride;
        if (cnt >= 0) total += cnt + 1;
    }
    GroupImpl* g = (GroupImpl*)rt_alloc(sizeof(GroupImpl) + sizeof(int32_t) * (uint64_t)(total + 1));
    if (!g) return MPI_ERR_OTHER;
    int32_t* rk = (int32_t*)(g + 1);
    int k = 0;
    for (int i = 0; i < n; i++) {
        int first = ranges[i][0], last = ranges[i][1], stride = ranges[i][2];
        int cnt = (last - first) / stride;
        for (int j = 0; j <= cnt; j++) {
            int r = first + j * stride;
            rk[k++] = group->ranks ? group->ranks[r] : r;
        }
    }
    g->size = total;
    g->refcount = 1;
    g->ranks = rk;
    *newgroup = g;
    return MPI_SUCCESS;
}
__device__ int MPI_Group_free(MPI_Group* group) {  // group.cu:95-108
    GPUMPI_ENTER();
    if (!group || !*group) return MPI_SUCCESS;
    GroupImpl* g = *group;
    *group = MPI_GROUP_NULL;
    if (g->refcount > 0) {
        g->refcount = g->refcount - 1;
        if (g->refcount == 0) rt_free(g);
    }
    return MPI_SUCCESS;
}

// ---------------------------------------------------------------- attributes (stubs, as reference)
__device__ int MPI_NULL_COPY_FN(MPI_Comm, int, void*, void*, void*, int* flag) {
    if (flag) *flag = 0;
    return MPI_SUCCESS;
}
__device__ int MPI_NULL_DELETE_FN(MPI_Comm, int, void*, void*) { return MPI_SUCCESS; }
__device__ int MPI_Keyval_create(MPI_Copy_function*, MPI_Delete_function*, int* keyval, void*) {
    if (keyval) *keyval = 0;
    return MPI_SUCCESS;
}
__device__ int MPI_Attr_get(MPI_Comm, int, void*, int* flag) {
    if (flag) *flag = 0;
    return MPI_SUCCESS;
}
__device__ int MPI_Attr_put(MPI_Comm, int, void*) { return MPI_SUCCESS; }

// ---------------------------------------------------------------- device heap API
// A rank that runs a served job allocates from the job's arena (256-byte granules, one
// atomic bump shared by the job's ranks); frees inside the arena are no-ops because the
// scheduler releases the arena in one piece when the job ends.
__device__ void* __gpu_malloc(size_t size) {
    JobArena* a = my_state()->arena;
    if (a) {
        if (size == 0) return nullptr;
        const uint64_t n = ((uint64_t)size + 255) & ~255ull;
        uint64_t off = 0;
        if (lane() == 0) off = fetch_add_relaxed(&a->bump, n);
        off = bcast(off);
        return off + n <= a->size ? a->base + off : nullptr;
    }
    return heap_alloc(&rt()->heap, &my_state()->cache, size);
}
__device__ void __gpu_free(void* p) {
    JobArena* a = my_state()->arena;
    if (a && (char*)p >= a->base && (char*)p < a->base + a->size) return;
    heap_free(&rt()->heap, &my_state()->cache, p);
}
__device__ void* __gpu_calloc(size_t nmemb, size_t size) {
    uint64_t n = (uint64_t)nmemb * size;
    void* p = __gpu_malloc(n);
    if (p) {
        char* c = (char*)p;
        uint64_t i = (uint64_t)lane() * 16;
        for (; i + 16 <= n; i += kWave * 16) *(u32x4*)(c + i) = u32x4{0u, 0u, 0u, 0u};
        for (uint64_t j = (n & ~(uint64_t)15) + lane(); j < n; j += kWave) c[j] = 0;
    }
    return p;
}
__device__ void* __gpu_realloc(void* ptr, size_t size) {
    if (!ptr) return __gpu_malloc(size);
    uint64_t old = heap_block_bytes(&rt()->heap, ptr);
    if (size <= old) return ptr;
    void* p = __gpu_malloc(size);
    if (p) {
        wave_copy(p, ptr, old);
        __gpu_free(ptr);
    }
    return p;
}
#define GPUMPI_MALLOC_VARIANT(suffix)                                                     \
    __device__ void* __gpu_malloc_##suffix(size_t size, bool) { return __gpu_malloc(size); } \
    __device__ void __gpu_free_##suffix(void* p) { __gpu_free(p); }
GPUMPI_MALLOC_VARIANT(coalesce)
GPUMPI_MALLOC_VARIANT(v1)
GPUMPI_MALLOC_VARIANT(v2)
GPUMPI_MALLOC_VARIANT(v3)
GPUMPI_MALLOC_VARIANT(v4)
GPUMPI_MALLOC_VARIANT(v5)
GPUMPI_MALLOC_VARIANT(v6)
#undef GPUMPI_MALLOC_VARIANT
__device__ void __gpu_init_malloc() {}
__device__ void __gpu_clean_malloc() {}
__device__ void __gpu_init_malloc_v6() {}
__device__ void __gpu_clean_malloc_v6() {}
__device__ long long __gpu_heap_in_use() { return (long long)ld_relaxed(&rt()->heap.bytes_in_use); }
__device__ long long __gpu_heap_capacity() { return (long long)rt()->heap.bytes; }
