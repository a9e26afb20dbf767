// gpumpi device MPI API for AMD Instinct MI355X (gfx950).
//
// Drop-in surface for converted MPI programs: the names, C types and argument meanings
// follow the reference device library, so a program converted for it compiles against
// this header unchanged.  Every function is a __device__ function called wave-uniformly
// by the wavefront that is the calling rank.
//
// Reference interface replaced, declaration by declaration:
//   constants, MPI_Status, MPI_Request ......... gpu_libs/gpu_mpi/mpi.cuh:1-36, 145-154,
//                                                gpu_libs/gpu_mpi/mpi_common.cuh:1-11
//   MPI_Datatype (enum, same order) ............ gpu_libs/gpu_mpi/datatypes.cuh:12-14,
//                                                gpu_libs/gpu_mpi/mpi_types_list.cuh:4-33
//   MPI_Op and the twelve built-in operators ... gpu_libs/gpu_mpi/operators.cuh:7-28,
//                                                gpu_libs/gpu_mpi/mpi_operators_list.cuh:4-16
//   MPI_Comm, MPI_COMM_WORLD/NULL, Comm_* ...... gpu_libs/gpu_mpi/communicator.cuh:6-49
//   MPI_Group, Group_* ......................... gpu_libs/gpu_mpi/group.cuh:1-31
//   point-to-point, collectives, misc .......... gpu_libs/gpu_mpi/mpi.cuh:38-145
//   MPI_Alltoall(v)_{coalesce,v1..v6} .......... gpu_libs/gpu_mpi/mpi.cuh:80-95 (all share
//                                                one implementation: the coalescing variants
//                                                only differed in which allocator served the
//                                                request arrays, mpi.cu:410-482)
//
// Differences on purpose, where the reference deviates from MPI semantics: standard sends complete only once the receiver no longer needs the
// sender's buffer; MPI_ANY_SOURCE / MPI_ANY_TAG match already-arrived messages;
// MPI_Sendrecv, Barrier on any communicator, Testall, Waitsome and Dims_create work;
// Wait(MPI_REQUEST_NULL) and Waitall(MPI_STATUSES_IGNORE) are legal; FP64 reductions use
// a fixed rank order and are deterministic.  Unsupported entry points record an error
// and return MPI_ERR_OTHER instead of trapping the whole GPU.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define MPI_MAX_PROCESSOR_NAME 256
#define MPI_ANY_TAG (-1)
#define MPI_ANY_SOURCE (-1)
#define MPI_TAG_UB ((1 << 15) - 1)

#define MPI_SUCCESS 0
#define MPI_FAILURE -1
#define MPI_ERR_OTHER 1
#define MPI_ERR_TRUNCATE 2
#define MPI_UNDEFINED -1

#define MPI_THREAD_SINGLE 1
#define MPI_THREAD_FUNNELED 2
#define MPI_THREAD_SERIALIZED 3
#define MPI_THREAD_MULTIPLE 4

enum MPI_Datatype {
    MPI_CHAR, MPI_SHORT, MPI_INT, MPI_LONG, MPI_LONG_LONG_INT, MPI_LONG_LONG,
    MPI_SIGNED_CHAR, MPI_UNSIGNED_CHAR, MPI_UNSIGNED_SHORT, MPI_UNSIGNED, MPI_UNSIGNED_LONG,
    MPI_UNSIGNED_LONG_LONG, MPI_FLOAT, MPI_DOUBLE, MPI_LONG_DOUBLE, MPI_WCHAR, MPI_C_BOOL,
    MPI_INT8_T, MPI_INT16_T, MPI_INT32_T, MPI_INT64_T, MPI_UINT8_T, MPI_UINT16_T,
    MPI_UINT32_T, MPI_UINT64_T, MPI_C_COMPLEX, MPI_C_FLOAT_COMPLEX, MPI_C_DOUBLE_COMPLEX,
    MPI_C_LONG_DOUBLE_COMPLEX, MPI_BYTE
};

struct MPI_Status {
    int MPI_SOURCE = MPI_ANY_SOURCE;
    int MPI_TAG = MPI_ANY_TAG;
    int MPI_ERROR = MPI_SUCCESS;
    long long gpumpi_bytes = 0;  // received byte count (MPI_Get_count)
};
#define MPI_STATUSES_IGNORE ((MPI_Status*)nullptr)
#define MPI_STATUS_IGNORE ((MPI_Status*)nullptr)

namespace gpumpi { struct Request; struct CommImpl; struct GroupImpl; struct OpImpl; }
typedef gpumpi::Request* MPI_Request;
#define MPI_REQUEST_NULL ((MPI_Request)nullptr)

typedef gpumpi::CommImpl* MPI_Comm;
typedef gpumpi::GroupImpl* MPI_Group;
typedef gpumpi::OpImpl* MPI_Op;

typedef void MPI_User_function(void* invec, void* inoutvec, int* len, MPI_Datatype* datatype);
#define MPI_OP_NULL ((MPI_Op)nullptr)

namespace gpumpi {
struct OpImpl {
    int builtin;                 // 0..11 for MPI_MAX..MPI_MINLOC, -1 for user ops
    MPI_User_function* user_fn;
    int commute;
};
__device__ MPI_Comm world_comm();
__device__ MPI_Comm null_comm();
__device__ MPI_Group world_group();
__device__ MPI_Group empty_group();
}  // namespace gpumpi

// The built-in operators are static objects, valid before MPI_Init.
extern __device__ gpumpi::OpImpl gpumpi_builtin_ops[12];
#define MPI_MAX (&gpumpi_builtin_ops[0])
#define MPI_MIN (&gpumpi_builtin_ops[1])
#define MPI_SUM (&gpumpi_builtin_ops[2])
#define MPI_PROD (&gpumpi_builtin_ops[3])
#define MPI_LAND (&gpumpi_builtin_ops[4])
#define MPI_BAND (&gpumpi_builtin_ops[5])
#define MPI_LOR (&gpumpi_builtin_ops[6])
#define MPI_BOR (&gpumpi_builtin_ops[7])
#define MPI_LXOR (&gpumpi_builtin_ops[8])
#define MPI_BXOR (&gpumpi_builtin_ops[9])
#define MPI_MAXLOC (&gpumpi_builtin_ops[10])
#define MPI_MINLOC (&gpumpi_builtin_ops[11])

// Communicators and groups of the running rank grid.
#define MPI_COMM_WORLD (gpumpi::world_comm())
#define MPI_COMM_NULL ((MPI_Comm)nullptr)
#define MPI_GROUP_WORLD (gpumpi::world_group())
#define MPI_GROUP_EMPTY (gpumpi::empty_group())
#define MPI_GROUP_NULL ((MPI_Group)nullptr)

// Prints once per rank (a rank is a wavefront; plain printf would print from all 64 lanes).
#define gpumpi_printf(...)                           \
    do {                                             \
        if (__lane_id() == 0) printf(__VA_ARGS__);   \
    } while (0)

// ---- environment ----
__device__ int MPI_Init(int* argc, char*** argv);
__device__ int MPI_Init_thread(int* argc, char*** argv, int required, int* provided);
__device__ int MPI_Initialized(int* flag);
__device__ int MPI_Finalize(void);
__device__ int MPI_Abort(MPI_Comm comm, int errorcode);
// exit(status) of a converted program (gpu_libs/gpu_libc/stdlib.cu:55-58): ends the
// calling rank at once; a non-zero status makes the program's exit status 1, as a
// non-zero return from main does.  Ranks still waiting on this one are not released.
__device__ __attribute__((noreturn)) void gpumpi_exit(int status);
// File I/O of converted programs, delegated to the launcher's host thread: one request
// (IoOp codes of the runtime) with up to 4096 bytes each way; returns the C library
// call's result, -1 on failure.
__device__ int64_t gpumpi_io(int op, int64_t a0, int64_t a1, int64_t a2, const void* in1,
                             int len1, const void* in2, int len2, void* out, int outcap);
__device__ int MPI_Get_processor_name(char* name, int* resultlen);
__device__ double MPI_Wtime(void);
__device__ double MPI_Wtick(void);

// ---- point to point ----
__device__ int MPI_Send(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                        MPI_Comm comm);
__device__ int MPI_Recv(void* buf, int count, MPI_Datatype datatype, int source, int tag,
                        MPI_Comm comm, MPI_Status* status);
__device__ int MPI_Isend(const void* buf, int count, MPI_Datatype datatype, int dest, int tag,
                         MPI_Comm comm, MPI_Request* request);
__device__ int MPI_Irecv(void* buf, int count, MPI_Datatype datatype, int source, int tag,
                         MPI_Comm comm, MPI_Request* request);
__device__ int MPI_Sendrecv(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                            int dest, int sendtag, void* recvbuf, int recvcount,
                            MPI_Datatype recvtype, int source, int recvtag, MPI_Comm comm,
                            MPI_Status* status);
__device__ int MPI_Wait(MPI_Request* request, MPI_Status* status);
__device__ int MPI_Waitall(int count, MPI_Request array_of_requests[],
                           MPI_Status* array_of_statuses);
__device__ int MPI_Waitsome(int incount, MPI_Request array_of_requests[], int* outcount,
                            int array_of_indices[], MPI_Status array_of_statuses[]);
__device__ int MPI_Test(MPI_Request* request, int* flag, MPI_Status* status);
__device__ int MPI_Testall(int count, MPI_Request array_of_requests[], int* flag,
                           MPI_Status array_of_statuses[]);
__device__ int MPI_Request_free(MPI_Request* request);
__device__ int MPI_Get_count(const MPI_Status* status, MPI_Datatype datatype, int* count);

// ---- collectives ----
__device__ int MPI_Barrier(MPI_Comm comm);
__device__ int MPI_Bcast(void* buffer, int count, MPI_Datatype datatype, int root,
                         MPI_Comm comm);
__device__ int MPI_Reduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype datatype,
                          MPI_Op op, int root, MPI_Comm comm);
__device__ int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count,
                             MPI_Datatype datatype, MPI_Op op, MPI_Comm comm);
__device__ int MPI_Gather(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                          void* recvbuf, int recvcount, MPI_Datatype recvtype, int root,
                          MPI_Comm comm);
__device__ int MPI_Gatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                           void* recvbuf, const int recvcounts[], const int displs[],
                           MPI_Datatype recvtype, int root, MPI_Comm comm);
__device__ int MPI_Allgather(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                             void* recvbuf, int recvcount, MPI_Datatype recvtype,
                             MPI_Comm comm);
__device__ int MPI_Allgatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                              void* recvbuf, const int recvcounts[], const int displs[],
                              MPI_Datatype recvtype, MPI_Comm comm);
__device__ int MPI_Scatter(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                           void* recvbuf, int recvcount, MPI_Datatype recvtype, int root,
                           MPI_Comm comm);
__device__ int MPI_Scatterv(const void* sendbuf, const int sendcounts[], const int displs[],
                            MPI_Datatype sendtype, void* recvbuf, int recvcount,
                            MPI_Datatype recvtype, int root, MPI_Comm comm);
__device__ int MPI_Alltoall(const void* sendbuf, int sendcount, MPI_Datatype sendtype,
                            void* recvbuf, int recvcount, MPI_Datatype recvtype,
                            MPI_Comm comm);
__device__ int MPI_Alltoallv(const void* sendbuf, const int sendcounts[], const int sdispls[],
                             MPI_Datatype sendtype, void* recvbuf, const int recvcounts[],
                             const int rdispls[], MPI_Datatype recvtype, MPI_Comm comm);

#define GPUMPI_ATA_VARIANT(suffix)                                                          \
    __device__ int MPI_Alltoall_##suffix(const void* sendbuf, int sendcount,                \
                                         MPI_Datatype sendtype, void* recvbuf,              \
                                         int recvcount, MPI_Datatype recvtype,              \
                                         MPI_Comm comm, bool __coalesced = false);          \
    __device__ int MPI_Alltoallv_##suffix(const void* sendbuf, const int sendcounts[],      \
                                          const int sdispls[], MPI_Datatype sendtype,       \
                                          void* recvbuf, const int recvcounts[],            \
                                          const int rdispls[], MPI_Datatype recvtype,       \
                                          MPI_Comm comm, bool __coalesced = false);
GPUMPI_ATA_VARIANT(coalesce)
GPUMPI_ATA_VARIANT(v1)
GPUMPI_ATA_VARIANT(v2)
GPUMPI_ATA_VARIANT(v3)
GPUMPI_ATA_VARIANT(v4)
GPUMPI_ATA_VARIANT(v5)
GPUMPI_ATA_VARIANT(v6)
#undef GPUMPI_ATA_VARIANT

// ---- operators and datatypes ----
__device__ int MPI_Op_create(MPI_User_function* user_fn, int commute, MPI_Op* op);
__device__ int MPI_Op_free(MPI_Op* op);
__device__ int MPI_Type_size(MPI_Datatype datatype, int* size);
__device__ int MPI_Type_contiguous(int count, MPI_Datatype oldtype, MPI_Datatype* newtype);
__device__ int MPI_Type_commit(MPI_Datatype* datatype);

// ---- communicators and groups ----
__device__ int MPI_Comm_size(MPI_Comm comm, int* size);
__device__ int MPI_Comm_rank(MPI_Comm comm, int* rank);
__device__ int MPI_Comm_group(MPI_Comm comm, MPI_Group* group);
__device__ int MPI_Comm_create(MPI_Comm comm, MPI_Group group, MPI_Comm* newcomm);
__device__ int MPI_Comm_dup(MPI_Comm comm, MPI_Comm* newcomm);
__device__ int MPI_Comm_split(MPI_Comm comm, int color, int key, MPI_Comm* newcomm);
__device__ int MPI_Comm_free(MPI_Comm* comm);
__device__ int MPI_Cart_create(MPI_Comm comm_old, int ndims, const int dims[],
                               const int periods[], int reorder, MPI_Comm* comm_cart);
__device__ int MPI_Cart_sub(MPI_Comm comm, const int remain_dims[], MPI_Comm* comm_new);
__device__ int MPI_Dims_create(int nnodes, int ndims, int dims[]);
__device__ int MPI_Group_incl(MPI_Group group, int n, const int ranks[], MPI_Group* newgroup);
__device__ int MPI_Group_range_incl(MPI_Group group, int n, int ranges[][3],
                                    MPI_Group* newgroup);
__device__ int MPI_Group_free(MPI_Group* group);
__device__ int MPI_Group_size(MPI_Group group, int* size);
__device__ int MPI_Group_rank(MPI_Group group, int* rank);

typedef int MPI_Copy_function(MPI_Comm oldcomm, int keyval, void* extra_state,
                              void* attribute_val_in, void* attribute_val_out, int* flag);
typedef int MPI_Delete_function(MPI_Comm comm, int keyval, void* attribute_val,
                                void* extra_state);
__device__ int MPI_NULL_COPY_FN(MPI_Comm oldcomm, int keyval, void* extra_state,
                                void* attribute_val_in, void* attribute_val_out, int* flag);
__device__ int MPI_NULL_DELETE_FN(MPI_Comm comm, int keyval, void* attribute_val,
                                  void* extra_state);
__device__ int MPI_Keyval_create(MPI_Copy_function* copy_fn, MPI_Delete_function* delete_fn,
                                 int* keyval, void* extra_state);
__device__ int MPI_Attr_get(MPI_Comm comm, int keyval, void* attribute_val, int* flag);
__device__ int MPI_Attr_put(MPI_Comm comm, int keyval, void* attribute_val);
