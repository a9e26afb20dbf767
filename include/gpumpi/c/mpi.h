/* C declarations of the gpumpi MPI API, for the source converter's parse step.
 *
 * The converter reads a C program through clang's AST; the program's `#include "mpi.h"`
 * resolves here during that step (the converted program is compiled against
 * include/gpumpi/mpi.h).  Names, argument types and constants are those of
 * include/gpumpi/mpi.h; handles are opaque pointers as there.
 */
#ifndef GPUMPI_C_MPI_H
#define GPUMPI_C_MPI_H

#include <stddef.h>
#include <stdint.h>

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

typedef enum MPI_Datatype { MPI_CHAR, MPI_SHORT, MPI_INT, MPI_LONG, MPI_LONG_LONG_INT, MPI_LONG_LONG, MPI_SIGNED_CHAR, MPI_UNSIGNED_CHAR, MPI_UNSIGNED_SHORT, MPI_UNSIGNED, MPI_UNSIGNED_LONG, MPI_UNSIGNED_LONG_LONG, MPI_FLOAT, MPI_DOUBLE, MPI_LONG_DOUBLE, MPI_WCHAR, MPI_C_BOOL, MPI_INT8_T, MPI_INT16_T, MPI_INT32_T, MPI_INT64_T, MPI_UINT8_T, MPI_UINT16_T, MPI_UINT32_T, MPI_UINT64_T, MPI_C_COMPLEX, MPI_C_FLOAT_COMPLEX, MPI_C_DOUBLE_COMPLEX, MPI_C_LONG_DOUBLE_COMPLEX, MPI_BYTE } MPI_Datatype;

typedef struct MPI_Status {
    int MPI_SOURCE;
    int MPI_TAG;
    int MPI_ERROR;
    long long gpumpi_bytes;
} MPI_Status;
#define MPI_STATUSES_IGNORE ((MPI_Status*)0)
#define MPI_STATUS_IGNORE ((MPI_Status*)0)

typedef struct gpumpi_Request* MPI_Request;
typedef struct gpumpi_CommImpl* MPI_Comm;
typedef struct gpumpi_GroupImpl* MPI_Group;
typedef struct gpumpi_OpImpl* MPI_Op;
#define MPI_REQUEST_NULL ((MPI_Request)0)
#define MPI_OP_NULL ((MPI_Op)0)
#define MPI_COMM_NULL ((MPI_Comm)0)
#define MPI_GROUP_NULL ((MPI_Group)0)
extern struct gpumpi_CommImpl* gpumpi_c_world_comm;
extern struct gpumpi_GroupImpl* gpumpi_c_world_group;
extern struct gpumpi_GroupImpl* gpumpi_c_empty_group;
#define MPI_COMM_WORLD gpumpi_c_world_comm
#define MPI_GROUP_WORLD gpumpi_c_world_group
#define MPI_GROUP_EMPTY gpumpi_c_empty_group
extern MPI_Op gpumpi_c_builtin_ops[12];
#define MPI_MAX (gpumpi_c_builtin_ops[0])
#define MPI_MIN (gpumpi_c_builtin_ops[1])
#define MPI_SUM (gpumpi_c_builtin_ops[2])
#define MPI_PROD (gpumpi_c_builtin_ops[3])
#define MPI_LAND (gpumpi_c_builtin_ops[4])
#define MPI_BAND (gpumpi_c_builtin_ops[5])
#define MPI_LOR (gpumpi_c_builtin_ops[6])
#define MPI_BOR (gpumpi_c_builtin_ops[7])
#define MPI_LXOR (gpumpi_c_builtin_ops[8])
#define MPI_BXOR (gpumpi_c_builtin_ops[9])
#define MPI_MAXLOC (gpumpi_c_builtin_ops[10])
#define MPI_MINLOC (gpumpi_c_builtin_ops[11])

typedef void MPI_User_function(void* invec, void* inoutvec, int* len, MPI_Datatype* datatype);
typedef int MPI_Copy_function(MPI_Comm oldcomm, int keyval, void* extra_state,
                              void* attribute_val_in, void* attribute_val_out, int* flag);
typedef int MPI_Delete_function(MPI_Comm comm, int keyval, void* attribute_val,
                                void* extra_state);
int MPI_NULL_COPY_FN(MPI_Comm oldcomm, int keyval, void* extra_state, void* attribute_val_in,
                     void* attribute_val_out, int* flag);
int MPI_NULL_DELETE_FN(MPI_Comm comm, int keyval, void* attribute_val, void* extra_state);

int MPI_Init(int* argc, char*** argv);
int MPI_Init_thread(int* argc, char*** argv, int required, int* provided);
int MPI_Initialized(int* flag);
int MPI_Finalize(void);
int MPI_Abort(MPI_Comm comm, int errorcode);
int MPI_Get_processor_name(char* name, int* resultlen);
double MPI_Wtime(void);
double MPI_Wtick(void);
int MPI_Send(const void* buf, int count, MPI_Datatype datatype, int dest, int tag, MPI_Comm comm);
int MPI_Recv(void* buf, int count, MPI_Datatype datatype, int source, int tag, MPI_Comm comm, MPI_Status* status);
int MPI_Isend(const void* buf, int count, MPI_Datatype datatype, int dest, int tag, MPI_Comm comm, MPI_Request* request);
int MPI_Irecv(void* buf, int count, MPI_Datatype datatype, int source, int tag, MPI_Comm comm, MPI_Request* request);
int MPI_Sendrecv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, int dest, int sendtag, void* recvbuf, int recvcount, MPI_Datatype recvtype, int source, int recvtag, MPI_Comm comm, MPI_Status* status);
int MPI_Wait(MPI_Request* request, MPI_Status* status);
int MPI_Waitall(int count, MPI_Request array_of_requests[], MPI_Status* array_of_statuses);
int MPI_Waitsome(int incount, MPI_Request array_of_requests[], int* outcount, int array_of_indices[], MPI_Status array_of_statuses[]);
int MPI_Test(MPI_Request* request, int* flag, MPI_Status* status);
int MPI_Testall(int count, MPI_Request array_of_requests[], int* flag, MPI_Status array_of_statuses[]);
int MPI_Request_free(MPI_Request* request);
int MPI_Get_count(const MPI_Status* status, MPI_Datatype datatype, int* count);
int MPI_Barrier(MPI_Comm comm);
int MPI_Bcast(void* buffer, int count, MPI_Datatype datatype, int root, MPI_Comm comm);
int MPI_Reduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype datatype, MPI_Op op, int root, MPI_Comm comm);
int MPI_Allreduce(const void* sendbuf, void* recvbuf, int count, MPI_Datatype datatype, MPI_Op op, MPI_Comm comm);
int MPI_Gather(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf, int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm);
int MPI_Gatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf, const int recvcounts[], const int displs[], MPI_Datatype recvtype, int root, MPI_Comm comm);
int MPI_Allgather(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf, int recvcount, MPI_Datatype recvtype, MPI_Comm comm);
int MPI_Allgatherv(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf, const int recvcounts[], const int displs[], MPI_Datatype recvtype, MPI_Comm comm);
int MPI_Scatter(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf, int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm);
int MPI_Scatterv(const void* sendbuf, const int sendcounts[], const int displs[], MPI_Datatype sendtype, void* recvbuf, int recvcount, MPI_Datatype recvtype, int root, MPI_Comm comm);
int MPI_Alltoall(const void* sendbuf, int sendcount, MPI_Datatype sendtype, void* recvbuf, int recvcount, MPI_Datatype recvtype, MPI_Comm comm);
int MPI_Alltoallv(const void* sendbuf, const int sendcounts[], const int sdispls[], MPI_Datatype sendtype, void* recvbuf, const int recvcounts[], const int rdispls[], MPI_Datatype recvtype, MPI_Comm comm);
int MPI_Op_create(MPI_User_function* user_fn, int commute, MPI_Op* op);
int MPI_Op_free(MPI_Op* op);
int MPI_Type_size(MPI_Datatype datatype, int* size);
int MPI_Type_contiguous(int count, MPI_Datatype oldtype, MPI_Datatype* newtype);
int MPI_Type_commit(MPI_Datatype* datatype);
int MPI_Comm_size(MPI_Comm comm, int* size);
int MPI_Comm_rank(MPI_Comm comm, int* rank);
int MPI_Comm_group(MPI_Comm comm, MPI_Group* group);
int MPI_Comm_create(MPI_Comm comm, MPI_Group group, MPI_Comm* newcomm);
int MPI_Comm_dup(MPI_Comm comm, MPI_Comm* newcomm);
int MPI_Comm_split(MPI_Comm comm, int color, int key, MPI_Comm* newcomm);
int MPI_Comm_free(MPI_Comm* comm);
int MPI_Cart_create(MPI_Comm comm_old, int ndims, const int dims[], const int periods[], int reorder, MPI_Comm* comm_cart);
int MPI_Cart_sub(MPI_Comm comm, const int remain_dims[], MPI_Comm* comm_new);
int MPI_Dims_create(int nnodes, int ndims, int dims[]);
int MPI_Group_incl(MPI_Group group, int n, const int ranks[], MPI_Group* newgroup);
int MPI_Group_range_incl(MPI_Group group, int n, int ranges[][3], MPI_Group* newgroup);
int MPI_Group_free(MPI_Group* group);
int MPI_Group_size(MPI_Group group, int* size);
int MPI_Group_rank(MPI_Group group, int* rank);
int MPI_NULL_COPY_FN(MPI_Comm oldcomm, int keyval, void* extra_state, void* attribute_val_in, void* attribute_val_out, int* flag);
int MPI_NULL_DELETE_FN(MPI_Comm comm, int keyval, void* attribute_val, void* extra_state);
int MPI_Keyval_create(MPI_Copy_function* copy_fn, MPI_Delete_function* delete_fn, int* keyval, void* extra_state);
int MPI_Attr_get(MPI_Comm comm, int keyval, void* attribute_val, int* flag);
int MPI_Attr_put(MPI_Comm comm, int keyval, void* attribute_val);

#endif
