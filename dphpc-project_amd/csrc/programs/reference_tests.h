// Rank programs restating the reference's device unit tests (Catch2 kernels run with one
// rank per GPU thread).  Each rank writes ok[rank] = 1 on success.  The known answers
// are the ones the reference tests assert:
//   tests/cuda_mpi_test.cu:5-317      transport: single int 3456, 16-int arrays, crossed
//                                     send/recv, repeated sizes, 3-rank flood, all-to-all
//   tests/collectives_test.cu:6-288   Reduce PROD int16, Allreduce SUM double (|err|<=1e-5),
//                                     Allgather, Scatter (root 2), Scatterv (root 2)
//   tests/communicator_test.cu:5-152  Comm_rank, Comm_dup context isolation, Comm_split
#pragma once

#include "../runtime/mpi_api.h"

namespace gpumpi {
namespace progs {

struct OkArgs { int* ok; };

__device__ __forceinline__ void set_ok(bool v) {
    const OkArgs* a = (const OkArgs*)rt()->user_args;
    a->ok[world_rank()] = v ? 1 : 0;
}

// cuda_mpi_test.cu:5-36
__device__ int t_single_int(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank = -1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = true;
    if (rank == 0) {
        int x = 3456;
        MPI_Send(&x, 1, MPI_INT, 1, 15, MPI_COMM_WORLD);
    } else if (rank == 1) {
        int x = 0;
        MPI_Recv(&x, 1, MPI_INT, 0, 15, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        ok = x == 3456;
    }
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// cuda_mpi_test.cu:38-73
__device__ int t_array(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank = -1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = true;
    if (rank == 0) {
        int x[16];
        for (int i = 0; i < 16; i++) x[i] = i;
        MPI_Request r;
        MPI_Isend(x, 16, MPI_INT, 1, 15, MPI_COMM_WORLD, &r);
        MPI_Wait(&r, MPI_STATUS_IGNORE);
    } else if (rank == 1) {
        int x[16] = {};
        MPI_Request r;
        MPI_Irecv(x, 16, MPI_INT, 0, 15, MPI_COMM_WORLD, &r);
        MPI_Wait(&r, MPI_STATUS_IGNORE);
        for (int i = 0; i < 16; i++) ok = ok && x[i] == i;
    }
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// cuda_mpi_test.cu:75-127
__device__ int t_sendrecv(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank = -1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    bool ok = true;
    if (rank < 2) {
        int other = 1 - rank;
        int x[16], y[16] = {};
        for (int i = 0; i < 16; i++) x[i] = rank == 0 ? i : -i;
        MPI_Request op[2];
        MPI_Irecv(y, 16, MPI_INT, other, 15, MPI_COMM_WORLD, &op[0]);
        MPI_Isend(x, 16, MPI_INT, other, 15, MPI_COMM_WORLD, &op[1]);
        MPI_Wait(&op[0], MPI_STATUS_IGNORE);
        for (int i = 0; i < 16; i++) ok = ok && y[i] == (rank == 0 ? -i : i);
        MPI_Wait(&op[1], MPI_STATUS_IGNORE);
    }
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// cuda_mpi_test.cu:129-174
__device__ int t_repeat(int, char**) {
    MPI_Init(nullptr, nullptr);
    const int numRanks = 2, numRepeats = 5, dataSize = 1 << (numRepeats - 1);
    int me = world_rank();
    int other = (me + 1) % numRanks;
    int local[dataSize], remote[dataSize] = {};
    for (int i = 0; i < dataSize; i++) local[i] = (me + 1) * i;
    MPI_Request s[numRepeats], r[numRepeats];
    for (int i = 0; i < numRepeats; i++) {
        MPI_Isend(local, 1 << i, MPI_INT, other, i + 10, MPI_COMM_WORLD, &s[i]);
        MPI_Irecv(remote, 1 << i, MPI_INT, other, i + 10, MPI_COMM_WORLD, &r[i]);
    }
    for (int i = 0; i < numRepeats; i++) {
        MPI_Wait(&s[i], MPI_STATUS_IGNORE);
        MPI_Wait(&r[i], MPI_STATUS_IGNORE);
    }
    bool ok = true;
    for (int i = 0; i < dataSize; i++) ok = ok && remote[i] == (other + 1) * i;
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// cuda_mpi_test.cu:177-233
__device__ int t_flood(int, char**) {
    MPI_Init(nullptr, nullptr);
    const int numRanks = 3, numRepeats = 5, dataSize = 1 << (numRepeats - 1);
    int me = world_rank();
    int next = (me + 1) % numRanks, prev = (numRanks + me - 1) % numRanks;
    int local[dataSize], prevData[dataSize] = {}, nextData[dataSize] = {};
    for (int i = 0; i < dataSize; i++) local[i] = (me + 1) * i;
    MPI_Request sn[numRepeats], rn[numRepeats], sp[numRepeats], rp[numRepeats];
    for (int i = 0; i < numRepeats; i++) {
        int tag = i + 10;
        MPI_Isend(local, 1 << i, MPI_INT, next, tag, MPI_COMM_WORLD, &sn[i]);
        MPI_Irecv(nextData, 1 << i, MPI_INT, next, tag, MPI_COMM_WORLD, &rn[i]);
        MPI_Isend(local, 1 << i, MPI_INT, prev, tag, MPI_COMM_WORLD, &sp[i]);
        MPI_Irecv(prevData, 1 << i, MPI_INT, prev, tag, MPI_COMM_WORLD, &rp[i]);
    }
    for (int i = 0; i < numRepeats; i++) {
        MPI_Wait(&rp[i], MPI_STATUS_IGNORE);
        MPI_Wait(&sn[i], MPI_STATUS_IGNORE);
        MPI_Wait(&sp[i], MPI_STATUS_IGNORE);
        MPI_Wait(&rn[i], MPI_STATUS_IGNORE);
    }
    bool ok = true;
    for (int i = 0; i < dataSize; i++) {
        ok = ok && prevData[i] == (prev + 1) * i;
        ok = ok && nextData[i] == (next + 1) * i;
    }
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// cuda_mpi_test.cu:235-317 (heap buffers, 20 repeats, every pair in flight at once)
__device__ int t_p2p_all_to_all(int, char**) {
    MPI_Init(nullptr, nullptr);
    int me = world_rank();
    const int numRepeats = 20, dataSize = 16;
    int P;
    MPI_Comm_size(MPI_COMM_WORLD, &P);
    int* localData = (int*)__gpu_malloc(sizeof(int) * dataSize * numRepeats * P);
    int* remoteData = (int*)__gpu_malloc(sizeof(int) * dataSize * numRepeats * P);
    MPI_Request* sreq = (MPI_Request*)__gpu_malloc(sizeof(MPI_Request) * numRepeats * P);
    MPI_Request* rreq = (MPI_Request*)__gpu_malloc(sizeof(MPI_Request) * numRepeats * P);
    for (int r = 0; r < P; r++)
        for (int rep = 0; rep < numRepeats; rep++)
            for (int i = lane(); i < dataSize; i += kWave) {
                int idx = i + rep * dataSize + r * dataSize * numRepeats;
                localData[idx] = i * rep * r;
                remoteData[idx] = 0;
            }
    for (int rep = 0; rep < numRepeats; rep++)
        for (int r = 0; r < P; r++)
            if (r != me) {
                int idx = rep * dataSize + r * dataSize * numRepeats;
                MPI_Isend(localData + idx, dataSize, MPI_INT, r, 15, MPI_COMM_WORLD,
                          &sreq[r + rep * P]);
                MPI_Irecv(remoteData + idx, dataSize, MPI_INT, r, 15, MPI_COMM_WORLD,
                          &rreq[r + rep * P]);
            }
    for (int rep = 0; rep < numRepeats; rep++)
        for (int r = 0; r < P; r++)
            if (r != me) {
                MPI_Wait(&sreq[r + rep * P], MPI_STATUS_IGNORE);
                MPI_Wait(&rreq[r + rep * P], MPI_STATUS_IGNORE);
            }
    bool ok = true;
    for (int rep = 0; rep < numRepeats; rep++)
        for (int r = 0; r < P; r++)
            if (r != me)
                for (int i = 0; i < dataSize; i++) {
                    int idx = i + rep * dataSize + r * dataSize * numRepeats;
                    ok = ok && remoteData[idx] == i * rep * me;
                }
    set_ok(ok);
    __gpu_free(sreq);
    __gpu_free(rreq);
    __gpu_free(localData);
    __gpu_free(remoteData);
    MPI_Finalize();
    return 0;
}

// collectives_test.cu:6-54
__device__ int t_reduce_prod_int16(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int n = 14;
    int16_t buffer[n] = {}, recv[n] = {};
    for (int i = 0; i < n; i++) buffer[i] = (int16_t)((rank + 1) * (i + 1));
    MPI_Reduce(buffer, recv, n, MPI_INT16_T, MPI_PROD, 0, MPI_COMM_WORLD);
    int16_t expected[n];
    for (int i = 0; i < n; i++) expected[i] = 1;
    for (int r = 0; r < size; r++)
        for (int i = 0; i < n; i++) expected[i] = (int16_t)(expected[i] * (r + 1) * (i + 1));
    bool ok = true;
    if (rank == 0)
        for (int i = 0; i < n; i++) ok = ok && recv[i] == expected[i];
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// collectives_test.cu:57-98
__device__ int t_allreduce_sum_double(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int n = 17;
    double buffer[n] = {}, recv[n] = {};
    for (int i = 0; i < n; i++) buffer[i] = (rank + 1) * (i + 1);
    MPI_Allreduce(buffer, recv, n, MPI_DOUBLE, MPI_SUM, MPI_COMM_WORLD);
    bool ok = true;
    for (int i = 0; i < n; i++) {
        double e = 0;
        for (int r = 0; r < size; r++) e += (r + 1) * (i + 1);
        ok = ok && fabs(recv[i] - e) <= 1e-5;
    }
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// collectives_test.cu:100-141
__device__ int t_allgather(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int n = 7;
    int* buffer = (int*)__gpu_malloc(sizeof(int) * n);
    int* recv = (int*)__gpu_malloc(sizeof(int) * n * size);
    for (int i = 0; i < n; i++) buffer[i] = (rank + 1) * (i + 1);
    MPI_Allgather(buffer, n, MPI_INT, recv, n, MPI_INT, MPI_COMM_WORLD);
    bool ok = true;
    for (int r = 0; r < size; r++)
        for (int i = 0; i < n; i++) ok = ok && recv[r * n + i] == (r + 1) * (i + 1);
    set_ok(ok);
    __gpu_free(buffer);
    __gpu_free(recv);
    MPI_Finalize();
    return 0;
}

// collectives_test.cu:144-196
__device__ int t_scatter(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int root = 2, n = 4;
    int* sendb = (int*)__gpu_malloc(sizeof(int) * n * size);
    int recv[n];
    for (int j = 0; j < size; j++)
        for (int i = 0; i < n; i++) sendb[i + j * n] = rank == root ? j : -2;
    for (int i = 0; i < n; i++) recv[i] = -1;
    MPI_Scatter(sendb, n, MPI_INT, recv, n, MPI_INT, root, MPI_COMM_WORLD);
    bool ok = true;
    for (int i = 0; i < n; i++) ok = ok && recv[i] == rank;
    set_ok(ok);
    __gpu_free(sendb);
    MPI_Finalize();
    return 0;
}

// collectives_test.cu:199-288
__device__ int t_scatterv(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank, size;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const int root = 2;
    int sizes[64], displs[64];
    for (int i = 0; i < size; i++) {
        sizes[i] = i + 1;
        displs[i] = i == 0 ? 0 : displs[i - 1] + sizes[i - 1];
    }
    int total = displs[size - 1] + sizes[size - 1];
    int* sendb = (int*)__gpu_malloc(sizeof(int) * total);
    int* recv = (int*)__gpu_malloc(sizeof(int) * sizes[rank]);
    for (int j = 0; j < size; j++)
        for (int i = 0; i < sizes[j]; i++) sendb[displs[j] + i] = rank == root ? j : -2;
    for (int i = 0; i < sizes[rank]; i++) recv[i] = -1;
    MPI_Scatterv(sendb, sizes, displs, MPI_INT, recv, sizes[rank], MPI_INT, root, MPI_COMM_WORLD);
    MPI_Barrier(MPI_COMM_WORLD);
    bool ok = true;
    for (int i = 0; i < sizes[rank]; i++) ok = ok && recv[i] == rank;
    set_ok(ok);
    __gpu_free(sendb);
    __gpu_free(recv);
    MPI_Finalize();
    return 0;
}

// communicator_test.cu:5-25
__device__ int t_comm_rank(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank = -1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    set_ok(rank == world_rank());
    MPI_Finalize();
    return 0;
}

// communicator_test.cu:27-90: three duplicated communicators keep messages apart
__device__ int t_comm_dup(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm comm1, comm2, comm3;
    MPI_Comm_dup(MPI_COMM_WORLD, &comm1);
    MPI_Comm_dup(MPI_COMM_WORLD, &comm2);
    MPI_Comm_dup(comm2, &comm3);
    MPI_Request request[3] = {MPI_REQUEST_NULL, MPI_REQUEST_NULL, MPI_REQUEST_NULL};
    int16_t data[30] = {};
    if (rank == 0) {
        for (int i = 0; i < 30; i++) data[i] = (int16_t)(i + 1);
        MPI_Isend(data + 0, 10, MPI_INT16_T, 1, 17, comm1, &request[0]);
        MPI_Irecv(data + 10, 10, MPI_INT16_T, 1, 17, comm2, &request[1]);
        MPI_Isend(data + 20, 10, MPI_INT16_T, 1, 17, comm3, &request[2]);
    } else if (rank == 1) {
        for (int i = 0; i < 30; i++) data[i] = (int16_t)-(i + 1);
        MPI_Irecv(data + 20, 10, MPI_INT16_T, 0, 17, comm3, &request[2]);
        MPI_Irecv(data + 0, 10, MPI_INT16_T, 0, 17, comm1, &request[0]);
        MPI_Isend(data + 10, 10, MPI_INT16_T, 0, 17, comm2, &request[1]);
    }
    MPI_Status status[3];
    MPI_Waitall(3, request, status);
    bool ok = true;
    if (rank == 0) {
        for (int i = 10; i < 20; i++) ok = ok && data[i] == -(i + 1);
    } else if (rank == 1) {
        for (int i = 0; i < 10; i++) ok = ok && data[i] == i + 1;
        for (int i = 20; i < 30; i++) ok = ok && data[i] == i + 1;
    }
    MPI_Comm_free(&comm1);
    MPI_Comm_free(&comm2);
    MPI_Comm_free(&comm3);
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

// communicator_test.cu:93-152
__device__ int t_comm_split(int, char**) {
    MPI_Init(nullptr, nullptr);
    int rank = 0;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    int color, key;
    if (rank < 3) { color = 0; key = rank; }
    else if (rank < 6) { color = 1; key = 6 - rank; }
    else if (rank < 9) { color = 2; key = 1; }
    else { color = 2; key = 0; }
    MPI_Comm comm;
    MPI_Comm_split(MPI_COMM_WORLD, color, key, &comm);
    int nr = -1, ns = -1;
    MPI_Comm_rank(comm, &nr);
    MPI_Comm_size(comm, &ns);
    bool ok = true;
    if (rank < 3) ok = rank == nr && ns == 3;
    else if (rank < 6) ok = (5 - rank) == nr && ns == 3;
    else if (rank < 9) ok = (rank - 5) == nr && ns == 4;
    else ok = nr == 0 && ns == 4;
    // the new communicator must also work: a barrier and an allreduce of comm ranks
    MPI_Barrier(comm);
    int sum = 0;
    MPI_Allreduce(&nr, &sum, 1, MPI_INT, MPI_SUM, comm);
    ok = ok && sum == ns * (ns - 1) / 2;
    MPI_Comm_free(&comm);
    set_ok(ok);
    MPI_Finalize();
    return 0;
}

}  // namespace progs
}  // namespace gpumpi
