/*
 * pra_mpi_single.h -- the MPI calls generated parallel-RA programs make in
 * main(), for the one-process-per-GPU backend.
 *
 * The reference's programs call MPI directly for timing and for reductions of
 * their own timers (e.g. tests/transitive_closure.cpp:14,107-110).  With one
 * process driving the GPU every communicator has one rank, so a reduction is a
 * copy of the input and a barrier returns at once.  Only the names those
 * programs use are provided.
 */
#pragma once

#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
typedef struct { int MPI_SOURCE, MPI_TAG, MPI_ERROR; } MPI_Status;

#define MPI_COMM_WORLD ((MPI_Comm)0)
#define MPI_SUCCESS 0

#define MPI_INT ((MPI_Datatype)4)
#define MPI_UNSIGNED ((MPI_Datatype)5)
#define MPI_DOUBLE ((MPI_Datatype)8)
#define MPI_UNSIGNED_LONG_LONG ((MPI_Datatype)9)
#define MPI_LONG_LONG ((MPI_Datatype)10)

#define MPI_SUM ((MPI_Op)1)
#define MPI_MAX ((MPI_Op)2)
#define MPI_MIN ((MPI_Op)3)
#define MPI_BAND ((MPI_Op)4)

static inline size_t pra_mpi_type_size(MPI_Datatype t)
{
    switch (t) {
    case MPI_INT:
    case MPI_UNSIGNED: return 4;
    default: return 8;
    }
}

static inline double MPI_Wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static inline int MPI_Init(int*, char***) { return MPI_SUCCESS; }
static inline int MPI_Finalize(void) { return MPI_SUCCESS; }
static inline int MPI_Comm_rank(MPI_Comm, int* r) { *r = 0; return MPI_SUCCESS; }
static inline int MPI_Comm_size(MPI_Comm, int* n) { *n = 1; return MPI_SUCCESS; }
static inline int MPI_Barrier(MPI_Comm) { return MPI_SUCCESS; }
static inline int MPI_Abort(MPI_Comm, int code) { exit(code); }

static inline int MPI_Reduce(const void* in, void* out, int count, MPI_Datatype t, MPI_Op, int, MPI_Comm)
{
    if (in != out) memmove(out, in, (size_t)count * pra_mpi_type_size(t));
    return MPI_SUCCESS;
}

static inline int MPI_Allreduce(const void* in, void* out, int count, MPI_Datatype t, MPI_Op, MPI_Comm)
{
    if (in != out) memmove(out, in, (size_t)count * pra_mpi_type_size(t));
    return MPI_SUCCESS;
}

static inline int MPI_Bcast(void*, int, MPI_Datatype, int, MPI_Comm) { return MPI_SUCCESS; }
