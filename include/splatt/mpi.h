/*
 * mpi.h -- the communication-info part of the SPLATT API (drop-in for the
 * reference's include/splatt/mpi.h:1-102). This library runs on one GPU
 * and has no MPI build, so only the non-MPI form exists: a one-rank
 * communicator description (src/mpi_dummy.c:15-30).
 */
#ifndef SPLATT_INCLUDE_MPI_H
#define SPLATT_INCLUDE_MPI_H

#include "cpd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* splatt_decomp_type and splatt_comm_type (include/splatt/mpi.h:19-51)
 * are in splatt/types.h. */

/* include/splatt/mpi.h:57: opaque; its non-MPI fields (world_rank,
 * world_npes) are in splatt/structs.h. */
typedef struct _splatt_comm_info splatt_comm_info;

/* include/splatt/mpi.h:76-77 (src/mpi_dummy.c:26-34): frees what
 * splatt_alloc_comm_info returned; NULL is ignored. */
void splatt_free_comm_info(
    splatt_comm_info * comm_info);

/* include/splatt/mpi.h:94 (src/mpi_dummy.c:17-23): rank 0 of 1. Returns
 * NULL when the host allocation fails. */
splatt_comm_info * splatt_alloc_comm_info(void);

#ifdef __cplusplus
}
#endif
#endif
