/*
 * splatt_amd.h -- extensions of the SPLATT API that only make sense with a
 * GPU underneath: device-resident operands, precision control, the algorithmic
 * cost model used for bandwidth figures, and error text. None of these has a
 * counterpart in the reference; everything in splatt.h works without them.
 */
#ifndef SPLATT_AMD_H
#define SPLATT_AMD_H

#include <stddef.h>
#include "splatt.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPLATT_AMD_F64 0
#define SPLATT_AMD_F32 1

/* Text of the last device-side failure. */
char const * splatt_amd_last_error(void);

/* Arithmetic type of MTTKRP inside splatt_mttkrp() and splatt_cpd():
 * SPLATT_AMD_F64 (default, the reference's val_t) or SPLATT_AMD_F32 (FP32
 * values, factors and accumulation; Grams, Cholesky and fit stay FP64). */
splatt_error_type splatt_amd_set_mttkrp_precision(int precision);

/* MTTKRP with operands already in GPU memory. d_mats[m] are device pointers
 * to row-major dims[m] x ncolumns arrays of `precision` type; d_out
 * (dims[mode] x ncolumns) is overwritten. The CSF must come from this
 * library (its GPU copy is used). stream is a hipStream_t or NULL for the
 * library stream; the call returns without synchronising. */
splatt_error_type splatt_amd_mttkrp_dev(
    splatt_idx_t const mode,
    splatt_idx_t const ncolumns,
    splatt_csf const * const tensors,
    void ** d_mats,
    void * d_out,
    int precision,
    double const * const options,
    void * stream);

/* Algorithmic bytes and flops of one MTTKRP of `mode` on the tensor the
 * options' CSF map selects (4-byte indices, value size of `precision`). */
splatt_error_type splatt_amd_mttkrp_cost(
    splatt_idx_t const mode,
    splatt_idx_t const ncolumns,
    splatt_csf const * const tensors,
    int precision,
    double const * const options,
    double * bytes,
    double * flops);

/* Kernel timing. While enabled (which also clears the record), a pair of HIP
 * events is recorded around the main kernel of every MTTKRP, on the stream
 * it runs on. last_ms returns the last such duration; kernel_times copies
 * the durations of the last `max` timed launches (oldest first, at most
 * 4096 kept) and returns how many. Both wait for the events. */
void splatt_amd_mttkrp_timing(int enable);
double splatt_amd_mttkrp_last_ms(void);
int splatt_amd_mttkrp_kernel_times(double * ms, int max);

/* Streaming CPD on a tensor in memory: splatt_cpd_stream without the file
 * and the output files. The stream is the tensor sorted by stream_mode (its
 * index is the time slice); the other modes are relabelled by first
 * appearance, so row i of factors[m] belongs to original index iperm[m][i].
 * cpd_options is modified as the reference modifies it (max-column-norm
 * constraints on the other modes, stream tolerances). Returns NULL on
 * error. */
typedef struct
{
  splatt_idx_t nmodes;
  splatt_idx_t rank;
  splatt_idx_t stream_mode;
  splatt_idx_t nbatches;                      /* time slices processed */
  splatt_idx_t dims[SPLATT_MAX_NMODES];       /* rows of factors[m] (time: nbatches) */
  splatt_val_t * factors[SPLATT_MAX_NMODES];  /* time mode: one row per slice */
  splatt_idx_t * iperm[SPLATT_MAX_NMODES];    /* relabelled row -> original index */
  double * local_err;                         /* per slice: squared error of the slice */
  double * batch_seconds;                     /* per slice: wall time, device synchronised */
  splatt_idx_t * outer_its;                   /* per slice: ALS iterations run */
  double final_err;                           /* squared error of the whole stream */
} splatt_amd_stream_result;

splatt_amd_stream_result * splatt_amd_cpd_stream_coo(
    struct _splatt_coord const * const tt,
    splatt_idx_t rank,
    splatt_idx_t stream_mode,
    double forget,
    splatt_cpd_opts * const cpd_options,
    splatt_global_opts const * const global_options);

void splatt_amd_free_stream_result(splatt_amd_stream_result * result);

/* Device memory helpers for callers that hold device-resident factors. */
void * splatt_amd_dev_alloc(size_t bytes);
void splatt_amd_dev_free(void * ptr);
splatt_error_type splatt_amd_h2d(void * dst, void const * src, size_t bytes);
splatt_error_type splatt_amd_d2h(void * dst, void const * src, size_t bytes);
splatt_error_type splatt_amd_sync(void);

#ifdef __cplusplus
}
#endif
#endif
