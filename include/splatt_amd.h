/*
 * splatt_amd.h -- extensions of the SPLATT API that only make sense with a
 * GPU underneath: device-resident operands, precision control, the algorithmic
 * cost model used for bandwidth figures, and error text. None of these has a
 * counterpart in the reference; everything in splatt.h works without them.
 */
#ifndef SPLATT_AMD_H
#define SPLATT_AMD_H

#include <stddef.h>
#include <stdint.h>
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

#define SPLATT_AMD_LAYOUT_AUTO 0
#define SPLATT_AMD_LAYOUT_CSF 1

/* How MTTKRP reaches internal and leaf output modes (no counterpart in the
 * reference, whose mttkrp_csf, src/mttkrp.c:1287-1343, always walks the
 * CSF with locks or privatisation). SPLATT_AMD_LAYOUT_AUTO (default): on
 * first use for a mode, a copy of the nonzeros ordered by that mode's row is
 * built on the device (4 bytes per other mode, 4 for the row and 12 for the
 * FP64 and FP32 values, per nonzero) when free device memory allows, and
 * the mode is computed like a root mode, without atomics on its rows.
 * SPLATT_AMD_LAYOUT_CSF: always walk the CSF, adding to output rows with
 * atomics (hot rows privatised in LDS). Results agree to rounding. */
splatt_error_type splatt_amd_set_mttkrp_layout(int layout);

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

/* splatt_amd_mttkrp_dev that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as splatt_amd_mttkrp_dev. Its partial-row
 * buffers are one fixed device array of splatt_amd_mttkrp_noalloc_bytes()
 * (16 MiB) per GPU, held until the process ends and shared by every call on
 * that GPU: calls that overlap on one GPU (other host threads, or streams
 * without an order between them) race and must not be issued. The CSF must
 * already carry what splatt_amd_mttkrp_dev builds on its first call for the
 * same mode, rank and precision (launch plans, hot-row tables, row-sorted
 * copies, the FP32 values): make that call once beforehand, outside the
 * allocation-free region. Returns SPLATT_ERROR_BADINPUT (with the reason in
 * splatt_amd_last_error) when something is missing or when the output's
 * partial rows exceed the array; every 3-mode root output on the
 * persistent kernel fits. */
splatt_error_type splatt_amd_mttkrp_dev_noalloc(
    splatt_idx_t const mode,
    splatt_idx_t const ncolumns,
    splatt_csf const * const tensors,
    void ** d_mats,
    void * d_out,
    int precision,
    double const * const options,
    void * stream);

size_t splatt_amd_mttkrp_noalloc_bytes(void);

/* splatt_coord_load (src/io.c:62-108, :273-348 through tt_read) with the
 * text parsed on the GPU: the same coordinates (0-based), values and dims
 * as the host reader, bit for bit. Lines whose fields the GPU conversion
 * does not take (hexadecimal or special values, more than 19 significant
 * digits, results outside the normal doubles, malformed fields) are
 * converted on the host with strtoull/strtod; *nhost (may be NULL) receives
 * their number. splatt_csf_load reads its input the same way. */
splatt_error_type splatt_amd_coord_load_dev(
    char const * const fname,
    splatt_coord * const coord,
    uint64_t * const nhost);

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
 * constraints on the other modes, stream tolerances). Every check_interval
 * slices and after the slice that consumes the last nonzero, the errors of
 * all slices so far and of the last 10 are computed, and with refit != 0 a
 * full CPD of all slices so far (its factors drawn from rand(), as in the
 * reference); splatt_cpd_stream uses 100 and refit, as the reference does
 * (StreamCPD.cc:517-528). check_interval 0 skips them. Returns NULL on
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
  double * global_err;                        /* per slice, -1 off checkpoints: all slices so far */
  double * local10_err;                       /* per slice, -1 off checkpoints: the last 10 */
  double * cpd_err;                           /* per slice, -1 off checkpoints: refit 1 - fit */
} splatt_amd_stream_result;

splatt_amd_stream_result * splatt_amd_cpd_stream_coo(
    struct _splatt_coord const * const tt,
    splatt_idx_t rank,
    splatt_idx_t stream_mode,
    double forget,
    splatt_idx_t check_interval,
    int refit,
    splatt_cpd_opts * const cpd_options,
    splatt_global_opts const * const global_options);

void splatt_amd_free_stream_result(splatt_amd_stream_result * result);

/* ---- serving many streams on one GPU ----------------------------------
 * No counterpart in the reference (one stream per process there). Each
 * stream follows the reference's streaming rules and is the same
 * computation as splatt_amd_cpd_stream_coo on its slices; its factor rows
 * are initialised from its own rand() sequence seeded with `seed`. Streams
 * are admitted while their device state fits the HBM budget; the others wait
 * until closed streams release memory. Up to max_inflight slices (of
 * different streams) run at once; the slices of one stream run in
 * submission order. */
typedef struct spd_server splatt_amd_server;

/* hbm_budget_bytes 0: 80% of the free device memory. */
splatt_amd_server * splatt_amd_server_create(size_t hbm_budget_bytes, int max_inflight);

/* dims_cap[m]: original indices of mode m stay below it (the streamed
 * mode's entry is ignored). cpd_options may be NULL. Returns the stream id,
 * or -1. */
int64_t splatt_amd_server_open(splatt_amd_server * srv, splatt_idx_t nmodes,
    splatt_idx_t const * dims_cap, splatt_idx_t rank, splatt_idx_t stream_mode,
    double forget, splatt_cpd_opts const * cpd_options, unsigned seed);

/* One time slice (original indices; inds[stream_mode] is not read). Returns
 * a ticket, or -1. */
int64_t splatt_amd_server_submit(splatt_amd_server * srv, int64_t stream,
    splatt_idx_t nnz, splatt_idx_t const * const * inds, splatt_val_t const * vals);

/* Wait for every submitted slice. */
splatt_error_type splatt_amd_server_drain(splatt_amd_server * srv);

/* 0 when done (outputs filled), 1 while pending, < 0 on error: latency from
 * submission to completion, device-side time, the slice's squared error,
 * ALS iterations. */
int splatt_amd_server_slice(splatt_amd_server * srv, int64_t ticket, double * latency_s,
    double * gpu_s, double * local_err, splatt_idx_t * outer_its);

/* Factor of `mode` after the stream's submitted slices (rows relabelled by
 * first appearance; the streamed mode: one row per slice). out may be NULL
 * to query *rows. */
splatt_error_type splatt_amd_server_factors(splatt_amd_server * srv, int64_t stream,
    splatt_idx_t mode, splatt_val_t * out, splatt_idx_t * rows);

/* No more slices for the stream; its memory returns once its queue drains. */
splatt_error_type splatt_amd_server_close(splatt_amd_server * srv, int64_t stream);

void splatt_amd_server_stats(splatt_amd_server * srv, size_t * used, size_t * budget,
    int * admitted, int * waiting);
void splatt_amd_server_free(splatt_amd_server * srv);

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
