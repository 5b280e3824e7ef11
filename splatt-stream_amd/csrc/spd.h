/*
 * spd.h -- internal C ABI between the C host library and the HIP (gfx950)
 * device code. Plain C types only: opaque handles, pointers, sizes, int
 * status. Not part of the public splatt.h surface.
 *
 * Conventions
 *   - Every function returns SPD_OK (0) or a negative SPD_E* code; the text
 *     of the last error is available from spd_last_error().
 *   - Device memory is owned by the handle that allocated it.
 *   - All work is issued on the library's stream (spd_stream()) unless a
 *     stream argument is given; functions that return host results
 *     synchronise that stream before returning.
 */
#ifndef SPD_ABI_H
#define SPD_ABI_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPD_OK          0
#define SPD_EHIP       -1   /* HIP runtime error */
#define SPD_ENOMEM     -2   /* device (or host bookkeeping) allocation failed */
#define SPD_EINPUT     -3   /* bad input (index out of range, sizes) */
#define SPD_ENODEV     -4   /* no usable GPU */
#define SPD_ENUMERIC   -5   /* non-SPD Gram matrix, NaN, ... */
#define SPD_EHOST      -6   /* input the device path leaves to the host path */

#define SPD_MAX_MODES 8

#define SPD_F64 0
#define SPD_F32 1

const char * spd_last_error(void);
int spd_init(void);
void * spd_stream(void);   /* hipStream_t of the library */
int spd_sync(void);

int spd_malloc(void ** ptr, size_t bytes);
int spd_free(void * ptr);
int spd_h2d(void * dst, void const * src, size_t bytes);
int spd_d2h(void * dst, void const * src, size_t bytes);
int spd_d2d(void * dst, void const * src, size_t bytes);
int spd_zero(void * ptr, size_t bytes);   /* fill kernel, graph-safe */

/* ---- coordinate tensor resident on the device ------------------------- */
typedef struct spd_coo spd_coo;

/* Upload nnz coordinates (inds[m][n], 0-based, < dims[m]) and values. */
int spd_coo_upload(int nmodes, uint64_t nnz, uint64_t const * dims,
    uint64_t const * const * inds, double const * vals, spd_coo ** out);
void spd_coo_free(spd_coo * coo);
void spd_coo_dims(spd_coo const * coo, int * nmodes, uint64_t * nnz, uint64_t * dims);
/* Copy the coordinates (inds[m]: nnz entries each, host) and values back. */
int spd_coo_download(spd_coo const * coo, uint64_t * const * inds, double * vals);

/* ---- text tensor input (src/io.c:62-108, :273-348) --------------------- */

/* Host conversion of one data line [line, line + len) (no '\n'): nmodes
 * strtoull fields then one strtod field, exactly as the host reader. 0 on
 * success. */
typedef int (*spd_line_fn)(char const * line, size_t len, int nmodes,
    uint64_t * inds, double * val);

/* Parse a whole .tns text of len bytes with nmodes index fields per line
 * on the device into a 0-based device COO (dims from the largest index; a
 * mode whose smallest index is 0 is 0-indexed, else 1-indexed). Lines whose
 * fields the device conversion does not take are converted by line_fn.
 * *nhost = their number. SPD_EHOST (and *out NULL) when more than
 * SPD_TNS_HOST_MAX lines would need line_fn: the host reader is then the
 * faster path. */
#define SPD_TNS_HOST_MAX (1u << 20)
int spd_tns_parse(char const * text, size_t len, int nmodes, spd_line_fn line_fn,
    spd_coo ** out, uint64_t * nhost);

/* Renumber each mode's occurring indices 0..k-1 in order and set dims to
 * k (src/sptensor.c:190-252). *removed = total indices removed. */
int spd_coo_remove_empty(spd_coo * coo, uint64_t * removed);

/* The device field conversions run on the host, for checking against the
 * C library. Field i is [buf + off[i], buf + off[i + 1]); kind 0 converts
 * it as a value (vals[i]), kind 1 as an index (idx[i]); ok[i] = 0 where the
 * conversion leaves the field to the C library. */
int spd_decimal_convert(char const * buf, uint64_t const * off, uint64_t n, int kind,
    double * vals, uint64_t * idx, int * ok);

/* Lexicographic order of the coordinates with perm[0] most significant:
 * order_host[k] (n entries, host) = position of the k-th smallest. Equal
 * coordinates keep their input order. */
int spd_coo_lexsort(spd_coo const * coo, int const * perm, uint32_t * order_host);

/* ---- CSF --------------------------------------------------------------- */
typedef struct spd_csf spd_csf;

/* Build one CSF from a device COO with level order perm[0..nmodes-1]
 * (perm[l] = mode stored at level l). tile_dims[m] > 1 requests dense
 * tiling of mode m (reference DENSETILE); all ones means untiled. The
 * untiled structure is always built (the kernels use it); a tiled one is
 * kept as well for the host mirror. */
int spd_csf_build(spd_coo const * coo, int const * perm,
    uint64_t const * tile_dims, spd_csf ** out);
void spd_csf_free(spd_csf * csf);

/* Shape of the host mirror: number of tiles, then per tile and level the
 * node count (ntiles x nmodes, tile-major). */
uint64_t spd_csf_ntiles(spd_csf const * csf);
int spd_csf_tile_nfibs(spd_csf const * csf, uint64_t * nfibs);

/* Copy the host mirror of tile t into host arrays sized from
 * spd_csf_tile_nfibs: fptr[l] (nfibs[l]+1 entries, l < nmodes-1),
 * fids[l] (nfibs[l] entries; fids[0] may be NULL to skip), vals (nfibs[N-1]).
 */
int spd_csf_export_tile(spd_csf const * csf, uint64_t t, uint64_t ** fptr,
    uint64_t ** fids, double * vals);

/* sum of squared values, FP64 accumulation */
int spd_csf_frobsq(spd_csf const * csf, double * out);

/* ---- MTTKRP ------------------------------------------------------------ */

/* out (dims[mode] x R, row-major, device) = MTTKRP of mode `mode` using this
 * CSF. mats[m] are device row-major dims[m] x R arrays of type dtype
 * (SPD_F64 / SPD_F32); mats[mode] is not read. root_rows_are_positions
 * selects the reference's untiled root-row convention. stream may be NULL
 * (library stream). */
int spd_mttkrp(spd_csf * csf, int mode, uint64_t R, void * const * mats,
    void * out, int dtype, int root_rows_are_positions, void * stream);

/* spd_mttkrp without any allocation: partial rows in a fixed device array of
 * spd_mttkrp_noalloc_bytes() bytes per GPU, shared by every call on that GPU
 * (overlapping calls race). The CSF must already hold what spd_mttkrp builds
 * on first use for this mode, rank and precision (plans, hot-row tables,
 * row-sorted copies, the FP32 values); outputs whose partial rows exceed the
 * array are rejected (SPD_EINPUT). */
int spd_mttkrp_noalloc(spd_csf * csf, int mode, uint64_t R, void * const * mats,
    void * out, int dtype, int rows_pos, void * stream);
size_t spd_mttkrp_noalloc_bytes(void);

/* Bytes and flops of one MTTKRP by the algorithmic model: compulsory reads of
 * values, indices, pointers and input factors plus one write of the output,
 * with 4-byte indices and sizeof(dtype) values. */
int spd_mttkrp_model(spd_csf const * csf, int mode, uint64_t R, int dtype,
    double * bytes, double * flops);

/* Record HIP events around the main MTTKRP kernel (on the stream it runs
 * on) when enabled; spd_mttkrp_last_kernel_ms() waits for and returns the
 * last such time in ms, or <0 if none was recorded. */
int spd_mttkrp_timing(int enable);
double spd_mttkrp_last_kernel_ms(void);
int spd_mttkrp_kernel_times(double * ms, int max);

/* Internal and leaf outputs: enable = 1 (default) walks a row-sorted copy of
 * the CSF, built on the device on first use for each output level when
 * device memory allows; enable = 0 always walks the CSF itself, adding to
 * the output rows atomically. */
int spd_mttkrp_set_copies(int enable);
/* 3-mode root outputs: variant 0 (default) the per-nonzero kernel k_root3;
 * 1 the persistent kernel with the hottest leaf rows in LDS (where its lane
 * layouts apply); 2 the persistent kernel with every row from HBM; 3
 * k_root3 gathering each fiber's A row once. */
int spd_mttkrp_set_root_variant(int variant);
/* Carry entries from which the root fix-up takes two passes (default
 * 131072; 0: always two). */
int spd_mttkrp_set_fix_two_level(uint64_t entries);

/* ---- dense / CPD ------------------------------------------------------- */

/* Convert FP64 device array to FP32 and back. */
int spd_f64_to_f32(double const * src, float * dst, uint64_t n, void * stream);
int spd_f32_to_f64(float const * src, double * dst, uint64_t n, void * stream);

/* Dense steps on device arrays (library stream, asynchronous), exported
 * for the tests:
 *   gram:     out (R x R) = X^T X, after X[:,r] /= div[r] when div != NULL;
 *             X32 (optional) receives an FP32 copy of X.
 *   chol_inv: G = Hadamard of aTa[m], m != mode, + diag_add (+ rho =
 *             trace/R when admm); ginv (RP x RP, RP = R rounded up to 16)
 *             = G^{-1} from its Cholesky factor; *info = failing pivot.
 *   apply:    X = M ginv (M FP64 or FP32 by mdtype); colnorm[r] = column
 *             2-norms of X when colnorm != NULL.
 *   admm_step: one ADMM inner iteration; sums needs 4 + R doubles:
 *             ||P||^2, ||P-aux||^2, ||P-init||^2, ||D||^2, then the column
 *             norms used by the prox.
 *   admm_solve: the ADMM inner loop to convergence (or max_inner); aux and
 *             init are I x R work matrices, sums 4 + R doubles, ctl[1] the
 *             iterations run.
 *   admm_chunked: ADMM in independent chunks of `chunk` rows (nonneg or no
 *             prox) to convergence; *its_rows = sum of iterations x rows. */
int spd_dense_gram(double * X, uint64_t I, uint64_t R, double const * div,
    float * X32, double * out);
int spd_dense_chol_inv(double const * const * aTa, int nmodes, int mode,
    uint64_t R, double diag_add, int admm, double * ginv, double * rho, int * info);
int spd_dense_apply(void const * M, int mdtype, double const * ginv,
    uint64_t I, uint64_t R, double * X, double * colnorm);
int spd_dense_admm_step(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint64_t R, double * P, double * aux, double * D,
    double * init, int prox, double * sums);
int spd_dense_admm_solve(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint64_t R, double * P, double * D, double * aux,
    double * init, int prox, double lam, uint64_t max_inner, double tol, double * sums,
    int * ctl);
int spd_dense_admm_chunked(void const * M, int mdtype, double const * ginv,
    double const * rho, uint64_t I, uint64_t R, uint64_t chunk, double * P, double * D,
    int prox, double lam, uint64_t max_inner, double tol, uint64_t * its_rows);

/* ---- CPD-ALS (src/cpd/cpd.c:161-332, src/cpd/admm.c:316-406) ---------- */

#define SPD_CON_NONE   0   /* unconstrained: closed form, columns normalised */
#define SPD_CON_CLOSED 1   /* closed form, Gram diagonal += diag_add */
#define SPD_CON_ADMM   2   /* ADMM with a proximal operator */

#define SPD_PROX_NONE 0
#define SPD_PROX_MAXCOLNORM 1
#define SPD_PROX_MAXCOLNORM_NONNEG 2
#define SPD_PROX_NONNEG 3
#define SPD_PROX_LASSO 4   /* soft threshold at lam / rho */
#define SPD_PROX_ROWSIMP 5 /* rows onto the probability simplex (chunked ADMM) */
#define SPD_PROX_SMOOTH 6  /* columns smoothed: rho (lam K + rho I)^-1 x */
#define SPD_PROX_COLSIMP 7 /* columns onto the probability simplex */
#define SPD_PROX_NTF_FROB 8  /* max(x rho / (lam + rho), 0) */
#define SPD_PROX_NTF_LASSO 9 /* max(x - lam / rho, 0) */

typedef struct {
  int kind;          /* SPD_CON_* */
  int prox;          /* SPD_PROX_* (ADMM) */
  double diag_add;   /* Frobenius regularisation added to the Gram diagonal */
  uint64_t chunk;    /* ADMM rows per independent chunk (row-separable prox
                        only, src/cpd/admm.c:355-400); 0: one block */
  double lam;        /* SPD_PROX_LASSO: the L1 multiplier */
} spd_con;

typedef struct {
  int nmodes;
  uint64_t rank;
  uint64_t max_its;
  double tol;
  uint64_t max_inner;
  double inner_tol;
  int dtype;           /* MTTKRP precision, SPD_F64 / SPD_F32 */
  int verbosity;       /* > 0: one line per iteration */
  int unconstrained;   /* every mode unconstrained: normalise, absorb into lambda */
  spd_con con[SPD_MAX_MODES];
} spd_cpd_cfg;

/* ALS sweeps on the GPU. csfs[k] are device CSFs, mode_csf[m] the one used
 * for mode m, rows_pos[k] the root-row convention of CSF k. factors[m]
 * (host, dims[m] x rank, FP64) are the initial factors on entry and the
 * result on return; lambda (host, rank) is the result. fits (host, max_its
 * entries, may be NULL) receives the fit of every iteration; *its the number
 * of iterations and *fit the final fit. */
int spd_cpd_als(spd_csf * const * csfs, int const * mode_csf,
    int const * rows_pos, spd_cpd_cfg const * cfg, double * const * factors,
    double * lambda, double * fits, uint64_t * its, double * fit);

/* ---- streaming CPD (src/cpd/stream/StreamCPD.cc:311-571) ------------- */

typedef struct {
  int stream_mode;
  uint64_t rank;
  double forget;
  uint64_t max_its;        /* outer ALS iterations per slice */
  double tol;              /* outer tolerance on the change measure */
  uint64_t max_inner;      /* ADMM */
  double inner_tol;
  int verbosity;
  spd_con con[SPD_MAX_MODES];   /* streamed mode: closed form only */
  /* values of new factor rows, in the order rows appear (the reference
   * draws them from rand(), StreamMatrix.cc:70-83) */
  void (*fill_rand)(void * ctx, double * vals, uint64_t n);
  void * rand_ctx;
  /* checkpoints (StreamCPD.cc:517-528): every check_interval slices and
   * after the slice that consumes the last nonzero, the squared errors of
   * all slices so far and of the last 10 against the current factors;
   * 0 disables them */
  uint64_t check_interval;
  /* at a checkpoint: a full CPD of the first nslices slices (host arrays,
   * time labels 0 .. nslices-1, dims[stream_mode] = nslices) returning its
   * relative squared error, 1 - fit (StreamCPD.cc:89-106). Called from the
   * thread running the stream; NULL: not computed (reported as -1). */
  double (*refit)(void * ctx, int nmodes, uint64_t nnz, uint64_t const * dims,
      uint32_t const * const * inds, double const * vals, uint64_t rank);
  void * refit_ctx;
} spd_stream_cfg;

typedef struct {
  uint64_t nbatches;
  uint64_t dims[SPD_MAX_MODES];      /* rows of each result (time: nbatches) */
  double * factors[SPD_MAX_MODES];   /* host buffers sized by the caller, or NULL */
  double * local_err;                /* per slice, squared; NULL to skip */
  double * batch_seconds;            /* per slice wall time; NULL to skip */
  uint64_t * outer_its;              /* per slice; NULL to skip */
  double * final_err;                /* squared; NULL to skip */
  /* per slice, -1 except at checkpoints; NULL to skip each */
  double * global_err;               /* squared, all slices so far */
  double * local10_err;              /* squared, the last 10 slices */
  double * cpd_err;                  /* the refit's 1 - fit */
} spd_stream_result;

/* The stream: N-mode coordinates (host, uint32) already sorted by time
 * (inds[stream_mode], the slice index) and relabelled, values (host), the
 * tensor's dims. Runs every slice on the GPU. */
int spd_stream_cpd(int N, uint64_t nnz, uint64_t const * dims,
    uint32_t const * const * inds, double const * vals, spd_stream_cfg const * cfg,
    spd_stream_result * res);

/* ---- serving: many streams on one GPU (no reference counterpart) ------ */

typedef struct spd_server spd_server;

/* budget: device bytes the streams' state may take (0: 80% of free memory);
 * max_inflight: slices processed at once (worker threads). */
int spd_server_create(size_t budget, int max_inflight, spd_server ** out);
/* A stream with row capacities cap[m] (original indices < cap[m]) and the
 * per-stream configuration (fill_rand is replaced by the stream's own rand()
 * sequence from `seed`). Admitted when its footprint fits the budget. */
int spd_server_open(spd_server * sv, int N, uint64_t const * cap,
    spd_stream_cfg const * cfg, unsigned seed, int64_t * id_out);
/* Queue one time slice (original labels; the streamed mode's array is not
 * read). Returns a ticket. */
int spd_server_submit(spd_server * sv, int64_t id, uint64_t nnz,
    uint64_t const * const * inds, double const * vals, int64_t * ticket_out);
/* Wait until every submitted slice is done. */
int spd_server_drain(spd_server * sv);
/* A slice's submit-to-done latency, device time, local error and outer
 * iterations; returns 1 while it is pending. */
int spd_server_slice(spd_server * sv, int64_t ticket, double * latency_s,
    double * gpu_s, double * local_err, uint64_t * outer);
/* Factor of `mode` (host, rows x rank; time mode: one row per slice) after
 * the stream's queued slices; *rows gets the row count (out may be NULL). */
int spd_server_factors(spd_server * sv, int64_t id, int mode, double * out,
    uint64_t * rows);
/* No more slices: memory returns to the budget once the queue drains. */
int spd_server_close(spd_server * sv, int64_t id);
void spd_server_stats(spd_server * sv, size_t * used, size_t * budget,
    int * admitted, int * waiting);
void spd_server_free(spd_server * sv);
/* The values a server stream seeded with `seed` draws for its first n new
 * factor entries (for tests). */
void spd_serve_rand_fill(unsigned seed, double * vals, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
