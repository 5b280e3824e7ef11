/*
 * splatt/structs.h -- the data structures the reference keeps in src/ but
 * whose fields its CLI, tests and callers read directly, plus the
 * internal-but-exported entry points on the CPD path.
 *
 * Field order is the reference's (src/csf.h:28-94, src/sptensor.h:16-61,
 * src/cpd/cpd.h:37-56, src/matrix.h:10-16). Fields this library adds sit at
 * the end of a struct, after every reference field.
 */
#ifndef SPLATT_INCLUDE_STRUCTS_H
#define SPLATT_INCLUDE_STRUCTS_H

/* Sparsity pattern of one CSF tile (src/csf.h:28-45). Level l has nfibs[l]
 * nodes; fptr[l][f]..fptr[l][f+1] are the children of node f at level l+1;
 * fids[l][f] is node f's index in mode dim_perm[l]. fids[0] is NULL for an
 * untiled tensor (the slice position is the index); vals has nfibs[N-1]. */
typedef struct
{
  splatt_idx_t nfibs[SPLATT_MAX_NMODES];
  splatt_idx_t * fptr[SPLATT_MAX_NMODES];
  splatt_idx_t * fids[SPLATT_MAX_NMODES];
  splatt_val_t * vals;
} csf_sparsity;

/* src/csf.h:56-94 plus a trailing device handle. */
struct _splatt_csf
{
  splatt_idx_t nnz;
  splatt_idx_t nmodes;
  splatt_idx_t dims[SPLATT_MAX_NMODES];
  splatt_idx_t dim_perm[SPLATT_MAX_NMODES];
  splatt_idx_t dim_iperm[SPLATT_MAX_NMODES];
  splatt_tile_type which_tile;
  splatt_idx_t ntiles;
  splatt_idx_t ntiled_modes;
  splatt_idx_t tile_dims[SPLATT_MAX_NMODES];
  csf_sparsity * pt;

  /* --- additions (not in the reference) --- */
  /* GPU-resident copy of this tensor (opaque), NULL if none. */
  void * dev;
  /* How many tensors the array this one belongs to holds, and which
   * SPLATT_OPTION_CSF_ALLOC produced it (lets splatt_cpd pick the right
   * tensor per mode without trusting default options). */
  splatt_idx_t alloc_count;
  splatt_csf_type alloc_type;
};

/* COO tensor, src/sptensor.h:16-61. */
struct _splatt_coord
{
  splatt_idx_t nmodes;
  splatt_idx_t dims[SPLATT_MAX_NMODES];
  splatt_idx_t nnz;
  splatt_idx_t * ind[SPLATT_MAX_NMODES];
  splatt_val_t * vals;
  int tiled;
  splatt_idx_t * indmap[SPLATT_MAX_NMODES];
};
typedef struct _splatt_coord sptensor_t;

/* Kruskal tensor, src/cpd/cpd.h:37-56. */
struct _splatt_kruskal
{
  splatt_idx_t rank;
  splatt_val_t * factors[SPLATT_MAX_NMODES];
  splatt_val_t * lambda;
  splatt_idx_t nmodes;
  splatt_idx_t dims[SPLATT_MAX_NMODES];
  double fit;
};

/* Communication info of the non-MPI build, src/splatt_mpi.h:36-47. */
struct _splatt_comm_info
{
  int world_rank;
  int world_npes;
};

/* Dense matrix, src/matrix.h:10-16. */
typedef struct
{
  splatt_idx_t I;
  splatt_idx_t J;
  splatt_val_t * vals;
  int rowmajor;
} matrix_t;

/* CSF mode orderings, src/csf.h:12-19. */
typedef enum
{
  CSF_SORTED_SMALLFIRST,
  CSF_SORTED_BIGFIRST,
  CSF_INORDER_MINUSONE,
  CSF_SORTED_MINUSONE,
  CSF_MODE_CUSTOM
} csf_mode_type;

#ifdef __cplusplus
extern "C" {
#endif

/* src/csf.h:121 (src/csf.c:695). Builds opts[CSF_ALLOC] tensors from tt on
 * the GPU. Like the reference, tt is left sorted (and, for DENSETILE,
 * tiled) in the order of the last tensor built. */
splatt_csf * splatt_csf_alloc(sptensor_t * const tt, double const * const opts);

/* src/csf.h:140 (src/csf.c:742): one tensor with a chosen ordering. */
void splatt_csf_alloc_mode(sptensor_t * const tt, csf_mode_type which_ordering,
    splatt_idx_t const mode_special, splatt_csf * const csf,
    double const * const opts);

/* src/csf.h:148-162 (src/csf.c:565-604). */
void splatt_csf_free(splatt_csf * const csf, double const * const opts);
void splatt_csf_free_mode(splatt_csf * const csf);

/* src/csf.h:188 (src/csf.c:753): squared Frobenius norm, FP64 accumulate. */
splatt_val_t splatt_csf_frobsq(splatt_csf const * const tensor);

/* src/csf.h:201 (src/csf.c:608). */
void splatt_csf_find_mode_order(splatt_idx_t const * const dims,
    splatt_idx_t const nmodes, csf_mode_type which, splatt_idx_t const mode,
    splatt_idx_t * const perm_dims);

/* src/csf.h:220-246. */
splatt_idx_t splatt_csf_mode_to_depth(splatt_csf const * const csf,
    splatt_idx_t const mode);
splatt_idx_t splatt_csf_depth_to_mode(splatt_csf const * const csf,
    splatt_idx_t const depth);

/* src/mttkrp.h:35 (src/mttkrp.c:1287): output written to
 * mats[SPLATT_MAX_NMODES]. thds is accepted for signature compatibility and
 * ignored (pass NULL). */
void splatt_mttkrp_csf(splatt_csf const * const tensors, matrix_t ** mats,
    splatt_idx_t const mode, void * const thds, splatt_mttkrp_ws * const ws,
    splatt_global_opts const * const global_opts);

/* COO helpers: src/sptensor.h:83-207 (src/sptensor.c). */
sptensor_t * splatt_tt_read(char const * const ifname);
sptensor_t * splatt_tt_alloc(splatt_idx_t const nnz, splatt_idx_t const nmodes);
void splatt_tt_free(sptensor_t * tt);
splatt_idx_t splatt_tt_remove_empty(sptensor_t * const tt);
splatt_val_t splatt_tt_normsq(sptensor_t const * const tt);

/* src/cpd/stream.h:24-30 (src/cpd/stream/stream.cc:48): the entry behind
 * `splatt cpd --stream=M`. Reads the tensor, runs the streaming CPD on the
 * GPU, prints the reference's summary lines and writes mode<N>.mat (rows at
 * their original indices) into the working directory. */
splatt_error_type splatt_cpd_stream(char const * const filename,
    splatt_idx_t rank, splatt_idx_t const stream_mode, double const forget,
    splatt_cpd_opts * const cpd_options,
    splatt_global_opts const * const global_options);

#ifdef __cplusplus
}
#endif
#endif
