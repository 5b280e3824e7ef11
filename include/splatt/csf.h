/*
 * splatt/csf.h -- compressed sparse fiber (CSF) tensors.
 *
 * A CSF built by this library is built on the GPU (sort, fiber compression,
 * tiling) and mirrored into the host structure the reference defines, with
 * identical arrays; the GPU copy stays resident for later kernels.
 */
#ifndef SPLATT_INCLUDE_CSF_H
#define SPLATT_INCLUDE_CSF_H

typedef struct _splatt_csf splatt_csf;

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/splatt/csf.h:55 (src/csf.c:18): read a tensor file,
 * compact empty slices, build opts[SPLATT_OPTION_CSF_ALLOC] tensors. */
splatt_error_type splatt_csf_load(
    char const * const fname,
    splatt_idx_t * nmodes,
    splatt_csf ** tensors,
    double const * const options);

/* The coordinate->CSF entry the reference declares but never defines
 * (include/splatt/csf.h:83-89 sits inside `#if 0`). Same signature. Nonzero
 * n sits at inds[0][n], ..., inds[nmodes-1][n] with value vals[n]; indices
 * are 0-based. The caller's arrays are read, not modified. */
splatt_error_type splatt_csf_convert(
    splatt_idx_t const nmodes,
    splatt_idx_t const nnz,
    splatt_idx_t ** const inds,
    splatt_val_t * const vals,
    splatt_csf ** tensors,
    double const * const options);

/* Replaces include/splatt/csf.h:100 (src/csf.c:39). options must carry the
 * same SPLATT_OPTION_CSF_ALLOC used to allocate. Frees GPU copies too. */
void splatt_free_csf(
    splatt_csf * tensors,
    double const * const options);

#ifdef __cplusplus
}
#endif
#endif
