/*
 * splatt/coord.h -- sparse tensors in coordinate (COO) form.
 */
#ifndef SPLATT_INCLUDE_COORD_H
#define SPLATT_INCLUDE_COORD_H

typedef struct _splatt_coord splatt_coord;

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/splatt/coord.h:53 (src/sptensor.c:351). Index and value
 * arrays are allocated but not initialised. */
splatt_coord * splatt_alloc_coord(
    splatt_idx_t const num_modes,
    splatt_idx_t const nnz);

/* Replaces include/splatt/coord.h:63. */
void splatt_free_coord(splatt_coord * coord);

/* Replaces include/splatt/coord.h:80. Reads a .tns/.coo text file or a .bin
 * file into an already allocated coord (its arrays are replaced). */
splatt_error_type splatt_coord_load(
    char const * const fname,
    splatt_coord * const coord);

/* Accessors: include/splatt/coord.h:95-128. */
splatt_idx_t splatt_coord_get_nnz(splatt_coord const * const coord);
splatt_idx_t splatt_coord_get_modes(splatt_coord const * const coord);
splatt_idx_t * splatt_coord_get_inds(
    splatt_coord const * const coord,
    splatt_idx_t const mode);
splatt_val_t * splatt_coord_get_vals(splatt_coord const * const coord);

#ifdef __cplusplus
}
#endif
#endif
