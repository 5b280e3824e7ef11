/*
 * splatt/api_kernels.h -- MTTKRP.
 */
#ifndef SPLATT_SPLATT_KERNELS_H
#define SPLATT_SPLATT_KERNELS_H

#include <stdbool.h>

/* Same layout as include/splatt/api_kernels.h:22-67. On the GPU the
 * partitions become per-CSF launch plans kept with the device tensor, so
 * tile_partition/tree_partition stay NULL and privatize_buffer is unused. */
typedef struct
{
  splatt_idx_t num_csf;
  splatt_idx_t mode_csf_map[SPLATT_MAX_NMODES];
  splatt_idx_t num_threads;
  splatt_idx_t * tile_partition[SPLATT_MAX_NMODES];
  splatt_idx_t * tree_partition[SPLATT_MAX_NMODES];
  bool is_privatized[SPLATT_MAX_NMODES];
  splatt_val_t * * privatize_buffer;
  double reduction_time;
} splatt_mttkrp_ws;

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/splatt/api_kernels.h:98 (src/mttkrp.c:1775).
 * matrices[m] are row-major dims[m] x ncolumns host arrays; matout
 * (dims[mode] x ncolumns) is overwritten. Runs on the GPU. */
int splatt_mttkrp(
    splatt_idx_t const mode,
    splatt_idx_t const ncolumns,
    splatt_csf const * const tensors,
    splatt_val_t ** matrices,
    splatt_val_t * const matout,
    double const * const options);

/* Replaces include/splatt/api_kernels.h:107 (src/mttkrp.c:1831): the
 * mode->CSF map and privatisation flags, computed as the reference does. */
splatt_mttkrp_ws * splatt_mttkrp_alloc_ws(
    splatt_csf const * const tensors,
    splatt_idx_t const ncolumns,
    double const * const options);

/* Replaces include/splatt/api_kernels.h:118 (src/mttkrp.c:1932). */
void splatt_mttkrp_free_ws(splatt_mttkrp_ws * const ws);

#ifdef __cplusplus
}
#endif
#endif
