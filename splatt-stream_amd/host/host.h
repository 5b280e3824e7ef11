/*
 * host.h -- internal declarations shared by the C host library.
 */
#ifndef SPLATT_AMD_HOST_H
#define SPLATT_AMD_HOST_H

#include "../../include/splatt.h"
#include "../../include/splatt/structs.h"
#include "../../include/splatt_amd.h"
#include "../csrc/spd.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef splatt_idx_t idx_t;
typedef splatt_val_t val_t;

#define SS_MIN(x,y) ((x) < (y) ? (x) : (y))
#define SS_MAX(x,y) ((x) > (y) ? (x) : (y))

void * splatt_malloc(size_t const bytes);
void splatt_free(void * ptr);

/* Print the device error (if any) with a context string, return the
 * splatt error code that maps to it. */
splatt_error_type spl_report(int spd_rc, char const * what);

/* Process-wide precision of MTTKRP inside splatt_mttkrp/splatt_cpd. */
int spl_mttkrp_dtype(void);

/* Constraint functions this library recognises and runs on the GPU. */
void splatt_frob_gram(val_t * gram, idx_t const N, void * data);
void splatt_maxcolnorm_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_maxcolnorm_nonneg_prox(val_t * primal, idx_t const nrows,
    idx_t const ncols, idx_t const offset, void * data, val_t const rho,
    bool const should_parallelize);
void splatt_colsimp_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_smooth_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_rowsimp_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_lasso_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_nonneg_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_ntf_frob_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);
void splatt_ntf_lasso_prox(val_t * primal, idx_t const nrows, idx_t const ncols,
    idx_t const offset, void * data, val_t const rho, bool const should_parallelize);

/* A registered constraint as the device drivers run it (0, or -1 when the
 * GPU path does not provide it; the reason is printed). */
int spl_map_constraint(splatt_cpd_constraint const * con, idx_t mode,
    idx_t chunk, spd_con * out);

/* glibc rand() initialisation (src/util.c:15-40). */
val_t spl_rand_val(void);
void spl_fill_rand(val_t * vals, idx_t nelems);

#endif
