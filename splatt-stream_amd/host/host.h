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

/* Pieces of the host text reader (coord.c). spl_slurp: the whole stream
 * in one '\0'-terminated buffer; spl_read_whole: the same for a regular
 * file, read by all host threads (NULL otherwise). spl_text_nmodes: modes
 * from the first data
 * line (0 if none). spl_line_convert: one data line converted with
 * strtoull/strtod (an spd_line_fn). spl_parse_text: the whole buffer (its
 * newlines are overwritten). spl_read_binary: the .bin format. */
char * spl_slurp(FILE * fin, size_t * out_len);
char * spl_read_whole(FILE * fin, size_t * out_len);
idx_t spl_text_nmodes(char const * buf, size_t len);
int spl_line_convert(char const * line, size_t len, int nmodes, uint64_t * inds,
    double * val);
sptensor_t * spl_parse_text(char * buf, size_t len);
sptensor_t * spl_read_binary(FILE * fin);

/* A tensor file into device memory (coordinates 0-based, dims as the host
 * reader's); *nhost = lines converted on the host. SPD_* status. */
int spl_coo_read_dev(char const * fname, spd_coo ** out, uint64_t * nhost);

/* glibc rand() initialisation (src/util.c:15-40). */
val_t spl_rand_val(void);
void spl_fill_rand(val_t * vals, idx_t nelems);

#endif
