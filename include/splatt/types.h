/*
 * splatt/types.h -- primitive types and enums of the SPLATT C API.
 *
 * Same names, widths and enumerator order as the reference API
 * (include/splatt/types_config.h:34-178, configured with its defaults
 * cmake/types.cmake:3-5: 64-bit indices, 64-bit values, 32-bit BLAS ints),
 * so that code written against the reference compiles and links unchanged.
 */
#ifndef SPLATT_SPLATT_TYPES_H
#define SPLATT_SPLATT_TYPES_H

#include <inttypes.h>
#include <stdbool.h>
#include <float.h>

#define SPLATT_IDX_TYPEWIDTH 64
#define SPLATT_VAL_TYPEWIDTH 64
#define SPLATT_BLAS_INTWIDTH 32

typedef uint64_t splatt_idx_t;
#define SPLATT_IDX_MAX UINT64_MAX
#define SPLATT_PF_IDX PRIu64

typedef double splatt_val_t;
#define SPLATT_VAL_MIN DBL_MIN
#define SPLATT_VAL_MAX DBL_MAX
#define SPLATT_PF_VAL "f"

typedef int32_t splatt_blas_int;

/* Index into the options array returned by splatt_default_opts(). */
typedef enum
{
  SPLATT_OPTION_NTHREADS,
  SPLATT_OPTION_TOLERANCE,
  SPLATT_OPTION_REGULARIZE,
  SPLATT_OPTION_NITER,
  SPLATT_OPTION_VERBOSITY,
  SPLATT_OPTION_RANDSEED,
  SPLATT_OPTION_CSF_ALLOC,
  SPLATT_OPTION_TILE,
  SPLATT_OPTION_TILELEVEL,
  SPLATT_OPTION_PRIVTHRESH,
  SPLATT_OPTION_DECOMP,
  SPLATT_OPTION_COMM,
  SPLATT_OPTION_NOPTIONS
} splatt_option_type;

/* Return codes. Success is 1, as in the reference. */
typedef enum
{
  SPLATT_SUCCESS = 1,
  SPLATT_ERROR_BADINPUT,
  SPLATT_ERROR_NOMEMORY
} splatt_error_type;

typedef enum
{
  SPLATT_VERBOSITY_NONE,
  SPLATT_VERBOSITY_LOW,
  SPLATT_VERBOSITY_HIGH,
  SPLATT_VERBOSITY_MAX
} splatt_verbosity_type;

typedef enum
{
  SPLATT_NOTILE,
  SPLATT_DENSETILE,
  SPLATT_SYNCTILE,   /* accepted for ABI compatibility; rejected by CSF */
  SPLATT_COOPTILE,   /* accepted for ABI compatibility; rejected by CSF */
} splatt_tile_type;

typedef enum
{
  SPLATT_CSF_ONEMODE,
  SPLATT_CSF_TWOMODE,
  SPLATT_CSF_ALLMODE,
} splatt_csf_type;

/* Distributed decompositions exist only so that option arrays keep their
 * reference values; this library runs on one GPU and ignores them. */
typedef enum
{
  SPLATT_DECOMP_COARSE,
  SPLATT_DECOMP_MEDIUM,
  SPLATT_DECOMP_FINE
} splatt_decomp_type;

typedef enum
{
  SPLATT_COMM_POINT2POINT,
  SPLATT_COMM_ALL2ALL
} splatt_comm_type;

#endif
