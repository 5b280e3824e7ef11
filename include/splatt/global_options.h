/*
 * splatt/global_options.h -- process-wide options
 * (replaces include/splatt/global_options.h:24-58; same struct layout).
 */
#ifndef SPLATT_API_GLOBAL_OPTIONS_H
#define SPLATT_API_GLOBAL_OPTIONS_H

typedef struct
{
  int num_threads;                 /* host threads (CPU-side helpers only) */
  splatt_verbosity_type verbosity;
  int random_seed;
} splatt_global_opts;

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/splatt/global_options.h:46 (src/opts.c:55). */
splatt_global_opts * splatt_alloc_global_opts(void);

/* Replaces include/splatt/global_options.h:54 (src/opts.c:77). */
void splatt_free_global_opts(splatt_global_opts * opts);

#ifdef __cplusplus
}
#endif
#endif
