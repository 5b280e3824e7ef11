/*
 * splatt/api_options.h -- the options array.
 */
#ifndef SPLATT_SPLATT_OPTION_H
#define SPLATT_SPLATT_OPTION_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces include/splatt/api_options.h:37 (src/opts.c:10). Returns a
 * double[SPLATT_OPTION_NOPTIONS] filled with the reference defaults. */
double * splatt_default_opts(void);

/* Replaces include/splatt/api_options.h:43 (src/opts.c:47). */
void splatt_free_opts(double * opts);

#ifdef __cplusplus
}
#endif
#endif
