/*
 * splatt/constants.h -- compile-time limits (include/splatt/constants.h:14-18).
 */
#ifndef SPLATT_SPLATT_CONSTANTS_H
#define SPLATT_SPLATT_CONSTANTS_H

#ifndef SPLATT_MAX_NMODES
#define SPLATT_MAX_NMODES ((splatt_idx_t) 8)
#endif

/* Sentinel for an unset entry of the options array. */
static double const SPLATT_VAL_OFF = -DBL_MAX;

#endif
