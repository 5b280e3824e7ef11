/*
 * splatt.h -- entry point of the SPLATT C API (drop-in for the reference's
 * include/splatt.h:1-58), served by an MI355X (gfx950) implementation.
 */
#ifndef SPLATT_SPLATT_H
#define SPLATT_SPLATT_H

#include "splatt/types.h"
#include "splatt/constants.h"

#include "splatt/global_options.h"
#include "splatt/coord.h"
#include "splatt/csf.h"
#include "splatt/mpi.h"
#include "splatt/cpd.h"
#include "splatt/api_kernels.h"
#include "splatt/api_options.h"
#include "splatt/api_version.h"

#endif
