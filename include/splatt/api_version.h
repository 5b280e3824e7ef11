/*
 * splatt/api_version.h -- version of the API this library is a drop-in for.
 */
#ifndef SPLATT_SPLATT_VERSION_H
#define SPLATT_SPLATT_VERSION_H

#define SPLATT_VER_MAJOR     2
#define SPLATT_VER_MINOR     0
#define SPLATT_VER_SUBMINOR  0

#ifdef __cplusplus
extern "C" {
#endif

/* Replace include/splatt/api_version.h:40-56 (src/base.c:16-30). */
int splatt_version_major(void);
int splatt_version_minor(void);
int splatt_version_subminor(void);

#ifdef __cplusplus
}
#endif
#endif
