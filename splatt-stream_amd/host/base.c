/*
 * base.c -- allocation, versions, error mapping, RNG helpers.
 */
#include "host.h"
#include <errno.h>

/* src/base.c:42-67: 64-byte aligned allocation, NULL (with a message) on
 * failure. */
void * splatt_malloc(size_t const bytes)
{
  void * ptr = NULL;
  int const rc = posix_memalign(&ptr, 64, bytes ? bytes : 64);
  if(rc != 0) {
    fprintf(stderr, "SPLATT: posix_memalign() of %zu bytes failed (%s).\n",
        bytes, rc == ENOMEM ? "ENOMEM" : "EINVAL");
    return NULL;
  }
  return ptr;
}

void splatt_free(void * ptr)
{
  free(ptr);
}

/* src/mpi_dummy.c:15-34: the non-MPI communicator is rank 0 of 1. */
splatt_comm_info * splatt_alloc_comm_info(void)
{
  splatt_comm_info * ci = (splatt_comm_info *) splatt_malloc(sizeof(*ci));
  if(ci == NULL) {
    return NULL;
  }
  ci->world_rank = 0;
  ci->world_npes = 1;
  return ci;
}

void splatt_free_comm_info(splatt_comm_info * comm_info)
{
  if(comm_info == NULL) {
    return;
  }
  splatt_free(comm_info);
}

int splatt_version_major(void) { return SPLATT_VER_MAJOR; }
int splatt_version_minor(void) { return SPLATT_VER_MINOR; }
int splatt_version_subminor(void) { return SPLATT_VER_SUBMINOR; }

splatt_error_type spl_report(int rc, char const * what)
{
  if(rc == SPD_OK) {
    return SPLATT_SUCCESS;
  }
  fprintf(stderr, "SPLATT: %s failed: %s\n", what, spd_last_error());
  if(rc == SPD_ENOMEM) {
    return SPLATT_ERROR_NOMEMORY;
  }
  return SPLATT_ERROR_BADINPUT;
}

char const * splatt_amd_last_error(void)
{
  return spd_last_error();
}

static int g_mttkrp_dtype = SPD_F64;

int spl_mttkrp_dtype(void) { return g_mttkrp_dtype; }

splatt_error_type splatt_amd_set_mttkrp_precision(int precision)
{
  if(precision != SPLATT_AMD_F64 && precision != SPLATT_AMD_F32) {
    return SPLATT_ERROR_BADINPUT;
  }
  g_mttkrp_dtype = (precision == SPLATT_AMD_F32) ? SPD_F32 : SPD_F64;
  return SPLATT_SUCCESS;
}

splatt_error_type splatt_amd_set_mttkrp_layout(int layout)
{
  if(layout != SPLATT_AMD_LAYOUT_AUTO && layout != SPLATT_AMD_LAYOUT_CSF) {
    return SPLATT_ERROR_BADINPUT;
  }
  spd_mttkrp_set_copies(layout == SPLATT_AMD_LAYOUT_AUTO);
  return SPLATT_SUCCESS;
}

/* src/util.c:15-24: magnitude 3*rand()/RAND_MAX, negated when a second
 * rand() is even. Drawn on the host so the sequence is glibc's. */
val_t spl_rand_val(void)
{
  val_t v = 3.0 * ((val_t) rand() / (val_t) RAND_MAX);
  if(rand() % 2 == 0) {
    v *= -1;
  }
  return v;
}

void spl_fill_rand(val_t * vals, idx_t nelems)
{
  for(idx_t i=0; i < nelems; ++i) {
    vals[i] = spl_rand_val();
  }
}
