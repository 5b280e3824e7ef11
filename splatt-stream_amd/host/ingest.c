/*
 * ingest.c -- tensor files into device memory: the text parsed on the GPU
 * (spd_tns_parse) with the host reader's line conversion for the lines the
 * GPU leaves to the C library, binary files read on the host and uploaded.
 */
#include "host.h"

/* A tensor file read into device memory. Text is parsed on the device
 * (spd_tns_parse), with spl_line_convert for the lines whose fields it leaves
 * to the C library; binary files, and texts with more such lines than the
 * device path takes, are read on the host and uploaded. Either way the
 * coordinates, values and dims equal the host reader's. */
int spl_coo_read_dev(char const * fname, spd_coo ** out, uint64_t * nhost)
{
  *out = NULL;
  if(nhost) {
    *nhost = 0;
  }
  FILE * fin = fopen(fname, "r");
  if(fin == NULL) {
    fprintf(stderr, "SPLATT ERROR: failed to open '%s'\n", fname);
    return SPD_EINPUT;
  }
  char const * suffix = strrchr(fname, '.');
  sptensor_t * tt = NULL;
  if(suffix != NULL && strcmp(suffix, ".bin") == 0) {
    tt = spl_read_binary(fin);
    fclose(fin);
  } else {
    size_t len = 0;
    char * buf = spl_read_whole(fin, &len);
    if(buf == NULL) {
      buf = spl_slurp(fin, &len);
    }
    fclose(fin);
    if(buf == NULL) {
      return SPD_ENOMEM;
    }
    idx_t const nmodes = spl_text_nmodes(buf, len);
    if(nmodes == 0 || nmodes > SPLATT_MAX_NMODES) {
      if(nmodes > SPLATT_MAX_NMODES) {
        fprintf(stderr, "SPLATT ERROR: maximum %d modes supported. Found %"
            SPLATT_PF_IDX ".\n", (int) SPLATT_MAX_NMODES, nmodes);
      }
      free(buf);
      return SPD_EINPUT;
    }
    uint64_t nh = 0;
    int rc = spd_tns_parse(buf, len, (int) nmodes, spl_line_convert, out, &nh);
    if(nhost) {
      *nhost = nh;
    }
    if(rc != SPD_EHOST) {
      free(buf);
      return rc;
    }
    tt = spl_parse_text(buf, len);
    free(buf);
  }
  if(tt == NULL) {
    return SPD_EINPUT;
  }
  int const rc = spd_coo_upload((int) tt->nmodes, tt->nnz, tt->dims,
      (uint64_t const * const *) tt->ind, tt->vals, out);
  splatt_tt_free(tt);
  return rc;
}

splatt_error_type splatt_amd_coord_load_dev(char const * const fname,
    splatt_coord * const coord, uint64_t * const nhost)
{
  spd_coo * d = NULL;
  int rc = spl_coo_read_dev(fname, &d, nhost);
  if(rc != SPD_OK) {
    return (rc == SPD_EINPUT) ? SPLATT_ERROR_BADINPUT : spl_report(rc, "tensor input");
  }
  int nm = 0;
  uint64_t nnz = 0;
  uint64_t dims[SPLATT_MAX_NMODES];
  spd_coo_dims(d, &nm, &nnz, dims);
  sptensor_t * tt = splatt_alloc_coord((idx_t) nm, nnz);
  if(tt == NULL) {
    spd_coo_free(d);
    return SPLATT_ERROR_NOMEMORY;
  }
  for(int m=0; m < nm; ++m) {
    tt->dims[m] = dims[m];
  }
  rc = spd_coo_download(d, (uint64_t * const *) tt->ind, tt->vals);
  spd_coo_free(d);
  if(rc != SPD_OK) {
    splatt_free_coord(tt);
    return spl_report(rc, "tensor download");
  }
  splatt_free(coord->vals);
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    splatt_free(coord->ind[m]);
    splatt_free(coord->indmap[m]);
  }
  *coord = *tt;
  free(tt);
  return SPLATT_SUCCESS;
}
