/*
 * coord.c -- COO tensors: allocation, file input (.tns/.coo text, .bin),
 * empty-slice compaction, norms.
 *
 * Text format (src/io.c:62-108, :273-348): one nonzero per line,
 * "i_1 ... i_N value"; lines that are empty or start with '#' are skipped;
 * each mode is 1-indexed unless its smallest index is 0; dims are the
 * largest index (+1 for 0-indexed modes).
 * Binary format (src/io.c:118-200, :477-560): int32 magic, uint64 index
 * width, uint64 value width, then nmodes, dims, nnz, each mode's indices
 * and the values.
 */
#define _GNU_SOURCE
#include "host.h"
#include <math.h>

sptensor_t * splatt_tt_alloc(splatt_idx_t const nnz, splatt_idx_t const nmodes)
{
  return splatt_alloc_coord(nmodes, nnz);
}

splatt_coord * splatt_alloc_coord(splatt_idx_t const num_modes,
    splatt_idx_t const nnz)
{
  splatt_coord * c = calloc(1, sizeof(*c));
  if(c == NULL) {
    return NULL;
  }
  c->nmodes = num_modes;
  c->nnz = nnz;
  c->vals = splatt_malloc((nnz ? nnz : 1) * sizeof(*c->vals));
  for(idx_t m=0; m < num_modes && m < SPLATT_MAX_NMODES; ++m) {
    c->ind[m] = splatt_malloc((nnz ? nnz : 1) * sizeof(**c->ind));
  }
  c->tiled = SPLATT_NOTILE;
  return c;
}

void splatt_free_coord(splatt_coord * c)
{
  if(c == NULL) {
    return;
  }
  splatt_free(c->vals);
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    splatt_free(c->ind[m]);
    splatt_free(c->indmap[m]);
  }
  free(c);
}

void splatt_tt_free(sptensor_t * tt)
{
  splatt_free_coord(tt);
}

splatt_idx_t splatt_coord_get_nnz(splatt_coord const * const c) { return c->nnz; }
splatt_idx_t splatt_coord_get_modes(splatt_coord const * const c) { return c->nmodes; }
splatt_idx_t * splatt_coord_get_inds(splatt_coord const * const c,
    splatt_idx_t const mode) { return c->ind[mode]; }
splatt_val_t * splatt_coord_get_vals(splatt_coord const * const c) { return c->vals; }

static sptensor_t * p_read_text(FILE * fin)
{
  char * line = NULL;
  size_t len = 0;
  ssize_t rd;
  idx_t nmodes = 0;
  while((rd = getline(&line, &len, fin)) != -1) {
    if(rd > 1 && line[0] != '#') {
      char * save = NULL;
      for(char * tok = strtok_r(line, " \t\n", &save); tok != NULL;
          tok = strtok_r(NULL, " \t\n", &save)) {
        ++nmodes;
      }
      break;
    }
  }
  if(nmodes < 2) {
    free(line);
    return NULL;
  }
  --nmodes;
  if(nmodes > SPLATT_MAX_NMODES) {
    fprintf(stderr, "SPLATT ERROR: maximum %d modes supported. Found %"
        SPLATT_PF_IDX ".\n", (int) SPLATT_MAX_NMODES, nmodes);
    free(line);
    return NULL;
  }
  idx_t dims[SPLATT_MAX_NMODES];
  idx_t offs[SPLATT_MAX_NMODES];
  for(idx_t m=0; m < nmodes; ++m) {
    dims[m] = 0;
    offs[m] = 1;
  }
  rewind(fin);
  idx_t nnz = 0;
  while((rd = getline(&line, &len, fin)) != -1) {
    if(rd > 1 && line[0] != '#') {
      char * p = line;
      for(idx_t m=0; m < nmodes; ++m) {
        idx_t const v = strtoull(p, &p, 10);
        dims[m] = v > dims[m] ? v : dims[m];
        offs[m] = v < offs[m] ? v : offs[m];
      }
      ++nnz;
    }
  }
  for(idx_t m=0; m < nmodes; ++m) {
    if(offs[m] != 0 && offs[m] != 1) {
      fprintf(stderr, "SPLATT: ERROR tensors must be 0 or 1 indexed. Mode %"
          SPLATT_PF_IDX " is %" SPLATT_PF_IDX " indexed.\n", m, offs[m]);
      free(line);
      return NULL;
    }
  }
  sptensor_t * tt = splatt_alloc_coord(nmodes, nnz);
  for(idx_t m=0; m < nmodes; ++m) {
    tt->dims[m] = dims[m] + (offs[m] == 0 ? 1 : 0);
  }
  rewind(fin);
  idx_t n = 0;
  while((rd = getline(&line, &len, fin)) != -1) {
    if(rd > 1 && line[0] != '#') {
      char * p = line;
      for(idx_t m=0; m < nmodes; ++m) {
        tt->ind[m][n] = strtoull(p, &p, 10) - offs[m];
      }
      tt->vals[n++] = strtod(p, &p);
    }
  }
  free(line);
  return tt;
}

static int p_read_idx(FILE * fin, uint64_t width, idx_t * out, idx_t count)
{
  if(width == sizeof(uint64_t)) {
    return fread(out, sizeof(uint64_t), count, fin) == count ? 0 : -1;
  }
  if(width != sizeof(uint32_t)) {
    return -1;
  }
  uint32_t buf[4096];
  for(idx_t n=0; n < count; ) {
    idx_t const k = SS_MIN((idx_t) 4096, count - n);
    if(fread(buf, sizeof(uint32_t), k, fin) != k) {
      return -1;
    }
    for(idx_t i=0; i < k; ++i) {
      out[n + i] = buf[i];
    }
    n += k;
  }
  return 0;
}

static sptensor_t * p_read_binary(FILE * fin)
{
  int32_t magic = 0;
  uint64_t iw = 0, vw = 0;
  if(fread(&magic, sizeof(magic), 1, fin) != 1 ||
     fread(&iw, sizeof(iw), 1, fin) != 1 ||
     fread(&vw, sizeof(vw), 1, fin) != 1 || magic != 0) {
    return NULL;
  }
  idx_t nmodes = 0, nnz = 0;
  idx_t dims[SPLATT_MAX_NMODES];
  if(p_read_idx(fin, iw, &nmodes, 1) || nmodes > SPLATT_MAX_NMODES ||
     p_read_idx(fin, iw, dims, nmodes) || p_read_idx(fin, iw, &nnz, 1)) {
    return NULL;
  }
  sptensor_t * tt = splatt_alloc_coord(nmodes, nnz);
  memcpy(tt->dims, dims, nmodes * sizeof(idx_t));
  for(idx_t m=0; m < nmodes; ++m) {
    if(p_read_idx(fin, iw, tt->ind[m], nnz)) {
      splatt_free_coord(tt);
      return NULL;
    }
  }
  if(vw == sizeof(double)) {
    if(fread(tt->vals, sizeof(double), nnz, fin) != nnz) {
      splatt_free_coord(tt);
      return NULL;
    }
  } else if(vw == sizeof(float)) {
    float buf[4096];
    for(idx_t n=0; n < nnz; ) {
      idx_t const k = SS_MIN((idx_t) 4096, nnz - n);
      if(fread(buf, sizeof(float), k, fin) != k) {
        splatt_free_coord(tt);
        return NULL;
      }
      for(idx_t i=0; i < k; ++i) {
        tt->vals[n + i] = buf[i];
      }
      n += k;
    }
  } else {
    splatt_free_coord(tt);
    return NULL;
  }
  return tt;
}

sptensor_t * splatt_tt_read(char const * const ifname)
{
  FILE * fin = fopen(ifname, "r");
  if(fin == NULL) {
    fprintf(stderr, "SPLATT ERROR: failed to open '%s'\n", ifname);
    return NULL;
  }
  char const * suffix = strrchr(ifname, '.');
  sptensor_t * tt = NULL;
  if(suffix != NULL && strcmp(suffix, ".bin") == 0) {
    tt = p_read_binary(fin);
  } else {
    tt = p_read_text(fin);
  }
  fclose(fin);
  return tt;
}

splatt_error_type splatt_coord_load(char const * const fname,
    splatt_coord * const coord)
{
  sptensor_t * tt = splatt_tt_read(fname);
  if(tt == NULL) {
    return SPLATT_ERROR_BADINPUT;
  }
  /* replace the arrays of the caller's structure */
  splatt_free(coord->vals);
  for(idx_t m=0; m < SPLATT_MAX_NMODES; ++m) {
    splatt_free(coord->ind[m]);
    splatt_free(coord->indmap[m]);
  }
  *coord = *tt;
  free(tt);
  return SPLATT_SUCCESS;
}

/* src/sptensor.c:190-252: per mode, relabel the present indices to
 * 0..k-1 keeping their order, record local->global in indmap. */
splatt_idx_t splatt_tt_remove_empty(sptensor_t * const tt)
{
  idx_t removed = 0;
  idx_t maxdim = 0;
  for(idx_t m=0; m < tt->nmodes; ++m) {
    maxdim = SS_MAX(maxdim, tt->dims[m]);
  }
  idx_t * map = splatt_malloc((maxdim ? maxdim : 1) * sizeof(idx_t));
  for(idx_t m=0; m < tt->nmodes; ++m) {
    idx_t const dim = tt->dims[m];
    memset(map, 0, dim * sizeof(idx_t));
    for(idx_t n=0; n < tt->nnz; ++n) {
      map[tt->ind[m][n]] = 1;
    }
    idx_t k = 0;
    for(idx_t i=0; i < dim; ++i) {
      k += map[i];
    }
    if(k == dim) {
      splatt_free(tt->indmap[m]);
      tt->indmap[m] = NULL;
      continue;
    }
    removed += dim - k;
    splatt_free(tt->indmap[m]);
    tt->indmap[m] = splatt_malloc((k ? k : 1) * sizeof(idx_t));
    idx_t next = 0;
    for(idx_t i=0; i < dim; ++i) {
      if(map[i]) {
        tt->indmap[m][next] = i;
        map[i] = next++;
      }
    }
    for(idx_t n=0; n < tt->nnz; ++n) {
      tt->ind[m][n] = map[tt->ind[m][n]];
    }
    tt->dims[m] = k;
  }
  splatt_free(map);
  return removed;
}

/* src/sptensor.c:71-80 (accumulates in val_t). */
splatt_val_t splatt_tt_normsq(sptensor_t const * const tt)
{
  val_t norm = 0;
  for(idx_t n=0; n < tt->nnz; ++n) {
    norm += tt->vals[n] * tt->vals[n];
  }
  return norm;
}
