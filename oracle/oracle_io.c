/*
 * oracle_io.c -- TEST INFRASTRUCTURE (the oracle; see oracle.h).
 * .tns reading and empty-slice compaction, restated on the CPU.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

or_coo * or_coo_alloc(uint64_t nmodes, uint64_t nnz)
{
  or_coo * tt = calloc(1, sizeof(*tt));
  tt->nmodes = nmodes;
  tt->nnz = nnz;
  for(uint64_t m=0; m < nmodes; ++m) {
    tt->ind[m] = malloc((nnz ? nnz : 1) * sizeof(uint64_t));
  }
  tt->vals = malloc((nnz ? nnz : 1) * sizeof(double));
  return tt;
}

void or_coo_free(or_coo * tt)
{
  if(tt == NULL) {
    return;
  }
  for(uint64_t m=0; m < tt->nmodes; ++m) {
    free(tt->ind[m]);
  }
  free(tt->vals);
  free(tt);
}

/* Two passes over the text, as src/io.c:273-348 then :62-108: the first
 * finds nmodes (tokens on the first data line minus one), nnz, the largest
 * index per mode and the smallest (0 or 1) as the offset; dims are the
 * largest index, plus one for 0-indexed modes. Lines that are empty or start
 * with '#' are skipped. */
or_coo * or_tns_read(char const * path)
{
  FILE * fin = fopen(path, "r");
  if(fin == NULL) {
    return NULL;
  }
  char * line = NULL;
  size_t len = 0;
  ssize_t rd;
  uint64_t nmodes = 0;
  while((rd = getline(&line, &len, fin)) != -1) {
    if(rd > 1 && line[0] != '#') {
      char * save = NULL;
      char * tok = strtok_r(line, " \t", &save);
      while(tok != NULL) {
        ++nmodes;
        tok = strtok_r(NULL, " \t", &save);
      }
      break;
    }
  }
  if(nmodes < 2) {
    free(line);
    fclose(fin);
    return NULL;
  }
  --nmodes;
  uint64_t dims[OR_MAX_NMODES];
  uint64_t offs[OR_MAX_NMODES];
  for(uint64_t m=0; m < nmodes; ++m) {
    dims[m] = 0;
    offs[m] = 1;
  }
  rewind(fin);
  uint64_t nnz = 0;
  while((rd = getline(&line, &len, fin)) != -1) {
    if(rd > 1 && line[0] != '#') {
      char * p = line;
      for(uint64_t m=0; m < nmodes; ++m) {
        uint64_t v = strtoull(p, &p, 10);
        if(v > dims[m]) dims[m] = v;
        if(v < offs[m]) offs[m] = v;
      }
      ++nnz;
    }
  }
  or_coo * tt = or_coo_alloc(nmodes, nnz);
  for(uint64_t m=0; m < nmodes; ++m) {
    tt->dims[m] = dims[m] + (offs[m] == 0 ? 1 : 0);
  }
  rewind(fin);
  uint64_t n = 0;
  while((rd = getline(&line, &len, fin)) != -1) {
    if(rd > 1 && line[0] != '#') {
      char * p = line;
      for(uint64_t m=0; m < nmodes; ++m) {
        tt->ind[m][n] = strtoull(p, &p, 10) - offs[m];
      }
      tt->vals[n++] = strtod(p, &p);
    }
  }
  free(line);
  fclose(fin);
  return tt;
}

/* src/sptensor.c:190-252: per mode, mark present indices, relabel them
 * 0..k-1 in increasing order, shrink the dim. */
uint64_t or_remove_empty(or_coo * tt)
{
  uint64_t removed = 0;
  for(uint64_t m=0; m < tt->nmodes; ++m) {
    uint64_t const dim = tt->dims[m];
    uint64_t * map = calloc(dim ? dim : 1, sizeof(uint64_t));
    for(uint64_t n=0; n < tt->nnz; ++n) {
      map[tt->ind[m][n]] = 1;
    }
    uint64_t k = 0;
    for(uint64_t i=0; i < dim; ++i) {
      if(map[i]) {
        map[i] = k++;
      } else {
        map[i] = UINT64_MAX;
      }
    }
    if(k != dim) {
      removed += dim - k;
      for(uint64_t n=0; n < tt->nnz; ++n) {
        tt->ind[m][n] = map[tt->ind[m][n]];
      }
      tt->dims[m] = k;
    }
    free(map);
  }
  return removed;
}
