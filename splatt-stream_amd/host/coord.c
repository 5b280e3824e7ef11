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
#include <omp.h>
#include <sys/stat.h>
#include <unistd.h>

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

/* The whole stream into one '\0'-terminated buffer. */
char * spl_slurp(FILE * fin, size_t * out_len)
{
  size_t cap = 1 << 20;
  size_t len = 0;
  char * buf = malloc(cap + 1);
  if(buf == NULL) {
    return NULL;
  }
  while(1) {
    if(len == cap) {
      char * nb = realloc(buf, 2 * cap + 1);
      if(nb == NULL) {
        free(buf);
        return NULL;
      }
      buf = nb;
      cap *= 2;
    }
    size_t const got = fread(buf + len, 1, cap - len, fin);
    len += got;
    if(got == 0) {
      break;
    }
  }
  buf[len] = '\0';
  *out_len = len;
  return buf;
}

/* A regular file read whole into one '\0'-terminated buffer: its size from
 * fstat, its bytes by all host threads (pread of one range each, so the
 * page-cache copy runs in parallel and the buffer is touched first by the
 * thread that fills it). NULL when fin is not a regular file or a read
 * comes up short (the caller then reads the stream with spl_slurp). */
char * spl_read_whole(FILE * fin, size_t * out_len)
{
  struct stat st;
  int const fd = fileno(fin);
  if(fd < 0 || fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
    return NULL;
  }
  size_t const len = (size_t) st.st_size;
  char * buf = malloc(len + 1);
  if(buf == NULL) {
    return NULL;
  }
  size_t const chunk = (size_t) 8 << 20;
  size_t const nch = (len + chunk - 1) / chunk;
  int bad = 0;
  #pragma omp parallel for schedule(dynamic, 1) reduction(|:bad)
  for(size_t c=0; c < nch; ++c) {
    size_t const lo = c * chunk;
    size_t const n = (lo + chunk < len) ? chunk : len - lo;
    size_t got = 0;
    while(got < n) {
      ssize_t const r = pread(fd, buf + lo + got, n - got, (off_t) (lo + got));
      if(r <= 0) {
        bad = 1;
        break;
      }
      got += (size_t) r;
    }
  }
  if(bad) {
    free(buf);
    return NULL;
  }
  buf[len] = '\0';
  *out_len = len;
  return buf;
}

/* End of the line starting at s (the '\n' or the buffer's end). */
static size_t p_line_end(char const * buf, size_t s, size_t len)
{
  char const * e = memchr(buf + s, '\n', len - s);
  return e ? (size_t) (e - buf) : len;
}

/* A data line as the reference's getline loop sees it (src/io.c:94-104):
 * its length with the '\n' above 1 and a first character other than '#'. */
static int p_is_data(char const * buf, size_t s, size_t e, size_t len)
{
  size_t const rd = e - s + (e < len ? 1 : 0);
  return rd > 1 && buf[s] != '#';
}

/* Modes of a text tensor (src/io.c:273-300): the fields of its first data
 * line, split at ' ' and '\t' with the line's '\n' kept (a trailing blank
 * before it makes the '\n' a field of its own, as in the reference), less
 * the value. 0 when there is no data line or fewer than two fields. */
idx_t spl_text_nmodes(char const * buf, size_t len)
{
  idx_t nmodes = 0;
  for(size_t s=0; s < len; ) {
    size_t const e = p_line_end(buf, s, len);
    if(p_is_data(buf, s, e, len)) {
      char * line = malloc(e - s + 2);
      if(line == NULL) {
        return 0;
      }
      memcpy(line, buf + s, e - s);
      line[e - s] = '\0';
      char * save = NULL;
      if(e < len) {
        line[e - s] = '\n';
        line[e - s + 1] = '\0';
      }
      for(char * tok = strtok_r(line, " \t", &save); tok != NULL;
          tok = strtok_r(NULL, " \t", &save)) {
        ++nmodes;
      }
      free(line);
      break;
    }
    s = e + 1;
  }
  return nmodes < 2 ? 0 : nmodes - 1;
}

/* One data line [line, line + len) converted as the host reader converts
 * it: strtoull per mode, then strtod, on a '\0'-terminated copy. */
int spl_line_convert(char const * line, size_t len, int nmodes,
    uint64_t * inds, double * val)
{
  char small[256];
  char * tmp = (len < sizeof(small)) ? small : malloc(len + 1);
  if(tmp == NULL) {
    return -1;
  }
  memcpy(tmp, line, len);
  tmp[len] = '\0';
  char * p = tmp;
  for(int m=0; m < nmodes; ++m) {
    inds[m] = strtoull(p, &p, 10);
  }
  *val = strtod(p, &p);
  if(tmp != small) {
    free(tmp);
  }
  return 0;
}

/* Text input, the same conversions as the reference's two getline passes
 * (strtoull per mode, then strtod) over the whole file read at once: the
 * buffer is cut into one range of whole lines per thread; a first pass
 * counts each range's data lines, a second converts each line (its '\n'
 * replaced by '\0', so no conversion runs past it, as with getline's
 * buffer) into its place and keeps per-thread index extremes. buf holds
 * len bytes and a '\0'; its newlines are overwritten. */
sptensor_t * spl_parse_text(char * buf, size_t len)
{
  idx_t const nmodes = spl_text_nmodes(buf, len);
  if(nmodes == 0) {
    return NULL;
  }
  if(nmodes > SPLATT_MAX_NMODES) {
    fprintf(stderr, "SPLATT ERROR: maximum %d modes supported. Found %"
        SPLATT_PF_IDX ".\n", (int) SPLATT_MAX_NMODES, nmodes);
    return NULL;
  }
  /* one range of whole lines per thread */
  int const T = omp_get_max_threads() > 0 ? omp_get_max_threads() : 1;
  size_t * cut = malloc((T + 1) * sizeof(size_t));
  idx_t * cnt = calloc(T + 1, sizeof(idx_t));
  idx_t * tmax = calloc((size_t) T * SPLATT_MAX_NMODES, sizeof(idx_t));
  idx_t * tmin = malloc((size_t) T * SPLATT_MAX_NMODES * sizeof(idx_t));
  cut[0] = 0;
  for(int k=1; k < T; ++k) {
    size_t c = (size_t) ((unsigned __int128) len * k / T);
    if(c < cut[k - 1]) {
      c = cut[k - 1];
    }
    /* the first line starting at or after c */
    if(c > 0 && c < len && buf[c - 1] != '\n') {
      c = p_line_end(buf, c, len);
      c = (c < len) ? c + 1 : len;
    }
    cut[k] = c;
  }
  cut[T] = len;
  #pragma omp parallel for schedule(static, 1) num_threads(T)
  for(int k=0; k < T; ++k) {
    idx_t n = 0;
    for(size_t s=cut[k]; s < cut[k + 1]; ) {
      size_t const e = p_line_end(buf, s, len);
      n += p_is_data(buf, s, e, len) ? 1 : 0;
      s = e + 1;
    }
    cnt[k + 1] = n;
  }
  for(int k=0; k < T; ++k) {
    cnt[k + 1] += cnt[k];
  }
  idx_t const nnz = cnt[T];
  sptensor_t * tt = splatt_alloc_coord(nmodes, nnz);
  #pragma omp parallel for schedule(static, 1) num_threads(T)
  for(int k=0; k < T; ++k) {
    idx_t * const mx = tmax + (size_t) k * SPLATT_MAX_NMODES;
    idx_t * const mn = tmin + (size_t) k * SPLATT_MAX_NMODES;
    for(idx_t m=0; m < nmodes; ++m) {
      mx[m] = 0;
      mn[m] = 1;
    }
    idx_t n = cnt[k];
    for(size_t s=cut[k]; s < cut[k + 1]; ) {
      size_t const e = p_line_end(buf, s, len);
      if(p_is_data(buf, s, e, len)) {
        buf[e] = '\0';
        char * p = buf + s;
        for(idx_t m=0; m < nmodes; ++m) {
          idx_t const v = strtoull(p, &p, 10);
          tt->ind[m][n] = v;
          mx[m] = v > mx[m] ? v : mx[m];
          mn[m] = v < mn[m] ? v : mn[m];
        }
        tt->vals[n] = strtod(p, &p);
        ++n;
      }
      s = e + 1;
    }
  }
  idx_t dims[SPLATT_MAX_NMODES];
  idx_t offs[SPLATT_MAX_NMODES];
  for(idx_t m=0; m < nmodes; ++m) {
    dims[m] = 0;
    offs[m] = 1;
    for(int k=0; k < T; ++k) {
      dims[m] = SS_MAX(dims[m], tmax[(size_t) k * SPLATT_MAX_NMODES + m]);
      offs[m] = SS_MIN(offs[m], tmin[(size_t) k * SPLATT_MAX_NMODES + m]);
    }
  }
  free(cut);
  free(cnt);
  free(tmax);
  free(tmin);
  for(idx_t m=0; m < nmodes; ++m) {
    if(offs[m] != 0 && offs[m] != 1) {
      fprintf(stderr, "SPLATT: ERROR tensors must be 0 or 1 indexed. Mode %"
          SPLATT_PF_IDX " is %" SPLATT_PF_IDX " indexed.\n", m, offs[m]);
      splatt_free_coord(tt);
      return NULL;
    }
  }
  for(idx_t m=0; m < nmodes; ++m) {
    tt->dims[m] = dims[m] + (offs[m] == 0 ? 1 : 0);
    if(offs[m] == 1) {
      idx_t * const ind = tt->ind[m];
      #pragma omp parallel for schedule(static)
      for(idx_t n=0; n < nnz; ++n) {
        ind[n] -= 1;
      }
    }
  }
  return tt;
}

static sptensor_t * p_read_text(FILE * fin)
{
  size_t len = 0;
  char * buf = spl_read_whole(fin, &len);
  if(buf == NULL) {
    buf = spl_slurp(fin, &len);
  }
  if(buf == NULL) {
    return NULL;
  }
  sptensor_t * tt = spl_parse_text(buf, len);
  free(buf);
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

sptensor_t * spl_read_binary(FILE * fin)
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
    tt = spl_read_binary(fin);
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
 * 0..k-1 keeping their order, record local->global in indmap. Marking
 * (atomic stores of 1) and relabelling run on all host threads; the
 * numbering of the present indices is a serial pass over the dimension. */
splatt_idx_t splatt_tt_remove_empty(sptensor_t * const tt)
{
  idx_t removed = 0;
  idx_t maxdim = 0;
  for(idx_t m=0; m < tt->nmodes; ++m) {
    maxdim = SS_MAX(maxdim, tt->dims[m]);
  }
  idx_t * map = splatt_malloc((maxdim ? maxdim : 1) * sizeof(idx_t));
  idx_t const nnz = tt->nnz;
  for(idx_t m=0; m < tt->nmodes; ++m) {
    idx_t const dim = tt->dims[m];
    idx_t * const ind = tt->ind[m];
    memset(map, 0, dim * sizeof(idx_t));
    #pragma omp parallel for schedule(static)
    for(idx_t n=0; n < nnz; ++n) {
      #pragma omp atomic write
      map[ind[n]] = 1;
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
    #pragma omp parallel for schedule(static)
    for(idx_t n=0; n < nnz; ++n) {
      ind[n] = map[ind[n]];
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
