/*
 * csf.c -- CSF allocation and host mirrors.
 *
 * The tensor is uploaded once; every CSF of an allocation (ONE/TWO/ALL-MODE,
 * src/csf.c:695-739) is sorted and compressed on the GPU, then mirrored into
 * the reference's host structure (src/csf.h:28-94) with identical arrays.
 * The GPU copy stays attached (ct->dev) for MTTKRP and CPD.
 */
#include "host.h"

/* ---- mode orderings (src/csf.c:93-218, :608-651) ----------------------- */

/* Modes sorted by dim, ascending (small-first) or descending (big-first);
 * equal dims keep increasing mode index, which is what the reference's
 * match-the-first-unmatched-mode loop yields. */
static void p_order_by_dim(idx_t const * dims, idx_t nmodes, int desc,
    idx_t * perm)
{
  for(idx_t m=0; m < nmodes; ++m) {
    perm[m] = m;
  }
  for(idx_t i=1; i < nmodes; ++i) {
    idx_t const v = perm[i];
    idx_t j = i;
    for(; j > 0; --j) {
      idx_t const a = dims[perm[j-1]];
      int const before = desc ? (a < dims[v]) : (a > dims[v]);
      if(!before) {
        break;
      }
      perm[j] = perm[j-1];
    }
    perm[j] = v;
  }
}

static void p_to_front(idx_t * perm, idx_t nmodes, idx_t mode)
{
  for(idx_t m=0; m < nmodes; ++m) {
    if(perm[m] == mode) {
      memmove(perm + 1, perm, m * sizeof(*perm));
      perm[0] = mode;
      return;
    }
  }
}

void splatt_csf_find_mode_order(splatt_idx_t const * const dims,
    splatt_idx_t const nmodes, csf_mode_type which, splatt_idx_t const mode,
    splatt_idx_t * const perm)
{
  switch(which) {
  case CSF_SORTED_SMALLFIRST:
    p_order_by_dim(dims, nmodes, 0, perm);
    break;
  case CSF_SORTED_BIGFIRST:
    p_order_by_dim(dims, nmodes, 1, perm);
    break;
  case CSF_INORDER_MINUSONE:
    for(idx_t m=0; m < nmodes; ++m) {
      perm[m] = m;
    }
    p_to_front(perm, nmodes, mode);
    break;
  case CSF_SORTED_MINUSONE:
    p_order_by_dim(dims, nmodes, 0, perm);
    p_to_front(perm, nmodes, mode);
    break;
  case CSF_MODE_CUSTOM: {
    int seen[SPLATT_MAX_NMODES] = {0};
    for(idx_t m=0; m < nmodes; ++m) {
      if(perm[m] >= nmodes || seen[perm[m]]++) {
        fprintf(stderr, "SPLATT: invalid permutation for CSF_MODE_CUSTOM.\n");
        break;
      }
    }
    break;
  }
  default:
    fprintf(stderr, "SPLATT: csf_mode_type '%d' not recognized.\n", which);
    break;
  }
}

splatt_idx_t splatt_csf_mode_to_depth(splatt_csf const * const csf,
    splatt_idx_t const mode)
{
  return csf->dim_iperm[mode];
}

splatt_idx_t splatt_csf_depth_to_mode(splatt_csf const * const csf,
    splatt_idx_t const depth)
{
  return csf->dim_perm[depth];
}

/* ---- one tensor ---------------------------------------------------------- */

static void p_free_pattern(splatt_csf * ct)
{
  if(ct->pt == NULL) {
    return;
  }
  for(idx_t t=0; t < ct->ntiles; ++t) {
    free(ct->pt[t].vals);
    for(idx_t l=0; l < ct->nmodes; ++l) {
      free(ct->pt[t].fptr[l]);
      free(ct->pt[t].fids[l]);
    }
  }
  free(ct->pt);
  ct->pt = NULL;
}

/* Fill the host mirror of ct from its device tensor. */
static int p_mirror(splatt_csf * ct)
{
  idx_t const N = ct->nmodes;
  idx_t const T = ct->ntiles;
  idx_t * nf = splatt_malloc(T * N * sizeof(idx_t));
  int rc = spd_csf_tile_nfibs((spd_csf *) ct->dev, nf);
  if(rc != SPD_OK) {
    free(nf);
    return rc;
  }
  ct->pt = calloc(T, sizeof(csf_sparsity));
  for(idx_t t=0; t < T && rc == SPD_OK; ++t) {
    csf_sparsity * pt = ct->pt + t;
    idx_t const * tn = nf + t * N;
    if(tn[N-1] == 0) {
      /* empty tile, src/csf.c:467-479 */
      pt->fptr[0] = splatt_malloc(2 * sizeof(idx_t));
      pt->fptr[0][0] = 0;
      pt->fptr[0][1] = 0;
      continue;
    }
    for(idx_t l=0; l < N; ++l) {
      pt->nfibs[l] = tn[l];
      if(l < N-1) {
        pt->fptr[l] = splatt_malloc((tn[l] + 1) * sizeof(idx_t));
      }
      if(l > 0 || T > 1) {
        pt->fids[l] = splatt_malloc((tn[l] ? tn[l] : 1) * sizeof(idx_t));
      }
    }
    pt->vals = splatt_malloc(tn[N-1] * sizeof(val_t));
    rc = spd_csf_export_tile((spd_csf *) ct->dev, t, pt->fptr, pt->fids, pt->vals);
  }
  free(nf);
  return rc;
}

/* Leave tt in the order of ct's nonzeros (tiles in order), as the
 * reference's in-place tt_sort/tt_densetile do. Root ids are taken from the
 * device copy so that slice positions never stand in for indices. */
static int p_reorder_coo(splatt_csf const * ct, sptensor_t * tt)
{
  idx_t const N = ct->nmodes;
  idx_t off = 0;
  for(idx_t t=0; t < ct->ntiles; ++t) {
    csf_sparsity const * pt = ct->pt + t;
    if(pt->vals == NULL) {
      continue;
    }
    idx_t const nn = pt->nfibs[N-1];
    idx_t * roots = pt->fids[0];
    if(roots == NULL) {
      roots = splatt_malloc(pt->nfibs[0] * sizeof(idx_t));
      idx_t * tmp[SPLATT_MAX_NMODES] = {0};
      tmp[0] = roots;
      int rc = spd_csf_export_tile((spd_csf *) ct->dev, t, NULL, tmp, NULL);
      if(rc != SPD_OK) {
        free(roots);
        return rc;
      }
    }
    /* leaf range of every node, bottom-up */
    idx_t * lo = splatt_malloc((nn + 1) * sizeof(idx_t));
    for(idx_t j=0; j <= nn; ++j) {
      lo[j] = j;
    }
    memcpy(tt->ind[ct->dim_perm[N-1]] + off, pt->fids[N-1], nn * sizeof(idx_t));
    memcpy(tt->vals + off, pt->vals, nn * sizeof(val_t));
    for(idx_t l=N-1; l-- != 0; ) {
      idx_t const nfl = pt->nfibs[l];
      idx_t * lo2 = splatt_malloc((nfl + 1) * sizeof(idx_t));
      for(idx_t f=0; f <= nfl; ++f) {
        lo2[f] = lo[pt->fptr[l][f]];
      }
      idx_t const * ids = (l == 0) ? roots : pt->fids[l];
      idx_t * dst = tt->ind[ct->dim_perm[l]] + off;
      for(idx_t f=0; f < nfl; ++f) {
        for(idx_t j=lo2[f]; j < lo2[f+1]; ++j) {
          dst[j] = ids[f];
        }
      }
      free(lo);
      lo = lo2;
    }
    free(lo);
    if(roots != pt->fids[0]) {
      free(roots);
    }
    off += nn;
  }
  return SPD_OK;
}

/* Build one CSF (src/csf.c:527-558) from an uploaded tensor. */
static int p_mk_csf(splatt_csf * ct, spd_coo const * dcoo, idx_t nmodes,
    idx_t nnz, idx_t const * dims, csf_mode_type mode_type, idx_t mode,
    double const * opts)
{
  ct->nnz = nnz;
  ct->nmodes = nmodes;
  for(idx_t m=0; m < nmodes; ++m) {
    ct->dims[m] = dims[m];
  }
  splatt_csf_find_mode_order(dims, nmodes, mode_type, mode, ct->dim_perm);
  for(idx_t l=0; l < nmodes; ++l) {
    ct->dim_iperm[ct->dim_perm[l]] = l;
  }
  ct->which_tile = (splatt_tile_type) opts[SPLATT_OPTION_TILE];
  ct->ntiles = 1;
  ct->ntiled_modes = 0;
  for(idx_t m=0; m < nmodes; ++m) {
    ct->tile_dims[m] = 1;
  }
  if(ct->which_tile == SPLATT_DENSETILE) {
    /* src/csf.c:434-451 */
    idx_t tl = (idx_t) opts[SPLATT_OPTION_TILELEVEL];
    ct->ntiled_modes = SS_MIN(tl, nmodes);
    idx_t const tile_depth = nmodes - ct->ntiled_modes;
    idx_t const nthreads = (idx_t) opts[SPLATT_OPTION_NTHREADS];
    for(idx_t m=0; m < nmodes; ++m) {
      ct->tile_dims[m] = (ct->dim_iperm[m] >= tile_depth) ? nthreads : 1;
      ct->ntiles *= ct->tile_dims[m];
    }
  } else if(ct->which_tile != SPLATT_NOTILE) {
    fprintf(stderr, "SPLATT: tiling '%d' unsupported for CSF tensors.\n",
        ct->which_tile);
    return SPD_EINPUT;
  }
  int perm[SPLATT_MAX_NMODES];
  for(idx_t l=0; l < nmodes; ++l) {
    perm[l] = (int) ct->dim_perm[l];
  }
  spd_csf * dev = NULL;
  int rc = spd_csf_build(dcoo, perm, ct->tile_dims, &dev);
  if(rc != SPD_OK) {
    return rc;
  }
  ct->dev = dev;
  return p_mirror(ct);
}

static idx_t p_count(double const * opts, idx_t nmodes)
{
  switch((splatt_csf_type) opts[SPLATT_OPTION_CSF_ALLOC]) {
  case SPLATT_CSF_ONEMODE: return 1;
  case SPLATT_CSF_TWOMODE: return 2;
  case SPLATT_CSF_ALLMODE: return nmodes;
  }
  return 0;
}

/* csf_alloc over an uploaded tensor. */
static splatt_csf * p_alloc_all(spd_coo const * dcoo, idx_t nmodes, idx_t nnz,
    idx_t const * dims, double const * opts)
{
  idx_t const count = p_count(opts, nmodes);
  if(count == 0) {
    fprintf(stderr, "SPLATT: CSF type '%d' not recognized.\n",
        (int) opts[SPLATT_OPTION_CSF_ALLOC]);
    return NULL;
  }
  splatt_csf * ret = calloc(count, sizeof(splatt_csf));
  int rc = SPD_OK;
  switch((splatt_csf_type) opts[SPLATT_OPTION_CSF_ALLOC]) {
  case SPLATT_CSF_ONEMODE:
    rc = p_mk_csf(ret, dcoo, nmodes, nnz, dims, CSF_SORTED_SMALLFIRST, 0, opts);
    break;
  case SPLATT_CSF_TWOMODE: {
    rc = p_mk_csf(ret, dcoo, nmodes, nnz, dims, CSF_SORTED_SMALLFIRST, 0, opts);
    if(rc == SPD_OK) {
      /* second tensor: root = leaf mode of the first, never tiled */
      double tmp[SPLATT_OPTION_NOPTIONS];
      memcpy(tmp, opts, sizeof(tmp));
      tmp[SPLATT_OPTION_TILE] = SPLATT_NOTILE;
      idx_t const last = ret[0].dim_perm[nmodes-1];
      rc = p_mk_csf(ret + 1, dcoo, nmodes, nnz, dims, CSF_SORTED_MINUSONE, last, tmp);
    }
    break;
  }
  case SPLATT_CSF_ALLMODE:
    for(idx_t m=0; m < nmodes && rc == SPD_OK; ++m) {
      rc = p_mk_csf(ret + m, dcoo, nmodes, nnz, dims, CSF_SORTED_MINUSONE, m, opts);
    }
    break;
  }
  for(idx_t i=0; i < count; ++i) {
    ret[i].alloc_count = count;
    ret[i].alloc_type = (splatt_csf_type) opts[SPLATT_OPTION_CSF_ALLOC];
  }
  if(rc != SPD_OK) {
    spl_report(rc, "CSF construction");
    for(idx_t i=0; i < count; ++i) {
      splatt_csf_free_mode(ret + i);
    }
    free(ret);
    return NULL;
  }
  return ret;
}

splatt_csf * splatt_csf_alloc(sptensor_t * const tt, double const * const opts)
{
  spd_coo * dcoo = NULL;
  int rc = spd_coo_upload((int) tt->nmodes, tt->nnz, tt->dims,
      (uint64_t const * const *) tt->ind, tt->vals, &dcoo);
  if(rc != SPD_OK) {
    spl_report(rc, "tensor upload");
    return NULL;
  }
  splatt_csf * ret = p_alloc_all(dcoo, tt->nmodes, tt->nnz, tt->dims, opts);
  spd_coo_free(dcoo);
  if(ret != NULL) {
    idx_t const count = ret[0].alloc_count;
    rc = p_reorder_coo(ret + count - 1, tt);
    if(rc != SPD_OK) {
      spl_report(rc, "tensor reorder");
    }
  }
  return ret;
}

void splatt_csf_alloc_mode(sptensor_t * const tt, csf_mode_type which_ordering,
    splatt_idx_t const mode_special, splatt_csf * const csf,
    double const * const opts)
{
  spd_coo * dcoo = NULL;
  int rc = spd_coo_upload((int) tt->nmodes, tt->nnz, tt->dims,
      (uint64_t const * const *) tt->ind, tt->vals, &dcoo);
  if(rc == SPD_OK) {
    rc = p_mk_csf(csf, dcoo, tt->nmodes, tt->nnz, tt->dims, which_ordering,
        mode_special, opts);
    spd_coo_free(dcoo);
  }
  csf->alloc_count = 1;
  csf->alloc_type = SPLATT_CSF_ONEMODE;
  if(rc == SPD_OK) {
    rc = p_reorder_coo(csf, tt);
  }
  if(rc != SPD_OK) {
    spl_report(rc, "CSF construction");
  }
}

splatt_error_type splatt_csf_convert(splatt_idx_t const nmodes,
    splatt_idx_t const nnz, splatt_idx_t ** const inds,
    splatt_val_t * const vals, splatt_csf ** tensors,
    double const * const options)
{
  *tensors = NULL;
  if(nmodes < 2 || nmodes > SPLATT_MAX_NMODES || nnz == 0 || inds == NULL ||
      vals == NULL) {
    return SPLATT_ERROR_BADINPUT;
  }
  idx_t dims[SPLATT_MAX_NMODES];
  for(idx_t m=0; m < nmodes; ++m) {
    idx_t mx = 0;
    for(idx_t n=0; n < nnz; ++n) {
      mx = SS_MAX(mx, inds[m][n]);
    }
    dims[m] = mx + 1;
  }
  spd_coo * dcoo = NULL;
  int rc = spd_coo_upload((int) nmodes, nnz, dims, (uint64_t const * const *) inds,
      vals, &dcoo);
  if(rc != SPD_OK) {
    return spl_report(rc, "tensor upload");
  }
  *tensors = p_alloc_all(dcoo, nmodes, nnz, dims, options);
  spd_coo_free(dcoo);
  return (*tensors == NULL) ? SPLATT_ERROR_BADINPUT : SPLATT_SUCCESS;
}

/* src/csf.c:18-37: read, drop empty slices, build. The tensor goes from the
 * file to device memory (the text parsed there), loses its empty slices
 * there and the CSFs are built from it; no host copy of the coordinates. */
splatt_error_type splatt_csf_load(char const * const fname,
    splatt_idx_t * nmodes, splatt_csf ** tensors, double const * const options)
{
  *tensors = NULL;
  spd_coo * dcoo = NULL;
  int rc = spl_coo_read_dev(fname, &dcoo, NULL);
  if(rc != SPD_OK) {
    return (rc == SPD_EINPUT) ? SPLATT_ERROR_BADINPUT : spl_report(rc, "tensor input");
  }
  rc = spd_coo_remove_empty(dcoo, NULL);
  if(rc != SPD_OK) {
    spd_coo_free(dcoo);
    return spl_report(rc, "empty-slice removal");
  }
  int nm = 0;
  uint64_t nnz = 0;
  idx_t dims[SPLATT_MAX_NMODES];
  spd_coo_dims(dcoo, &nm, &nnz, (uint64_t *) dims);
  *tensors = p_alloc_all(dcoo, (idx_t) nm, nnz, dims, options);
  *nmodes = (idx_t) nm;
  spd_coo_free(dcoo);
  return (*tensors == NULL) ? SPLATT_ERROR_BADINPUT : SPLATT_SUCCESS;
}

void splatt_csf_free_mode(splatt_csf * const csf)
{
  p_free_pattern(csf);
  if(csf->dev != NULL) {
    spd_csf_free((spd_csf *) csf->dev);
    csf->dev = NULL;
  }
}

void splatt_csf_free(splatt_csf * const csf, double const * const opts)
{
  if(csf == NULL) {
    return;
  }
  idx_t count = csf[0].alloc_count;
  if(count == 0) {
    count = p_count(opts, csf[0].nmodes);
  }
  for(idx_t i=0; i < count; ++i) {
    splatt_csf_free_mode(csf + i);
  }
  free(csf);
}

void splatt_free_csf(splatt_csf * tensors, double const * const options)
{
  splatt_csf_free(tensors, options);
}

splatt_val_t splatt_csf_frobsq(splatt_csf const * const tensor)
{
  double s = 0;
  int rc = SPD_EINPUT;
  if(tensor->dev != NULL) {
    rc = spd_csf_frobsq((spd_csf const *) tensor->dev, &s);
  }
  if(rc != SPD_OK) {
    fprintf(stderr, "SPLATT: csf_frobsq needs a tensor built by this library "
        "on the GPU (%s)\n", spd_last_error());
    return -1.0;
  }
  return s;
}
