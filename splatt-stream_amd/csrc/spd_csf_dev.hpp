// spd_csf_dev.hpp -- device-side layouts of the COO and CSF handles.
#pragma once
#include "spd_internal.hpp"

struct DevPtrs { uint64_t const * p[SPD_MAX_MODES]; };
struct DevPtrs32 { uint32_t * p[SPD_MAX_MODES]; };
struct LevelCounts { uint64_t v[SPD_MAX_MODES]; };

// One CSF structure in 32-bit device arrays. Level l < N-1 has nfibs[l]
// nodes with fptr[l] (nfibs[l]+1 entries, child ids) and fids[l] (index in
// mode perm[l]); level N-1 holds the nonzeros: fids[N-1] and vals.
struct CsfArrays {
  uint64_t nfibs[SPD_MAX_MODES];
  uint32_t * fptr[SPD_MAX_MODES];
  uint32_t * fids[SPD_MAX_MODES];
  double * vals;
};

struct MttkrpPlan;

struct spd_coo {
  int nmodes;
  uint64_t nnz;
  uint64_t dims[SPD_MAX_MODES];
  uint64_t * ind[SPD_MAX_MODES];   // device, 0-based
  double * vals;                   // device
};

struct spd_csf {
  int nmodes;
  uint64_t nnz;
  uint64_t dims[SPD_MAX_MODES];
  int perm[SPD_MAX_MODES];         // level -> mode
  uint64_t tile_dims[SPD_MAX_MODES];
  uint64_t ntiles;
  CsfArrays u;                     // untiled canonical structure (kernels)
  CsfArrays t;                     // tiled structure (host mirror only)
  uint32_t * tile_off;             // ntiles x N first node id per tile/level
  uint64_t * tile_nnz;             // host, ntiles+1 prefix of nnz per tile
  float * vals32;                  // FP32 copy of u.vals, made on demand
  MttkrpPlan * plans;              // launch plans, one per item size
  struct HotPlan * hot;            // output-row rankings, one per level
  struct RowCopy * copies;         // row-sorted copies, one per output level
  struct RootHot * roothot;        // hot-encoded leaf ids of the 3-mode root kernel
  uint64_t * startbits;            // node starts: per 64 nonzeros, one word per level < N-1
};

void spd_csf_drop_plans(spd_csf * c);
