// GAMG hierarchy: host setup routines and the device-side level descriptors.
#pragma once

#include "pfm_internal.h"

namespace pfm {

int gamg_aggregate(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei, int32_t target,
                   std::vector<int32_t> &f2c, int32_t &nagg);
int gamg_coarse_faces(int32_t nf, const int32_t *own, const int32_t *nei,
                      const std::vector<int32_t> &f2c, std::vector<int32_t> &c_own,
                      std::vector<int32_t> &c_nei, std::vector<int32_t> &face_map);
int pair_level(int32_t n_fine, int32_t n_cur, int32_t nf, const int32_t *own, const int32_t *nei,
               const std::vector<int32_t> &cell_map, std::vector<int32_t> &out, int32_t &n_new);

// Device pointers of one level, as the kernels see them (T = double, or
// float for the inner cycles of the mixed-precision solve).
template <typename T>
struct LevelDevT {
  int32_t n, nslices;
  const SliceInfo *slices;
  const int32_t *col;
  const uint32_t *cp;     // paired int16 column offsets (row_spmv_pair), or null
  const T *packed;
  const T *diag;
  const T *inv;           // 1 / max(|diag|, 1e-30)
  const int32_t *finite;  // 1 when diag and every off-diagonal value are finite
  T *x[2];                // Jacobi double buffer
  T *b;
  T *r;
  // restriction to the next coarser level (unused on the coarsest)
  const int32_t *f2c;
  const int32_t *agg_ptr;
  const int32_t *agg_cells;
  int32_t nc;
  int32_t nt;             // 1: coefficients beyond the MALL, streamed nontemporally
  // Gauss-Seidel smoothing: slot -> face map and the sweep level schedules
  const int32_t *map;
  const int32_t *gs_fptr, *gs_fcells, *gs_bptr, *gs_bcells;
  const GsEnt *gs_fent, *gs_bent;  // the schedules as gs_entry() entries
  int32_t gs_nfl, gs_nbl;
};
using LevelDev = LevelDevT<double>;

}  // namespace pfm

// One level of a hierarchy (host side).
struct GamgLevelHost {
  int32_t n = 0, nf = 0;
  pfm_ldu ldu = nullptr;          // level 0 borrows the caller's handle
  bool owns_ldu = false;
  std::vector<int32_t> own, nei;  // level topology (coarse levels)
  // maps to the next coarser level
  int32_t nc = 0, nfc = 0;
  std::vector<int32_t> f2c, face_map;
  int32_t *d_f2c = nullptr, *d_agg_ptr = nullptr, *d_agg_cells = nullptr;
  int32_t *d_cf_ptr = nullptr, *d_cf_faces = nullptr;
  int32_t *d_in_ptr = nullptr, *d_in_faces = nullptr;  // faces inside each aggregate
};

struct pfm_gamg_s {
  pfm_ldu fine = nullptr;
  pfm_gamg_params prm{};
  std::vector<GamgLevelHost> lv;
  int device = 0;
  pfm::PollPool poll;  // pinned stop-flag slots of the solves
};
