// GAMG solver on the device (reference solvers/gamg.py:104-585).
//
// Hierarchy topology is built once on the host (gamg_setup.cpp).  Every solve
// coarsens the coefficients on the device in the reference's summation order,
// packs each level into the row-slice layout and runs V-cycles:
//   * Jacobi sweep  x' = x + inv*(b - A x), inv = 1/max(|diag|, 1e-30)
//     (_smooth :385-429), one fused kernel per sweep, double-buffered x;
//   * restriction   b_c[c] = sum of r over the aggregate, fine-cell order
//     (_restrict :340-360, scatter_add order);
//   * prolongation  x[i] += e_c[f2c[i]] (_prolongate :363-379);
//   * the coarsest level smoothed n_coarsest (50) times (:458-461);
//   * optionally (PFM_SMOOTHER_GAUSS_SEIDEL, not in the reference GAMG) every
//     smoothing step is instead a symmetric Gauss-Seidel sweep in place,
//     level-scheduled like pfm_gs_smooth (one launch per sweep level on large
//     levels, the single-workgroup tail walks the levels of small ones).
// Levels at or below kTailCells run, together with everything coarser, in a
// single-workgroup kernel (one launch instead of dozens of tiny ones).  After
// each cycle r = b - A x is reduced and the residual monitor runs on the
// device (GAMG records no initial residual, :574-582).  All arithmetic is
// element-wise in a fixed order, so the solution is bit-identical to the
// reference's sequential evaluation; only ||r|| depends on the reduction tree.

#include <cmath>
#include <cstring>

#include "gamg_internal.h"
#include "solver_state.h"

using namespace pfm;

namespace {

constexpr int kTailCells = 4096;
constexpr int kTailThreads = 1024;
constexpr int kMaxLevels = 64;

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------ coarsening ---
__global__ __launch_bounds__(kBlock) void k_coarse_diag(int nc, const int32_t *__restrict__ agg_ptr,
                                                         const int32_t *__restrict__ agg_cells,
                                                         const int32_t *__restrict__ in_ptr,
                                                         const int32_t *__restrict__ in_faces,
                                                         const double *__restrict__ diag,
                                                         const double *__restrict__ lower,
                                                         const double *__restrict__ upper,
                                                         double *__restrict__ out) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nc) return;
  double s = 0.0;
  for (int k = agg_ptr[c]; k < agg_ptr[c + 1]; ++k) s += diag[agg_cells[k]];
  if (in_ptr) {  // Galerkin extension: couplings inside the aggregate
    for (int k = in_ptr[c]; k < in_ptr[c + 1]; ++k) {
      const int f = in_faces[k];
      s += lower[f] + upper[f];
    }
  }
  out[c] = s;
}

__global__ __launch_bounds__(kBlock) void k_coarse_faces(int nfc, const int32_t *__restrict__ cf_ptr,
                                                          const int32_t *__restrict__ cf_faces,
                                                          const double *__restrict__ lower,
                                                          const double *__restrict__ upper,
                                                          double *__restrict__ clo,
                                                          double *__restrict__ cup) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nfc) return;
  const int b = cf_ptr[c], e = cf_ptr[c + 1];
  double l = lower[cf_faces[b]], u = upper[cf_faces[b]];
  for (int k = b + 1; k < e; ++k) {
    l = l + lower[cf_faces[k]];
    u = u + upper[cf_faces[k]];
  }
  clo[c] = l;
  cup[c] = u;
}

__global__ __launch_bounds__(kBlock) void k_inv_diag(int n, const double *__restrict__ d,
                                                      double *__restrict__ inv) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double a = fabs(d[i]);
  inv[i] = 1.0 / (a < 1e-30 ? 1e-30 : a);
}

// ------------------------------------------------------------- V-cycle -----
__device__ __forceinline__ double level_ax(const LevelDev &L, int s, int lane, int row,
                                          const double *x) {
  const SliceInfo si = load_slice(L.slices, s);
  const double dx = row < L.n ? L.diag[row] * x[row] : 0.0;
  return row_spmv<double>(si, lane, dx, L.col, L.packed, x);
}

__global__ __launch_bounds__(kBlock) void k_jacobi(LevelDev L, int cur, const KState *st) {
  if (st->done) return;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= L.nslices) return;
  const int lane = threadIdx.x & 63, row = s * kSlice + lane;
  const double *x = L.x[cur];
  const double ax = level_ax(L, s, lane, row, x);
  if (row < L.n) L.x[cur ^ 1][row] = x[row] + L.inv[row] * (L.b[row] - ax);
}

// One Gauss-Seidel sweep level of a level's x[cur], in place.
__global__ __launch_bounds__(kBlock) void k_gs_level(LevelDev L, int cur,
                                                      const int32_t *__restrict__ cells, int count,
                                                      const KState *st) {
  if (st->done) return;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < count) gs_cell(cells[t], L.slices, L.col, L.map, L.packed, L.diag, L.b, L.x[cur], 1.0);
}

__global__ __launch_bounds__(kBlock) void k_resid(LevelDev L, int cur, const KState *st) {
  if (st->done) return;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= L.nslices) return;
  const int lane = threadIdx.x & 63, row = s * kSlice + lane;
  const double ax = level_ax(L, s, lane, row, L.x[cur]);
  if (row < L.n) L.r[row] = L.b[row] - ax;
}

// b_c = sum of r over each aggregate (fine-cell order); x_c[0] = 0.
__global__ __launch_bounds__(kBlock) void k_restrict(LevelDev F, LevelDev Cc, const KState *st) {
  if (st->done) return;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= F.nc) return;
  double s = 0.0;
  for (int k = F.agg_ptr[c]; k < F.agg_ptr[c + 1]; ++k) s += F.r[F.agg_cells[k]];
  Cc.b[c] = s;
  Cc.x[0][c] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_prolong(LevelDev F, int cur, const double *__restrict__ xc,
                                                     const KState *st) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= F.n) return;
  F.x[cur][i] = F.x[cur][i] + xc[F.f2c[i]];
}

// Single-workgroup V-cycle of levels k0..nlev-1 (small levels).  Waves map
// onto whole slices (rows t, t + 1024, ... with 64-aligned wave offsets).
__device__ void tail_jacobi(const LevelDev &L, int cur) {
  for (int row0 = threadIdx.x; row0 < L.nslices * kSlice; row0 += kTailThreads) {
    const int s = row0 >> 6, lane = row0 & 63;
    const double *x = L.x[cur];
    const double ax = level_ax(L, s, lane, row0, x);
    if (row0 < L.n) L.x[cur ^ 1][row0] = x[row0] + L.inv[row0] * (L.b[row0] - ax);
  }
  __syncthreads();
}

// One symmetric Gauss-Seidel sweep of a small level by the whole workgroup.
__device__ void tail_gs(const LevelDev &L, int cur) {
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t *ptr = pass ? L.gs_bptr : L.gs_fptr;
    const int32_t *cells = pass ? L.gs_bcells : L.gs_fcells;
    const int nl = pass ? L.gs_nbl : L.gs_nfl;
    for (int l = 0; l < nl; ++l) {
      for (int k = ptr[l] + (int)threadIdx.x; k < ptr[l + 1]; k += kTailThreads)
        gs_cell(cells[k], L.slices, L.col, L.map, L.packed, L.diag, L.b, L.x[cur], 1.0);
      __syncthreads();
    }
  }
}

// One smoothing step: Jacobi flips the level's buffer, Gauss-Seidel does not.
__device__ __forceinline__ void tail_smooth(const LevelDev &L, int &cur, bool gs) {
  if (gs) {
    tail_gs(L, cur);
  } else {
    tail_jacobi(L, cur);
    cur ^= 1;
  }
}

__global__ __launch_bounds__(kTailThreads) void k_tail(const LevelDev *__restrict__ Ls, int k0,
                                                        int nlev, int cur0, int n_pre, int n_post,
                                                        int n_coarsest, int gs, const KState *st) {
  if (st->done) return;
  int cur[kMaxLevels];
  cur[k0] = cur0;
  for (int k = k0; k < nlev - 1; ++k) {
    const LevelDev L = Ls[k];
    for (int it = 0; it < n_pre; ++it) tail_smooth(L, cur[k], gs);
    for (int row0 = threadIdx.x; row0 < L.nslices * kSlice; row0 += kTailThreads) {
      const double ax = level_ax(L, row0 >> 6, row0 & 63, row0, L.x[cur[k]]);
      if (row0 < L.n) L.r[row0] = L.b[row0] - ax;
    }
    __syncthreads();
    const LevelDev Cc = Ls[k + 1];
    for (int c = threadIdx.x; c < L.nc; c += kTailThreads) {
      double s = 0.0;
      for (int q = L.agg_ptr[c]; q < L.agg_ptr[c + 1]; ++q) s += L.r[L.agg_cells[q]];
      Cc.b[c] = s;
      Cc.x[0][c] = 0.0;
    }
    cur[k + 1] = 0;
    __syncthreads();
  }
  {
    const LevelDev L = Ls[nlev - 1];
    for (int it = 0; it < n_coarsest; ++it) tail_smooth(L, cur[nlev - 1], gs);
  }
  for (int k = nlev - 2; k >= k0; --k) {
    const LevelDev L = Ls[k];
    const double *xc = Ls[k + 1].x[cur[k + 1]];
    for (int i = threadIdx.x; i < L.n; i += kTailThreads)
      L.x[cur[k]][i] = L.x[cur[k]][i] + xc[L.f2c[i]];
    __syncthreads();
    for (int it = 0; it < n_post; ++it) tail_smooth(L, cur[k], gs);
  }
}

// r = b - A x on level 0 and the partial sums of r.r.
__global__ __launch_bounds__(kBlock) void k_res_norm(LevelDev L, int cur, double *__restrict__ part,
                                                      const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc = 0.0;
  const SliceRange sr = block_slices(L.nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    const double ax = level_ax(L, s, lane, row, L.x[cur]);
    if (row < L.n) {
      const double ri = L.b[row] - ax;
      acc += ri * ri;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kBlock) void k_gamg_fin(const double *__restrict__ part, int G, KState *st,
                                                      double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const double RR = sum_partials(part, G, lds);
  if (threadIdx.x == 0) {
    const int it = st->iter + 1;
    st->iter = it;
    if (monitor_update(st, hist, sqrt(RR), it))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (it >= st->max_iter)
      finish(st, PFM_SOLVE_MAXITER);
  }
}

// ----------------------------------------------------------- workspace -----
inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct GamgWork {
  KState *st;
  double *part;
  double *hist;
  LevelDev *d_levels;
  std::vector<LevelDev> lv;
  std::vector<double *> diag, lower, upper, packed;
};

inline int red_grid(const pfm_ldu_s *h) {
  return std::max(1, std::min(blocks_for_slices(h->nslices), kMaxRedBlocks));
}

size_t gamg_layout(const pfm_gamg_s *g, int max_iter, GamgWork *w, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  const int L = (int)g->lv.size();
  GamgWork t;
  t.lv.resize(L);
  t.diag.resize(L);
  t.lower.resize(L);
  t.upper.resize(L);
  t.packed.resize(L);
  for (int k = 0; k < L; ++k) {
    const auto &H = g->lv[k];
    const size_t n = (size_t)H.n, nf = (size_t)H.nf;
    LevelDev &D = t.lv[k];
    memset(&D, 0, sizeof(D));
    D.n = H.n;
    D.nslices = H.ldu->nslices;
    D.slices = H.ldu->d_slices;
    D.col = H.ldu->d_col;
    D.map = H.ldu->d_map;
    D.gs_fptr = H.ldu->d_gs_fwd_ptr;
    D.gs_fcells = H.ldu->d_gs_fwd_cells;
    D.gs_bptr = H.ldu->d_gs_bwd_ptr;
    D.gs_bcells = H.ldu->d_gs_bwd_cells;
    D.gs_nfl = (int32_t)H.ldu->gs_fwd_ptr.size() - 1;
    D.gs_nbl = (int32_t)H.ldu->gs_bwd_ptr.size() - 1;
    t.packed[k] = (double *)take(sizeof(double) * (size_t)std::max<int64_t>(H.ldu->nslots, 1));
    D.packed = t.packed[k];
    D.inv = (double *)take(sizeof(double) * n);
    D.r = (double *)take(sizeof(double) * n);
    D.x[1] = (double *)take(sizeof(double) * n);
    if (k > 0) {
      t.diag[k] = (double *)take(sizeof(double) * n);
      t.lower[k] = (double *)take(sizeof(double) * std::max<size_t>(nf, 1));
      t.upper[k] = (double *)take(sizeof(double) * std::max<size_t>(nf, 1));
      D.diag = t.diag[k];
      D.x[0] = (double *)take(sizeof(double) * n);
      D.b = (double *)take(sizeof(double) * n);
    }
    if (k + 1 < L) {
      D.f2c = H.d_f2c;
      D.agg_ptr = H.d_agg_ptr;
      D.agg_cells = H.d_agg_cells;
      D.nc = H.nc;
    }
  }
  // scalars last: the level arrays keep their offsets whatever max_iter is
  t.st = (KState *)take(sizeof(KState));
  t.part = (double *)take(sizeof(double) * (size_t)red_grid(g->fine));
  t.d_levels = (LevelDev *)take(sizeof(LevelDev) * (size_t)L);
  t.hist = (double *)take(sizeof(double) * ((size_t)std::max(max_iter, 0) + 1));
  if (w) *w = t;
  return off;
}

template <typename T>
int upload(T **dst, const std::vector<T> &src) {
  *dst = nullptr;
  if (src.empty()) return PFM_OK;
  PFM_CHECK_HIP(hipMalloc((void **)dst, src.size() * sizeof(T)));
  PFM_CHECK_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return PFM_OK;
}

// CSR of items grouped by key (stable: original order inside a group).
void group_by(int32_t nkeys, const std::vector<int32_t> &key, std::vector<int32_t> &ptr,
              std::vector<int32_t> &items, bool skip_negative) {
  ptr.assign(nkeys + 1, 0);
  for (int32_t k : key)
    if (k >= 0) ptr[k + 1]++;
  for (int32_t i = 0; i < nkeys; ++i) ptr[i + 1] += ptr[i];
  items.assign(ptr[nkeys], 0);
  std::vector<int32_t> c(ptr.begin(), ptr.end() - 1);
  for (int32_t i = 0; i < (int32_t)key.size(); ++i)
    if (key[i] >= 0) items[c[key[i]]++] = i;
  (void)skip_negative;
}

int finish_level_maps(GamgLevelHost &H, const int32_t *own, const int32_t *nei, bool galerkin) {
  std::vector<int32_t> ptr, items;
  int st;
  if ((st = upload(&H.d_f2c, H.f2c))) return st;
  group_by(H.nc, H.f2c, ptr, items, false);
  if ((st = upload(&H.d_agg_ptr, ptr)) || (st = upload(&H.d_agg_cells, items))) return st;
  group_by(H.nfc, H.face_map, ptr, items, true);
  if ((st = upload(&H.d_cf_ptr, ptr)) || (st = upload(&H.d_cf_faces, items))) return st;
  if (galerkin) {
    std::vector<int32_t> inkey(H.nf, -1);
    for (int32_t f = 0; f < H.nf; ++f)
      if (H.face_map[f] < 0) inkey[f] = H.f2c[own[f]];
    group_by(H.nc, inkey, ptr, items, true);
    if ((st = upload(&H.d_in_ptr, ptr)) || (st = upload(&H.d_in_faces, items))) return st;
  }
  return PFM_OK;
}

void free_level(GamgLevelHost &H) {
  (void)hipFree(H.d_f2c);
  (void)hipFree(H.d_agg_ptr);
  (void)hipFree(H.d_agg_cells);
  (void)hipFree(H.d_cf_ptr);
  (void)hipFree(H.d_cf_faces);
  (void)hipFree(H.d_in_ptr);
  (void)hipFree(H.d_in_faces);
  if (H.owns_ldu) pfm_ldu_destroy(H.ldu);
}

}  // namespace

extern "C" {

int pfm_gamg_destroy(pfm_gamg g) {
  if (!g) return PFM_OK;
  for (auto &H : g->lv) free_level(H);
  if (g->h_flag) (void)hipHostFree(g->h_flag);
  delete g;
  return PFM_OK;
}

int pfm_gamg_create(pfm_ldu fine, const pfm_gamg_params *prm, pfm_gamg *out) {
  PFM_REQUIRE(fine && prm && out, "null argument");
  PFM_REQUIRE(prm->max_levels >= 0 && prm->n_pre >= 0 && prm->n_post >= 0 && prm->n_coarsest >= 0,
              "negative GAMG parameter");
  PFM_REQUIRE(prm->agglomeration == PFM_AGGLOM_GREEDY || prm->agglomeration == PFM_AGGLOM_PAIR,
              "unknown agglomeration %d", prm->agglomeration);
  PFM_REQUIRE(prm->coarse == PFM_COARSE_REFERENCE || prm->coarse == PFM_COARSE_GALERKIN,
              "unknown coarse operator %d", prm->coarse);
  PFM_REQUIRE(prm->smoother == PFM_SMOOTHER_JACOBI || prm->smoother == PFM_SMOOTHER_GAUSS_SEIDEL,
              "unknown smoother %d", prm->smoother);
  *out = nullptr;
  pfm_gamg g = new pfm_gamg_s();
  g->fine = fine;
  g->prm = *prm;
  g->device = fine->device;
  const bool galerkin = prm->coarse == PFM_COARSE_GALERKIN;
  // level 0: the caller's topology
  {
    GamgLevelHost H;
    H.n = fine->n;
    H.nf = fine->nf;
    H.ldu = fine;
    H.own.resize(fine->nf);
    H.nei.resize(fine->nf);
    if (fine->nf) {
      hipError_t e1 = hipMemcpy(H.own.data(), fine->d_own, sizeof(int32_t) * fine->nf, hipMemcpyDeviceToHost);
      hipError_t e2 = hipMemcpy(H.nei.data(), fine->d_nei, sizeof(int32_t) * fine->nf, hipMemcpyDeviceToHost);
      if (e1 != hipSuccess || e2 != hipSuccess) {
        delete g;
        set_error("reading the fine topology failed");
        return PFM_ERR_HIP;
      }
    }
    g->lv.push_back(std::move(H));
  }
  int st = PFM_OK;
  std::vector<int32_t> pair_map;  // finest -> current level (pair agglomeration)
  if (prm->agglomeration == PFM_AGGLOM_PAIR) {
    pair_map.resize(fine->n);
    for (int32_t i = 0; i < fine->n; ++i) pair_map[i] = i;
  }
  // finest topology for the pair agglomeration (a copy: g->lv grows below)
  const std::vector<int32_t> own0 = g->lv[0].own, nei0 = g->lv[0].nei;
  g->lv.reserve((size_t)prm->max_levels + 1);
  for (int step = 0; step < prm->max_levels && st == PFM_OK; ++step) {
    GamgLevelHost &H = g->lv.back();
    int32_t nagg = 0;
    if (prm->agglomeration == PFM_AGGLOM_GREEDY) {
      if (H.n <= prm->min_cells_coarse) break;
      const int32_t target = std::max(H.n / 2, prm->min_cells_coarse);
      gamg_aggregate(H.n, H.nf, H.own.data(), H.nei.data(), target, H.f2c, nagg);
      if (nagg >= H.n) break;
    } else {
      // PairGamgAgglomeration.build_levels: pairs over the finest faces with the
      // cumulative map; stop when nothing pairs or the level shrinks < 10%
      if (H.n <= 4) break;
      std::vector<int32_t> next;
      pair_level(fine->n, H.n, fine->nf, own0.data(), nei0.data(), pair_map, next, nagg);
      if (nagg == 0 || nagg > 0.9 * H.n) break;
      H.f2c.assign(H.n, 0);
      for (int32_t i = 0; i < fine->n; ++i) H.f2c[pair_map[i]] = next[i];
      pair_map.swap(next);
    }
    H.nc = nagg;
    GamgLevelHost C;
    C.n = nagg;
    gamg_coarse_faces(H.nf, H.own.data(), H.nei.data(), H.f2c, C.own, C.nei, H.face_map);
    C.nf = (int32_t)C.own.size();
    H.nfc = C.nf;
    st = finish_level_maps(H, H.own.data(), H.nei.data(), galerkin);
    if (st != PFM_OK) break;
    st = pfm_ldu_create(C.n, C.nf, C.own.data(), C.nei.data(), &C.ldu);
    if (st != PFM_OK) break;
    C.owns_ldu = true;
    g->lv.push_back(std::move(C));  // H is not used past this point
  }
  if (st == PFM_OK && (int)g->lv.size() > kMaxLevels) {
    set_error("too many GAMG levels");
    st = PFM_ERR_ARG;
  }
  if (st == PFM_OK &&
      hipHostMalloc((void **)&g->h_flag, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    set_error("hipHostMalloc failed");
    st = PFM_ERR_HIP;
  }
  if (st != PFM_OK) {
    pfm_gamg_destroy(g);
    return st;
  }
  *out = g;
  return PFM_OK;
}

int pfm_gamg_num_levels(pfm_gamg g) { return g ? (int)g->lv.size() : -1; }

int pfm_gamg_level_info(pfm_gamg g, int32_t level, int32_t *n_cells, int32_t *n_faces) {
  PFM_REQUIRE(g && level >= 0 && level < (int)g->lv.size(), "bad level %d", level);
  if (n_cells) *n_cells = g->lv[level].n;
  if (n_faces) *n_faces = g->lv[level].nf;
  return PFM_OK;
}

int pfm_gamg_level_maps(pfm_gamg g, int32_t level, int32_t *f2c_h, int32_t *face_map_h,
                        int32_t *c_owner_h, int32_t *c_neighbour_h) {
  PFM_REQUIRE(g && level >= 0 && level + 1 < (int)g->lv.size(), "bad level %d", level);
  const auto &H = g->lv[level];
  const auto &C = g->lv[level + 1];
  if (f2c_h) memcpy(f2c_h, H.f2c.data(), sizeof(int32_t) * H.f2c.size());
  if (face_map_h) memcpy(face_map_h, H.face_map.data(), sizeof(int32_t) * H.face_map.size());
  if (c_owner_h) memcpy(c_owner_h, C.own.data(), sizeof(int32_t) * C.own.size());
  if (c_neighbour_h) memcpy(c_neighbour_h, C.nei.data(), sizeof(int32_t) * C.nei.size());
  return PFM_OK;
}

int64_t pfm_gamg_workspace_bytes(pfm_gamg g, int32_t max_iter) {
  if (!g) return -1;
  return (int64_t)gamg_layout(g, max_iter, nullptr, nullptr);
}

int pfm_gamg_setup_values(pfm_gamg g, const double *diag, const double *lower, const double *upper,
                          void *work, void *stream) {
  PFM_REQUIRE(g && work, "null argument");
  PFM_REQUIRE(g->lv[0].n == 0 || diag, "null diag");
  PFM_REQUIRE(g->lv[0].nf == 0 || (lower && upper), "null lower/upper");
  hipStream_t s = (hipStream_t)stream;
  GamgWork w;
  gamg_layout(g, 0, &w, (char *)work);
  const int L = (int)g->lv.size();
  const double *d = diag, *lo = lower, *up = upper;
  for (int k = 0; k < L; ++k) {
    const auto &H = g->lv[k];
    if (k > 0) {
      const auto &Fh = g->lv[k - 1];
      k_coarse_diag<<<grid1d(H.n), kBlock, 0, s>>>(H.n, Fh.d_agg_ptr, Fh.d_agg_cells, Fh.d_in_ptr,
                                                   Fh.d_in_faces, d, lo, up, w.diag[k]);
      if (H.nf)
        k_coarse_faces<<<grid1d(H.nf), kBlock, 0, s>>>(H.nf, Fh.d_cf_ptr, Fh.d_cf_faces, lo, up,
                                                       w.lower[k], w.upper[k]);
      d = w.diag[k];
      lo = w.lower[k];
      up = w.upper[k];
    }
    if (H.nf) {
      int st = pfm_ldu_pack(H.ldu, 0, lo, up, w.packed[k], stream);
      if (st != PFM_OK) return st;
    }
    if (H.n) k_inv_diag<<<grid1d(H.n), kBlock, 0, s>>>(H.n, d, (double *)w.lv[k].inv);
  }
  PFM_CHECK_HIP(hipGetLastError());
  return PFM_OK;
}

int pfm_gamg_level_values(pfm_gamg g, void *work, int32_t level, double *diag, double *lower,
                          double *upper, void *stream) {
  PFM_REQUIRE(g && work && level >= 1 && level < (int)g->lv.size(), "bad level %d", level);
  hipStream_t s = (hipStream_t)stream;
  GamgWork w;
  gamg_layout(g, 0, &w, (char *)work);
  const auto &H = g->lv[level];
  if (diag) PFM_CHECK_HIP(hipMemcpyAsync(diag, w.diag[level], sizeof(double) * H.n, hipMemcpyDeviceToDevice, s));
  if (lower && H.nf)
    PFM_CHECK_HIP(hipMemcpyAsync(lower, w.lower[level], sizeof(double) * H.nf, hipMemcpyDeviceToDevice, s));
  if (upper && H.nf)
    PFM_CHECK_HIP(hipMemcpyAsync(upper, w.upper[level], sizeof(double) * H.nf, hipMemcpyDeviceToDevice, s));
  return PFM_OK;
}

int pfm_gamg_solve(pfm_gamg g, const pfm_solve_params *prm, const double *diag, const double *lower,
                   const double *upper, const double *b, double *x, void *work, double *history_h,
                   pfm_solve_result *res, void *stream) {
  PFM_REQUIRE(g && prm && res && work, "null argument");
  PFM_REQUIRE(prm->max_iter >= 0, "max_iter must be >= 0");
  const int n0 = g->lv[0].n;
  PFM_REQUIRE(n0 == 0 || (diag && b && x), "null device pointer");
  memset(res, 0, sizeof(*res));
  if (n0 == 0) {
    res->status = PFM_SOLVE_MAXITER;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  int st = pfm_gamg_setup_values(g, diag, lower, upper, work, stream);
  if (st != PFM_OK) return st;
  GamgWork w;
  gamg_layout(g, prm->max_iter, &w, (char *)work);
  const int L = (int)g->lv.size();
  w.lv[0].diag = diag;
  w.lv[0].x[0] = x;
  w.lv[0].b = const_cast<double *>(b);
  PFM_CHECK_HIP(hipMemcpyAsync(w.d_levels, w.lv.data(), sizeof(LevelDev) * L, hipMemcpyHostToDevice, s));
  const int hist_cap = prm->max_iter + 1;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  const int npre = g->prm.n_pre, npost = g->prm.n_post, ncoarse = g->prm.n_coarsest;
  int tail_k = L - 1;  // first level handled by the single-workgroup kernel
  while (tail_k > 0 && g->lv[tail_k - 1].n <= kTailCells) --tail_k;
  int cur0 = 0;
  const int G = red_grid(g->fine);
  const bool gs = g->prm.smoother == PFM_SMOOTHER_GAUSS_SEIDEL;
  auto jac = [&](int k, int &cur) {
    if (gs) {  // one symmetric sweep, level by level, in place
      const LevelDev &D = w.lv[k];
      const pfm_ldu h = g->lv[k].ldu;
      for (int pass = 0; pass < 2; ++pass) {
        const std::vector<int32_t> &ptr = pass ? h->gs_bwd_ptr : h->gs_fwd_ptr;
        const int32_t *cells = pass ? h->d_gs_bwd_cells : h->d_gs_fwd_cells;
        for (size_t l = 0; l + 1 < ptr.size(); ++l) {
          const int cnt = ptr[l + 1] - ptr[l];
          k_gs_level<<<grid1d(cnt), kBlock, 0, s>>>(D, cur, cells + ptr[l], cnt, w.st);
        }
      }
      return;
    }
    k_jacobi<<<blocks_for_slices(w.lv[k].nslices), kBlock, 0, s>>>(w.lv[k], cur, w.st);
    cur ^= 1;
  };
  auto cycle = [&]() {
    int cur[kMaxLevels];
    cur[0] = cur0;
    int k = 0;
    for (; k < tail_k; ++k) {  // descend through the large levels
      for (int it = 0; it < npre; ++it) jac(k, cur[k]);
      k_resid<<<blocks_for_slices(w.lv[k].nslices), kBlock, 0, s>>>(w.lv[k], cur[k], w.st);
      k_restrict<<<grid1d(w.lv[k].nc), kBlock, 0, s>>>(w.lv[k], w.lv[k + 1], w.st);
      cur[k + 1] = 0;
    }
    // levels tail_k .. L-1 in one workgroup
    k_tail<<<1, kTailThreads, 0, s>>>(w.d_levels, tail_k, L, cur[tail_k], npre, npost, ncoarse,
                                      gs ? 1 : 0, w.st);
    if (!gs) cur[tail_k] ^= (tail_k == L - 1 ? ncoarse : npre + npost) & 1;
    for (k = tail_k - 1; k >= 0; --k) {  // ascend
      k_prolong<<<grid1d(w.lv[k].n), kBlock, 0, s>>>(w.lv[k], cur[k], w.lv[k + 1].x[cur[k + 1]],
                                                    w.st);
      for (int it = 0; it < npost; ++it) jac(k, cur[k]);
    }
    cur0 = cur[0];
    k_res_norm<<<G, kBlock, 0, s>>>(w.lv[0], cur0, w.part, w.st);
    k_gamg_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
  };
  // the buffer parity of level 0 advances by the same amount every cycle, so
  // enqueueing cycles ahead of the stop flag never changes which buffer holds x
  int32_t *flag = g->h_flag;
  hipEvent_t ev[2];
  PFM_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  PFM_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  int enq = 0, chunk = 2, slot = 0;
  bool pending = false;
  int rc = PFM_OK;
  // cur0 after a cycle that did not run (stop flag set) is unchanged on the
  // device, so track the value after every completed cycle separately
  std::vector<int> cur_after;
  cur_after.reserve(prm->max_iter + 1);
  double last_report = now_seconds();
  while (true) {
    bool queued = false;
    report_progress(g->fine, enq, last_report);
    if (enq < prm->max_iter) {
      const int c = std::min(chunk, prm->max_iter - enq);
      for (int k = 0; k < c; ++k) {
        cycle();
        cur_after.push_back(cur0);
      }
      enq += c;
      chunk = std::min(chunk * 2, 16);
      if (hipMemcpyAsync(&flag[slot], &w.st->done, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipEventRecord(ev[slot], s) != hipSuccess) {
        rc = PFM_ERR_HIP;
        break;
      }
      queued = true;
    }
    if (pending) {
      if (hipEventSynchronize(ev[slot ^ 1]) != hipSuccess) {
        rc = PFM_ERR_HIP;
        break;
      }
      if (flag[slot ^ 1]) break;
    }
    if (!queued) break;
    pending = true;
    slot ^= 1;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc != PFM_OK) {
    set_error("HIP error while driving the GAMG loop");
    return rc;
  }
  KState hs;
  PFM_CHECK_HIP(hipMemcpyAsync(&hs, w.st, sizeof(KState), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  // x after the last executed cycle lives in buffer cur_after[iter-1]
  const int done_cycles = hs.iter;
  const int final_cur = done_cycles > 0 ? cur_after[done_cycles - 1] : 0;
  if (final_cur == 1)
    PFM_CHECK_HIP(hipMemcpyAsync(x, w.lv[0].x[1], sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
  const int nh = std::min(hs.nhist, hist_cap);
  if (history_h && nh > 0)
    PFM_CHECK_HIP(hipMemcpyAsync(history_h, w.hist, sizeof(double) * nh, hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  res->iterations = hs.nhist;
  res->status = hs.status == 0 ? PFM_SOLVE_MAXITER : hs.status;
  res->converged = res->status == PFM_SOLVE_CONVERGED;
  res->initial_residual = hs.nhist > 0 ? hs.res0 : 0.0;
  res->final_residual = hs.nhist > 0 ? hs.res : 0.0;
  return PFM_OK;
}

}  // extern "C"
