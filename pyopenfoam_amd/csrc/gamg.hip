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
// single-workgroup kernel (one launch instead of dozens of tiny ones); inside
// it, each row is held by one thread in registers (x exchanged through LDS
// between barriers), and the levels of at most 64 cells by one wave
// (cross-lane reads, no barriers).  Larger levels run as grid launches:
// measured at 128^3, a 4096-cell level costs less as seven ~4.6 us launches
// than as barrier-separated phases of one workgroup walking its rows with
// chains of dependent loads (the tail's fallback when a level does not fit
// the registers, and the Gauss-Seidel path).  After
// each cycle r = b - A x is reduced and the residual monitor runs on the
// device (GAMG records no initial residual, :574-582).  All arithmetic is
// element-wise in a fixed order, so the solution is bit-identical to the
// reference's sequential evaluation; only ||r|| depends on the reduction tree.

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "gamg_internal.h"
#include "solver_state.h"

using namespace pfm;

namespace {

constexpr int kTailCells = 1024;
constexpr int kFp32MinCells = 4096;
constexpr int kFuseCells = 262144;  // levels whose launches cost more than their traffic
constexpr int kTailThreads = 1024;
constexpr int kMaxLevels = 64;
constexpr int64_t kMallBytes = int64_t(256) << 20;

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// The paired int16 columns of a level's handle for the level kernels' SpMV,
// unless PFM_GAMG_COLS=int32 keeps them on the int32 columns (for comparing).
const uint32_t *level_pairs(const pfm_ldu_s *h) {
  static const bool int32_only = [] {
    const char *e = std::getenv("PFM_GAMG_COLS");
    return e && std::strcmp(e, "int32") == 0;
  }();
  return int32_only ? nullptr : h->d_colp2;
}

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

// flag = 0 if any of a0[0..n0), a1[0..n1), a2[0..n2) is not finite (the flag
// is set to 1 beforehand by k_flag_one).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_finite(int64_t n0, const T *__restrict__ a0,
                                                    int64_t n1, const T *__restrict__ a1,
                                                    int64_t n2, const T *__restrict__ a2,
                                                    int32_t *flag) {
  const int64_t tot = n0 + n1 + n2;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * kBlock) {
    const T v = i < n0 ? a0[i] : i < n0 + n1 ? a1[i - n0] : a2[i - n0 - n1];
    if (!isfinite(v)) *flag = 0;
  }
}

__global__ void k_flag_one(int32_t *flag) { *flag = 1; }

// ------------------------------------------------------------- V-cycle -----
// A level whose coefficients do not fit the 256 MB MALL (the finest levels
// at 256^3) streams them, with its per-row vectors, nontemporally, so that
// the gathered x stays cached: 7.46 -> 7.22 s for the 256^3 FP64 GAMG solve,
// 4.33 -> 4.12 s mixed.  Values and arithmetic are the same either way.
template <typename T>
__device__ __forceinline__ T lvld(const LevelDevT<T> &L, const T *a, int i) {
  return L.nt ? __builtin_nontemporal_load(a + i) : a[i];
}

// kPairs = false keeps the int32 columns (the one-workgroup tail, whose tiny
// levels gain nothing from the pairs and whose registers are all in use).
template <typename T, bool kPairs = true>
__device__ __forceinline__ T level_ax(const LevelDevT<T> &L, int s, int lane, int row, const T *x) {
  const SliceInfo si = load_slice(L.slices, s);
  auto gx = [x](int c) { return x[c]; };
  if (kPairs && L.cp) {  // paired int16 columns: half the index bytes
    if (L.nt) {
      const T dx = row < L.n ? __builtin_nontemporal_load(L.diag + row) * x[row] : T(0);
      return row_spmv_pair<T, decltype(gx), false, true>(si, s, lane, row, dx, L.cp, L.packed, gx);
    }
    const T dx = row < L.n ? L.diag[row] * x[row] : T(0);
    return row_spmv_pair<T, decltype(gx)>(si, s, lane, row, dx, L.cp, L.packed, gx);
  }
  if (L.nt) {
    const T dx = row < L.n ? __builtin_nontemporal_load(L.diag + row) * x[row] : T(0);
    return row_spmv_g<T, decltype(gx), false, ColAbsT<true>, true>(si, lane, dx,
                                                                   ColAbsT<true>{L.col}, L.packed, gx);
  }
  const T dx = row < L.n ? L.diag[row] * x[row] : T(0);
  return row_spmv<T>(si, lane, dx, L.col, L.packed, x);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_jacobi(LevelDevT<T> L, int cur, const KState *st) {
  if (st->done) return;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= L.nslices) return;
  const int lane = threadIdx.x & 63, row = s * kSlice + lane;
  const T *x = L.x[cur];
  const T ax = level_ax(L, s, lane, row, x);
  if (row < L.n) L.x[cur ^ 1][row] = x[row] + lvld(L, L.inv, row) * (lvld<T>(L, L.b, row) - ax);
}

// One Gauss-Seidel sweep level of a level's x[cur], in place.
// The stop flag is read beside the cell's operands and checked before the
// store, so its load does not head the chain of dependent loads.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gs_level(LevelDevT<T> L, int cur,
                                                      const GsEnt *__restrict__ ents, int count,
                                                      const KState *st) {
  const bool live = !st->done;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < count)
    gs_entry<T>(ents[t], L.col, L.map, L.packed, L.diag, L.b, L.x[cur], T(1), live);
}

// One symmetric Gauss-Seidel sweep of a whole level by one workgroup, which
// walks the level's dependency levels with a barrier each (a level with few
// cells per dependency level: a launch per dependency level costs more than
// its work).  Each thread loads its first schedule entry of the next
// dependency level while the current one runs.  Same cell updates, in place.
constexpr int kGsWgThreads = 1024;
template <typename T>
__global__ __launch_bounds__(kGsWgThreads) void k_gs_wg(LevelDevT<T> L, int cur,
                                                         const KState *st) {
  if (st->done) return;
  const int t = (int)threadIdx.x;
  T *x = L.x[cur];
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t *ptr = pass ? L.gs_bptr : L.gs_fptr;
    const GsEnt *ents = pass ? L.gs_bent : L.gs_fent;
    const int nl = pass ? L.gs_nbl : L.gs_nfl;
    if (nl <= 0) continue;
    int b = ptr[0], e = ptr[1];
    bool has = b + t < e;
    GsEnt f = has ? ents[b + t] : GsEnt{};
    for (int l = 0; l < nl; ++l) {
      const int en = l + 1 < nl ? ptr[l + 2] : e;  // dependency level l + 1: [e, en)
      const bool hn = e + t < en;
      const GsEnt fn = hn ? ents[e + t] : GsEnt{};
      if (has) gs_entry<T>(f, L.col, L.map, L.packed, L.diag, L.b, x, T(1), true);
      for (int k = b + t + kGsWgThreads; k < e; k += kGsWgThreads)
        gs_entry<T>(ents[k], L.col, L.map, L.packed, L.diag, L.b, x, T(1), true);
      __syncthreads();
      b = e;
      e = en;
      has = hn;
      f = fn;
    }
  }
}

// Levels with at most this many cells per Gauss-Seidel dependency level (one
// per thread of k_gs_wg) run their sweeps in k_gs_wg: a dependency level
// there still costs ~5 us (the schedule entry, the couplings and the gathered
// x are dependent round trips, then the barrier), so only levels whose cells
// fit one pass of the workgroup gain over a launch each (128^3 pair GAMG:
// 90.1 -> 86.4 ms per cycle; 2048 cells: 105 ms, 4096: 138 ms).
// PFM_GS_WG_CELLS overrides, read at each enqueue; 0 = never.
int gs_wg_cells() {
  const char *e = std::getenv("PFM_GS_WG_CELLS");
  return e ? std::atoi(e) : kGsWgThreads;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_resid(LevelDevT<T> L, int cur, const KState *st) {
  if (st->done) return;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= L.nslices) return;
  const int lane = threadIdx.x & 63, row = s * kSlice + lane;
  const T ax = level_ax(L, s, lane, row, L.x[cur]);
  if (row < L.n) L.r[row] = lvld<T>(L, L.b, row) - ax;
}

// Row c of a level's first Jacobi sweep from x = 0, given b_c, written to
// x[1] (k_jacobi's value).  With finite coefficients every product in A 0 is
// a signed zero and the sums start from +0, so A 0 is exactly +0 and the
// sweep is 0 + inv*(b - 0), element-wise; otherwise the row's products are
// formed as k_jacobi forms them (NaN and infinity propagate the same way).
// c runs over whole slices: one wave = 64 consecutive rows = one slice.
template <typename T>
__device__ __forceinline__ void sweep_from_zero(const LevelDevT<T> &L, int c, T bc) {
  if (*L.finite) {
    if (c < L.n) L.x[1][c] = T(0) + L.inv[c] * (bc - T(0));
    return;
  }
  if (c >= L.nslices * kSlice) return;
  const SliceInfo si = load_slice(L.slices, c >> 6);
  const T dx = c < L.n ? L.diag[c] * T(0) : T(0);
  const T ax = row_spmv_g<T>(si, c & 63, dx, ColAbs{L.col}, L.packed, [](int) { return T(0); });
  if (c < L.n) L.x[1][c] = T(0) + L.inv[c] * (bc - ax);
}

// b_c = sum of r over each aggregate (fine-cell order, accumulated in the
// coarse level's type); then x_c[0] = 0, or (kSweep) the coarse level's first
// Jacobi sweep from zero into x_c[1].
template <typename TF, typename TC, bool kSweep = false>
__global__ __launch_bounds__(kBlock) void k_restrict(LevelDevT<TF> F, LevelDevT<TC> Cc,
                                                      const KState *st) {
  if (st->done) return;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  TC s = TC(0);
  if (c < F.nc) {
    for (int k = F.agg_ptr[c]; k < F.agg_ptr[c + 1]; ++k) s += (TC)F.r[F.agg_cells[k]];
    Cc.b[c] = s;
    if constexpr (!kSweep) Cc.x[0][c] = TC(0);
  }
  if constexpr (kSweep) sweep_from_zero(Cc, c, s);
}

template <typename TF, typename TC>
__global__ __launch_bounds__(kBlock) void k_prolong(LevelDevT<TF> F, int cur, const TC *__restrict__ xc,
                                                     const KState *st) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= F.n) return;
  F.x[cur][i] = F.x[cur][i] + (TF)xc[F.f2c[i]];
}

// Prolongation fused with the first post-smoothing Jacobi sweep: the sweep
// reads x_j + e[f2c[j]] (k_prolong's value) wherever it reads x_j and writes
// the swept values to x[cur ^ 1].
template <typename TF, typename TC>
__global__ __launch_bounds__(kBlock) void k_prolong_jacobi(LevelDevT<TF> F, int cur,
                                                            const TC *__restrict__ xc,
                                                            const KState *st) {
  if (st->done) return;
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= F.nslices) return;
  const int lane = threadIdx.x & 63, row = s * kSlice + lane;
  const TF *x = F.x[cur];
  const int32_t *f2c = F.f2c;
  auto xp = [=](int j) { return x[j] + (TF)xc[f2c[j]]; };
  const SliceInfo si = load_slice(F.slices, s);
  const TF xi = row < F.n ? xp(row) : TF(0);
  const TF dx = row < F.n ? F.diag[row] * xi : TF(0);
  const TF ax = F.cp ? row_spmv_pair<TF, decltype(xp)>(si, s, lane, row, dx, F.cp, F.packed, xp)
                     : row_spmv_g<TF>(si, lane, dx, ColAbs{F.col}, F.packed, xp);
  if (row < F.n) F.x[cur ^ 1][row] = xi + F.inv[row] * (F.b[row] - ax);
}

// x[c] of a level by selection (no dynamic index into the descriptor).
__device__ __forceinline__ double *xbuf(const LevelDev &L, int c) { return c ? L.x[1] : L.x[0]; }

// Single-workgroup V-cycle of levels k0..nlev-1 (small levels, always FP64).
// Waves map onto whole slices (rows t, t + 1024, ... with 64-aligned wave
// offsets).
__device__ void tail_jacobi(const LevelDev &L, int cur) {
  for (int row0 = threadIdx.x; row0 < L.nslices * kSlice; row0 += kTailThreads) {
    const int s = row0 >> 6, lane = row0 & 63;
    const double *x = L.x[cur];
    const double ax = level_ax<double, false>(L, s, lane, row0, x);
    if (row0 < L.n) L.x[cur ^ 1][row0] = x[row0] + L.inv[row0] * (L.b[row0] - ax);
  }
  __syncthreads();
}

// One symmetric Gauss-Seidel sweep of a small level by the whole workgroup.
__device__ void tail_gs(const LevelDev &L, int cur) {
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t *ptr = pass ? L.gs_bptr : L.gs_fptr;
    const GsEnt *ents = pass ? L.gs_bent : L.gs_fent;
    const int nl = pass ? L.gs_nbl : L.gs_nfl;
    for (int l = 0; l < nl; ++l) {
      for (int k = ptr[l] + (int)threadIdx.x; k < ptr[l + 1]; k += kTailThreads)
        gs_entry<double>(ents[k], L.col, L.map, L.packed, L.diag, L.b, L.x[cur], 1.0, true);
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

// ---- levels of at most one slice, held by one wave ----
// A level with n <= 64 cells is one slice: lane i is cell i, its couplings
// sit in registers and the x value of column c is lane c's register
// (__shfl), so Jacobi sweeps, the residual, restriction and prolongation
// need neither memory round trips nor workgroup barriers.  Every sum keeps
// the association of row_spmv_g and of the block path: (diag*x + sum_lower)
// + sum_upper, each sum from 0 in slot order, padding slots skipped;
// aggregate sums from 0 in fine-cell order.  Shuffles run on all 64 lanes
// (loops with wave-uniform bounds); lanes >= n carry zeros.
constexpr int kWaveSlots = 8;     // slots per group kept in registers
constexpr int kMaxWaveLevels = 8;   // levels of the one-wave part (LDS rows kept)
constexpr int kRegCells = 2048;     // cells of the one-row-per-thread levels (LDS rows kept)

struct WaveRow {
  int cl[kWaveSlots], cu[kWaveSlots];
  double vl[kWaveSlots], vu[kWaveSlots];
  double d, inv;
  int n;
  int wl, wu;  // the slice's group widths (uniform across its 64 rows)
};

// Whether level L can be held by one wave.
__device__ __forceinline__ bool wave_level(const LevelDev &L) {
  if (L.n > kSlice) return false;
  if (L.nslices == 0) return true;
  const SliceInfo si = load_slice(L.slices, 0);
  return si.wl <= kWaveSlots && si.wu <= kWaveSlots;
}

// Row t's couplings (slice t / 64, lane t % 64; t is wave-uniform per slice).
__device__ __forceinline__ void wave_load(const LevelDev &L, int t, WaveRow &R) {
  R.n = L.n;
  SliceInfo si{0, 0, 0};
  if ((t >> 6) < L.nslices) si = load_slice(L.slices, t >> 6);
  R.wl = si.wl;
  R.wu = si.wu;
  const int lane = t & 63;
  const int bl = si.ptr + lane, bu = si.ptr + si.wl * kSlice + lane;
#pragma unroll
  for (int j = 0; j < kWaveSlots; ++j) {
    R.cl[j] = j < si.wl ? L.col[bl + j * kSlice] : -1;
    R.vl[j] = j < si.wl ? L.packed[bl + j * kSlice] : 0.0;
    R.cu[j] = j < si.wu ? L.col[bu + j * kSlice] : -1;
    R.vu[j] = j < si.wu ? L.packed[bu + j * kSlice] : 0.0;
  }
  R.d = t < L.n ? L.diag[t] : 0.0;
  R.inv = t < L.n ? L.inv[t] : 0.0;
}

__device__ __forceinline__ double wave_ax(const WaveRow &R, int lane, double x) {
  const double dx = lane < R.n ? R.d * x : 0.0;
  double accL = 0.0, accU = 0.0;
  // Slots at or beyond the slice's width are padding on every lane: skip
  // their cross-lane reads (the width is wave-uniform).
#pragma unroll
  for (int j = 0; j < kWaveSlots; ++j) {
    if (j < R.wl) {
      const double xv = __shfl(x, R.cl[j] >= 0 ? R.cl[j] : 0);
      if (R.cl[j] >= 0) accL += R.vl[j] * xv;
    }
  }
#pragma unroll
  for (int j = 0; j < kWaveSlots; ++j) {
    if (j < R.wu) {
      const double xv = __shfl(x, R.cu[j] >= 0 ? R.cu[j] : 0);
      if (R.cu[j] >= 0) accU += R.vu[j] * xv;
    }
  }
  return (dx + accL) + accU;
}

__device__ __forceinline__ double wave_jacobi(const WaveRow &R, int lane, double x, double b) {
  const double ax = wave_ax(R, lane, x);
  return lane < R.n ? x + R.inv * (b - ax) : 0.0;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off));
  return v;
}

// The V-cycle below level kw (levels kw..nlev-1, all one-wave levels), run by
// the 64 lanes of one wave: lane i enters with level kw's x and b of cell i
// and returns its x after the cycle.  The other levels' b and x stay in
// registers and in `save` (lane-private LDS rows, indexed by k - kw).
__device__ double wave_cycle(const LevelDev *__restrict__ Ls, int kw, int nlev, double x,
                             double b, int n_pre, int n_post, int n_coarsest,
                             double (*save)[2][kSlice]) {
  const int lane = threadIdx.x;
  WaveRow R;
  for (int k = kw; k < nlev - 1; ++k) {
    const LevelDev &L = Ls[k];
    wave_load(L, lane, R);
    for (int it = 0; it < n_pre; ++it) x = wave_jacobi(R, lane, x, b);
    const double ax = wave_ax(R, lane, x);
    const double r = lane < L.n ? b - ax : 0.0;
    save[k - kw][0][lane] = x;
    save[k - kw][1][lane] = b;
    int a0 = 0, len = 0;
    if (lane < L.nc) {
      a0 = L.agg_ptr[lane];
      len = L.agg_ptr[lane + 1] - a0;
    }
    const int maxlen = wave_max(len);
    double s = 0.0;
    for (int q = 0; q < maxlen; ++q) {
      const double v = __shfl(r, q < len ? L.agg_cells[a0 + q] : 0);
      if (q < len) s += v;
    }
    b = s;
    x = 0.0;
  }
  wave_load(Ls[nlev - 1], lane, R);
  for (int it = 0; it < n_coarsest; ++it) x = wave_jacobi(R, lane, x, b);
  for (int k = nlev - 2; k >= kw; --k) {
    const LevelDev &L = Ls[k];
    const double xc = x;
    const int f = lane < L.n ? L.f2c[lane] : 0;
    const double e = __shfl(xc, f);
    x = lane < L.n ? save[k - kw][0][lane] + e : 0.0;
    b = save[k - kw][1][lane];
    wave_load(L, lane, R);
    for (int it = 0; it < n_post; ++it) x = wave_jacobi(R, lane, x, b);
  }
  return x;
}

// ---- levels of at most kTailThreads cells, one row per thread ----
// Thread t holds row t's couplings in registers; the x values a sweep reads
// are exchanged through LDS (xs) between two workgroup barriers, the
// residual through rs.  Same association as everywhere: (diag*x + sum_lower)
// + sum_upper, aggregate sums from 0 in fine-cell order.
__device__ __forceinline__ double reg_ax(const WaveRow &R, int t, double x,
                                         const double *__restrict__ xs) {
  const double dx = t < R.n ? R.d * x : 0.0;
  double accL = 0.0, accU = 0.0;
#pragma unroll
  for (int j = 0; j < kWaveSlots; ++j)
    if (R.cl[j] >= 0) accL += R.vl[j] * xs[R.cl[j]];
#pragma unroll
  for (int j = 0; j < kWaveSlots; ++j)
    if (R.cu[j] >= 0) accU += R.vu[j] * xs[R.cu[j]];
  return (dx + accL) + accU;
}

__device__ __forceinline__ double reg_jacobi(const WaveRow &R, int t, double x, double b,
                                             double *xs) {
  xs[t] = x;
  __syncthreads();
  const double ax = reg_ax(R, t, x, xs);
  __syncthreads();
  return t < R.n ? x + R.inv * (b - ax) : 0.0;
}

// Whether level L can run one row per thread (every slice's groups fit the
// registers).  Called by the whole workgroup.
__device__ bool reg_level(const LevelDev &L) {
  if (L.n > kTailThreads) return false;
  int ok = 1;
  for (int s = threadIdx.x; s < L.nslices; s += kTailThreads) {
    const SliceInfo si = load_slice_lane(L.slices, s);
    ok &= si.wl <= kWaveSlots && si.wu <= kWaveSlots;
  }
  return __syncthreads_and(ok) != 0;
}

__global__ __launch_bounds__(kTailThreads) void k_tail(const LevelDev *__restrict__ Ls, int k0,
                                                        int nlev, int cur0, int n_pre, int n_post,
                                                        int n_coarsest, int gs, const KState *st) {
  __shared__ double save[kMaxWaveLevels][2][kSlice];
  __shared__ double rsave[2][kRegCells];
  __shared__ double xs[kTailThreads], rs[kTailThreads];
  if (st->done) return;
  // Jacobi buffer parity of every level, one bit per level (an array indexed
  // by level would live in private memory).
  uint64_t par = (uint64_t)(cur0 & 1) << k0;
  auto parity = [&](int k) { return (int)((par >> k) & 1); };
  auto set_parity = [&](int k, int c) { par = (par & ~(1ull << k)) | ((uint64_t)(c & 1) << k); };
  auto smooth = [&](const LevelDev &L, int k) {
    int c = parity(k);
    tail_smooth(L, c, gs);
    set_parity(k, c);
  };
  // kw: first level of the one-wave part (every level from kw on fits one
  // wave); nlev when there is none.  kt: first level of the one-row-per-
  // thread part [kt, kw) just above it (kt = kw when there is none).  Both
  // for Jacobi smoothing only; levels [k0, kb) take the block path.
  int kw = nlev;
  if (!gs)
    while (kw > k0 && nlev - (kw - 1) <= kMaxWaveLevels && wave_level(Ls[kw - 1])) --kw;
  int kt = kw;
  if (kw < nlev) {
    int cells = 0;
    while (kt > k0 && cells + Ls[kt - 1].n <= kRegCells && reg_level(Ls[kt - 1])) {
      cells += Ls[kt - 1].n;
      --kt;
    }
  }
  const int kb = kw < nlev ? kt : nlev - 1;
  for (int k = k0; k < kb; ++k) {
    const LevelDev &L = Ls[k];
    for (int it = 0; it < n_pre; ++it) smooth(L, k);
    const double *xk = xbuf(L, parity(k));
    for (int row0 = threadIdx.x; row0 < L.nslices * kSlice; row0 += kTailThreads) {
      const double ax = level_ax<double, false>(L, row0 >> 6, row0 & 63, row0, xk);
      if (row0 < L.n) L.r[row0] = L.b[row0] - ax;
    }
    __syncthreads();
    const LevelDev &Cc = Ls[k + 1];
    for (int c = threadIdx.x; c < L.nc; c += kTailThreads) {
      double s = 0.0;
      for (int q = L.agg_ptr[c]; q < L.agg_ptr[c + 1]; ++q) s += L.r[L.agg_cells[q]];
      Cc.b[c] = s;
      Cc.x[0][c] = 0.0;
    }
    set_parity(k + 1, 0);
    __syncthreads();
  }
  if (kw < nlev) {
    // Levels kt..nlev-1 in registers.  Level kt enters from memory (x[cur],
    // b) and leaves in x[pout], the buffer the block path would leave it in
    // (every Jacobi step flips the parity).
    const int t = threadIdx.x;
    const int flips = kt == nlev - 1 ? n_coarsest : n_pre + n_post;
    const int pout = (parity(kt) + flips) & 1;
    double x = t < Ls[kt].n ? xbuf(Ls[kt], parity(kt))[t] : 0.0;
    double b = t < Ls[kt].n ? Ls[kt].b[t] : 0.0;
    WaveRow R;
    int off = 0;
    for (int k = kt; k < kw; ++k) {  // one row per thread, descending
      const LevelDev &L = Ls[k];
      wave_load(L, t, R);
      for (int it = 0; it < n_pre; ++it) x = reg_jacobi(R, t, x, b, xs);
      xs[t] = x;
      __syncthreads();
      const double ax = reg_ax(R, t, x, xs);
      rs[t] = t < L.n ? b - ax : 0.0;
      if (t < L.n) {
        rsave[0][off + t] = x;
        rsave[1][off + t] = b;
      }
      __syncthreads();
      double sb = 0.0;
      if (t < L.nc)
        for (int q = L.agg_ptr[t]; q < L.agg_ptr[t + 1]; ++q) sb += rs[L.agg_cells[q]];
      b = sb;
      x = 0.0;
      off += L.n;
      __syncthreads();
    }
    if (t < kSlice) x = wave_cycle(Ls, kw, nlev, x, b, n_pre, n_post, n_coarsest, save);
    for (int k = kw - 1; k >= kt; --k) {  // ascending
      const LevelDev &L = Ls[k];
      off -= L.n;
      xs[t] = t < Ls[k + 1].n ? x : 0.0;
      wave_load(L, t, R);
      const int f = t < L.n ? L.f2c[t] : 0;
      __syncthreads();
      x = t < L.n ? rsave[0][off + t] + xs[f] : 0.0;
      b = t < L.n ? rsave[1][off + t] : 0.0;
      __syncthreads();
      for (int it = 0; it < n_post; ++it) x = reg_jacobi(R, t, x, b, xs);
    }
    if (t < Ls[kt].n) xbuf(Ls[kt], pout)[t] = x;
    set_parity(kt, pout);
    __syncthreads();
  } else {
    const LevelDev &L = Ls[nlev - 1];
    for (int it = 0; it < n_coarsest; ++it) smooth(L, nlev - 1);
  }
  for (int k = kb - 1; k >= k0; --k) {
    const LevelDev &L = Ls[k];
    const double *xc = xbuf(Ls[k + 1], parity(k + 1));
    double *xk = xbuf(L, parity(k));
    for (int i = threadIdx.x; i < L.n; i += kTailThreads) xk[i] = xk[i] + xc[L.f2c[i]];
    __syncthreads();
    for (int it = 0; it < n_post; ++it) smooth(L, k);
  }
}

// r = b - A x on level 0 and the partial sums of r.r (accumulated in FP64).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_res_norm(LevelDevT<T> L, int cur,
                                                      double *__restrict__ part, const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc = 0.0;
  const SliceRange sr = block_slices(L.nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    const T ax = level_ax(L, s, lane, row, L.x[cur]);
    if (row < L.n) {
      const double ri = (double)(L.b[row] - ax);
      acc += ri * ri;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// k_res_norm fused with the next cycle's first Jacobi sweep of level 0, which
// forms the same r = b - A x: partials of r.r as k_res_norm (same slices per
// block, same order) and x[cur ^ 1] = x + inv*r as k_jacobi.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_res_norm_jacobi(LevelDevT<T> L, int cur,
                                                             double *__restrict__ part,
                                                             const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc = 0.0;
  const SliceRange sr = block_slices(L.nslices);
  const int lane = threadIdx.x & 63;
  const T *x = L.x[cur];
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    const T ax = level_ax(L, s, lane, row, x);
    if (row < L.n) {
      const T rt = lvld<T>(L, L.b, row) - ax;
      const double ri = (double)rt;
      acc += ri * ri;
      L.x[cur ^ 1][row] = x[row] + lvld(L, L.inv, row) * rt;
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

// ------------------------------------------------- mixed precision -----
// Outer FP64 residual r = b - A x of level 0 (the caller's x and b) and its
// partial sums of r.r.
__global__ __launch_bounds__(kBlock) void k_gm_resid(LevelDev L, double *__restrict__ r,
                                                      double *__restrict__ part, const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc = 0.0;
  const SliceRange sr = block_slices(L.nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    const double ax = level_ax(L, s, lane, row, L.x[0]);
    if (row < L.n) {
      const double ri = L.b[row] - ax;
      r[row] = ri;
      acc += ri * ri;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Outer monitor (the reference rule on the FP64 residual, the initial one
// included) and the inner state: stop when the inner residual falls below
// tau times this outer residual or the remaining cycle budget is spent.
__global__ __launch_bounds__(kBlock) void k_gm_outer_fin(const double *__restrict__ part, int G,
                                                          KState *st, double *__restrict__ hist,
                                                          int outer, int max_outer, int budget,
                                                          double tau, KState *in) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const double RR = sum_partials(part, G, lds);
  if (threadIdx.x == 0) {
    const double res = sqrt(RR);
    st->iter = outer;
    if (monitor_update(st, hist, res, outer))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (outer >= max_outer || budget <= 0)
      finish(st, PFM_SOLVE_MAXITER);
    in->tol = -1.0;
    in->rel_tol = tau;
    in->res0 = res;
    in->res = res;
    in->nhist = 1;
    in->iter = 0;
    in->max_iter = budget;
    in->min_iter = 0;
    in->hist_cap = 0;
    in->status = 0;
    in->done = st->done;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_cast(int n, const double *__restrict__ r,
                                                     T *__restrict__ b, T *__restrict__ x,
                                                     const KState *st) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  b[i] = (T)r[i];
  x[i] = T(0);
}

// b = (T) r on level L, x[0] = 0, and the level's first Jacobi sweep from
// x = 0 into x[1] (sweep_from_zero); grid over whole slices of L.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_cast_sweep(LevelDevT<T> L,
                                                           const double *__restrict__ r,
                                                           const KState *st) {
  if (st->done) return;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  T bc = T(0);
  if (c < L.n) {
    bc = (T)r[c];
    L.b[c] = bc;
    L.x[0][c] = T(0);  // the solution when no cycle runs
  }
  sweep_from_zero(L, c, bc);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gm_correct(int n, const T *__restrict__ e,
                                                        double *__restrict__ x, const KState *st) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + (double)e[i];
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const double *__restrict__ a,
                                                    float *__restrict__ b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ----------------------------------------- PCG with a V-cycle preconditioner ---
// The PCG of solvers/pcg.py:97-187 with z = M^-1 r one V-cycle from zero on
// A z = r (the hierarchy's levels, FP64 or FP32 on the finest ones).  Partial
// regions: A = [0, G) p.Ap, B = [G, 2G) r.r, C = [2G, 3G) r.z.

// Ap = A p on level 0; partial p.Ap (region A).
__global__ __launch_bounds__(kBlock) void k_gp_spmv(LevelDev L, const double *__restrict__ p,
                                                     double *__restrict__ Ap,
                                                     double *__restrict__ part, const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc = 0.0;
  const SliceRange sr = block_slices(L.nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    const double pi = row < L.n ? p[row] : 0.0;
    const double v = level_ax(L, s, lane, row, p);
    if (row < L.n) {
      Ap[row] = v;
      acc += pi * v;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// alpha = rz / pAp (breakdown if |pAp| < 1e-30); x += alpha p; r -= alpha Ap;
// partial r.r (region B).
__global__ __launch_bounds__(kBlock) void k_gp_update(int n, int nslices, int it,
                                                       const double *__restrict__ p,
                                                       const double *__restrict__ Ap,
                                                       double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       double *__restrict__ part, KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double PAP = reduce_partials(part, G, lds);
  if (fabs(PAP) < 1e-30) {
    if (blockIdx.x == 0 && threadIdx.x == 0) finish(st, PFM_SOLVE_BREAKDOWN);
    return;
  }
  const double alpha = st->rzs[(it - 1) & 1] / PAP;
  double acc = 0.0;
  const SliceRange sr = block_slices(nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    if (row < n) {
      x[row] = x[row] + alpha * p[row];
      const double ri = r[row] - alpha * Ap[row];
      r[row] = ri;
      acc += ri * ri;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[G + blockIdx.x] = acc;
}

// Residual monitor of iteration it (region B); stop before the preconditioner.
__global__ __launch_bounds__(kBlock) void k_gp_decide(const double *__restrict__ part, int G,
                                                       int it, KState *st,
                                                       double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const double RR = reduce_partials(part + G, G, lds);
  if (threadIdx.x == 0) {
    st->iter = it;
    if (monitor_update(st, hist, sqrt(RR), it))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (it >= st->max_iter)
      finish(st, PFM_SOLVE_MAXITER);
  }
}

// z = M^-1 r from the cycle's level-0 solution; partial r.z (region C).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gp_z(int n, int nslices, const T *__restrict__ e,
                                                  const double *__restrict__ r,
                                                  double *__restrict__ z, double *__restrict__ part,
                                                  const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  double acc = 0.0;
  const SliceRange sr = block_slices(nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    if (row < n) {
      const double zi = (double)e[row];
      z[row] = zi;
      acc += r[row] * zi;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[2 * G + blockIdx.x] = acc;
}

// rz' = r.z (region C); beta = rz' / rz (0 on the first pass: p = z);
// p = z + beta p.  Workgroup 0 records rz' by iteration parity.
__global__ __launch_bounds__(kBlock) void k_gp_pupdate(int n, int nslices, int it,
                                                        const double *__restrict__ z,
                                                        double *__restrict__ p,
                                                        const double *__restrict__ part,
                                                        KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double RZ = reduce_partials(part + 2 * G, G, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rzs[it & 1] = RZ;
  const double beta = it == 0 ? 0.0 : RZ / st->rzs[(it - 1) & 1];
  const SliceRange sr = block_slices(nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const int row = s * kSlice + lane;
    if (row < n) p[row] = it == 0 ? z[row] : z[row] + beta * p[row];
  }
}

// Initial monitor (r = b - A x0 from k_gm_resid, region A of its partials).
__global__ __launch_bounds__(kBlock) void k_gp_init_fin(const double *__restrict__ part, int G,
                                                         KState *st, double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  const double RR = sum_partials(part, G, lds);
  if (threadIdx.x == 0) {
    if (monitor_update(st, hist, sqrt(RR), 0))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (st->max_iter < 1)
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
    D.nt = (int64_t)H.ldu->nslots * 12 >= kMallBytes;  // 8-byte value + 4-byte column
    D.slices = H.ldu->d_slices;
    D.col = H.ldu->d_col;
    D.cp = level_pairs(H.ldu);
    D.map = H.ldu->d_map;
    D.gs_fptr = H.ldu->d_gs_fwd_ptr;
    D.gs_fcells = H.ldu->d_gs_fwd_cells;
    D.gs_fent = H.ldu->d_gs_fwd_ent;
    D.gs_bptr = H.ldu->d_gs_bwd_ptr;
    D.gs_bcells = H.ldu->d_gs_bwd_cells;
    D.gs_bent = H.ldu->d_gs_bwd_ent;
    D.gs_nfl = (int32_t)H.ldu->gs_fwd_ptr.size() - 1;
    D.gs_nbl = (int32_t)H.ldu->gs_bwd_ptr.size() - 1;
    t.packed[k] = (double *)take(sizeof(double) * (size_t)std::max<int64_t>(H.ldu->nslots, 1));
    D.packed = t.packed[k];
    D.inv = (double *)take(sizeof(double) * n);
    D.r = (double *)take(sizeof(double) * n);
    D.x[1] = (double *)take(sizeof(double) * n);
    D.finite = (int32_t *)take(sizeof(int32_t));
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

// Mixed-precision workspace: the FP64 layout (every level's coefficients,
// FP64 tail levels, outer state and history), then FP32 copies of levels
// [0, tail_k), the inner state, the outer residual and FP64 level-0 inner
// buffers (used when every level is a tail level).
struct GamgMixedWork {
  GamgWork w64;
  std::vector<LevelDevT<float>> lv32;
  std::vector<float *> packed32, diag32, inv32;
  KState *in;
  double *r64, *e64, *b64;
  LevelDev *d_small;
};

// First level handled by the single-workgroup tail kernel.
inline int tail_level(const pfm_gamg_s *g) {
  int k = (int)g->lv.size() - 1;
  while (k > 0 && g->lv[k - 1].n <= kTailCells) --k;
  return k;
}

// First level with at most kFp32MinCells cells (it and all coarser levels
// stay FP64 in the mixed-precision cycles).
inline int fp64_level(const pfm_gamg_s *g) {
  int k = (int)g->lv.size() - 1;
  while (k > 0 && g->lv[k - 1].n <= kFp32MinCells) --k;
  return k;
}

// Levels computed in FP32: the first fp32_levels, and never one of at most
// kFp32MinCells cells.
inline int fp32_count(const pfm_gamg_s *g, int fp32_levels) {
  return std::min(fp32_levels, fp64_level(g));
}

size_t gamg_mixed_layout(const pfm_gamg_s *g, int max_outer, int fp32_levels, GamgMixedWork *m,
                         char *base) {
  GamgWork w64;
  size_t off = gamg_layout(g, max_outer, &w64, base);
  auto take = [&](size_t bytes) -> void * {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  const int nbig = fp32_count(g, fp32_levels);
  GamgMixedWork t;
  t.lv32.resize(nbig);
  t.packed32.resize(nbig);
  t.diag32.resize(nbig);
  t.inv32.resize(nbig);
  for (int k = 0; k < nbig; ++k) {
    const auto &H = g->lv[k];
    const size_t n = (size_t)H.n;
    const LevelDev &D64 = w64.lv[k];
    LevelDevT<float> &D = t.lv32[k];
    memset(&D, 0, sizeof(D));
    D.n = D64.n;
    D.nslices = D64.nslices;
    D.nt = (int64_t)H.ldu->nslots * 8 >= kMallBytes;  // 4-byte value + 4-byte column
    D.slices = D64.slices;
    D.col = D64.col;
    D.cp = D64.cp;
    D.map = D64.map;
    D.gs_fptr = D64.gs_fptr;
    D.gs_fcells = D64.gs_fcells;
    D.gs_fent = D64.gs_fent;
    D.gs_bptr = D64.gs_bptr;
    D.gs_bcells = D64.gs_bcells;
    D.gs_bent = D64.gs_bent;
    D.gs_nfl = D64.gs_nfl;
    D.gs_nbl = D64.gs_nbl;
    D.f2c = D64.f2c;
    D.agg_ptr = D64.agg_ptr;
    D.agg_cells = D64.agg_cells;
    D.nc = D64.nc;
    t.packed32[k] = (float *)take(sizeof(float) * (size_t)std::max<int64_t>(H.ldu->nslots, 1));
    t.diag32[k] = (float *)take(sizeof(float) * n);
    t.inv32[k] = (float *)take(sizeof(float) * n);
    D.packed = t.packed32[k];
    D.diag = t.diag32[k];
    D.inv = t.inv32[k];
    D.finite = (int32_t *)take(sizeof(int32_t));
    D.x[0] = (float *)take(sizeof(float) * n);
    D.x[1] = (float *)take(sizeof(float) * n);
    D.b = (float *)take(sizeof(float) * n);
    D.r = (float *)take(sizeof(float) * n);
  }
  const size_t n0 = (size_t)g->lv[0].n;
  t.in = (KState *)take(sizeof(KState));
  t.r64 = (double *)take(sizeof(double) * n0);
  t.e64 = (double *)take(sizeof(double) * n0);
  t.b64 = (double *)take(sizeof(double) * n0);
  t.d_small = (LevelDev *)take(sizeof(LevelDev) * g->lv.size());
  if (m) {
    t.w64 = w64;
    *m = t;
  }
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

// Set *flag (device) to whether a0[0..n0), a1[0..n1) and a2[0..n2) are all
// finite.
template <typename T>
void mark_finite(const int32_t *flag, int64_t n0, const T *a0, int64_t n1, const T *a1,
                 int64_t n2, const T *a2, hipStream_t s) {
  int32_t *f = const_cast<int32_t *>(flag);
  k_flag_one<<<1, 1, 0, s>>>(f);
  const int64_t tot = n0 + n1 + n2;
  if (tot > 0)
    k_finite<T><<<(int)std::min<int64_t>((tot + kBlock - 1) / kBlock, 1024), kBlock, 0, s>>>(
        n0, a0, n1, a1, n2, a2, f);
}

// The levels one V-cycle runs on: levels [0, nbig) with values of type T,
// the others in FP64: [nbig, tail_k) by per-level launches, [tail_k, L) in
// the one-workgroup tail (host descriptors and their device copy).
template <typename T>
struct CycleLevels {
  std::vector<LevelDevT<T>> big;  // levels [0, nbig) in T
  std::vector<LevelDev> small;    // FP64 descriptors of every level: [nbig, L) use them
  const LevelDev *d_small = nullptr;
  int nbig = 0;                   // levels in T (nbig <= tail_k)
  int tail_k = 0;                 // levels [tail_k, L) run in the one-workgroup tail
};

// Enqueue one V-cycle (solvers/gamg.py:437-496); cur0 is the parity of
// level 0's Jacobi buffer before and after the cycle.
template <typename T>
void enqueue_cycle(const pfm_gamg_s *g, const CycleLevels<T> &C, const KState *st, int &cur0,
                   hipStream_t s, bool swept0 = false) {
  const int L = (int)g->lv.size();
  const int npre = g->prm.n_pre, npost = g->prm.n_post, ncoarse = g->prm.n_coarsest;
  const bool gs = g->prm.smoother == PFM_SMOOTHER_GAUSS_SEIDEL;
  const int tail_k = C.tail_k, nbig = C.nbig;
  auto smooth_level = [&](const auto &D, int k, int &cur) {
    using V = typename std::remove_const<typename std::remove_pointer<decltype(D.packed)>::type>::type;
    if (gs) {  // one symmetric sweep, level by level, in place
      const pfm_ldu h = g->lv[k].ldu;
      const int64_t ndep = std::max<int64_t>(
          std::max<int64_t>((int64_t)h->gs_fwd_ptr.size() - 1, (int64_t)h->gs_bwd_ptr.size() - 1), 1);
      if ((int64_t)h->n <= (int64_t)gs_wg_cells() * ndep) {  // few cells per dependency level
        k_gs_wg<V><<<1, kGsWgThreads, 0, s>>>(D, cur, st);
        return;
      }
      for (int pass = 0; pass < 2; ++pass) {
        const std::vector<int32_t> &ptr = pass ? h->gs_bwd_ptr : h->gs_fwd_ptr;
        const GsEnt *ents = pass ? h->d_gs_bwd_ent : h->d_gs_fwd_ent;
        for (size_t l = 0; l + 1 < ptr.size(); ++l) {
          const int cnt = ptr[l + 1] - ptr[l];
          k_gs_level<V><<<grid1d(cnt), kBlock, 0, s>>>(D, cur, ents + ptr[l], cnt, st);
        }
      }
      return;
    }
    k_jacobi<V><<<blocks_for_slices(D.nslices), kBlock, 0, s>>>(D, cur, st);
    cur ^= 1;
  };
  auto smooth = [&](int k, int &cur) {
    if (k < nbig)
      smooth_level(C.big[k], k, cur);
    else
      smooth_level(C.small[k], k, cur);
  };
  // A launched coarse level starts from x = 0: its first Jacobi pre-sweep is
  // done by the restriction (sweep_from_zero), and likewise level 0's when
  // the caller's kernel did it (swept0, with cur0 = 1).
  int cur[kMaxLevels];
  bool swept[kMaxLevels] = {};
  cur[0] = cur0;
  swept[0] = swept0;
  auto restrict_to = [&](const auto &F, const auto &Cc, bool sweep) {
    using TF = typename std::remove_const<typename std::remove_pointer<decltype(F.packed)>::type>::type;
    using TC = typename std::remove_const<typename std::remove_pointer<decltype(Cc.packed)>::type>::type;
    if (sweep)
      k_restrict<TF, TC, true><<<grid1d(F.nc), kBlock, 0, s>>>(F, Cc, st);
    else
      k_restrict<TF, TC, false><<<grid1d(F.nc), kBlock, 0, s>>>(F, Cc, st);
  };
  for (int k = 0; k < tail_k; ++k) {  // descend through the large levels
    for (int it = swept[k] ? 1 : 0; it < npre; ++it) smooth(k, cur[k]);
    const bool sweep = !gs && npre >= 1 && k + 1 < tail_k;
    if (k < nbig) {
      const LevelDevT<T> &F = C.big[k];
      k_resid<T><<<blocks_for_slices(F.nslices), kBlock, 0, s>>>(F, cur[k], st);
      if (k + 1 < nbig)
        restrict_to(F, C.big[k + 1], sweep);
      else
        restrict_to(F, C.small[k + 1], sweep);
    } else {
      const LevelDev &F = C.small[k];
      k_resid<double><<<blocks_for_slices(F.nslices), kBlock, 0, s>>>(F, cur[k], st);
      restrict_to(F, C.small[k + 1], sweep);
    }
    cur[k + 1] = sweep ? 1 : 0;
    swept[k + 1] = sweep;
  }
  // levels tail_k .. L-1 in one workgroup
  k_tail<<<1, kTailThreads, 0, s>>>(C.d_small, tail_k, L, cur[tail_k], npre, npost, ncoarse,
                                    gs ? 1 : 0, st);
  if (!gs) cur[tail_k] ^= (tail_k == L - 1 ? ncoarse : npre + npost) & 1;
  for (int k = tail_k - 1; k >= 0; --k) {  // ascend
    // On levels of at most kFuseCells cells the prolongation is folded into
    // the first post-sweep (one launch less; at 128^3 0.655 ms per cycle with
    // 65,536, 0.652 with 262,144, 0.656 on every level: on the finest levels
    // the sweep's extra gathers of the coarse correction cost what the
    // separate pass saves).
    const bool pj = !gs && npost >= 1 && g->lv[k].n <= kFuseCells;
    auto prolong = [&](const auto &F, const auto *xc) {
      using TF = typename std::remove_const<typename std::remove_pointer<decltype(F.packed)>::type>::type;
      using TC = typename std::remove_const<typename std::remove_pointer<decltype(xc)>::type>::type;
      if (pj) {
        k_prolong_jacobi<TF, TC><<<blocks_for_slices(F.nslices), kBlock, 0, s>>>(F, cur[k], xc, st);
        cur[k] ^= 1;
      } else {
        k_prolong<TF, TC><<<grid1d(F.n), kBlock, 0, s>>>(F, cur[k], xc, st);
      }
    };
    if (k < nbig) {
      const LevelDevT<T> &F = C.big[k];
      if (k + 1 < nbig)
        prolong(F, C.big[k + 1].x[cur[k + 1]]);
      else
        prolong(F, C.small[k + 1].x[cur[k + 1]]);
    } else {
      prolong(C.small[k], C.small[k + 1].x[cur[k + 1]]);
    }
    for (int it = pj ? 1 : 0; it < npost; ++it) smooth(k, cur[k]);
  }
  cur0 = cur[0];
}

// Chunked enqueue of V-cycles until the device stop flag of `st` is seen (one
// chunk queued ahead of the poll).  cycle() enqueues one cycle and its
// residual check; cur_after[i] is level 0's buffer parity after cycle i+1.
template <typename Cycle>
int drive_cycles(const pfm_gamg_s *g, const KState *st, int max_cycles, hipStream_t s,
                 std::vector<int> &cur_after, int &cur0, Cycle cycle) {
  PollLease lease;
  int rc0 = lease.acquire(const_cast<PollPool *>(&g->poll), true);
  if (rc0 != PFM_OK) return rc0;
  int32_t *flag = lease.flag;
  hipEvent_t *ev = lease.ev;
  int enq = 0, chunk = 2, slot = 0, rc = PFM_OK;
  bool pending = false;
  double last_report = now_seconds();
  while (true) {
    bool queued = false;
    report_progress(g->fine, enq, last_report);
    if (enq < max_cycles) {
      const int c = std::min(chunk, max_cycles - enq);
      for (int k = 0; k < c; ++k) {
        cycle();
        cur_after.push_back(cur0);
      }
      enq += c;
      chunk = std::min(chunk * 2, 16);
      if (hipMemcpyAsync(&flag[slot], &st->done, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
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
  if (rc != PFM_OK) set_error("HIP error while driving the GAMG cycles");
  return rc;
}

extern "C" {

int pfm_gamg_destroy(pfm_gamg g) {
  if (!g) return PFM_OK;
  for (auto &H : g->lv) free_level(H);
  poll_pool_free(g->poll);
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
  if (st == PFM_OK) st = poll_pool_init(g->poll);
  if (st == PFM_OK && g->prm.smoother == PFM_SMOOTHER_GAUSS_SEIDEL)  // every level sweeps
    for (size_t k = 0; k < g->lv.size() && st == PFM_OK; ++k)
      st = ensure_gs_entries(g->lv[k].ldu, nullptr);
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
    mark_finite<double>(w.lv[k].finite, H.n, d, H.nf, lo, H.nf, up, s);
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
  CycleLevels<double> C;
  C.tail_k = tail_level(g);
  C.nbig = C.tail_k;
  C.big.assign(w.lv.begin(), w.lv.begin() + C.tail_k);
  C.small = w.lv;
  C.d_small = w.d_levels;
  const int G = red_grid(g->fine);
  int cur0 = 0;
  // the buffer parity of level 0 advances by the same amount every cycle, so
  // enqueueing cycles ahead of the stop flag never changes which buffer holds
  // x; cur0 after a cycle that did not run is unchanged on the device, so the
  // value after every enqueued cycle is kept
  std::vector<int> cur_after;
  cur_after.reserve(prm->max_iter + 1);
  // With a launched level 0 and Jacobi smoothing, cycle i's residual is
  // formed by cycle i+1's first level-0 sweep (k_res_norm_jacobi) and
  // monitored there; the remaining kernels of cycle i+1 return once the
  // monitor has stopped the solve, and the last cycle's residual is formed
  // after the loop.  Otherwise each cycle ends with k_res_norm.
  const bool fold = C.tail_k > 0 && g->prm.smoother != PFM_SMOOTHER_GAUSS_SEIDEL &&
                    g->prm.n_pre >= 1;
  int ncyc = 0;
  int rc = drive_cycles(g, w.st, prm->max_iter, s, cur_after, cur0, [&]() {
    if (fold && ncyc > 0) {
      k_res_norm_jacobi<double><<<G, kBlock, 0, s>>>(w.lv[0], cur0, w.part, w.st);
      k_gamg_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
      cur0 ^= 1;
      enqueue_cycle<double>(g, C, w.st, cur0, s, true);
    } else {
      enqueue_cycle<double>(g, C, w.st, cur0, s);
      if (!fold) {
        k_res_norm<double><<<G, kBlock, 0, s>>>(w.lv[0], cur0, w.part, w.st);
        k_gamg_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
      }
    }
    ++ncyc;
  });
  if (rc != PFM_OK) return rc;
  if (fold && ncyc > 0) {  // the last enqueued cycle's residual (no-op once stopped)
    k_res_norm<double><<<G, kBlock, 0, s>>>(w.lv[0], cur0, w.part, w.st);
    k_gamg_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
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

int64_t pfm_gamg_mixed_workspace_bytes(pfm_gamg g, int32_t max_outer, int32_t fp32_levels) {
  if (!g || max_outer < 0 || fp32_levels < 0) return -1;
  return (int64_t)gamg_mixed_layout(g, max_outer, fp32_levels, nullptr, nullptr);
}

int pfm_gamg_mixed_solve(pfm_gamg g, const pfm_solve_params *prm, double inner_rel_tol,
                         int32_t max_outer, int32_t fp32_levels, const double *diag,
                         const double *lower,
                         const double *upper, const double *b, double *x, void *work,
                         double *history_h, pfm_solve_result *res, int32_t *inner_cycles,
                         void *stream) {
  PFM_REQUIRE(g && prm && res && work, "null argument");
  PFM_REQUIRE(prm->max_iter >= 0 && max_outer >= 0 && fp32_levels >= 0,
              "negative iteration or level count");
  PFM_REQUIRE(inner_rel_tol > 0.0 && inner_rel_tol < 1.0, "inner_rel_tol must be in (0, 1)");
  const int n0 = g->lv[0].n;
  PFM_REQUIRE(n0 == 0 || (diag && b && x), "null device pointer");
  memset(res, 0, sizeof(*res));
  if (inner_cycles) *inner_cycles = 0;
  if (n0 == 0) {
    res->status = PFM_SOLVE_MAXITER;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  int st = pfm_gamg_setup_values(g, diag, lower, upper, work, stream);
  if (st != PFM_OK) return st;
  GamgMixedWork m;
  gamg_mixed_layout(g, max_outer, fp32_levels, &m, (char *)work);
  GamgWork &w = m.w64;
  const int L = (int)g->lv.size();
  const int tail_k = tail_level(g);
  const int nbig = fp32_count(g, fp32_levels);
  w.lv[0].diag = diag;
  w.lv[0].x[0] = x;
  w.lv[0].b = const_cast<double *>(b);
  // FP32 values of the large levels: coefficients packed from the FP64
  // (coarsened in the reference order) ones, diagonal and 1/|diag| cast
  for (int k = 0; k < nbig; ++k) {
    const auto &H = g->lv[k];
    const double *lo = k == 0 ? lower : w.lower[k], *up = k == 0 ? upper : w.upper[k];
    const double *d64 = k == 0 ? diag : w.diag[k];
    if (H.nf && (st = pfm_ldu_pack_f32(H.ldu, 0, lo, up, m.packed32[k], stream)) != PFM_OK)
      return st;
    k_to_f32<<<grid1d(H.n), kBlock, 0, s>>>(H.n, d64, m.diag32[k]);
    k_to_f32<<<grid1d(H.n), kBlock, 0, s>>>(H.n, w.lv[k].inv, m.inv32[k]);
    mark_finite<float>(m.lv32[k].finite, H.n, m.diag32[k], H.nf ? H.ldu->nslots : 0,
                       m.packed32[k], 0, (const float *)nullptr, s);
  }
  // FP64 descriptors; with no FP32 level the inner cycle starts on the FP64
  // level 0, whose solution and right-hand side are the inner buffers
  std::vector<LevelDev> small = w.lv;
  small[0].x[0] = m.e64;
  small[0].b = m.b64;
  PFM_CHECK_HIP(hipMemcpyAsync(m.d_small, small.data(), sizeof(LevelDev) * L,
                               hipMemcpyHostToDevice, s));
  CycleLevels<float> C;
  C.tail_k = tail_k;
  C.nbig = nbig;
  C.big = m.lv32;
  C.small = small;
  C.d_small = m.d_small;
  const int hist_cap = max_outer + 1;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  PFM_CHECK_HIP(hipGetLastError());
  const int G = red_grid(g->fine);
  int inner_total = 0;
  std::vector<int> cur_after;
  int cur0_ = 0;  // level 0's buffer parity in the inner cycles
  for (int outer = 0;; ++outer) {
    k_gm_resid<<<G, kBlock, 0, s>>>(w.lv[0], m.r64, w.part, w.st);
    const int budget = prm->max_iter - inner_total;
    k_gm_outer_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist, outer, max_outer, budget,
                                        inner_rel_tol, m.in);
    int32_t done = 0;
    PFM_CHECK_HIP(hipMemcpyAsync(&done, &w.st->done, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PFM_CHECK_HIP(hipStreamSynchronize(s));
    if (done) break;
    // As in pfm_gamg_solve: with a launched level 0 and Jacobi smoothing, the
    // first inner cycle's first level-0 sweep (from e = 0) is done where r is
    // cast, and inner cycle i's residual is formed by cycle i+1's first
    // level-0 sweep; the last one after the loop.
    const bool fold = C.tail_k > 0 && g->prm.smoother != PFM_SMOOTHER_GAUSS_SEIDEL &&
                      g->prm.n_pre >= 1;
    auto res_norm = [&](bool sweep) {
      if (nbig > 0 && sweep)
        k_res_norm_jacobi<float><<<G, kBlock, 0, s>>>(m.lv32[0], cur0_, w.part, m.in);
      else if (nbig > 0)
        k_res_norm<float><<<G, kBlock, 0, s>>>(m.lv32[0], cur0_, w.part, m.in);
      else if (sweep)
        k_res_norm_jacobi<double><<<G, kBlock, 0, s>>>(small[0], cur0_, w.part, m.in);
      else
        k_res_norm<double><<<G, kBlock, 0, s>>>(small[0], cur0_, w.part, m.in);
      k_gamg_fin<<<1, kBlock, 0, s>>>(w.part, G, m.in, nullptr);
    };
    if (fold && nbig > 0)
      k_gm_cast_sweep<float><<<grid1d(n0), kBlock, 0, s>>>(m.lv32[0], m.r64, w.st);
    else if (fold)
      k_gm_cast_sweep<double><<<grid1d(n0), kBlock, 0, s>>>(small[0], m.r64, w.st);
    else if (nbig > 0)
      k_gm_cast<float><<<grid1d(n0), kBlock, 0, s>>>(n0, m.r64, m.lv32[0].b, m.lv32[0].x[0], w.st);
    else
      k_gm_cast<double><<<grid1d(n0), kBlock, 0, s>>>(n0, m.r64, m.b64, m.e64, w.st);
    cur0_ = fold ? 1 : 0;
    int ncyc = 0;
    cur_after.clear();
    int rc = drive_cycles(g, m.in, budget, s, cur_after, cur0_, [&]() {
      if (fold && ncyc > 0) {
        res_norm(true);
        cur0_ ^= 1;
      }
      enqueue_cycle<float>(g, C, m.in, cur0_, s, fold);
      if (!fold) res_norm(false);
      ++ncyc;
    });
    if (rc != PFM_OK) return rc;
    if (fold && ncyc > 0) res_norm(false);  // the last cycle's residual (no-op once stopped)
    int32_t executed = 0;
    PFM_CHECK_HIP(hipMemcpyAsync(&executed, &m.in->iter, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PFM_CHECK_HIP(hipStreamSynchronize(s));
    const int par = executed > 0 ? cur_after[executed - 1] : 0;
    if (nbig > 0)
      k_gm_correct<float><<<grid1d(n0), kBlock, 0, s>>>(n0, m.lv32[0].x[par], x, w.st);
    else
      k_gm_correct<double><<<grid1d(n0), kBlock, 0, s>>>(n0, par ? small[0].x[1] : m.e64, x,
                                                        w.st);
    inner_total += executed;
  }
  KState hs;
  PFM_CHECK_HIP(hipMemcpyAsync(&hs, w.st, sizeof(KState), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  const int nh = std::min(hs.nhist, hist_cap);
  if (history_h && nh > 0)
    PFM_CHECK_HIP(hipMemcpyAsync(history_h, w.hist, sizeof(double) * nh, hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  res->iterations = hs.nhist;
  res->status = hs.status == 0 ? PFM_SOLVE_MAXITER : hs.status;
  res->converged = res->status == PFM_SOLVE_CONVERGED;
  res->initial_residual = hs.nhist > 0 ? hs.res0 : 0.0;
  res->final_residual = hs.nhist > 0 ? hs.res : 0.0;
  if (inner_cycles) *inner_cycles = inner_total;
  return PFM_OK;
}

namespace {
struct GamgPcgWork {
  GamgMixedWork m;
  double *r, *p, *Ap, *z, *part;
};

size_t gamg_pcg_layout(const pfm_gamg_s *g, int max_iter, int fp32_levels, GamgPcgWork *w,
                       char *base) {
  GamgMixedWork m;
  size_t off = gamg_mixed_layout(g, max_iter, fp32_levels, &m, base);
  auto take = [&](size_t bytes) -> void * {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  GamgPcgWork t;
  const size_t n0 = (size_t)g->lv[0].n;
  t.r = (double *)take(sizeof(double) * n0);
  t.p = (double *)take(sizeof(double) * n0);
  t.Ap = (double *)take(sizeof(double) * n0);
  t.z = (double *)take(sizeof(double) * n0);
  t.part = (double *)take(sizeof(double) * 3 * (size_t)red_grid(g->fine));
  if (w) {
    t.m = m;
    *w = t;
  }
  return off;
}
}  // namespace

int64_t pfm_gamg_pcg_workspace_bytes(pfm_gamg g, int32_t max_iter, int32_t fp32_levels) {
  if (!g || max_iter < 0 || fp32_levels < 0) return -1;
  return (int64_t)gamg_pcg_layout(g, max_iter, fp32_levels, nullptr, nullptr);
}

int pfm_gamg_pcg_solve(pfm_gamg g, const pfm_solve_params *prm, int32_t fp32_levels,
                       const double *diag, const double *lower, const double *upper,
                       const double *b, double *x, void *work, double *history_h,
                       pfm_solve_result *res, void *stream) {
  PFM_REQUIRE(g && prm && res && work, "null argument");
  PFM_REQUIRE(prm->max_iter >= 0 && fp32_levels >= 0, "negative iteration or level count");
  const int n0 = g->lv[0].n;
  PFM_REQUIRE(n0 == 0 || (diag && b && x), "null device pointer");
  memset(res, 0, sizeof(*res));
  if (n0 == 0) {
    res->status = PFM_SOLVE_CONVERGED;
    res->converged = 1;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  int st = pfm_gamg_setup_values(g, diag, lower, upper, work, stream);
  if (st != PFM_OK) return st;
  GamgPcgWork W;
  gamg_pcg_layout(g, prm->max_iter, fp32_levels, &W, (char *)work);
  GamgMixedWork &m = W.m;
  GamgWork &w = m.w64;
  const int L = (int)g->lv.size();
  const int tail_k = tail_level(g);
  const int nbig = fp32_count(g, fp32_levels);
  w.lv[0].diag = diag;
  w.lv[0].x[0] = x;
  w.lv[0].b = const_cast<double *>(b);
  for (int k = 0; k < nbig; ++k) {  // FP32 copies of the finest levels
    const auto &H = g->lv[k];
    const double *lo = k == 0 ? lower : w.lower[k], *up = k == 0 ? upper : w.upper[k];
    const double *d64 = k == 0 ? diag : w.diag[k];
    if (H.nf && (st = pfm_ldu_pack_f32(H.ldu, 0, lo, up, m.packed32[k], stream)) != PFM_OK)
      return st;
    k_to_f32<<<grid1d(H.n), kBlock, 0, s>>>(H.n, d64, m.diag32[k]);
    k_to_f32<<<grid1d(H.n), kBlock, 0, s>>>(H.n, w.lv[k].inv, m.inv32[k]);
    mark_finite<float>(m.lv32[k].finite, H.n, m.diag32[k], H.nf ? H.ldu->nslots : 0,
                       m.packed32[k], 0, (const float *)nullptr, s);
  }
  // the preconditioner's level 0 solves for e64 (FP64 path) with rhs b64
  std::vector<LevelDev> small = w.lv;
  small[0].x[0] = m.e64;
  small[0].b = m.b64;
  PFM_CHECK_HIP(hipMemcpyAsync(m.d_small, small.data(), sizeof(LevelDev) * L,
                               hipMemcpyHostToDevice, s));
  CycleLevels<float> C;
  C.tail_k = tail_k;
  C.nbig = nbig;
  C.big = m.lv32;
  C.small = small;
  C.d_small = m.d_small;
  const int hist_cap = prm->max_iter + 1;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  const int G = red_grid(g->fine);
  const int ns0 = g->fine->nslices;
  // z = M^-1 r: level-0 rhs = r, solution from zero, one cycle
  int zpar = 0;
  // Level 0 starts from zero: with launched levels (tail_k > 0) and Jacobi
  // smoothing its first sweep is done where r is cast (k_gm_cast_sweep).
  const bool fold0 = tail_k > 0 && g->prm.smoother != PFM_SMOOTHER_GAUSS_SEIDEL &&
                     g->prm.n_pre >= 1;
  auto precondition = [&](int it) {
    if (fold0 && nbig > 0)
      k_gm_cast_sweep<float><<<grid1d(n0), kBlock, 0, s>>>(m.lv32[0], W.r, w.st);
    else if (fold0)
      k_gm_cast_sweep<double><<<grid1d(n0), kBlock, 0, s>>>(small[0], W.r, w.st);
    else if (nbig > 0)
      k_gm_cast<float><<<grid1d(n0), kBlock, 0, s>>>(n0, W.r, m.lv32[0].b, m.lv32[0].x[0], w.st);
    else
      k_gm_cast<double><<<grid1d(n0), kBlock, 0, s>>>(n0, W.r, m.b64, m.e64, w.st);
    int cur0 = fold0 ? 1 : 0;
    enqueue_cycle<float>(g, C, w.st, cur0, s, fold0);
    zpar = cur0;
    if (nbig > 0)
      k_gp_z<float><<<G, kBlock, 0, s>>>(n0, ns0, m.lv32[0].x[zpar], W.r, W.z, W.part, w.st);
    else
      k_gp_z<double><<<G, kBlock, 0, s>>>(n0, ns0, zpar ? small[0].x[1] : m.e64, W.r, W.z, W.part,
                                          w.st);
    k_gp_pupdate<<<G, kBlock, 0, s>>>(n0, ns0, it, W.z, W.p, W.part, w.st);
  };
  k_gm_resid<<<G, kBlock, 0, s>>>(w.lv[0], W.r, W.part, w.st);
  k_gp_init_fin<<<1, kBlock, 0, s>>>(W.part, G, w.st, w.hist);
  precondition(0);
  PFM_CHECK_HIP(hipGetLastError());
  std::vector<int> unused;
  int cur_unused = 0, it = 0;
  int rc = drive_cycles(g, w.st, prm->max_iter, s, unused, cur_unused, [&]() {
    ++it;
    k_gp_spmv<<<G, kBlock, 0, s>>>(w.lv[0], W.p, W.Ap, W.part, w.st);
    k_gp_update<<<G, kBlock, 0, s>>>(n0, ns0, it, W.p, W.Ap, x, W.r, W.part, w.st);
    k_gp_decide<<<1, kBlock, 0, s>>>(W.part, G, it, w.st, w.hist);
    precondition(it);
  });
  if (rc != PFM_OK) return rc;
  KState hs;
  PFM_CHECK_HIP(hipMemcpyAsync(&hs, w.st, sizeof(KState), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
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
