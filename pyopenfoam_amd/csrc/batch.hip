// Batched PCG-DIC: many independent LDU systems of uneven size solved by one
// sequence of kernels, replayed as a hipGraph.
//
// The systems are concatenated block-diagonally; each one is padded to whole
// row slices (dummy rows: diag 1, b 0, x 0, no faces, so they contribute 0 to
// every dot product), so every slice belongs to one system.  Work is cut into
// tiles of `tile_slices` slices of one system; one workgroup per tile.  Each
// system keeps its own scalars and stopping state (the reference's monitor,
// residual.py:112-145, per system), reduced in a fixed order over its tiles,
// so every system's solve is identical to solving it alone except for the
// summation order of its dot products.  Systems that have stopped skip their
// tiles.  Iterations are captured into a hipGraph in chunks; the host replays
// the graph and polls a device counter of finished systems.  The tile size is
// picked by an autotuner that times whole iterations on the actual batch.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "solver_state.h"

using namespace pfm;

struct BSys {
  double res0, res;
  double rz[2];
  double alpha[2];  // alpha by iteration parity (k_b_update -> k_b_pupdate)
  int32_t nhist, done, status;
  int32_t last_it;  // the last iteration the system runs (INT32_MAX until it stops)
};

struct alignas(256) BGlobal {
  double tol, rel_tol;
  int32_t it;        // iterations scheduled so far (chunks started)
  int32_t base;      // iterations before the current chunk
  int32_t ndone;     // systems finished
  int32_t max_iter, min_iter;
  int32_t nact;      // entries of the active-tile list
};

struct pfm_batch_s {
  int32_t ns = 0, n = 0, nf = 0, np = 0;
  std::vector<int32_t> cell_off, face_off, pad_off;
  pfm_ldu ldu = nullptr;  // padded concatenation
  int32_t ntiles = 0, tile_slices = 0;
  int32_t *d_tile_sys = nullptr, *d_tile_s0 = nullptr, *d_tile_s1 = nullptr, *d_sys_t0 = nullptr;
  int32_t *d_cell_map = nullptr;  // unpadded cell -> padded row
  pfm::PollPool poll;  // pinned stop-flag slots of the solves
  float tuned_us[8] = {0};
};

namespace {

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
int upload(T **dst, const std::vector<T> &src) {
  *dst = nullptr;
  if (src.empty()) return PFM_OK;
  PFM_CHECK_HIP(hipMalloc((void **)dst, src.size() * sizeof(T)));
  PFM_CHECK_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return PFM_OK;
}

// Tile kernels run 256-thread workgroups; a tile's slices are shared among
// the waves.
constexpr int kMaxWaves = kBlock / kSlice;

#define TILE_ROWS(...)                                                          \
  {                                                                             \
    const int lane = threadIdx.x & 63;                                          \
    const int nw = blockDim.x >> 6;                                             \
    for (int s = s0 + (threadIdx.x >> 6); s < s1; s += nw) {                    \
      const int row = s * kSlice + lane;                                        \
      __VA_ARGS__                                                               \
    }                                                                           \
  }

// Deterministic workgroup sum for any wave count (waves added in order).
__device__ __forceinline__ double wg_sum(double v, double *lds /*[kMaxWaves]*/) {
  v = wave_sum(v);
  const int nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = lds[0];
  for (int i = 1; i < nw; ++i) s += lds[i];
  __syncthreads();
  return s;
}

// Sum of this system's tile partials (fixed tile order), on every thread.
__device__ __forceinline__ double sys_sum(const double *__restrict__ part, int t0, int t1,
                                          double *lds) {
  double s = 0.0;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += (int)blockDim.x) s += part[t];
  return wg_sum(s, lds);
}

__global__ __launch_bounds__(kBlock) void k_b_fill(int np, double *__restrict__ d,
                                                    double *__restrict__ x, double *__restrict__ b) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= np) return;
  d[i] = 1.0;
  x[i] = 0.0;
  b[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_b_scatter(int n, const int32_t *__restrict__ map,
                                                       const double *__restrict__ d,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ b,
                                                       double *__restrict__ dp,
                                                       double *__restrict__ xp,
                                                       double *__restrict__ bp) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int j = map[i];
  dp[j] = d[i];
  xp[j] = x[i];
  bp[j] = b[i];
}

__global__ __launch_bounds__(kBlock) void k_b_gather(int n, const int32_t *__restrict__ map,
                                                      const double *__restrict__ xp,
                                                      double *__restrict__ x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = xp[map[i]];
}

__global__ void k_b_state_init(BGlobal *g, BSys *sys, int ns, double tol, double rel_tol,
                               int max_iter, int min_iter) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    g->tol = tol;
    g->rel_tol = rel_tol;
    g->it = 0;
    g->base = 0;
    g->nact = 0;
    g->ndone = 0;
    g->max_iter = max_iter;
    g->min_iter = min_iter;
  }
  if (k < ns) {
    BSys z;
    memset(&z, 0, sizeof(z));
    z.last_it = 0x7fffffff;
    sys[k] = z;
  }
}

// The fields of a system's state that iteration `it`'s direction update
// reads, loaded one by one (a copy of the whole struct, indexed by parity,
// would live in private memory).
struct SysView {
  double res0, alpha, rz_prev;
  int32_t nhist, done, last_it;
};
__device__ __forceinline__ SysView sys_view(const BSys &s, int it) {
  SysView v;
  v.res0 = s.res0;
  v.alpha = (it & 1) ? s.alpha[1] : s.alpha[0];
  v.rz_prev = (it & 1) ? s.rz[0] : s.rz[1];
  v.nhist = s.nhist;
  v.done = s.done;
  v.last_it = s.last_it;
  return v;
}

__device__ __forceinline__ bool sys_decision(const BGlobal *g, double res0, double res, int it) {
  if (it >= g->min_iter) {
    if (res < g->tol) return true;
    if (res0 > 0.0 && res / res0 < g->rel_tol) return true;
  }
  return false;
}

// r = b - A x; z = r/rD; p = z; tile partials r.r, r.z.
__global__ __launch_bounds__(kBlock) void k_b_init(const int32_t *__restrict__ tsys,
                                                    const int32_t *__restrict__ ts0,
                                                    const int32_t *__restrict__ ts1, int ntiles,
                                                    const SliceInfo *__restrict__ sl,
                                                    const int32_t *__restrict__ col,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ diag,
                                                    const double *__restrict__ rD,
                                                    const double *__restrict__ b,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ r, double *__restrict__ p,
                                                    double *__restrict__ part) {
  __shared__ double lds[kMaxWaves];
  const int t = blockIdx.x;
  const int s0 = ts0[t], s1 = ts1[t];
  double rr = 0.0, rz = 0.0;
  TILE_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double ax = row_spmv<double>(si, lane, diag[row] * x[row], col, val, x);
    const double ri = b[row] - ax;
    const double zi = rD ? ri / rD[row] : ri;
    r[row] = ri;
    p[row] = zi;
    rr += ri * ri;
    rz += ri * zi;
  })
  rr = wg_sum(rr, lds);
  rz = wg_sum(rz, lds);
  if (threadIdx.x == 0) {
    part[t] = rr;
    part[ntiles + t] = rz;
  }
  (void)tsys;
}

// One thread per system: initial residual, monitor(0), rz.
// One workgroup per system: the tile partials are added in the order of the
// iteration kernels (sys_sum), so every dot product of a system's solve has
// one summation order.
__global__ __launch_bounds__(kBlock) void k_b_init_fin(int ns, const int32_t *__restrict__ sys_t0,
                                                        int ntiles,
                                                        const double *__restrict__ part,
                                                        BGlobal *g, BSys *sys) {
  __shared__ double lds[kMaxWaves];
  const int k = blockIdx.x;
  if (k >= ns) return;
  const double RR = sys_sum(part, sys_t0[k], sys_t0[k + 1], lds);
  const double RZ = sys_sum(part + ntiles, sys_t0[k], sys_t0[k + 1], lds);
  if (threadIdx.x != 0) return;
  BSys S = sys[k];
  const double res = sqrt(RR);
  S.res0 = res;
  S.res = res;
  S.nhist = 1;
  S.rz[0] = RZ;
  if (sys_decision(g, S.res0, res, 0)) {
    S.done = 1;
    S.status = PFM_SOLVE_CONVERGED;
    S.last_it = 0;
  } else if (g->max_iter < 1) {
    S.done = 1;
    S.status = PFM_SOLVE_MAXITER;
    S.last_it = 0;
  }
  sys[k] = S;
  if (S.done) atomicAdd(&g->ndone, 1);
}

// Active-tile list: descriptors {tile, system, first slice, end slice} of
// the tiles of every system not yet stopped, in system order.  Rebuilt by one
// workgroup at the start of each chunk of iterations; tiles of systems that
// stop inside a chunk skip themselves.  It also fixes the chunk's first
// iteration index: the kernels of the chunk's j-th iteration (a launch
// argument, so a captured chunk replays unchanged) run iteration base + j + 1.
__global__ __launch_bounds__(1024) void k_b_compact(int ns, const int32_t *__restrict__ sys_t0,
                                                     const int32_t *__restrict__ ts0,
                                                     const int32_t *__restrict__ ts1,
                                                     const BSys *__restrict__ sys,
                                                     int4 *__restrict__ act, BGlobal *g,
                                                     int chunk) {
  __shared__ int32_t scan[1024];
  const int tid = threadIdx.x;
  int base = 0;
  for (int k0 = 0; k0 < ns; k0 += 1024) {
    const int k = k0 + tid;
    const int cnt = (k < ns && !sys[k].done) ? sys_t0[k + 1] - sys_t0[k] : 0;
    scan[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    const int start = base + scan[tid] - cnt;
    for (int j = 0; j < cnt; ++j) {
      const int t = sys_t0[k] + j;
      act[start + j] = make_int4(t, k, ts0[t], ts1[t]);
    }
    base += scan[1023];
    __syncthreads();
  }
  if (tid == 0) {
    g->nact = base;
    g->base = g->it;  // iterations of this chunk are base+1 .. base+chunk
    g->it += chunk;
  }
}

// Workgroup b runs on XCD b mod 8.  With a grid that is a multiple of 8, each
// XCD walks its own contiguous eighth of the active list, so the tiles of a
// system (whose rows read each other's vector entries) share one XCD's L2.
struct XcdRange {
  int begin, end, step;
};

__device__ __forceinline__ XcdRange xcd_range(int nact) {
  const int G = gridDim.x, b = blockIdx.x;
  XcdRange r;
  if (G % kXcds) {
    r.begin = b;
    r.end = nact;
    r.step = G;
    return r;
  }
  const int x = b % kXcds;
  r.begin = (int)((int64_t)nact * x / kXcds) + b / kXcds;
  r.end = (int)((int64_t)nact * (x + 1) / kXcds);
  r.step = G / kXcds;
  return r;
}

// The iteration kernels run a fixed grid that walks the active-tile list, so
// the cost of an iteration follows the work still left in the batch.  A
// system's stop flag can be set by another workgroup while this one reads it,
// so thread 0 reads it once per tile and the workgroup takes that value: every
// wave then runs the same sequence of tiles and barriers.  (Either value is
// correct: a tile that sees the flag just set would have taken the same stop
// decision itself.)
#define FOR_ACTIVE_TILES(CHECK_DONE, ...)                                       \
  {                                                                             \
    __shared__ int s_done;                                                      \
    const XcdRange rg = xcd_range(g->nact);                                     \
    for (int i = rg.begin; i < rg.end; i += rg.step) {                          \
      const int4 dsc = act[i];                                                  \
      const int t = dsc.x, k = dsc.y, s0 = dsc.z, s1 = dsc.w;                   \
      if (CHECK_DONE) {                                                         \
        if (threadIdx.x == 0) s_done = sys[k].done;                             \
        __syncthreads();                                                        \
        const int done_k = s_done;                                              \
        __syncthreads();                                                        \
        if (done_k) continue;                                                   \
      }                                                                         \
      __VA_ARGS__                                                               \
    }                                                                           \
  }

// Latency: in the tail of a batch only a few tiles remain active and each
// wave's slices are the whole of its work, so the update kernels' row loops
// keep two slices' loads in flight (the next slice's operands are loaded
// before the current slice's stores; the tail-mode SpMV forms two slices side
// by side), and a tile's first operands are loaded before its stop flag and
// partial sums are read.  Per-row arithmetic and summation order are those of
// one slice at a time.

// The vector streams of the 256-thread update kernels (r, Ap, rD, p, x: 150 MB
// each for the 512-system batch) are read nontemporally, which keeps the
// gathered p and the small per-system data cached (3.58 -> 3.41 s).
__device__ __forceinline__ double ldnt(const double *a) { return __builtin_nontemporal_load(a); }

// Ap = A p; tile partial p.Ap.  Tiles of a system that stopped earlier in the
// chunk compute values nobody reads, so this kernel skips the stop check.
template <bool kPair>
__global__ __launch_bounds__(kBlock) void k_b_spmv(const int4 *__restrict__ act,
                                                    const SliceInfo *__restrict__ sl,
                                                    const int32_t *__restrict__ col,
                                                    const uint32_t *__restrict__ cp,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ diag,
                                                    const double *__restrict__ p,
                                                    double *__restrict__ Ap,
                                                    double *__restrict__ part,
                                                    const BGlobal *__restrict__ g,
                                                    const BSys *sys) {
  __shared__ double lds[kMaxWaves];
  // The batch's coefficients (1.3 GB for the 512-system batch) and diagonal
  // are read once per iteration: nontemporal loads keep the vectors cached
  // (3.73 -> 3.64 s; the diagonal's, 611 -> 598 us per full-load iteration).
  auto gx = [p](int c) { return p[c]; };
  FOR_ACTIVE_TILES(false, {
    (void)k;
    const int lane = threadIdx.x & 63;
    const int nw = blockDim.x >> 6;
    double pap = 0.0;
    // one slice per wave step: each lane adds its rows' products in slice
    // order s0 + w, s0 + w + nw, ...  (forming two slices side by side, as the
    // tail-mode kernel does, was 3.30 against 3.24 s for the 512-system batch)
    for (int s = s0 + (threadIdx.x >> 6); s < s1; s += nw) {
      const SliceInfo sa_ = load_slice(sl, s);
      const int ra = s * kSlice + lane;
      const double pa = p[ra];
      double va;
      if constexpr (kPair)  // paired int16 column offsets
        va = row_spmv_pair<double, decltype(gx), false, true>(sa_, s, lane, ra, ldnt(diag + ra) * pa,
                                                              cp, val, gx);
      else
        va = row_spmv_g<double, decltype(gx), false, ColAbsT<true>, true>(
            sa_, lane, ldnt(diag + ra) * pa, ColAbsT<true>{col}, val, gx);
      Ap[ra] = va;
      pap += pa * va;
    }
    pap = wg_sum(pap, lds);
    if (threadIdx.x == 0) part[t] = pap;
  })
}


// alpha per system (kept for k_b_pupdate); r -= alpha Ap; tile partials r.r,
// r.z.  A breakdown ends the system before this iteration's x update.
__global__ __launch_bounds__(kBlock) void k_b_update(const int4 *__restrict__ act, int ntiles,
                                                      const int32_t *__restrict__ sys_t0,
                                                      const double *__restrict__ rD,
                                                      const double *__restrict__ Ap,
                                                      double *__restrict__ r,
                                                      double *__restrict__ part,
                                                      BGlobal *g, BSys *sys, int j) {
  __shared__ double lds[kMaxWaves];
  __shared__ int s_done;
  const int it = g->base + j + 1;
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  const XcdRange rg = xcd_range(g->nact);
  for (int i = rg.begin; i < rg.end; i += rg.step) {
    const int4 dsc = act[i];
    const int t = dsc.x, k = dsc.y, s0 = dsc.z, s1 = dsc.w;
    int s = s0 + (threadIdx.x >> 6);
    int row = s * kSlice + lane;
    double rv = 0.0, av = 0.0, dv = 1.0;
    if (s < s1) {
      rv = ldnt(r + row);
      av = ldnt(Ap + row);
      if (rD) dv = ldnt(rD + row);
    }
    // The tile range and r.z of the previous iteration depend only on k and
    // are not written by this kernel: load them with the stop flag.
    const int t0 = sys_t0[k], t1 = sys_t0[k + 1];
    const double rz_old = sys[k].rz[(it - 1) & 1];
    if (threadIdx.x == 0) s_done = sys[k].done;  // see FOR_ACTIVE_TILES
    __syncthreads();
    const int done_k = s_done;
    __syncthreads();
    if (done_k) continue;
    const double PAP = sys_sum(part, t0, t1, lds);
    if (fabs(PAP) < 1e-30) {
      if (t == t0 && threadIdx.x == 0) {
        sys[k].done = 1;
        sys[k].status = PFM_SOLVE_BREAKDOWN;
        sys[k].last_it = it - 1;
        atomicAdd((int *)&g->ndone, 1);
      }
      continue;
    }
    const double alpha = rz_old / PAP;
    if (t == t0 && threadIdx.x == 0) sys[k].alpha[it & 1] = alpha;
    double rr = 0.0, rz = 0.0;
    for (; s < s1; s += nw) {
      const int sn = s + nw, rn = sn * kSlice + lane;
      double rnx = 0.0, an = 0.0, dn = 1.0;
      if (sn < s1) {
        rnx = ldnt(r + rn);
        an = ldnt(Ap + rn);
        if (rD) dn = ldnt(rD + rn);
      }
      const double ri = rv - alpha * av;
      r[row] = ri;
      const double zi = rD ? ri / dv : ri;
      rr += ri * ri;
      rz += ri * zi;
      row = rn;
      rv = rnx;
      av = an;
      dv = dn;
    }
    rr = wg_sum(rr, lds);
    rz = wg_sum(rz, lds);
    if (threadIdx.x == 0) {
      part[ntiles + t] = rr;
      part[2 * ntiles + t] = rz;
    }
  }
}

// monitor per system; x += alpha p (where p is read anyway); beta; p = z +
// beta p (not after the system's last iteration).
__global__ __launch_bounds__(kBlock) void k_b_pupdate(const int4 *__restrict__ act, int ntiles,
                                                       const int32_t *__restrict__ sys_t0,
                                                       const double *__restrict__ rD,
                                                       const double *__restrict__ r,
                                                       double *__restrict__ p,
                                                       double *__restrict__ x,
                                                       const double *__restrict__ part,
                                                       BGlobal *g, BSys *sys, int j) {
  __shared__ double lds[kMaxWaves];
  __shared__ int s_skip;
  const int it = g->base + j + 1;
  const int lane = threadIdx.x & 63;
  const int nw = blockDim.x >> 6;
  const XcdRange rg = xcd_range(g->nact);
  for (int i = rg.begin; i < rg.end; i += rg.step) {
    const int4 dsc = act[i];
    const int t = dsc.x, k = dsc.y, s0 = dsc.z, s1 = dsc.w;
    int s = s0 + (threadIdx.x >> 6);
    int row = s * kSlice + lane;
    double rv = 0.0, pv = 0.0, xv = 0.0, dv = 1.0;
    if (s < s1) {
      rv = ldnt(r + row);
      pv = ldnt(p + row);
      xv = ldnt(x + row);
      if (rD) dv = ldnt(rD + row);
    }
    // The system's state is read once, before the partial sums: within this
    // kernel only the system's first tile writes it (W below), and the fields
    // the other tiles use (res0, alpha and r.z of the previous iterations) are
    // not among those it changes.  A tile skips the system only if it stopped
    // in an earlier iteration (done with last_it < it): the first tile's write
    // of done and last_it may be seen in part, and every mix of old and new
    // values keeps a system that stops in this iteration (whose x update every
    // tile must still make).
    const int t0 = sys_t0[k], t1 = sys_t0[k + 1];
    const SysView S = sys_view(sys[k], it);
    if (threadIdx.x == 0) s_skip = S.done && S.last_it < it;
    __syncthreads();
    const int skip = s_skip;
    __syncthreads();
    if (skip) continue;
    const double RR = sys_sum(part + ntiles, t0, t1, lds);
    const double RZ = sys_sum(part + 2 * ntiles, t0, t1, lds);
    const double res = sqrt(RR);
    const bool conv = sys_decision(g, S.res0, res, it);
    const bool stop = conv || it >= g->max_iter;
    if (t == t0 && threadIdx.x == 0) {  // field by field: no private copy of the state
      BSys &W = sys[k];
      W.nhist = S.nhist + 1;
      W.res = res;
      if (it & 1)
        W.rz[1] = RZ;
      else
        W.rz[0] = RZ;
      if (stop) {
        W.done = 1;
        W.status = conv ? PFM_SOLVE_CONVERGED : PFM_SOLVE_MAXITER;
        W.last_it = it;
      }
      if (stop) atomicAdd((int *)&g->ndone, 1);
    }
    const double alpha = S.alpha;
    const double beta = RZ / S.rz_prev;
    for (; s < s1; s += nw) {
      const int sn = s + nw, rn = sn * kSlice + lane;
      double rnx = 0.0, pn = 0.0, xn = 0.0, dn = 1.0;
      if (sn < s1) {
        rnx = ldnt(r + rn);
        pn = ldnt(p + rn);
        xn = ldnt(x + rn);
        if (rD) dn = ldnt(rD + rn);
      }
      x[row] = xv + alpha * pv;
      if (!stop) {
        const double zi = rD ? rv / dv : rv;
        p[row] = zi + beta * pv;
      }
      row = rn;
      rv = rnx;
      pv = pn;
      xv = xn;
      dv = dn;
    }
  }
}

// ---------------------------------------------------------- tail mode -----
// Once few systems remain, an iteration is a chain of three launches whose
// cost is each tile's walk over its slices: four waves take a 32-slice tile
// in eight steps.  In tail mode a tile is run by a 1024-thread workgroup
// whose 16 waves hold every slice of the tile at once (at most two per wave).
// On the 512-system batch (n ~ U{8..48}) switching at 32 systems left takes
// the solve from 3.87 s to 3.73 s; at 8 left, 3.75 s; from the start, 3.71 s
// (the autotuned 256-thread kernels are kept while the batch is full).
// The tile partials keep the summation order of the 256-thread kernels: each
// row's product goes to LDS, then waves 0-3 add, lane by lane, the products
// of slices v, v + 4, v + 8, ... in that order (the running sums of the
// 256-thread kernels' four waves) and the four wave sums are added as wg_sum
// adds them; a system's sum over its tiles is formed by threads 0-255 as
// sys_sum forms it with 256 threads.  Results are bit-identical to the
// 256-thread kernels'.
constexpr int kTailThreads = 1024;
constexpr int kTailWaves = kTailThreads / kSlice;  // 16
constexpr int kTailTileSlices = 2 * kTailWaves;    // tiles of at most this many slices
constexpr int kTailGrid = 512;                     // a multiple of the XCD count
constexpr int kTailSystems = 32;                   // tail mode once at most this many remain

// wg_sum of a 256-thread workgroup, on a 1024-thread one: v of waves 0-3.
__device__ __forceinline__ double wg_sum_first4(double v, double *lds /*[kMaxWaves]*/) {
  const int w = threadIdx.x >> 6;
  if (w < kMaxWaves) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) lds[w] = v;
  }
  __syncthreads();
  double s = lds[0];
  for (int i = 1; i < kMaxWaves; ++i) s += lds[i];
  __syncthreads();
  return s;
}

// sys_sum as 256 threads form it.
__device__ __forceinline__ double sys_sum_first4(const double *__restrict__ part, int t0, int t1,
                                                 double *lds) {
  double s = 0.0;
  if (threadIdx.x < kBlock)
    for (int t = t0 + (int)threadIdx.x; t < t1; t += kBlock) s += part[t];
  return wg_sum_first4(s, lds);
}

// The tile partial of the 256-thread kernels from the products prod[q * 64 +
// lane] of the tile's slices q = 0..nsl-1.  Called by every thread after the
// products are in LDS and a barrier.
__device__ __forceinline__ double ordered_tile_sum(const double *prod, int nsl, double *lds) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double acc = 0.0;
  if (w < kMaxWaves)
    for (int q = w; q < nsl; q += kMaxWaves) acc += prod[q * kSlice + lane];
  return wg_sum_first4(acc, lds);
}

// k_b_spmv in tail mode.
template <bool kPair>
__global__ __launch_bounds__(kTailThreads) void k_bt_spmv(const int4 *__restrict__ act,
                                                           const SliceInfo *__restrict__ sl,
                                                           const int32_t *__restrict__ col,
                                                           const uint32_t *__restrict__ cp,
                                                           const double *__restrict__ val,
                                                           const double *__restrict__ diag,
                                                           const double *__restrict__ p,
                                                           double *__restrict__ Ap,
                                                           double *__restrict__ part,
                                                           const BGlobal *__restrict__ g) {
  __shared__ double prod[kTailTileSlices * kSlice];
  __shared__ double lds[kMaxWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const XcdRange rg = xcd_range(g->nact);
  for (int i = rg.begin; i < rg.end; i += rg.step) {
    const int4 dsc = act[i];
    const int t = dsc.x, s0 = dsc.z, nsl = dsc.w - dsc.z;
    if (w < nsl) {
      const bool two = w + kTailWaves < nsl;
      const int sa = s0 + w, sb = two ? sa + kTailWaves : sa;
      const SliceInfo ia = load_slice(sl, sa), ib = load_slice(sl, sb);
      const int ra = sa * kSlice + lane, rb = sb * kSlice + lane;
      const double pa = p[ra], pb = p[rb];
      double va, vb;
      if constexpr (kPair) {
        auto gx = [p](int c) { return p[c]; };
        va = row_spmv_pair<double, decltype(gx)>(ia, sa, lane, ra, diag[ra] * pa, cp, val, gx);
        vb = row_spmv_pair<double, decltype(gx)>(ib, sb, lane, rb, diag[rb] * pb, cp, val, gx);
      } else {
        va = row_spmv<double>(ia, lane, diag[ra] * pa, col, val, p);
        vb = row_spmv<double>(ib, lane, diag[rb] * pb, col, val, p);
      }
      Ap[ra] = va;
      prod[w * kSlice + lane] = pa * va;
      if (two) {
        Ap[rb] = vb;
        prod[(w + kTailWaves) * kSlice + lane] = pb * vb;
      }
    }
    __syncthreads();
    const double pap = ordered_tile_sum(prod, nsl, lds);
    if (threadIdx.x == 0) part[t] = pap;
  }
}

// k_b_update in tail mode.
__global__ __launch_bounds__(kTailThreads) void k_bt_update(const int4 *__restrict__ act,
                                                             int ntiles,
                                                             const int32_t *__restrict__ sys_t0,
                                                             const double *__restrict__ rD,
                                                             const double *__restrict__ Ap,
                                                             double *__restrict__ r,
                                                             double *__restrict__ part,
                                                             BGlobal *g, BSys *sys, int j) {
  __shared__ double prr[kTailTileSlices * kSlice], prz[kTailTileSlices * kSlice];
  __shared__ double lds[kMaxWaves];
  __shared__ int s_done;
  const int it = g->base + j + 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const XcdRange rg = xcd_range(g->nact);
  for (int i = rg.begin; i < rg.end; i += rg.step) {
    const int4 dsc = act[i];
    const int t = dsc.x, k = dsc.y, s0 = dsc.z, nsl = dsc.w - dsc.z;
    const bool has = w < nsl, two = w + kTailWaves < nsl;
    const int ra = (s0 + w) * kSlice + lane, rb = ra + kTailWaves * kSlice;
    double rva = 0.0, ava = 0.0, dva = 1.0, rvb = 0.0, avb = 0.0, dvb = 1.0;
    if (has) {
      rva = r[ra];
      ava = Ap[ra];
      if (rD) dva = rD[ra];
    }
    if (two) {
      rvb = r[rb];
      avb = Ap[rb];
      if (rD) dvb = rD[rb];
    }
    const int t0 = sys_t0[k], t1 = sys_t0[k + 1];
    const double rz_old = sys[k].rz[(it - 1) & 1];
    if (threadIdx.x == 0) s_done = sys[k].done;  // see FOR_ACTIVE_TILES
    __syncthreads();
    const int done_k = s_done;
    __syncthreads();
    if (done_k) continue;
    const double PAP = sys_sum_first4(part, t0, t1, lds);
    if (fabs(PAP) < 1e-30) {
      if (t == t0 && threadIdx.x == 0) {
        sys[k].done = 1;
        sys[k].status = PFM_SOLVE_BREAKDOWN;
        sys[k].last_it = it - 1;
        atomicAdd((int *)&g->ndone, 1);
      }
      continue;
    }
    const double alpha = rz_old / PAP;
    if (t == t0 && threadIdx.x == 0) sys[k].alpha[it & 1] = alpha;
    if (has) {
      const double ri = rva - alpha * ava;
      r[ra] = ri;
      const double zi = rD ? ri / dva : ri;
      prr[w * kSlice + lane] = ri * ri;
      prz[w * kSlice + lane] = ri * zi;
    }
    if (two) {
      const double ri = rvb - alpha * avb;
      r[rb] = ri;
      const double zi = rD ? ri / dvb : ri;
      prr[(w + kTailWaves) * kSlice + lane] = ri * ri;
      prz[(w + kTailWaves) * kSlice + lane] = ri * zi;
    }
    __syncthreads();
    const double rr = ordered_tile_sum(prr, nsl, lds);
    const double rz = ordered_tile_sum(prz, nsl, lds);
    if (threadIdx.x == 0) {
      part[ntiles + t] = rr;
      part[2 * ntiles + t] = rz;
    }
  }
}

// k_b_pupdate in tail mode.
__global__ __launch_bounds__(kTailThreads) void k_bt_pupdate(const int4 *__restrict__ act,
                                                              int ntiles,
                                                              const int32_t *__restrict__ sys_t0,
                                                              const double *__restrict__ rD,
                                                              const double *__restrict__ r,
                                                              double *__restrict__ p,
                                                              double *__restrict__ x,
                                                              const double *__restrict__ part,
                                                              BGlobal *g, BSys *sys, int j) {
  __shared__ double lds[kMaxWaves];
  __shared__ int s_skip;
  const int it = g->base + j + 1;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const XcdRange rg = xcd_range(g->nact);
  for (int i = rg.begin; i < rg.end; i += rg.step) {
    const int4 dsc = act[i];
    const int t = dsc.x, k = dsc.y, s0 = dsc.z, nsl = dsc.w - dsc.z;
    const bool has = w < nsl, two = w + kTailWaves < nsl;
    const int ra = (s0 + w) * kSlice + lane, rb = ra + kTailWaves * kSlice;
    double rva = 0.0, pva = 0.0, xva = 0.0, dva = 1.0, rvb = 0.0, pvb = 0.0, xvb = 0.0, dvb = 1.0;
    if (has) {
      rva = r[ra];
      pva = p[ra];
      xva = x[ra];
      if (rD) dva = rD[ra];
    }
    if (two) {
      rvb = r[rb];
      pvb = p[rb];
      xvb = x[rb];
      if (rD) dvb = rD[rb];
    }
    // as in k_b_pupdate: the state is read once, before the partial sums
    const int t0 = sys_t0[k], t1 = sys_t0[k + 1];
    const SysView S = sys_view(sys[k], it);
    if (threadIdx.x == 0) s_skip = S.done && S.last_it < it;
    __syncthreads();
    const int skip = s_skip;
    __syncthreads();
    if (skip) continue;
    const double RR = sys_sum_first4(part + ntiles, t0, t1, lds);
    const double RZ = sys_sum_first4(part + 2 * ntiles, t0, t1, lds);
    const double res = sqrt(RR);
    const bool conv = sys_decision(g, S.res0, res, it);
    const bool stop = conv || it >= g->max_iter;
    if (t == t0 && threadIdx.x == 0) {  // field by field: no private copy of the state
      BSys &W = sys[k];
      W.nhist = S.nhist + 1;
      W.res = res;
      if (it & 1)
        W.rz[1] = RZ;
      else
        W.rz[0] = RZ;
      if (stop) {
        W.done = 1;
        W.status = conv ? PFM_SOLVE_CONVERGED : PFM_SOLVE_MAXITER;
        W.last_it = it;
      }
      if (stop) atomicAdd((int *)&g->ndone, 1);
    }
    const double alpha = S.alpha;
    const double beta = RZ / S.rz_prev;
    if (has) {
      x[ra] = xva + alpha * pva;
      if (!stop) p[ra] = (rD ? rva / dva : rva) + beta * pva;
    }
    if (two) {
      x[rb] = xvb + alpha * pvb;
      if (!stop) p[rb] = (rD ? rvb / dvb : rvb) + beta * pvb;
    }
  }
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct BWork {
  BGlobal *g;
  BSys *sys;
  int4 *act;
  double *part;
  double *dp, *xp, *bp, *rD, *r, *p, *Ap, *packed;
};

size_t batch_layout(const pfm_batch_s *B, int ntiles, BWork *w, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  BWork t;
  t.g = (BGlobal *)take(sizeof(BGlobal));
  t.sys = (BSys *)take(sizeof(BSys) * (size_t)B->ns);
  t.act = (int4 *)take(sizeof(int4) * (size_t)std::max(ntiles, 1));
  t.part = (double *)take(sizeof(double) * 3 * (size_t)std::max(ntiles, 1));
  const size_t np = (size_t)B->np;
  t.dp = (double *)take(sizeof(double) * np);
  t.xp = (double *)take(sizeof(double) * np);
  t.bp = (double *)take(sizeof(double) * np);
  t.rD = (double *)take(sizeof(double) * np);
  t.r = (double *)take(sizeof(double) * np);
  t.p = (double *)take(sizeof(double) * np);
  t.Ap = (double *)take(sizeof(double) * np);
  t.packed = (double *)take(sizeof(double) * (size_t)std::max<int64_t>(B->ldu->nslots, 1));
  if (w) *w = t;
  return off;
}

int build_tiles(pfm_batch_s *B, int tile_slices) {
  std::vector<int32_t> tsys, ts0, ts1, sys_t0(B->ns + 1, 0);
  for (int k = 0; k < B->ns; ++k) {
    sys_t0[k] = (int32_t)tsys.size();
    const int s_begin = B->pad_off[k] / kSlice, s_end = B->pad_off[k + 1] / kSlice;
    for (int s = s_begin; s < s_end; s += tile_slices) {
      tsys.push_back(k);
      ts0.push_back(s);
      ts1.push_back(std::min(s + tile_slices, s_end));
    }
  }
  sys_t0[B->ns] = (int32_t)tsys.size();
  (void)hipFree(B->d_tile_sys);
  (void)hipFree(B->d_tile_s0);
  (void)hipFree(B->d_tile_s1);
  (void)hipFree(B->d_sys_t0);
  B->ntiles = (int32_t)tsys.size();
  B->tile_slices = tile_slices;
  int st;
  if ((st = upload(&B->d_tile_sys, tsys)) || (st = upload(&B->d_tile_s0, ts0)) ||
      (st = upload(&B->d_tile_s1, ts1)) || (st = upload(&B->d_sys_t0, sys_t0)))
    return st;
  return PFM_OK;
}


// Workgroups of the iteration kernels (they walk the active-tile list).
constexpr int kBatchGrid = 2048;

inline int iter_grid(const pfm_batch_s *B) {
  const int G = std::min(B->ntiles, kBatchGrid);
  return G >= kXcds ? G / kXcds * kXcds : std::max(G, 1);
}

void launch_compact(const pfm_batch_s *B, const BWork &w, int chunk, hipStream_t q) {
  k_b_compact<<<1, 1024, 0, q>>>(B->ns, B->d_sys_t0, B->d_tile_s0, B->d_tile_s1, w.sys, w.act,
                                 w.g, chunk);
}

// The iteration SpMVs read the paired int16 column offsets where the batch's
// handle has them, unless PFM_BATCH_COLS=int32 (for comparing the layouts).
bool batch_pairs(const pfm_ldu_s *L) {
  static const bool int32_only = [] {
    const char *e = std::getenv("PFM_BATCH_COLS");
    return e && std::strcmp(e, "int32") == 0;
  }();
  return L->d_colp2 && !int32_only;
}

// Iteration j of a chunk: three kernels.
void launch_iteration(const pfm_batch_s *B, const BWork &w, const double *pd, int j,
                      hipStream_t q) {
  const pfm_ldu L = B->ldu;
  const int T = B->ntiles, G = iter_grid(B);
  if (batch_pairs(L))
    k_b_spmv<true><<<G, kBlock, 0, q>>>(w.act, L->d_slices, L->d_col, L->d_colp2, w.packed, w.dp,
                                        w.p, w.Ap, w.part, w.g, w.sys);
  else
    k_b_spmv<false><<<G, kBlock, 0, q>>>(w.act, L->d_slices, L->d_col, nullptr, w.packed, w.dp,
                                         w.p, w.Ap, w.part, w.g, w.sys);
  k_b_update<<<G, kBlock, 0, q>>>(w.act, T, B->d_sys_t0, pd, w.Ap, w.r, w.part, w.g, w.sys, j);
  k_b_pupdate<<<G, kBlock, 0, q>>>(w.act, T, B->d_sys_t0, pd, w.r, w.p, w.xp, w.part, w.g, w.sys,
                                   j);
}

// Iteration j of a chunk in tail mode (tiles of at most kTailTileSlices slices).
void launch_iteration_tail(const pfm_batch_s *B, const BWork &w, const double *pd, int j,
                           hipStream_t q) {
  const pfm_ldu L = B->ldu;
  const int T = B->ntiles;
  if (batch_pairs(L))
    k_bt_spmv<true><<<kTailGrid, kTailThreads, 0, q>>>(w.act, L->d_slices, L->d_col, L->d_colp2,
                                                       w.packed, w.dp, w.p, w.Ap, w.part, w.g);
  else
    k_bt_spmv<false><<<kTailGrid, kTailThreads, 0, q>>>(w.act, L->d_slices, L->d_col, nullptr,
                                                        w.packed, w.dp, w.p, w.Ap, w.part, w.g);
  k_bt_update<<<kTailGrid, kTailThreads, 0, q>>>(w.act, T, B->d_sys_t0, pd, w.Ap, w.r, w.part, w.g,
                                                 w.sys, j);
  k_bt_pupdate<<<kTailGrid, kTailThreads, 0, q>>>(w.act, T, B->d_sys_t0, pd, w.r, w.p, w.xp,
                                                  w.part, w.g, w.sys, j);
}

constexpr int kTileCand[6] = {1, 2, 4, 8, 16, 32};

// Times `kTuneIters` full iterations per candidate tile size on synthetic
// values (unit diagonal, no off-diagonal coefficients, p = 1, tolerances that
// never stop a system) and keeps the largest within 5 % of the fastest.  The values go to NaN after the
// first iteration; the kernels' memory traffic and launch shapes do not
// depend on them.
int autotune(pfm_batch_s *B) {
  constexpr int kTuneIters = 10;
  const int max_tiles = std::max(B->np / kSlice, 1);
  const size_t bytes = batch_layout(B, max_tiles, nullptr, nullptr);
  char *base = nullptr;
  PFM_CHECK_HIP(hipMalloc((void **)&base, bytes));
  BWork w;
  batch_layout(B, max_tiles, &w, base);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool ok = hipMemset(base, 0, bytes) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
            hipEventCreate(&e1) == hipSuccess;
  const int np = B->np;
  if (ok) {
    k_b_fill<<<grid1d(np), kBlock>>>(np, w.dp, w.r, w.bp);  // diag 1
    k_b_fill<<<grid1d(np), kBlock>>>(np, w.p, w.xp, w.Ap);  // p 1
    k_b_fill<<<grid1d(np), kBlock>>>(np, w.rD, w.r, w.bp);  // rD 1
    ok = hipGetLastError() == hipSuccess;
  }
  int best = 8;
  float best_t = 1e30f;
  for (int ci = 0; ok && ci < 6; ++ci) {
    if (build_tiles(B, kTileCand[ci]) != PFM_OK) {
      ok = false;
      break;
    }
    k_b_state_init<<<(B->ns + 255) / 256, 256>>>(w.g, w.sys, B->ns, -1.0, -1.0, 0x7fffffff, 0);
    launch_compact(B, w, kTuneIters + 2, nullptr);
    for (int rep = 0; rep < 2; ++rep) launch_iteration(B, w, w.rD, rep, nullptr);
    ok = hipEventRecord(e0, nullptr) == hipSuccess;
    for (int rep = 0; ok && rep < kTuneIters; ++rep) launch_iteration(B, w, w.rD, rep + 2, nullptr);
    float ms = 0.f;
    ok = ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    B->tuned_us[ci] = ms * 1000.f / kTuneIters;  // us per iteration
    if (ok) best_t = std::min(best_t, ms);
  }
  // The largest tile within 5 % of the fastest: close candidates would
  // otherwise trade places from run to run (each tile size sums the dot
  // products in its own order), and the tail-mode kernels fill their waves
  // with large tiles.
  for (int ci = 0; ok && ci < 6; ++ci)
    if (B->tuned_us[ci] * kTuneIters <= 1.05f * best_t * 1000.f) best = kTileCand[ci];
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(base);
  if (!ok) {
    set_error("batch autotuning failed");
    return PFM_ERR_HIP;
  }
  return build_tiles(B, best);
}

}  // namespace

extern "C" {

int pfm_batch_destroy(pfm_batch B) {
  if (!B) return PFM_OK;
  (void)hipFree(B->d_tile_sys);
  (void)hipFree(B->d_tile_s0);
  (void)hipFree(B->d_tile_s1);
  (void)hipFree(B->d_sys_t0);
  (void)hipFree(B->d_cell_map);
  poll_pool_free(B->poll);
  if (B->ldu) pfm_ldu_destroy(B->ldu);
  delete B;
  return PFM_OK;
}

int pfm_batch_create(int32_t ns, const int32_t *cell_offsets, const int32_t *face_offsets,
                     const int32_t *owner, const int32_t *neighbour, int32_t tile_slices,
                     pfm_batch *out) {
  PFM_REQUIRE(ns > 0 && cell_offsets && face_offsets && out, "bad batch description");
  PFM_REQUIRE(tile_slices >= 0 && tile_slices <= 4096, "bad tile size");
  *out = nullptr;
  pfm_batch B = new pfm_batch_s();
  B->ns = ns;
  B->cell_off.assign(cell_offsets, cell_offsets + ns + 1);
  B->face_off.assign(face_offsets, face_offsets + ns + 1);
  B->n = B->cell_off[ns];
  B->nf = B->face_off[ns];
  B->pad_off.assign(ns + 1, 0);
  for (int k = 0; k < ns; ++k) {
    const int nk = B->cell_off[k + 1] - B->cell_off[k];
    if (nk < 0 || B->face_off[k + 1] < B->face_off[k]) {
      delete B;
      set_error("offsets must be non-decreasing");
      return PFM_ERR_ARG;
    }
    B->pad_off[k + 1] = B->pad_off[k] + (nk + kSlice - 1) / kSlice * kSlice;
  }
  B->np = B->pad_off[ns];
  // padded concatenated topology (local indices of system k shifted to its block)
  std::vector<int32_t> own(B->nf), nei(B->nf), cmap(B->n);
  for (int k = 0; k < ns; ++k) {
    const int nk = B->cell_off[k + 1] - B->cell_off[k];
    for (int f = B->face_off[k]; f < B->face_off[k + 1]; ++f) {
      if (owner[f] < 0 || owner[f] >= nk || neighbour[f] < 0 || neighbour[f] >= nk) {
        delete B;
        set_error("system %d face %d addresses a cell outside the system", k, f);
        return PFM_ERR_ARG;
      }
      own[f] = owner[f] + B->pad_off[k];
      nei[f] = neighbour[f] + B->pad_off[k];
    }
    for (int i = 0; i < nk; ++i) cmap[B->cell_off[k] + i] = B->pad_off[k] + i;
  }
  int st = pfm_ldu_create(B->np, B->nf, own.data(), nei.data(), &B->ldu);
  if (st == PFM_OK) st = upload(&B->d_cell_map, cmap);
  if (st == PFM_OK) st = poll_pool_init(B->poll);
  if (st != PFM_OK) {
    pfm_batch_destroy(B);
    return st;
  }
  if (tile_slices > 0) {
    st = build_tiles(B, tile_slices);
  } else {
    // autotune: time whole iterations for each candidate tile size
    st = autotune(B);
  }
  if (st != PFM_OK) {
    pfm_batch_destroy(B);
    return st;
  }
  *out = B;
  return PFM_OK;
}

int pfm_batch_info(pfm_batch B, int32_t *n_padded, int32_t *n_tiles, int32_t *tile_slices,
                   float *tuned_us) {
  PFM_REQUIRE(B, "null batch");
  if (n_padded) *n_padded = B->np;
  if (n_tiles) *n_tiles = B->ntiles;
  if (tile_slices) *tile_slices = B->tile_slices;
  if (tuned_us) memcpy(tuned_us, B->tuned_us, sizeof(B->tuned_us[0]) * 6);
  return PFM_OK;
}

int64_t pfm_batch_workspace_bytes(pfm_batch B) {
  if (!B) return -1;
  return (int64_t)batch_layout(B, B->ntiles, nullptr, nullptr);
}

int pfm_batch_pcg_solve(pfm_batch B, const pfm_solve_params *prm, int transpose,
                        const double *diag, const double *lower, const double *upper,
                        const double *b, double *x, void *work, int32_t graph_chunk,
                        int32_t *iterations_h, int32_t *status_h, double *initial_res_h,
                        double *final_res_h, void *stream) {
  PFM_REQUIRE(B && prm && work && diag && b && x, "null argument");
  PFM_REQUIRE(B->nf == 0 || (lower && upper), "null lower/upper");
  PFM_REQUIRE(prm->max_iter >= 0, "max_iter must be >= 0");
  PFM_REQUIRE(prm->precond == PFM_PRECOND_NONE || prm->precond == PFM_PRECOND_DIAG,
              "unknown preconditioner");
  hipStream_t s = (hipStream_t)stream;
  BWork w;
  batch_layout(B, B->ntiles, &w, (char *)work);
  pfm_ldu L = B->ldu;
  const int np = B->np, ns = B->ns, T = B->ntiles;
  k_b_fill<<<grid1d(np), kBlock, 0, s>>>(np, w.dp, w.xp, w.bp);
  if (B->n) k_b_scatter<<<grid1d(B->n), kBlock, 0, s>>>(B->n, B->d_cell_map, diag, x, b, w.dp, w.xp, w.bp);
  int st = pfm_ldu_pack(L, transpose, lower, upper, w.packed, stream);
  if (st != PFM_OK) return st;
  const bool pre = prm->precond == PFM_PRECOND_DIAG;
  if (pre && (st = pfm_dic_factor(L, w.dp, w.packed, w.rD, stream)) != PFM_OK) return st;
  const double *pd = pre ? w.rD : nullptr;
  k_b_state_init<<<(ns + 255) / 256, 256, 0, s>>>(w.g, w.sys, ns, prm->tolerance, prm->rel_tol,
                                                   prm->max_iter, prm->min_iter);
  k_b_init<<<T, kBlock, 0, s>>>(B->d_tile_sys, B->d_tile_s0, B->d_tile_s1, T, L->d_slices, L->d_col,
                                w.packed, w.dp, pd, w.bp, w.xp, w.r, w.p, w.part);
  k_b_init_fin<<<ns, kBlock, 0, s>>>(ns, B->d_sys_t0, T, w.part, w.g, w.sys);
  PFM_CHECK_HIP(hipGetLastError());
  // Tail mode (1024-thread tiles, same results) once at most kTailSystems
  // systems are left, as seen at the last poll.
  const bool tail_ok = B->tile_slices <= kTailTileSlices;
  auto iteration = [&](int j, hipStream_t q, bool tail) {
    if (tail)
      launch_iteration_tail(B, w, pd, j, q);
    else
      launch_iteration(B, w, pd, j, q);
  };
  const int chunk = std::max(1, graph_chunk);
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // chunk graphs: normal, tail mode
  hipStream_t cap = nullptr;
  auto release = [&] {
    for (hipGraphExec_t &e : exec)
      if (e) (void)hipGraphExecDestroy(e);
    if (cap) (void)hipStreamDestroy(cap);
  };
  auto capture = [&](bool tail) -> int {
    if (!cap && hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
      set_error("stream creation failed");
      return PFM_ERR_HIP;
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      launch_compact(B, w, chunk, cap);
      for (int k = 0; k < chunk; ++k) iteration(k, cap, tail);
      e = hipStreamEndCapture(cap, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec[tail], graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return PFM_ERR_HIP;
    }
    return PFM_OK;
  };
  // replay chunks until every system has stopped (one chunk queued ahead)
  PollLease lease;
  int rc = lease.acquire(&B->poll, true);
  if (rc != PFM_OK) return rc;
  int32_t *flag = lease.flag;
  hipEvent_t *ev = lease.ev;
  int enq = 0, slot = 0, seen = 0;
  bool pending = false;
  while (true) {
    bool queued = false;
    if (enq < prm->max_iter) {
      const bool tail = tail_ok && ns - seen <= kTailSystems;
      if (graph_chunk > 0) {
        if (!exec[tail] && (rc = capture(tail)) != PFM_OK) break;
        if (hipGraphLaunch(exec[tail], s) != hipSuccess) {
          rc = PFM_ERR_HIP;
          break;
        }
        enq += chunk;
      } else {
        const int c = std::min(8, prm->max_iter - enq);
        launch_compact(B, w, c, s);
        for (int k = 0; k < c; ++k) iteration(k, s, tail);
        enq += c;
      }
      if (hipMemcpyAsync(&flag[slot], &w.g->ndone, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
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
      seen = flag[slot ^ 1];
      if (seen >= ns) break;
    }
    if (!queued) break;
    pending = true;
    slot ^= 1;
  }
  if (rc != PFM_OK) (void)hipStreamSynchronize(s);  // queued chunks may still use the graphs
  release();
  if (rc != PFM_OK) {
    if (rc == PFM_ERR_HIP) set_error("HIP error while driving the batched solve");
    return rc;
  }
  if (B->n) k_b_gather<<<grid1d(B->n), kBlock, 0, s>>>(B->n, B->d_cell_map, w.xp, x);
  std::vector<BSys> hs(ns);
  PFM_CHECK_HIP(hipMemcpyAsync(hs.data(), w.sys, sizeof(BSys) * ns, hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < ns; ++k) {
    if (iterations_h) iterations_h[k] = hs[k].nhist;
    if (status_h) status_h[k] = hs[k].status == 0 ? PFM_SOLVE_MAXITER : hs[k].status;
    if (initial_res_h) initial_res_h[k] = hs[k].res0;
    if (final_res_h) final_res_h[k] = hs[k].res;
  }
  return PFM_OK;
}

}  // extern "C"
