// Device-resident Krylov solvers: PCG (reference solvers/pcg.py:97-187) and
// PBiCGStab (solvers/pbicgstab.py:103-239) with the stopping rule of
// solvers/residual.py:112-145.
//
// Each iteration is a fixed sequence of kernels.  Bandwidth kernels (the
// fused SpMV + dot products, the fused vector updates) run on a fixed grid of
// at most kMaxRedBlocks workgroups and write one partial sum per workgroup;
// a one-workgroup finalize kernel adds the partials in a fixed order and takes
// the scalar decisions the reference takes on the host (alpha, beta, omega,
// breakdown tests, the residual monitor).  Scalars, the residual history and
// the stop flag live in device memory, so nothing synchronises with the host
// inside the loop; results are bit-reproducible run to run.  The host
// enqueues iterations in chunks and polls the stop flag of the previous chunk
// through pinned memory while the next chunk is already queued; kernels that
// start after the stop flag is set return immediately.

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "solver_state.h"

using namespace pfm;

namespace {

// Write this block's partials (part[k*G + b]).
template <int K>
__device__ __forceinline__ void store_partials(double (&v)[K], double *part, double *lds) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = block_sum(v[k], lds);
    if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = s;
  }
}

// Rows of a producer kernel: a grid-stride walk over slices in which the
// workgroups of one XCD (consecutive logical blocks after xcd_remap) cover one
// contiguous chunk of rows per step, so neighbouring rows' x[] values are
// served from that XCD's L2 while the chunk moves through the matrix.
#define FOR_ROWS(...)                                                           \
  {                                                                             \
    const int lb_ = xcd_remap(blockIdx.x, gridDim.x);                           \
    const int lane = threadIdx.x & 63;                                          \
    for (int s = lb_ * kWavesPerBlock + (threadIdx.x >> 6); s < nslices;        \
         s += gridDim.x * kWavesPerBlock) {                                     \
      const int row = s * kSlice + lane;                                        \
      __VA_ARGS__                                                               \
    }                                                                           \
  }

// Which PCG streams are loaded nontemporally.  Across iterations the 256 MB
// MALL keeps much of the matrix (about 142 MB at 128^3) and the vectors, and
// an iteration moves more than the MALL holds: nontemporal matrix loads
// (59.6 -> 68.8 us per iteration at 128^3) or x / p_{it-1} loads (-> 60.2 us)
// cost more than they save; Ap, read once after the SpMV writes it, gains
// 0.7 % (59.6 -> 59.2 us).
constexpr bool kNtMat = false;
constexpr bool kNtAp = true;

// One row of A v in column form kCF (s: the row's slice).
template <int kCF, bool kNT = false>
__device__ __forceinline__ double row_av(const SliceInfo si, int s, int lane, int row, double dx,
                                         const int32_t *__restrict__ col,
                                         const int16_t *__restrict__ col16,
                                         const uint32_t *__restrict__ cp,
                                         const double *__restrict__ val,
                                         const double *__restrict__ v) {
  auto gx = [v](int c) { return v[c]; };
  if constexpr (kCF == kColsPair)
    return row_spmv_pair<double, decltype(gx), false, kNT>(si, s, lane, row, dx, cp, val, gx);
  else if constexpr (kCF == kCols16)
    return row_spmv_g<double, decltype(gx), false, ColRelT<kNT>, kNT>(
        si, lane, dx, ColRelT<kNT>{col16, row}, val, gx);
  else
    return row_spmv_g<double, decltype(gx), false, ColAbsT<kNT>, kNT>(si, lane, dx,
                                                                      ColAbsT<kNT>{col}, val, gx);
}

// One row of A v with the handle's paired column offsets when cp is given,
// else its int32 columns (kUpperFirst: the association of the reference's
// Ax_T, for the transposed packed values).
template <bool kUpperFirst = false>
__device__ __forceinline__ double row_any(const SliceInfo si, int s, int lane, int row, double dx,
                                          const int32_t *__restrict__ col,
                                          const uint32_t *__restrict__ cp,
                                          const double *__restrict__ val,
                                          const double *__restrict__ v) {
  auto gx = [v](int c) { return v[c]; };
  if (cp) return row_spmv_pair<double, decltype(gx), kUpperFirst>(si, s, lane, row, dx, cp, val, gx);
  return row_spmv_g<double, decltype(gx), kUpperFirst>(si, lane, dx, ColAbs{col}, val, gx);
}

// ---------------------------------------------------------------- PCG ------
// Per iteration i: k_pcg_spmv (Ap, p.Ap partials) -> k_pcg_update (every
// workgroup reduces p.Ap itself: breakdown test, alpha; r; r.r and r.z
// partials) -> k_pcg_pupdate (every workgroup reduces r.r, r.z: monitor
// decision, beta; x += alpha p, then p).  x is updated where p is read anyway
// (one vector less per iteration); since every workgroup of k_pcg_pupdate
// must update x even when iteration i converges, its stop decision is left in
// `pending` and the next k_pcg_spmv turns it into `done` (or collect() reads
// it after the last iteration).  Workgroup 0 alone writes the state; scalars
// another workgroup of the same kernel still reads (r.z, alpha) are kept per
// iteration parity.  Partial regions: A = part[0, G) (p.Ap), B = part[G, 3G)
// (r.r, r.z).

// r = b - A x; z = M^-1 r; p = z; partials r.r and r.z (region B).
template <int kCF>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, int nslices, const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const int16_t *__restrict__ col16,
                                                      const uint32_t *__restrict__ cp,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ rD,
                                                      const double *__restrict__ b,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ r, double *__restrict__ p,
                                                      double *__restrict__ part) {
  __shared__ double lds[kWavesPerBlock];
  double acc[2] = {0.0, 0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double dx = row < n ? diag[row] * x[row] : 0.0;
    const double ax = row_av<kCF>(si, s, lane, row, dx, col, col16, cp, val, x);
    if (row < n) {
      const double ri = b[row] - ax;
      const double zi = rD ? ri / rD[row] : ri;
      r[row] = ri;
      p[row] = zi;
      acc[0] += ri * ri;
      acc[1] += ri * zi;
    }
  })
  store_partials<2>(acc, part + gridDim.x, lds);
}

__global__ __launch_bounds__(kBlock) void k_pcg_init_fin(const double *__restrict__ part, int G,
                                                          KState *st, double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  const double RR = reduce_partials(part + G, G, lds);
  const double RZ = reduce_partials(part + 2 * G, G, lds);
  if (threadIdx.x == 0) {
    if (monitor_update(st, hist, sqrt(RR), 0))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (st->max_iter < 1)
      finish(st, PFM_SOLVE_MAXITER);
    st->rzs[0] = RZ;
  }
}

// Ap = A p; partial p.Ap (region A).  Ends the solve if the previous
// iteration decided so.
template <int kCF>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(int n, int nslices, const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const int16_t *__restrict__ col16,
                                                      const uint32_t *__restrict__ cp,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ p,
                                                      double *__restrict__ Ap,
                                                      double *__restrict__ part, KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  if (st->pending) {  // the previous iteration ended the solve
    if (blockIdx.x == 0 && threadIdx.x == 0) finish(st, st->pending);
    return;
  }
  double acc[1] = {0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double pi = row < n ? p[row] : 0.0;
    const double dx = row < n ? ld<kNtMat>(diag + row) * pi : 0.0;
    const double v = row_av<kCF, kNtMat>(si, s, lane, row, dx, col, col16, cp, val, p);
    if (row < n) {
      Ap[row] = v;
      acc[0] += pi * v;
    }
  })
  store_partials<1>(acc, part, lds);
}

// alpha = rz / pAp (breakdown if |pAp| < 1e-30), kept for k_pcg_pupdate;
// r -= alpha Ap; partials r.r and r.z (region B).  On a breakdown at an even
// iteration (pprev = p_{it-1} given) every workgroup applies the x update
// deferred at iteration it-1 to its rows.  The breakdown is recorded in
// `pending`, which this kernel never tests, and becomes `done` in the next
// k_pcg_spmv: a `done` stored here would make a workgroup dispatched after
// the store return at the entry test and skip its rows' update.
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, int nslices, int it,
                                                        const double *__restrict__ rD,
                                                        const double *__restrict__ Ap,
                                                        double *__restrict__ r,
                                                        double *__restrict__ x,
                                                        const double *__restrict__ pprev,
                                                        double *__restrict__ part, KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  // The first round's loads do not depend on alpha: issue them before the
  // reduction so both latencies overlap.
  const int lane = threadIdx.x & 63;
  const int stride = G * kWavesPerBlock;
  int s = xcd_remap(blockIdx.x, G) * kWavesPerBlock + (threadIdx.x >> 6);
  int row = s * kSlice + lane;
  double rp = 0.0, ap = 0.0, dp = 1.0;
  if (s < nslices && row < n) {
    rp = r[row];
    ap = ld<kNtAp>(Ap + row);
    if (rD) dp = rD[row];
  }
  const double PAP = reduce_partials(part, G, lds);
  if (fabs(PAP) < 1e-30) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->pending = PFM_SOLVE_BREAKDOWN;
    if (pprev) {  // the x update deferred at iteration it-1 (see k_pcg_pupdate)
      const double ap_ = st->alphas[(it - 1) & 1];
      FOR_ROWS({
        if (row < n) x[row] = x[row] + ap_ * pprev[row];
      })
    }
    return;
  }
  const double alpha = st->rzs[(it - 1) & 1] / PAP;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alphas[it & 1] = alpha;
  double acc[2] = {0.0, 0.0};
  for (; s < nslices; s += stride, row = s * kSlice + lane) {
    if (row < n) {
      const double ri = rp - alpha * ap;
      r[row] = ri;
      const double zi = rD ? ri / dp : ri;
      acc[0] += ri * ri;
      acc[1] += ri * zi;
    }
    const int sn = s + stride, rn = sn * kSlice + lane;
    if (sn < nslices && rn < n) {
      rp = r[rn];
      ap = ld<kNtAp>(Ap + rn);
      if (rD) dp = rD[rn];
    }
  }
  store_partials<2>(acc, part + G, lds);
}

// monitor(||r||, i); x update; beta = r.z' / r.z; p' = M^-1 r + beta p (not
// after the last iteration).  x is updated every second iteration, which
// saves one x read and write per pair: at even `it` as
// (x + alpha_{it-1} p_{it-1}) + alpha_it p_it, which rounds exactly as the two
// single updates, with p_{it-1} read from pprev; at odd `it` only when the
// solve ends there.  p' goes to pn, a buffer other than p; at even `it` pn is
// pprev's buffer (each row reads p_{it-1} before p_{it+1} is written over it),
// so neither is restrict-qualified.
__global__ __launch_bounds__(kBlock) void k_pcg_pupdate(int n, int nslices, int it,
                                                         const double *__restrict__ rD,
                                                         const double *__restrict__ r,
                                                         const double *__restrict__ p,
                                                         const double *pprev, double *pn,
                                                         double *__restrict__ x,
                                                         const double *__restrict__ part, KState *st,
                                                         double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  // A breakdown in this iteration's k_pcg_update ended the solve.  This
  // kernel's workgroup 0 writes `pending` only as CONVERGED or MAXITER, so a
  // workgroup that reads it after that store still proceeds, as it must.
  if (st->done || st->pending == PFM_SOLVE_BREAKDOWN) return;
  const int G = gridDim.x;
  const bool pair = pprev != nullptr;
  // First round's operands loaded before the reductions (they do not depend
  // on beta), so their latency overlaps the partial sums'.
  const int lane = threadIdx.x & 63;
  const int stride = G * kWavesPerBlock;
  int s = xcd_remap(blockIdx.x, G) * kWavesPerBlock + (threadIdx.x >> 6);
  int row = s * kSlice + lane;
  double rp = 0.0, dp = 1.0, pp = 0.0, qp = 0.0, xp = 0.0;
  if (s < nslices && row < n) {
    rp = r[row];
    if (rD) dp = rD[row];
    pp = p[row];
    if (pair) {
      qp = pprev[row];
      xp = x[row];
    }
  }
  const double RR = reduce_partials(part + G, G, lds);
  const double RZ = reduce_partials(part + 2 * G, G, lds);
  const double res = sqrt(RR);
  const bool conv = monitor_decision(st, res, it);
  const bool last = conv || it >= st->max_iter;
  const double rz_old = st->rzs[(it - 1) & 1];
  const double alpha = st->alphas[it & 1];
  const double alpha_prev = pair ? st->alphas[(it - 1) & 1] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    monitor_update(st, hist, res, it);
    st->iter = it;
    st->rzs[it & 1] = RZ;
    if (conv)
      st->pending = PFM_SOLVE_CONVERGED;
    else if (it >= st->max_iter)
      st->pending = PFM_SOLVE_MAXITER;
  }
  const double beta = RZ / rz_old;
  for (; s < nslices; s += stride, row = s * kSlice + lane) {
    if (row < n) {
      if (pair)
        x[row] = (xp + alpha_prev * qp) + alpha * pp;
      else if (last)
        x[row] = x[row] + alpha * pp;
      if (!last) {
        const double zi = rD ? rp / dp : rp;
        pn[row] = zi + beta * pp;
      }
    }
    const int sn = s + stride, rn = sn * kSlice + lane;
    if (sn < nslices && rn < n) {
      rp = r[rn];
      if (rD) dp = rD[rn];
      pp = p[rn];
      if (pair) {
        qp = pprev[rn];
        xp = x[rn];
      }
    }
  }
}

// ---------------------------------------------------------- PBiCGStab ------
// r = b - A x; rt = r; p = r; y = M^-1 p; partial r.r (= rt.r).
__global__ __launch_bounds__(kBlock) void k_bs_init(int n, int nslices, const SliceInfo *__restrict__ sl,
                                                     const int32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ cp,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ diag,
                                                     const double *__restrict__ rD,
                                                     const double *__restrict__ b,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ r, double *__restrict__ rt,
                                                     double *__restrict__ p, double *__restrict__ y,
                                                     double *__restrict__ part) {
  __shared__ double lds[kWavesPerBlock];
  double acc[1] = {0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double dx = row < n ? diag[row] * x[row] : 0.0;
    const double ax = row_any(si, s, lane, row, dx, col, cp, val, x);
    if (row < n) {
      const double ri = b[row] - ax;
      r[row] = ri;
      rt[row] = ri;
      p[row] = ri;
      y[row] = rD ? ri / rD[row] : ri;
      acc[0] += ri * ri;
    }
  })
  store_partials<1>(acc, part, lds);
}

__global__ __launch_bounds__(kBlock) void k_bs_init_fin(const double *__restrict__ part, int G,
                                                         KState *st, double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  const double RR = sum_partials(part, G, lds);
  if (threadIdx.x == 0) {
    if (monitor_update(st, hist, sqrt(RR), 0))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (st->max_iter < 1)
      finish(st, PFM_SOLVE_MAXITER);
    st->rhos[0] = RR;  // rho = rt . r with rt == r
    if (!st->done && fabs(RR) < 1e-30) finish(st, PFM_SOLVE_BREAKDOWN);
  }
}

// v = A y; partial rt.v.
__global__ __launch_bounds__(kBlock) void k_bs_spmv1(int n, int nslices, const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ cp,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ y,
                                                      const double *__restrict__ rt,
                                                      double *__restrict__ v,
                                                      double *__restrict__ part, const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc[1] = {0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double dx = row < n ? diag[row] * y[row] : 0.0;
    const double av = row_any(si, s, lane, row, dx, col, cp, val, y);
    if (row < n) {
      v[row] = av;
      acc[0] += rt[row] * av;
    }
  })
  store_partials<1>(acc, part, lds);
}

// Every workgroup: rt.v from region A, breakdown test, alpha = rho / rt.v;
// s = r - alpha v; z = M^-1 s; partial s.s (region B).  Workgroup 0 records
// alpha for the later kernels of the iteration.
__global__ __launch_bounds__(kBlock) void k_bs_s(int n, int nslices, int it,
                                                  const double *__restrict__ rD,
                                                  const double *__restrict__ r,
                                                  const double *__restrict__ v,
                                                  double *__restrict__ sv, double *__restrict__ z,
                                                  double *__restrict__ part, KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double RTV = reduce_partials(part, G, lds);
  if (fabs(RTV) < 1e-30) {
    if (blockIdx.x == 0 && threadIdx.x == 0) finish(st, PFM_SOLVE_BREAKDOWN);
    return;
  }
  const double alpha = st->rhos[(it - 1) & 1] / RTV;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
  double acc[1] = {0.0};
  FOR_ROWS({
    if (row < n) {
      const double si = r[row] - alpha * v[row];
      sv[row] = si;
      z[row] = rD ? si / rD[row] : si;
      acc[0] += si * si;
    }
  })
  store_partials<1>(acc, part + G, lds);
}

// Every workgroup: monitor decision on ||s|| (converged -> branch 2, the
// half-step exit); t = A z; partials t.t, t.z (regions C).  Workgroup 0
// records the monitor and the branch.
__global__ __launch_bounds__(kBlock) void k_bs_spmv2(int n, int nslices, int it,
                                                      const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ cp,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ z,
                                                      double *__restrict__ t,
                                                      double *__restrict__ part, KState *st,
                                                      double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double res = sqrt(reduce_partials(part + G, G, lds));
  const bool conv = monitor_decision(st, res, it);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    monitor_update(st, hist, res, it);
    st->iter = it;
    st->branch = conv ? 2 : 0;
  }
  if (conv) return;
  double acc[2] = {0.0, 0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double zi = row < n ? z[row] : 0.0;
    const double dx = row < n ? diag[row] * zi : 0.0;
    const double at = row_any(si, s, lane, row, dx, col, cp, val, z);
    if (row < n) {
      t[row] = at;
      acc[0] += at * at;
      acc[1] += at * zi;
    }
  })
  store_partials<2>(acc, part + 2 * G, lds);
}

// Every workgroup: omega = t.z / t.t or the t ~ 0 branch (region C); the
// solution/residual update of the branch; partials r.r and rt.r (region D).
__global__ __launch_bounds__(kBlock) void k_bs_update(int n, int nslices, const double *__restrict__ y,
                                                       const double *__restrict__ z,
                                                       const double *__restrict__ sv,
                                                       const double *__restrict__ t,
                                                       const double *__restrict__ rt,
                                                       double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       double *__restrict__ part, KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double alpha = st->alpha;
  if (st->branch == 2) {  // converged on s: x += alpha y and stop
    FOR_ROWS({
      if (row < n) x[row] = x[row] + alpha * y[row];
    })
    return;
  }
  const double TT = reduce_partials(part + 2 * G, G, lds);
  const double TZ = reduce_partials(part + 3 * G, G, lds);
  const int branch = fabs(TT) < 1e-30 ? 1 : 0;
  const double omega = branch ? 0.0 : TZ / TT;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->branch = branch;
    if (!branch) st->omega = omega;
  }
  double acc[2] = {0.0, 0.0};
  FOR_ROWS({
    if (row < n) {
      double ri;
      if (branch == 1) {
        x[row] = x[row] + alpha * y[row];
        ri = sv[row];
      } else {
        x[row] = (x[row] + alpha * y[row]) + omega * z[row];
        ri = sv[row] - omega * t[row];
      }
      r[row] = ri;
      acc[0] += ri * ri;
      acc[1] += rt[row] * ri;
    }
  })
  store_partials<2>(acc, part + 4 * G, lds);
}

// Every workgroup: monitor decision on ||r||, rho_new = rt.r, beta and the
// breakdown / max-iteration tests (region D); p = r + beta (p - omega v)
// (branch 1: p = r + beta (p - v)); y = M^-1 p.  Workgroup 0 records the
// monitor, rho (by iteration parity) and the stop.
__global__ __launch_bounds__(kBlock) void k_bs_pupdate(int n, int nslices, int it,
                                                        const double *__restrict__ rD,
                                                        const double *__restrict__ r,
                                                        const double *__restrict__ v,
                                                        double *__restrict__ p,
                                                        double *__restrict__ y,
                                                        const double *__restrict__ part, KState *st,
                                                        double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int branch = st->branch;
  if (branch == 2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) finish(st, PFM_SOLVE_CONVERGED);
    return;
  }
  const int G = gridDim.x;
  const double RR = reduce_partials(part + 4 * G, G, lds);
  const double RTR = reduce_partials(part + 5 * G, G, lds);
  const double res = sqrt(RR);
  const double alpha = st->alpha, omega = st->omega;
  const double rho_old = st->rhos[(it - 1) & 1];
  int status = 0;
  double beta = 0.0;
  if (monitor_decision(st, res, it)) {
    status = PFM_SOLVE_CONVERGED;
  } else if (branch == 1) {
    beta = (RTR / rho_old) * alpha;
  } else if (fabs(omega) < 1e-30) {
    status = PFM_SOLVE_BREAKDOWN;
  } else {
    beta = (RTR / rho_old) * (alpha / omega);
  }
  if (!status && it >= st->max_iter) status = PFM_SOLVE_MAXITER;
  // breakdown test at the top of the next iteration (pbicgstab.py:152-156)
  if (!status && fabs(RTR) < 1e-30) status = PFM_SOLVE_BREAKDOWN;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    monitor_update(st, hist, res, it);
    if (status != PFM_SOLVE_CONVERGED && !(branch == 0 && fabs(omega) < 1e-30))
      st->rhos[it & 1] = RTR;
    if (status) finish(st, status);
  }
  if (status) return;
  FOR_ROWS({
    if (row < n) {
      const double vi = branch == 1 ? v[row] : omega * v[row];
      const double pi = r[row] + beta * (p[row] - vi);
      p[row] = pi;
      y[row] = rD ? pi / rD[row] : pi;
    }
  })
}

// ------------------------------------------------------------ PBiCG --------
// solvers/pbicg.py:107-220.  Per iteration: v = A p and t = A^T p with
// z~.v partials (region A) | rho/z~.v breakdown, alpha, x, r, r~ and the
// partials r.r, z~.r (region B) | monitor, rho_new, beta, p = z + beta p.
// z = r / rD and z~ = r~ / rD are formed where they are read.

// r = b - A x; r~ = r; p = z = r / rD; partials r.r and z~.r (= z.r).
__global__ __launch_bounds__(kBlock) void k_bicg_init(int n, int nslices,
                                                       const SliceInfo *__restrict__ sl,
                                                       const int32_t *__restrict__ col,
                                                       const uint32_t *__restrict__ cp,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ diag,
                                                       const double *__restrict__ rD,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       double *__restrict__ rt,
                                                       double *__restrict__ p,
                                                       double *__restrict__ part) {
  __shared__ double lds[kWavesPerBlock];
  double acc[2] = {0.0, 0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double dx = row < n ? diag[row] * x[row] : 0.0;
    const double ax = row_any(si, s, lane, row, dx, col, cp, val, x);
    if (row < n) {
      const double ri = b[row] - ax;
      const double zi = rD ? ri / rD[row] : ri;
      r[row] = ri;
      rt[row] = ri;
      p[row] = zi;
      acc[0] += ri * ri;
      acc[1] += zi * ri;
    }
  })
  store_partials<2>(acc, part + gridDim.x, lds);
}

__global__ __launch_bounds__(kBlock) void k_bicg_init_fin(const double *__restrict__ part, int G,
                                                           KState *st, double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  const double RR = reduce_partials(part + G, G, lds);
  const double RZ = reduce_partials(part + 2 * G, G, lds);
  if (threadIdx.x == 0) {
    st->rhos[0] = RZ;
    if (monitor_update(st, hist, sqrt(RR), 0))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (st->max_iter < 1)
      finish(st, PFM_SOLVE_MAXITER);
    else if (fabs(RZ) < 1e-30)  // the rho test at the top of iteration 1
      finish(st, PFM_SOLVE_BREAKDOWN);
  }
}

// v = A p, t = A^T p; partial z~.v (region A).
__global__ __launch_bounds__(kBlock) void k_bicg_spmv(int n, int nslices,
                                                       const SliceInfo *__restrict__ sl,
                                                       const int32_t *__restrict__ col,
                                                       const uint32_t *__restrict__ cp,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ val_t,
                                                       const double *__restrict__ diag,
                                                       const double *__restrict__ rD,
                                                       const double *__restrict__ p,
                                                       const double *__restrict__ rt,
                                                       double *__restrict__ v,
                                                       double *__restrict__ t,
                                                       double *__restrict__ part, const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc[1] = {0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double dx = row < n ? diag[row] * p[row] : 0.0;
    const double av = row_any(si, s, lane, row, dx, col, cp, val, p);
    // A^T p in the association of the reference's Ax_T
    const double at = row_any<true>(si, s, lane, row, dx, col, cp, val_t, p);
    if (row < n) {
      v[row] = av;
      t[row] = at;
      const double zt = rD ? rt[row] / rD[row] : rt[row];
      acc[0] += zt * av;
    }
  })
  store_partials<1>(acc, part, lds);
}

// alpha = rho / z~.v (breakdown if |z~.v| < 1e-30); x += alpha p;
// r -= alpha v; r~ -= alpha t; partials r.r, z~.r (region B).
__global__ __launch_bounds__(kBlock) void k_bicg_update(int n, int nslices, int it,
                                                         const double *__restrict__ rD,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ v,
                                                         const double *__restrict__ t,
                                                         double *__restrict__ x,
                                                         double *__restrict__ r,
                                                         double *__restrict__ rt,
                                                         double *__restrict__ part, KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double ZTV = reduce_partials(part, G, lds);
  if (fabs(ZTV) < 1e-30) {
    if (blockIdx.x == 0 && threadIdx.x == 0) finish(st, PFM_SOLVE_BREAKDOWN);
    return;
  }
  const double alpha = st->rhos[(it - 1) & 1] / ZTV;
  double acc[2] = {0.0, 0.0};
  FOR_ROWS({
    if (row < n) {
      x[row] = x[row] + alpha * p[row];
      const double ri = r[row] - alpha * v[row];
      r[row] = ri;
      const double rti = rt[row] - alpha * t[row];
      rt[row] = rti;
      const double zt = rD ? rti / rD[row] : rti;
      acc[0] += ri * ri;
      acc[1] += zt * ri;
    }
  })
  store_partials<2>(acc, part + G, lds);
}

// monitor(||r||, it); rho_new = z~.r; beta = rho_new / rho; p = z + beta p.
__global__ __launch_bounds__(kBlock) void k_bicg_pupdate(int n, int nslices, int it,
                                                          const double *__restrict__ rD,
                                                          const double *__restrict__ r,
                                                          double *__restrict__ p,
                                                          const double *__restrict__ part,
                                                          KState *st, double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const int G = gridDim.x;
  const double RR = reduce_partials(part + G, G, lds);
  const double RZ = reduce_partials(part + 2 * G, G, lds);
  const double res = sqrt(RR);
  const double rho_old = st->rhos[(it - 1) & 1];
  int status = 0;
  if (monitor_decision(st, res, it))
    status = PFM_SOLVE_CONVERGED;
  else if (it >= st->max_iter)
    status = PFM_SOLVE_MAXITER;
  else if (fabs(RZ) < 1e-30)  // the rho test at the top of the next iteration
    status = PFM_SOLVE_BREAKDOWN;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    monitor_update(st, hist, res, it);
    st->iter = it;
    st->rhos[it & 1] = RZ;
    if (status) finish(st, status);
  }
  if (status) return;
  const double beta = RZ / rho_old;
  FOR_ROWS({
    if (row < n) {
      const double ri = r[row];
      const double zi = rD ? ri / rD[row] : ri;
      p[row] = zi + beta * p[row];
    }
  })
}

// ------------------------------------------------------------ workspace ----
struct Work {
  KState *st;
  double *part;
  double *hist;
  double *vec[8];
};

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Workgroups of the reduction grid: one per slice quartet, at most
// kMaxRedBlocks, or fewer when PFM_RED_BLOCKS (8 .. kMaxRedBlocks) says so.
inline int red_cap() {
  static const int cap = [] {
    const char *e = std::getenv("PFM_RED_BLOCKS");
    const int v = e ? std::atoi(e) : kMaxRedBlocks;
    return v >= 8 && v <= kMaxRedBlocks ? v : kMaxRedBlocks;
  }();
  return cap;
}

inline int red_grid(const pfm_ldu_s *h) {
  return std::max(1, std::min(blocks_for_slices(h->nslices), red_cap()));
}

inline int pcg_col_form(const pfm_ldu_s *h) { return krylov_col_form(h); }

size_t work_layout(const pfm_ldu_s *h, int max_iter, int nvec, Work *w, char *base) {
  const int G = red_grid(h);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  Work tmp;
  tmp.st = (KState *)take(sizeof(KState));
  tmp.part = (double *)take(sizeof(double) * 6 * (size_t)G);
  tmp.hist = (double *)take(sizeof(double) * (2 * (size_t)std::max(max_iter, 0) + 2));
  for (int k = 0; k < nvec; ++k) tmp.vec[k] = (double *)take(sizeof(double) * (size_t)h->n);
  if (w) *w = tmp;
  return off;
}

int check_solve_args(pfm_ldu h, const pfm_solve_params *prm, const double *diag,
                     const double *packed, const double *rD, const double *b, double *x, void *work,
                     pfm_solve_result *res) {
  PFM_REQUIRE(h && prm && res, "null handle, params or result");
  PFM_REQUIRE(prm->max_iter >= 0, "max_iter must be >= 0");
  PFM_REQUIRE(prm->precond == PFM_PRECOND_NONE || prm->precond == PFM_PRECOND_DIAG,
              "unknown preconditioner code %d", prm->precond);
  PFM_REQUIRE(prm->precond == PFM_PRECOND_NONE || rD, "preconditioner diagonal missing");
  PFM_REQUIRE(h->n == 0 || (diag && b && x && work), "null device pointer");
  PFM_REQUIRE(h->nslots == 0 || packed, "null packed coefficients");
  return PFM_OK;
}

// Enqueue `body()` (one iteration) until the device stop flag is seen.
// One chunk stays queued ahead of the poll so the GPU never idles on it.
template <typename Body>
int drive(pfm_ldu h, const KState *st_dev, int max_iter, hipStream_t s, Body body,
          bool may_allocate = true) {
  PollLease lease;
  int rc0 = lease.acquire(&h->poll, may_allocate);
  if (rc0 != PFM_OK) return rc0;
  int32_t *flag = lease.flag;
  hipEvent_t *ev = lease.ev;
  int enq = 0, chunk = 4, slot = 0, rc = PFM_OK;
  bool pending = false;
  double last_report = now_seconds();
  while (true) {
    bool queued = false;
    report_progress(h, enq, last_report);
    if (enq < max_iter) {
      const int c = std::min(chunk, max_iter - enq);
      for (int k = 0; k < c; ++k) body(enq + k + 1);
      enq += c;
      chunk = std::min(chunk * 2, 32);
      if (hipMemcpyAsync(&flag[slot], &st_dev->done, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
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
  if (rc != PFM_OK) set_error("HIP error while driving the solver loop");
  return rc;
}

int collect(const KState *st_dev, const double *hist_dev, double *history_h, int hist_cap,
            pfm_solve_result *res, hipStream_t s);

int collect(const KState *st_dev, const double *hist_dev, double *history_h, int hist_cap,
            pfm_solve_result *res, hipStream_t s) {
  KState st;
  PFM_CHECK_HIP(hipMemcpyAsync(&st, st_dev, sizeof(KState), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  const int nh = std::min(st.nhist, hist_cap);
  if (history_h && nh > 0) {
    PFM_CHECK_HIP(hipMemcpyAsync(history_h, hist_dev, sizeof(double) * nh, hipMemcpyDeviceToHost, s));
    PFM_CHECK_HIP(hipStreamSynchronize(s));
  }
  res->iterations = st.nhist;
  res->status = st.status != 0 ? st.status : st.pending != 0 ? st.pending : PFM_SOLVE_MAXITER;
  res->converged = res->status == PFM_SOLVE_CONVERGED;
  res->initial_residual = st.nhist > 0 ? st.res0 : 0.0;
  res->final_residual = st.nhist > 0 ? st.res : 0.0;
  return PFM_OK;
}

}  // namespace

extern "C" {

int64_t pfm_pcg_workspace_bytes(pfm_ldu h, int32_t max_iter) {
  if (!h) return -1;
  return (int64_t)work_layout(h, max_iter, 4, nullptr, nullptr);
}

int64_t pfm_pbicgstab_workspace_bytes(pfm_ldu h, int32_t max_iter) {
  if (!h) return -1;
  return (int64_t)work_layout(h, max_iter, 8, nullptr, nullptr);
}

int pfm_pcg_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag, const double *packed,
                  const double *rD, const double *b, double *x, void *work, double *history_h,
                  pfm_solve_result *res, void *stream) {
  return pfm::pcg_solve(h, prm, diag, packed, rD, b, x, work, history_h, res, stream, true);
}

}  // extern "C"

namespace pfm {

int pcg_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag, const double *packed,
              const double *rD, const double *b, double *x, void *work, double *history_h,
              pfm_solve_result *res, void *stream, bool may_allocate) {
  int rc = check_solve_args(h, prm, diag, packed, rD, b, x, work, res);
  if (rc != PFM_OK) return rc;
  memset(res, 0, sizeof(*res));
  if (h->n == 0) {
    res->status = PFM_SOLVE_CONVERGED;
    res->converged = 1;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  Work w;
  work_layout(h, prm->max_iter, 4, &w, (char *)work);
  // Two direction buffers: p_it lives in P[(it + 1) & 1] (k_pcg_init writes
  // p_1 to P[0]); k_pcg_pupdate writes p_{it+1} to P[it & 1].
  double *r = w.vec[0], *P[2] = {w.vec[1], w.vec[3]}, *p = P[0], *Ap = w.vec[2];
  const double *pd = prm->precond == PFM_PRECOND_DIAG ? rD : nullptr;
  const int hist_cap = prm->max_iter + 1;
  const int G = red_grid(h);
  const int n = h->n, ns = h->nslices;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  const int16_t *c16 = h->d_col16;
  const uint32_t *cp = h->d_colp2;
  const int cf = pcg_col_form(h);
  if (cf == kColsPair)
    k_pcg_init<kColsPair><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, c16, cp, packed, diag,
                                               pd, b, x, r, p, w.part);
  else if (cf == kCols16)
    k_pcg_init<kCols16><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, c16, cp, packed, diag,
                                             pd, b, x, r, p, w.part);
  else
    k_pcg_init<kCols32><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, c16, cp, packed, diag,
                                             pd, b, x, r, p, w.part);
  k_pcg_init_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
  PFM_CHECK_HIP(hipGetLastError());
  rc = drive(h, w.st, prm->max_iter, s, [&](int it) {
    double *pc = P[(it + 1) & 1], *po = P[it & 1];  // p_it; p_{it+1} (and p_{it-1})
    const double *pprev = (it & 1) ? nullptr : po;  // pairs of x updates end at even it
    if (cf == kColsPair)
      k_pcg_spmv<kColsPair><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, c16, cp, packed,
                                                 diag, pc, Ap, w.part, w.st);
    else if (cf == kCols16)
      k_pcg_spmv<kCols16><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, c16, cp, packed,
                                               diag, pc, Ap, w.part, w.st);
    else
      k_pcg_spmv<kCols32><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, c16, cp, packed,
                                               diag, pc, Ap, w.part, w.st);
    k_pcg_update<<<G, kBlock, 0, s>>>(n, ns, it, pd, Ap, r, x, pprev, w.part, w.st);
    k_pcg_pupdate<<<G, kBlock, 0, s>>>(n, ns, it, pd, r, pc, pprev, po, x, w.part, w.st, w.hist);
  }, may_allocate);
  if (rc != PFM_OK) return rc;
  PFM_CHECK_HIP(hipGetLastError());
  return collect(w.st, w.hist, history_h, hist_cap, res, s);
}

}  // namespace

extern "C" {

int pfm_pbicgstab_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag,
                        const double *packed, const double *rD, const double *b, double *x,
                        void *work, double *history_h, pfm_solve_result *res, void *stream) {
  int rc = check_solve_args(h, prm, diag, packed, rD, b, x, work, res);
  if (rc != PFM_OK) return rc;
  memset(res, 0, sizeof(*res));
  if (h->n == 0) {
    res->status = PFM_SOLVE_CONVERGED;
    res->converged = 1;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  Work w;
  work_layout(h, prm->max_iter, 8, &w, (char *)work);
  double *r = w.vec[0], *rt = w.vec[1], *p = w.vec[2], *v = w.vec[3], *y = w.vec[4],
         *sv = w.vec[5], *z = w.vec[6], *t = w.vec[7];
  const double *pd = prm->precond == PFM_PRECOND_DIAG ? rD : nullptr;
  const int hist_cap = 2 * prm->max_iter + 2;
  const int G = red_grid(h);
  const int n = h->n, ns = h->nslices;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  const uint32_t *cp = pcg_col_form(h) == kColsPair ? h->d_colp2 : nullptr;
  k_bs_init<<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, cp, packed, diag, pd, b, x, r, rt, p,
                                 y, w.part);
  k_bs_init_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
  PFM_CHECK_HIP(hipGetLastError());
  // Per iteration: v = A y | alpha, s, z | monitor(s), t = A z | omega, x, r |
  // monitor(r), rho, beta, p, y.  Partial regions: A = [0, G) rt.v, B = [G, 2G)
  // s.s, C = [2G, 4G) t.t t.z, D = [4G, 6G) r.r rt.r.
  rc = drive(h, w.st, prm->max_iter, s, [&](int it) {
    k_bs_spmv1<<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, cp, packed, diag, y, rt, v, w.part,
                                    w.st);
    k_bs_s<<<G, kBlock, 0, s>>>(n, ns, it, pd, r, v, sv, z, w.part, w.st);
    k_bs_spmv2<<<G, kBlock, 0, s>>>(n, ns, it, h->d_slices, h->d_col, cp, packed, diag, z, t, w.part,
                                    w.st, w.hist);
    k_bs_update<<<G, kBlock, 0, s>>>(n, ns, y, z, sv, t, rt, x, r, w.part, w.st);
    k_bs_pupdate<<<G, kBlock, 0, s>>>(n, ns, it, pd, r, v, p, y, w.part, w.st, w.hist);
  });
  if (rc != PFM_OK) return rc;
  PFM_CHECK_HIP(hipGetLastError());
  return collect(w.st, w.hist, history_h, hist_cap, res, s);
}


int64_t pfm_pbicg_workspace_bytes(pfm_ldu h, int32_t max_iter) {
  if (!h || max_iter < 0) return -1;
  return (int64_t)work_layout(h, max_iter, 5, nullptr, nullptr);
}

int pfm_pbicg_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag,
                    const double *packed, const double *packed_t, const double *rD,
                    const double *b, double *x, void *work, double *history_h,
                    pfm_solve_result *res, void *stream) {
  int rc = check_solve_args(h, prm, diag, packed, rD, b, x, work, res);
  if (rc != PFM_OK) return rc;
  PFM_REQUIRE(h->nslots == 0 || packed_t, "null transposed packed coefficients");
  memset(res, 0, sizeof(*res));
  if (h->n == 0) {
    res->status = PFM_SOLVE_CONVERGED;
    res->converged = 1;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  Work w;
  work_layout(h, prm->max_iter, 5, &w, (char *)work);
  double *r = w.vec[0], *rt = w.vec[1], *p = w.vec[2], *v = w.vec[3], *t = w.vec[4];
  const double *pd = prm->precond == PFM_PRECOND_DIAG ? rD : nullptr;
  const int hist_cap = 2 * prm->max_iter + 2;
  const int G = red_grid(h);
  const int n = h->n, ns = h->nslices;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  const uint32_t *cp = pcg_col_form(h) == kColsPair ? h->d_colp2 : nullptr;
  k_bicg_init<<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, cp, packed, diag, pd, b, x, r, rt, p,
                                   w.part);
  k_bicg_init_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.hist);
  PFM_CHECK_HIP(hipGetLastError());
  rc = drive(h, w.st, prm->max_iter, s, [&](int it) {
    k_bicg_spmv<<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, cp, packed, packed_t, diag, pd, p,
                                     rt, v, t, w.part, w.st);
    k_bicg_update<<<G, kBlock, 0, s>>>(n, ns, it, pd, p, v, t, x, r, rt, w.part, w.st);
    k_bicg_pupdate<<<G, kBlock, 0, s>>>(n, ns, it, pd, r, p, w.part, w.st, w.hist);
  });
  if (rc != PFM_OK) return rc;
  PFM_CHECK_HIP(hipGetLastError());
  return collect(w.st, w.hist, history_h, hist_cap, res, s);
}

}  // extern "C"
