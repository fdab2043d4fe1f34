// Mixed-precision PCG: FP32 matrix, FP32 SpMV and FP32 diagonal DIC apply in
// the inner Krylov loop, wrapped in FP64 iterative refinement.
//
//   r = b - A x                       FP64 SpMV, FP64 ||r|| (outer monitor)
//   while not converged:
//     e = 0; inner PCG on A32 e = (float) r until ||r_in|| < tau ||r||
//        (vectors FP32, every dot product / norm accumulated in FP64,
//         alpha and beta computed in FP64 and rounded once to FP32)
//     x += (double) e; r = b - A x    FP64
//
// The outer stopping rule is the reference's (residual.py:112-145) applied to
// the true FP64 residual b - A x, so the returned x meets the same relative
// tolerance as the FP64 solve.  The host runs the outer loop (one
// synchronisation per refinement step, typically 2-4 per solve); each inner
// loop is a device-resident PCG exactly like the FP64 one.

#include <cmath>
#include <cstring>

#include "solver_state.h"

using namespace pfm;

namespace {

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

template <int K>
__device__ __forceinline__ void store_partials(double (&v)[K], double *part, double *lds) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = block_sum(v[k], lds);
    if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = s;
  }
}

// Inner-loop state (separate from the outer KState).
struct alignas(256) IState {
  double target;   // stop when ||r_in|| < target
  double rzs[2];
  double alphas[2];  // alpha by iteration parity (k_pcg32_update -> k_pcg32_pupdate)
  int32_t done;
  int32_t iters;   // inner iterations run
  int32_t max_iter;
  int32_t pending; // stop decided by k_pcg32_pupdate, made `done` by the next k_pcg32_spmv
};

__global__ __launch_bounds__(kBlock) void k_cast(int n, const double *__restrict__ a,
                                                  float *__restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = (float)a[i];
}

// r = b - A x in FP64; partial r.r (region A).
__global__ __launch_bounds__(kBlock) void k_ir_resid(int n, int nslices, const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ b,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ r,
                                                      double *__restrict__ part) {
  __shared__ double lds[kWavesPerBlock];
  double acc[1] = {0.0};
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const double dx = row < n ? diag[row] * x[row] : 0.0;
    const double ax = row_spmv<double>(si, lane, dx, col, val, x);
    if (row < n) {
      const double ri = b[row] - ax;
      r[row] = ri;
      acc[0] += ri * ri;
    }
  })
  store_partials<1>(acc, part, lds);
}

// Outer monitor: record ||b - A x||, stop or set the inner target.
__global__ __launch_bounds__(kBlock) void k_ir_outer_fin(const double *__restrict__ part, int G,
                                                          int outer, double tau, int inner_budget,
                                                          KState *st, IState *in,
                                                          double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  const double RR = reduce_partials(part, G, lds);
  if (threadIdx.x == 0) {
    const double res = sqrt(RR);
    if (monitor_update(st, hist, res, outer))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (inner_budget <= 0)
      finish(st, PFM_SOLVE_MAXITER);
    in->target = tau * res;
    in->done = 0;
    in->pending = 0;
    in->iters = 0;
    in->max_iter = inner_budget;
  }
}

// Inner start: r32 = (float) r; z = r32 / rD32; p = z; e = 0; partial r.z (region B).
__global__ __launch_bounds__(kBlock) void k_ir_inner_init(int n, int nslices,
                                                           const double *__restrict__ r64,
                                                           const float *__restrict__ rD,
                                                           float *__restrict__ r, float *__restrict__ p,
                                                           float *__restrict__ e,
                                                           double *__restrict__ part,
                                                           const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc[1] = {0.0};
  FOR_ROWS({
    if (row < n) {
      const float ri = (float)r64[row];
      const float zi = rD ? ri / rD[row] : ri;
      r[row] = ri;
      p[row] = zi;
      e[row] = 0.0f;
      acc[0] += (double)ri * (double)zi;
    }
  })
  store_partials<1>(acc, part + gridDim.x, lds);
}

__global__ __launch_bounds__(kBlock) void k_ir_inner_init_fin(const double *__restrict__ part, int G,
                                                               const KState *st, IState *in) {
  __shared__ double lds[kWavesPerBlock];
  const double RZ = reduce_partials(part + G, G, lds);
  if (threadIdx.x == 0) {
    in->rzs[0] = RZ;
    if (st->done) in->done = 1;
  }
}

// Ap = A32 p (FP32); partial p.Ap in FP64 (region A).  kRel: columns from the
// handle's int16 row offsets.  Ends the inner loop if the previous iteration
// decided so.
template <int kCF>
__global__ __launch_bounds__(kBlock) void k_pcg32_spmv(int n, int nslices, const SliceInfo *__restrict__ sl,
                                                        const int32_t *__restrict__ col,
                                                        const int16_t *__restrict__ col16,
                                                        const uint32_t *__restrict__ cp,
                                                        const float *__restrict__ val,
                                                        const float *__restrict__ diag,
                                                        const float *__restrict__ p,
                                                        float *__restrict__ Ap,
                                                        double *__restrict__ part, IState *in) {
  __shared__ double lds[kWavesPerBlock];
  if (in->done) return;
  if (in->pending) {
    if (blockIdx.x == 0 && threadIdx.x == 0) in->done = 1;
    return;
  }
  double acc[1] = {0.0};
  auto gx = [p](int c) { return p[c]; };
  FOR_ROWS({
    const SliceInfo si = load_slice(sl, s);
    const float pi = row < n ? p[row] : 0.0f;
    const float dx = row < n ? diag[row] * pi : 0.0f;
    float v;
    if constexpr (kCF == kColsPair)
      v = row_spmv_pair<float, decltype(gx)>(si, s, lane, row, dx, cp, val, gx);
    else if constexpr (kCF == kCols16)
      v = row_spmv_g<float, decltype(gx), false, ColRel>(si, lane, dx, ColRel{col16, row}, val,
                                                         gx);
    else
      v = row_spmv_g<float>(si, lane, dx, ColAbs{col}, val, gx);
    if (row < n) {
      Ap[row] = v;
      acc[0] += (double)pi * (double)v;
    }
  })
  store_partials<1>(acc, part, lds);
}

// alpha (kept for k_pcg32_pupdate); r -= alpha Ap; partials r.r, r.z (region B).
__global__ __launch_bounds__(kBlock) void k_pcg32_update(int n, int nslices, int it,
                                                          const float *__restrict__ rD,
                                                          const float *__restrict__ Ap,
                                                          float *__restrict__ r,
                                                          double *__restrict__ part, IState *in) {
  __shared__ double lds[kWavesPerBlock];
  if (in->done) return;
  const int G = gridDim.x;
  // first round's loads issued before the reduction (independent of alpha)
  const int lane = threadIdx.x & 63;
  const int stride = G * kWavesPerBlock;
  int s = xcd_remap(blockIdx.x, G) * kWavesPerBlock + (threadIdx.x >> 6);
  int row = s * kSlice + lane;
  float rp = 0.0f, ap = 0.0f, dp = 1.0f;
  if (s < nslices && row < n) {
    rp = r[row];
    ap = Ap[row];
    if (rD) dp = rD[row];
  }
  const double PAP = reduce_partials(part, G, lds);
  if (fabs(PAP) < 1e-30) {
    if (blockIdx.x == 0 && threadIdx.x == 0) in->done = 1;
    return;
  }
  const double alpha64 = in->rzs[(it - 1) & 1] / PAP;
  if (blockIdx.x == 0 && threadIdx.x == 0) in->alphas[it & 1] = alpha64;
  const float alpha = (float)alpha64;
  double acc[2] = {0.0, 0.0};
  for (; s < nslices; s += stride, row = s * kSlice + lane) {
    if (row < n) {
      const float ri = rp - alpha * ap;
      r[row] = ri;
      const float zi = rD ? ri / dp : ri;
      acc[0] += (double)ri * (double)ri;
      acc[1] += (double)ri * (double)zi;
    }
    const int sn = s + stride, rn = sn * kSlice + lane;
    if (sn < nslices && rn < n) {
      rp = r[rn];
      ap = Ap[rn];
      if (rD) dp = rD[rn];
    }
  }
  store_partials<2>(acc, part + G, lds);
}

// inner stop test; e += alpha p; beta; p = z + beta p (not after the last
// inner iteration).  The stop is left in `pending`: every workgroup still
// updates e in this kernel.
__global__ __launch_bounds__(kBlock) void k_pcg32_pupdate(int n, int nslices, int it,
                                                           const float *__restrict__ rD,
                                                           const float *__restrict__ r,
                                                           float *__restrict__ p,
                                                           float *__restrict__ e,
                                                           const double *__restrict__ part,
                                                           IState *in) {
  __shared__ double lds[kWavesPerBlock];
  if (in->done) return;  // set by an earlier kernel only (a breakdown in k_pcg32_update)
  const int G = gridDim.x;
  const int lane = threadIdx.x & 63;
  const int stride = G * kWavesPerBlock;
  int s = xcd_remap(blockIdx.x, G) * kWavesPerBlock + (threadIdx.x >> 6);
  int row = s * kSlice + lane;
  float rp = 0.0f, dp = 1.0f, pp = 0.0f, ep = 0.0f;
  if (s < nslices && row < n) {
    rp = r[row];
    if (rD) dp = rD[row];
    pp = p[row];
    ep = e[row];
  }
  const double RR = reduce_partials(part + G, G, lds);
  const double RZ = reduce_partials(part + 2 * G, G, lds);
  const bool stop = sqrt(RR) < in->target || it >= in->max_iter;
  const double rz_old = in->rzs[(it - 1) & 1];
  const float alpha = (float)in->alphas[it & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    in->iters = it;
    in->rzs[it & 1] = RZ;
    if (stop) in->pending = 1;
  }
  const float beta = (float)(RZ / rz_old);
  for (; s < nslices; s += stride, row = s * kSlice + lane) {
    if (row < n) {
      e[row] = ep + alpha * pp;
      if (!stop) {
        const float zi = rD ? rp / dp : rp;
        p[row] = zi + beta * pp;
      }
    }
    const int sn = s + stride, rn = sn * kSlice + lane;
    if (sn < nslices && rn < n) {
      rp = r[rn];
      if (rD) dp = rD[rn];
      pp = p[rn];
      ep = e[rn];
    }
  }
}

// x += (double) e
__global__ __launch_bounds__(kBlock) void k_ir_correct(int n, const float *__restrict__ e,
                                                        double *__restrict__ x, const KState *st) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + (double)e[i];
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct MixedWork {
  KState *st;
  IState *in;
  double *part, *hist, *r64;
  float *diag32, *rD32, *r, *p, *Ap, *e;
};

size_t mixed_layout(const pfm_ldu_s *h, int max_outer, MixedWork *w, char *base) {
  const int G = std::max(1, std::min(blocks_for_slices(h->nslices), kMaxRedBlocks));
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  MixedWork t;
  t.st = (KState *)take(sizeof(KState));
  t.in = (IState *)take(sizeof(IState));
  t.part = (double *)take(sizeof(double) * 3 * (size_t)G);
  t.hist = (double *)take(sizeof(double) * ((size_t)std::max(max_outer, 0) + 2));
  const size_t n = (size_t)h->n;
  t.r64 = (double *)take(sizeof(double) * n);
  t.diag32 = (float *)take(sizeof(float) * n);
  t.rD32 = (float *)take(sizeof(float) * n);
  t.r = (float *)take(sizeof(float) * n);
  t.p = (float *)take(sizeof(float) * n);
  t.Ap = (float *)take(sizeof(float) * n);
  t.e = (float *)take(sizeof(float) * n);
  if (w) *w = t;
  return off;
}

}  // namespace

extern "C" {

int64_t pfm_pcg_mixed_workspace_bytes(pfm_ldu h, int32_t max_outer) {
  if (!h) return -1;
  return (int64_t)mixed_layout(h, max_outer, nullptr, nullptr);
}

int pfm_pcg_mixed_solve(pfm_ldu h, const pfm_solve_params *prm, double inner_rel_tol,
                        int32_t max_outer, const double *diag, const double *packed,
                        const float *packed32, const double *rD, const double *b, double *x,
                        void *work, double *history_h, pfm_solve_result *res,
                        int32_t *inner_iterations, void *stream) {
  PFM_REQUIRE(h && prm && res && work, "null argument");
  PFM_REQUIRE(prm->max_iter >= 0 && max_outer >= 0, "negative iteration limit");
  PFM_REQUIRE(inner_rel_tol > 0.0 && inner_rel_tol < 1.0, "inner_rel_tol must be in (0, 1)");
  PFM_REQUIRE(prm->precond == PFM_PRECOND_NONE || rD, "preconditioner diagonal missing");
  PFM_REQUIRE(h->n == 0 || (diag && b && x), "null device pointer");
  PFM_REQUIRE(h->nslots == 0 || (packed && packed32), "null packed coefficients");
  memset(res, 0, sizeof(*res));
  if (inner_iterations) *inner_iterations = 0;
  if (h->n == 0) {
    res->status = PFM_SOLVE_CONVERGED;
    res->converged = 1;
    return PFM_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  MixedWork w;
  mixed_layout(h, max_outer, &w, (char *)work);
  const int n = h->n, ns = h->nslices;
  const int G = std::max(1, std::min(blocks_for_slices(ns), kMaxRedBlocks));
  const int g1 = (n + kBlock - 1) / kBlock;
  const int cf = krylov_col_form(h);
  const bool pre = prm->precond == PFM_PRECOND_DIAG;
  k_cast<<<g1, kBlock, 0, s>>>(n, diag, w.diag32);
  if (pre) k_cast<<<g1, kBlock, 0, s>>>(n, rD, w.rD32);
  const float *pd = pre ? w.rD32 : nullptr;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, max_outer, prm->min_iter,
                               max_outer + 1);
  PFM_CHECK_HIP(hipGetLastError());
  int inner_total = 0, rc = PFM_OK;
  KState hs;
  IState is;
  PollLease lease;
  if ((rc = lease.acquire(&h->poll, true)) != PFM_OK) return rc;
  for (int outer = 0;; ++outer) {
    k_ir_resid<<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, packed, diag, b, x, w.r64,
                                    w.part);
    const int budget = std::min(prm->max_iter - inner_total, prm->max_iter);
    k_ir_outer_fin<<<1, kBlock, 0, s>>>(w.part, G, outer, inner_rel_tol,
                                        outer >= max_outer ? 0 : budget, w.st, w.in, w.hist);
    PFM_CHECK_HIP(hipMemcpyAsync(&hs, w.st, sizeof(KState), hipMemcpyDeviceToHost, s));
    PFM_CHECK_HIP(hipStreamSynchronize(s));
    if (hs.done) break;
    // inner FP32 PCG
    k_ir_inner_init<<<G, kBlock, 0, s>>>(n, ns, w.r64, pd, w.r, w.p, w.e, w.part, w.st);
    k_ir_inner_init_fin<<<1, kBlock, 0, s>>>(w.part, G, w.st, w.in);
    int32_t *flag = lease.flag;
    hipEvent_t *ev = lease.ev;
    int enq = 0, chunk = 4, slot = 0;
    bool pending = false;
    double last_report = now_seconds();
    while (true) {
      bool queued = false;
      report_progress(h, enq, last_report);
      if (enq < budget) {
        const int c = std::min(chunk, budget - enq);
        for (int k = 0; k < c; ++k) {
          const int it = enq + k + 1;
          if (cf == kColsPair)
            k_pcg32_spmv<kColsPair><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, h->d_col16,
                                                         h->d_colp2, packed32, w.diag32, w.p, w.Ap,
                                                         w.part, w.in);
          else if (cf == kCols16)
            k_pcg32_spmv<kCols16><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, h->d_col16,
                                                       nullptr, packed32, w.diag32, w.p, w.Ap,
                                                       w.part, w.in);
          else
            k_pcg32_spmv<kCols32><<<G, kBlock, 0, s>>>(n, ns, h->d_slices, h->d_col, nullptr,
                                                       nullptr, packed32, w.diag32, w.p, w.Ap,
                                                       w.part, w.in);
          k_pcg32_update<<<G, kBlock, 0, s>>>(n, ns, it, pd, w.Ap, w.r, w.part, w.in);
          k_pcg32_pupdate<<<G, kBlock, 0, s>>>(n, ns, it, pd, w.r, w.p, w.e, w.part, w.in);
        }
        enq += c;
        chunk = std::min(chunk * 2, 32);
        if (hipMemcpyAsync(&flag[slot], &w.in->done, sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
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
    if (rc != PFM_OK) {
      set_error("HIP error while driving the inner loop");
      return rc;
    }
    k_ir_correct<<<g1, kBlock, 0, s>>>(n, w.e, x, w.st);
    PFM_CHECK_HIP(hipMemcpyAsync(&is, w.in, sizeof(IState), hipMemcpyDeviceToHost, s));
    PFM_CHECK_HIP(hipStreamSynchronize(s));
    inner_total += is.iters;
    if (is.iters == 0) {  // breakdown at the first inner step: no progress possible
      (void)hipMemcpyAsync(&hs, w.st, sizeof(KState), hipMemcpyDeviceToHost, s);
      break;
    }
  }
  PFM_CHECK_HIP(hipMemcpyAsync(&hs, w.st, sizeof(KState), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  const int nh = std::min(hs.nhist, max_outer + 1);
  if (history_h && nh > 0) {
    PFM_CHECK_HIP(hipMemcpyAsync(history_h, w.hist, sizeof(double) * nh, hipMemcpyDeviceToHost, s));
    PFM_CHECK_HIP(hipStreamSynchronize(s));
  }
  res->iterations = hs.nhist;
  res->status = hs.status == 0 ? PFM_SOLVE_BREAKDOWN : hs.status;
  res->converged = res->status == PFM_SOLVE_CONVERGED;
  res->initial_residual = hs.nhist > 0 ? hs.res0 : 0.0;
  res->final_residual = hs.nhist > 0 ? hs.res : 0.0;
  if (inner_iterations) *inner_iterations = inner_total;
  return PFM_OK;
}

}  // extern "C"
