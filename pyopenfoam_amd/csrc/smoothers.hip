// Smoothers on the LDU handle: symmetric Gauss-Seidel (level-scheduled, the
// sequential sweeps' exact arithmetic) and damped Jacobi.

#include "device_common.h"
#include "pfm_internal.h"
#include "solver_state.h"

using namespace pfm;

namespace {

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// One sweep level: its cells are mutually independent.
__global__ __launch_bounds__(kBlock) void k_gs_level(const GsEnt *__restrict__ ents, int count,
                                                      const int32_t *__restrict__ col,
                                                      const int32_t *__restrict__ map,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ b, double *x,
                                                      double omega) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < count) gs_entry(ents[t], col, map, val, diag, b, x, omega, true);
}

constexpr int kGsThreads = 1024;

// Small meshes: one workgroup walks every level of n_sweeps symmetric sweeps,
// with a barrier between levels.
__global__ __launch_bounds__(kGsThreads) void k_gs_sweeps_wg(
    int n_sweeps, const int32_t *__restrict__ fptr, int nfl, const GsEnt *__restrict__ fents,
    const int32_t *__restrict__ bptr, int nbl, const GsEnt *__restrict__ bents,
    const int32_t *__restrict__ col,
    const int32_t *__restrict__ map, const double *__restrict__ val,
    const double *__restrict__ diag, const double *__restrict__ b, double *x, double omega) {
  for (int s = 0; s < n_sweeps; ++s) {
    for (int pass = 0; pass < 2; ++pass) {
      const int32_t *ptr = pass ? bptr : fptr;
      const GsEnt *ents = pass ? bents : fents;
      const int nl = pass ? nbl : nfl;
      for (int l = 0; l < nl; ++l) {
        for (int k = ptr[l] + (int)threadIdx.x; k < ptr[l + 1]; k += kGsThreads)
          gs_entry(ents[k], col, map, val, diag, b, x, omega, true);
        __syncthreads();
      }
    }
  }
}

// One damped Jacobi step (solvers/smoothers.py:196-198):
// xout = xin + (omega * 1/max(|d|, 1e-30)) * (b - A xin).
__global__ __launch_bounds__(kBlock) void k_jacobi_step(int n, int nslices,
                                                         const SliceInfo *__restrict__ sl,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ diag,
                                                         const double *__restrict__ b,
                                                         double omega,
                                                         const double *__restrict__ xin,
                                                         double *__restrict__ xout) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  const double dx = row < n ? diag[row] * xin[row] : 0.0;
  const double ax = row_spmv<double>(si, lane, dx, col, val, xin);
  if (row < n) {
    const double r = b[row] - ax;
    const double inv = 1.0 / fmax(fabs(diag[row]), 1e-30);
    xout[row] = xin[row] + (omega * inv) * r;
  }
}


// ---------------------------------------------------------------------------
// DIC / DILU apply_full (preconditioners.py:153-201, :290-337): forward face
// sweep z[n] -= (upper_f z[p]) / rD[p], z /= rD, reverse face sweep
// z[p] -= (c_f z[n]) / rD[n] with c = upper (DIC) or lower (DILU); DIC skips
// faces with owner > neighbour.  |rD| <= 1e-30 skips a face.

// Faces oriented owner < neighbour with owners non-decreasing: a cell's
// forward updates are its neighbour-faces (the row's upper slots, face
// order) and its backward updates its owner-faces in reverse face order;
// cells are scheduled by dependency level.
__global__ __launch_bounds__(kBlock) void k_af_fwd(const int32_t *__restrict__ cells, int count,
                                                    const SliceInfo *__restrict__ sl,
                                                    const int32_t *__restrict__ col,
                                                    const int32_t *__restrict__ map,
                                                    const double *__restrict__ upper,
                                                    const double *__restrict__ rD, double *z) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int n = cells[t];
  const SliceInfo si = load_slice_lane(sl, n / kSlice);
  int slot = si.ptr + si.wl * kSlice + (n % kSlice);
  double zn = z[n];
  for (int j = 0; j < si.wu; ++j, slot += kSlice) {
    const int p = col[slot];
    if (p < 0) break;
    const double d = rD[p];
    if (fabs(d) > 1e-30) zn = zn - (upper[map[slot]] * z[p]) / d;
  }
  z[n] = zn;
}

__global__ __launch_bounds__(kBlock) void k_af_bwd(const int32_t *__restrict__ cells, int count,
                                                    const SliceInfo *__restrict__ sl,
                                                    const int32_t *__restrict__ col,
                                                    const int32_t *__restrict__ map,
                                                    const double *__restrict__ coef,
                                                    const double *__restrict__ rD, double *z) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int p = cells[t];
  const SliceInfo si = load_slice_lane(sl, p / kSlice);
  const int base = si.ptr + (p % kSlice);
  int k = si.wl - 1;
  while (k >= 0 && col[base + k * kSlice] < 0) --k;  // last owner-face of the row
  double zp = z[p];
  for (; k >= 0; --k) {
    const int slot = base + k * kSlice;
    const int n = col[slot];
    const double d = rD[n];
    if (fabs(d) > 1e-30) zp = zp - (coef[map[slot]] * z[n]) / d;
  }
  z[p] = zp;
}

__global__ __launch_bounds__(kBlock) void k_af_scale(int n, const double *__restrict__ rD,
                                                      double *__restrict__ z) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) z[i] = z[i] / rD[i];
}

// Any face order: one workgroup walks the face levels (faces of a level touch
// distinct cells; a face waits for every earlier face touching its cells).
__global__ __launch_bounds__(kGsThreads) void k_af_faces_wg(
    const int32_t *__restrict__ lptr, int nl, const int32_t *__restrict__ faces,
    const int32_t *__restrict__ own, const int32_t *__restrict__ nei,
    const double *__restrict__ upper, const double *__restrict__ coef_b,
    const double *__restrict__ rD, double *z, int n, int dic) {
  for (int l = 0; l < nl; ++l) {
    for (int k = lptr[l] + (int)threadIdx.x; k < lptr[l + 1]; k += kGsThreads) {
      const int f = faces[k];
      const int p = own[f], q = nei[f];
      if (dic && p >= q) continue;
      const double d = rD[p];
      if (fabs(d) > 1e-30) z[q] = z[q] - (upper[f] * z[p]) / d;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += kGsThreads) z[i] = z[i] / rD[i];
  __syncthreads();
  for (int l = nl - 1; l >= 0; --l) {
    for (int k = lptr[l] + (int)threadIdx.x; k < lptr[l + 1]; k += kGsThreads) {
      const int f = faces[k];
      const int p = own[f], q = nei[f];
      if (dic && p >= q) continue;
      const double d = rD[q];
      if (fabs(d) > 1e-30) z[p] = z[p] - (coef_b[f] * z[q]) / d;
    }
    __syncthreads();
  }
}

// x = b / max(|diag|, 1e-30); partial of r.r with r = b - diag x
// (solvers/diagonal.py:74-84).
__global__ __launch_bounds__(kBlock) void k_diag_solve(int n, const double *__restrict__ diag,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ x,
                                                        double *__restrict__ part) {
  __shared__ double lds[kWavesPerBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
    const double xi = b[i] / fmax(fabs(diag[i]), 1e-30);
    x[i] = xi;
    const double ri = b[i] - diag[i] * xi;
    acc = ri * ri;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kBlock) void k_sum_sqrt(const double *__restrict__ part, int np,
                                                      double *__restrict__ out) {
  __shared__ double lds[kWavesPerBlock];
  const double s = sum_partials(part, np, lds);
  if (threadIdx.x == 0) *out = sqrt(s);
}

// ---------------------------------------------------------------------------
// SmoothSolver (smooth_solver.py:96-232).

// One level of a forward Gauss-Seidel sweep with direct assignment.
__global__ __launch_bounds__(kBlock) void k_sm_gs_level(const GsEnt *__restrict__ ents, int count,
                                                         const int32_t *__restrict__ col,
                                                         const int32_t *__restrict__ map,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ diag,
                                                         const double *__restrict__ b, double *x,
                                                         const KState *st) {
  const bool live = !st->done;  // checked before the store, off the load chain
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < count) gs_entry<double, true>(ents[t], col, map, val, diag, b, x, 1.0, live);
}

__global__ __launch_bounds__(kGsThreads) void k_sm_gs_wg(int n_sweeps, const int32_t *__restrict__ fptr,
                                                          int nfl, const GsEnt *__restrict__ fents,
                                                          const int32_t *__restrict__ col,
                                                          const int32_t *__restrict__ map,
                                                          const double *__restrict__ val,
                                                          const double *__restrict__ diag,
                                                          const double *__restrict__ b, double *x,
                                                          const KState *st) {
  if (st->done) return;
  for (int sw = 0; sw < n_sweeps; ++sw)
    for (int l = 0; l < nfl; ++l) {
      for (int k = fptr[l] + (int)threadIdx.x; k < fptr[l + 1]; k += kGsThreads)
        gs_entry<double, true>(fents[k], col, map, val, diag, b, x, 1.0, true);
      __syncthreads();
    }
}

// r = b - A x; optional partials of r.r.
__global__ __launch_bounds__(kBlock) void k_sm_resid(int n, int nslices,
                                                      const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ b,
                                                      double *__restrict__ r,
                                                      double *__restrict__ part, const KState *st) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  double acc = 0.0;
  const SliceRange sr = block_slices(nslices);
  const int lane = threadIdx.x & 63;
  for (int s = sr.begin + (threadIdx.x >> 6); s < sr.end; s += kWavesPerBlock) {
    const SliceInfo si = load_slice(sl, s);
    const int row = s * kSlice + lane;
    const double dx = row < n ? diag[row] * x[row] : 0.0;
    const double v = row_spmv<double>(si, lane, dx, col, val, x);
    if (row < n) {
      const double ri = b[row] - v;
      if (r) r[row] = ri;
      acc += ri * ri;
    }
  }
  if (part) {
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_sm_fin(const double *__restrict__ part, int G, int it,
                                                    KState *st, double *__restrict__ hist) {
  __shared__ double lds[kWavesPerBlock];
  if (st->done) return;
  const double RR = sum_partials(part, G, lds);
  if (threadIdx.x == 0) {
    st->iter = it;
    if (monitor_update(st, hist, sqrt(RR), it))
      finish(st, PFM_SOLVE_CONVERGED);
    else if (it >= st->max_iter)
      finish(st, PFM_SOLVE_MAXITER);
  }
}

__global__ __launch_bounds__(kBlock) void k_sm_axpy(int n, double omega, const double *__restrict__ z,
                                                     double *__restrict__ x, const KState *st) {
  if (st->done) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + omega * z[i];
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct SmoothWork {
  KState *st;
  double *part, *hist, *r, *z;
};

inline int sm_grid(const pfm_ldu_s *h) {
  return std::max(1, std::min(blocks_for_slices(h->nslices), kMaxRedBlocks));
}

size_t smooth_layout(const pfm_ldu_s *h, int max_iter, SmoothWork *w, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> void * {
    size_t o = off;
    off += align256(bytes);
    return base ? (void *)(base + o) : nullptr;
  };
  SmoothWork t;
  t.st = (KState *)take(sizeof(KState));
  t.part = (double *)take(sizeof(double) * (size_t)sm_grid(h));
  t.hist = (double *)take(sizeof(double) * ((size_t)std::max(max_iter, 0) + 1));
  t.r = (double *)take(sizeof(double) * (size_t)h->n);
  t.z = (double *)take(sizeof(double) * (size_t)h->n);
  if (w) *w = t;
  return off;
}
}  // namespace

extern "C" {

int pfm_gs_smooth(pfm_ldu h, const double *diag, const double *packed, const double *b, double *x,
                  int32_t n_sweeps, double omega, void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(n_sweeps >= 0, "n_sweeps must be >= 0");
  if (h->n == 0 || n_sweeps == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(b);
  PFM_NONNULL(x);
  PFM_REQUIRE(h->nslots == 0 || packed, "null packed coefficients");
  hipStream_t s = (hipStream_t)stream;
  int st = ensure_gs_entries(h, s);
  if (st != PFM_OK) return st;
  const int nfl = (int)h->gs_fwd_ptr.size() - 1, nbl = (int)h->gs_bwd_ptr.size() - 1;
  const int levels = std::max(nfl, nbl);
  if (h->n < 2048 * levels) {
    k_gs_sweeps_wg<<<1, kGsThreads, 0, s>>>(n_sweeps, h->d_gs_fwd_ptr, nfl, h->d_gs_fwd_ent,
                                            h->d_gs_bwd_ptr, nbl, h->d_gs_bwd_ent, h->d_col,
                                            h->d_map, packed, diag, b, x, omega);
  } else {
    for (int sw = 0; sw < n_sweeps; ++sw) {
      for (int pass = 0; pass < 2; ++pass) {
        const std::vector<int32_t> &ptr = pass ? h->gs_bwd_ptr : h->gs_fwd_ptr;
        const GsEnt *ents = pass ? h->d_gs_bwd_ent : h->d_gs_fwd_ent;
        for (size_t l = 0; l + 1 < ptr.size(); ++l) {
          const int cnt = ptr[l + 1] - ptr[l];
          k_gs_level<<<grid1d(cnt), kBlock, 0, s>>>(ents + ptr[l], cnt, h->d_col, h->d_map, packed,
                                                    diag, b, x, omega);
        }
      }
    }
  }
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_jacobi_smooth(pfm_ldu h, const double *diag, const double *packed, const double *b,
                      double *x, int32_t n_iter, double omega, double *work, void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(n_iter >= 0, "n_iter must be >= 0");
  if (h->n == 0 || n_iter == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(b);
  PFM_NONNULL(x);
  PFM_NONNULL(work);
  PFM_REQUIRE(h->nslots == 0 || packed, "null packed coefficients");
  hipStream_t s = (hipStream_t)stream;
  double *buf[2] = {x, work};
  for (int k = 0; k < n_iter; ++k)
    k_jacobi_step<<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
        h->n, h->nslices, h->d_slices, h->d_col, packed, diag, b, omega, buf[k & 1],
        buf[(k + 1) & 1]);
  if (n_iter & 1)
    PFM_CHECK_HIP(hipMemcpyAsync(x, work, sizeof(double) * (size_t)h->n,
                                 hipMemcpyDeviceToDevice, s));
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}


int pfm_precond_apply_full(pfm_ldu h, int dilu, const double *rD, const double *lower,
                           const double *upper, const double *r, double *z, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(rD);
  PFM_NONNULL(r);
  PFM_NONNULL(z);
  PFM_REQUIRE(h->nf == 0 || (upper && (lower || !dilu)), "null face coefficients");
  hipStream_t s = (hipStream_t)stream;
  if (z != r)
    PFM_CHECK_HIP(hipMemcpyAsync(z, r, sizeof(double) * (size_t)h->n, hipMemcpyDeviceToDevice, s));
  const double *coef_b = dilu ? lower : upper;
  if (h->nf == 0) {
    k_af_scale<<<grid1d(h->n), kBlock, 0, s>>>(h->n, rD, z);
  } else if (h->oriented_sorted) {
    for (size_t l = 0; l + 1 < h->dic_level_ptr.size(); ++l) {
      const int b = h->dic_level_ptr[l], cnt = h->dic_level_ptr[l + 1] - b;
      k_af_fwd<<<grid1d(cnt), kBlock, 0, s>>>(h->d_dic_cells + b, cnt, h->d_slices, h->d_col,
                                               h->d_map, upper, rD, z);
    }
    k_af_scale<<<grid1d(h->n), kBlock, 0, s>>>(h->n, rD, z);
    for (size_t l = 0; l + 1 < h->gs_bwd_ptr.size(); ++l) {
      const int b = h->gs_bwd_ptr[l], cnt = h->gs_bwd_ptr[l + 1] - b;
      k_af_bwd<<<grid1d(cnt), kBlock, 0, s>>>(h->d_gs_bwd_cells + b, cnt, h->d_slices, h->d_col,
                                               h->d_map, coef_b, rD, z);
    }
  } else {
    const int nl = (int)h->dilu_level_ptr.size() - 1;
    k_af_faces_wg<<<1, kGsThreads, 0, s>>>(h->d_dilu_level_ptr, nl, h->d_dilu_faces, h->d_own,
                                           h->d_nei, upper, coef_b, rD, z, h->n, dilu ? 0 : 1);
  }
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_diagonal_solve(pfm_ldu h, const double *diag, const double *b, double *x,
                       double *work, double *residual_h, void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(residual_h, "null residual output");
  *residual_h = 0.0;
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(b);
  PFM_NONNULL(x);
  PFM_NONNULL(work);
  hipStream_t s = (hipStream_t)stream;
  const int G = grid1d(h->n);
  k_diag_solve<<<G, kBlock, 0, s>>>(h->n, diag, b, x, work);
  k_sum_sqrt<<<1, kBlock, 0, s>>>(work, G, work + G);
  PFM_CHECK_HIP(hipMemcpyAsync(residual_h, work + G, sizeof(double), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  return PFM_OK;
}


int64_t pfm_smooth_workspace_bytes(pfm_ldu h, int32_t max_iter) {
  if (!h || max_iter < 0) return -1;
  return (int64_t)smooth_layout(h, max_iter, nullptr, nullptr);
}

int pfm_smooth_solve(pfm_ldu h, const pfm_solve_params *prm, int smoother, int32_t n_sweeps,
                     double omega, const double *diag, const double *lower, const double *upper,
                     const double *packed, const double *rD, const double *b, double *x,
                     void *work, double *history_h, pfm_solve_result *res, void *stream) {
  PFM_REQUIRE(h && prm && res && work, "null argument");
  PFM_REQUIRE(prm->max_iter >= 0 && n_sweeps >= 0, "negative iteration or sweep count");
  PFM_REQUIRE(smoother >= PFM_SMOOTH_GAUSS_SEIDEL && smoother <= PFM_SMOOTH_DILU,
              "unknown smoother %d", smoother);
  memset(res, 0, sizeof(*res));
  if (h->n == 0) {
    res->status = PFM_SOLVE_CONVERGED;
    res->converged = 1;
    return PFM_OK;
  }
  PFM_NONNULL(diag);
  PFM_NONNULL(b);
  PFM_NONNULL(x);
  PFM_REQUIRE(h->nslots == 0 || packed, "null packed coefficients");
  PFM_REQUIRE(smoother == PFM_SMOOTH_GAUSS_SEIDEL || rD, "preconditioner diagonal missing");
  PFM_REQUIRE(h->nf == 0 || (lower && upper), "null face coefficients");
  hipStream_t s = (hipStream_t)stream;
  if (smoother == PFM_SMOOTH_GAUSS_SEIDEL) {
    const int est = ensure_gs_entries(h, s);
    if (est != PFM_OK) return est;
  }
  SmoothWork w;
  smooth_layout(h, prm->max_iter, &w, (char *)work);
  const int hist_cap = prm->max_iter + 1;
  k_state_init<<<1, 1, 0, s>>>(w.st, prm->tolerance, prm->rel_tol, prm->max_iter, prm->min_iter,
                               hist_cap);
  const int G = sm_grid(h), nsl = h->nslices, n = h->n;
  auto residual_check = [&](int it) {
    k_sm_resid<<<G, kBlock, 0, s>>>(n, nsl, h->d_slices, h->d_col, packed, diag, x, b, nullptr,
                                    w.part, w.st);
    k_sm_fin<<<1, kBlock, 0, s>>>(w.part, G, it, w.st, w.hist);
  };
  residual_check(0);
  PFM_CHECK_HIP(hipGetLastError());
  const int nfl = (int)h->gs_fwd_ptr.size() - 1;
  const bool wg = h->n < 2048 * std::max(nfl, 1);
  PollLease lease;
  int rc = lease.acquire(&h->poll, true);
  if (rc != PFM_OK) return rc;
  int32_t *flag = lease.flag;
  hipEvent_t *ev = lease.ev;
  int enq = 0, chunk = 2, slot = 0;
  bool pending = false;
  double last_report = now_seconds();
  while (rc == PFM_OK) {
    bool queued = false;
    report_progress(h, enq, last_report);
    if (enq < prm->max_iter) {
      const int c = std::min(chunk, prm->max_iter - enq);
      for (int k = 0; k < c; ++k) {
        const int it = enq + k + 1;
        if (smoother == PFM_SMOOTH_GAUSS_SEIDEL) {
          if (wg) {
            k_sm_gs_wg<<<1, kGsThreads, 0, s>>>(n_sweeps, h->d_gs_fwd_ptr, nfl, h->d_gs_fwd_ent,
                                                h->d_col, h->d_map, packed, diag, b, x, w.st);
          } else {
            for (int sw = 0; sw < n_sweeps; ++sw)
              for (int l = 0; l < nfl; ++l) {
                const int bb = h->gs_fwd_ptr[l], cnt = h->gs_fwd_ptr[l + 1] - bb;
                k_sm_gs_level<<<grid1d(cnt), kBlock, 0, s>>>(h->d_gs_fwd_ent + bb, cnt, h->d_col,
                                                             h->d_map, packed, diag, b, x, w.st);
              }
          }
        } else {
          for (int sw = 0; sw < n_sweeps; ++sw) {
            k_sm_resid<<<G, kBlock, 0, s>>>(n, nsl, h->d_slices, h->d_col, packed, diag, x, b, w.r,
                                            nullptr, w.st);
            int st = pfm_precond_apply_full(h, smoother == PFM_SMOOTH_DILU, rD, lower, upper, w.r,
                                            w.z, stream);
            if (st != PFM_OK) {
              rc = st;
              break;
            }
            k_sm_axpy<<<grid1d(n), kBlock, 0, s>>>(n, omega, w.z, x, w.st);
          }
        }
        residual_check(it);
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
