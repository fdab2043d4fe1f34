// Smoothers on the LDU handle: symmetric Gauss-Seidel (level-scheduled, the
// sequential sweeps' exact arithmetic) and damped Jacobi.

#include "device_common.h"
#include "pfm_internal.h"

using namespace pfm;

namespace {

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// One sweep level: its cells are mutually independent.
__global__ __launch_bounds__(kBlock) void k_gs_level(const int32_t *__restrict__ cells, int count,
                                                      const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const int32_t *__restrict__ map,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ b, double *x,
                                                      double omega) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t < count) gs_cell(cells[t], sl, col, map, val, diag, b, x, omega);
}

constexpr int kGsThreads = 1024;

// Small meshes: one workgroup walks every level of n_sweeps symmetric sweeps,
// with a barrier between levels.
__global__ __launch_bounds__(kGsThreads) void k_gs_sweeps_wg(
    int n_sweeps, const int32_t *__restrict__ fptr, int nfl, const int32_t *__restrict__ fcells,
    const int32_t *__restrict__ bptr, int nbl, const int32_t *__restrict__ bcells,
    const SliceInfo *__restrict__ sl, const int32_t *__restrict__ col,
    const int32_t *__restrict__ map, const double *__restrict__ val,
    const double *__restrict__ diag, const double *__restrict__ b, double *x, double omega) {
  for (int s = 0; s < n_sweeps; ++s) {
    for (int pass = 0; pass < 2; ++pass) {
      const int32_t *ptr = pass ? bptr : fptr;
      const int32_t *cells = pass ? bcells : fcells;
      const int nl = pass ? nbl : nfl;
      for (int l = 0; l < nl; ++l) {
        for (int k = ptr[l] + (int)threadIdx.x; k < ptr[l + 1]; k += kGsThreads)
          gs_cell(cells[k], sl, col, map, val, diag, b, x, omega);
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
  const int nfl = (int)h->gs_fwd_ptr.size() - 1, nbl = (int)h->gs_bwd_ptr.size() - 1;
  const int levels = std::max(nfl, nbl);
  if (h->n < 2048 * levels) {
    k_gs_sweeps_wg<<<1, kGsThreads, 0, s>>>(n_sweeps, h->d_gs_fwd_ptr, nfl, h->d_gs_fwd_cells,
                                            h->d_gs_bwd_ptr, nbl, h->d_gs_bwd_cells, h->d_slices,
                                            h->d_col, h->d_map, packed, diag, b, x, omega);
  } else {
    for (int sw = 0; sw < n_sweeps; ++sw) {
      for (int pass = 0; pass < 2; ++pass) {
        const std::vector<int32_t> &ptr = pass ? h->gs_bwd_ptr : h->gs_fwd_ptr;
        const int32_t *cells = pass ? h->d_gs_bwd_cells : h->d_gs_fwd_cells;
        for (size_t l = 0; l + 1 < ptr.size(); ++l) {
          const int cnt = ptr[l + 1] - ptr[l];
          k_gs_level<<<grid1d(cnt), kBlock, 0, s>>>(cells + ptr[l], cnt, h->d_slices, h->d_col,
                                                    h->d_map, packed, diag, b, x, omega);
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

}  // extern "C"
