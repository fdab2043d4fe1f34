// Device-resident solver state shared by the Krylov and GAMG drivers: the
// residual monitor of the reference (solvers/residual.py:112-145) evaluated
// by one device thread, and the stop flag polled by the host.
#pragma once

#include <cstring>

#include "device_common.h"

namespace pfm {

struct alignas(256) KState {
  double tol, rel_tol;
  double res0, res;
  double rz, alpha, beta;    // PCG
  double rzs[2];             // PCG r.z by iteration parity (written while read)
  double alphas[2];          // PCG alpha by iteration parity
  double rho, omega;         // PBiCGStab (rho kept per iteration parity in rhos)
  double rhos[2];
  int32_t max_iter, min_iter;
  int32_t iter;              // current outer iteration i
  int32_t nhist;             // residual history length
  int32_t done;
  int32_t status;
  int32_t branch;            // PBiCGStab: 0 normal, 1 t.t ~ 0, 2 converged on s
  int32_t hist_cap;
  int32_t pending;           // PCG: status decided by the last kernel of an
                             // iteration, made `done` by the next kernel
};

// ResidualMonitor.update (residual.py:112-145), evaluated by one thread.
__device__ inline bool monitor_update(KState *st, double *hist, double res, int it) {
  if (st->nhist < st->hist_cap) hist[st->nhist] = res;
  st->nhist += 1;
  if (st->nhist == 1) st->res0 = res;
  st->res = res;
  if (it >= st->min_iter) {
    if (res < st->tol) return true;
    if (st->res0 > 0.0) {
      if (res / st->res0 < st->rel_tol) return true;
    }
  }
  return false;
}

// The convergence decision of ResidualMonitor.update alone (no side effects),
// so that every workgroup of a kernel can take it; st->res0 must be set.
__device__ __forceinline__ bool monitor_decision(const KState *st, double res, int it) {
  if (it >= st->min_iter) {
    if (res < st->tol) return true;
    if (st->res0 > 0.0) {
      if (res / st->res0 < st->rel_tol) return true;
    }
  }
  return false;
}

// Partials of at most kMaxRedBlocks workgroups summed in a fixed order by
// every workgroup that needs the scalar: all loads issued before the adds.
__device__ __forceinline__ double reduce_partials(const double *__restrict__ part, int np,
                                                  double *lds) {
  constexpr int K = kMaxRedBlocks / kBlock;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * kBlock;
    v[k] = i < np ? part[i] : 0.0;
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) s += v[k];
  return block_sum(s, lds);
}

__device__ __forceinline__ void finish(KState *st, int status) {
  st->done = 1;
  st->status = status;
}

__global__ inline void k_state_init(KState *st, double tol, double rel_tol, int max_iter, int min_iter,
                             int hist_cap) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  KState s;
  memset(&s, 0, sizeof(s));
  s.tol = tol;
  s.rel_tol = rel_tol;
  s.max_iter = max_iter;
  s.min_iter = min_iter;
  s.hist_cap = hist_cap;
  *st = s;
}

}  // namespace pfm
