// The pressure-equation step around the solve (solvers/pressure_equation.py):
// assembly of laplacian(1/A_p, p) = div(phiHbyA) per unit volume, the face
// flux and cell velocity corrections, and the closed-domain flux adjustment.
//
// Face sums follow the reference's scatter order: torch's scatter_add and
// index_add_ on the CPU add a face's contribution to its cell in face order,
// so a cell's sum runs over its owner faces (ascending), then its neighbour
// faces (ascending), then its boundary faces (ascending) -- exactly the
// lower slots, the upper slots and the boundary CSR of a row here.  Built
// with -ffp-contract=off; 1/A_p = 1 / clamp(|A_p|, 1e-30) everywhere.

#include "device_common.h"
#include "pfm_internal.h"

using namespace pfm;

namespace {

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// torch's clamp(min=lo) keeps NaN.
__device__ inline double clamp_min(double x, double lo) { return x < lo ? lo : x; }
__device__ inline double inv_ap(double a) { return 1.0 / clamp_min(fabs(a), 1e-30); }

// Face-interpolated 1/A_p: w*inv_P + (1 - w)*inv_N (pressure_equation.py:102-111, 313)
// of a face with weight wf, owner o and neighbour q.
__device__ inline double inv_face(const double *__restrict__ ap, double wf, int o, int q) {
  return wf * inv_ap(ap[o]) + (1.0 - wf) * inv_ap(ap[q]);
}

// lower = -c/V_P, upper = -c/V_N with c = (1/A_p)_f |S_f| delta_f (:117-129).
__global__ __launch_bounds__(kBlock) void k_pe_faces(int nf, const int32_t *__restrict__ own,
                                                      const int32_t *__restrict__ nei,
                                                      const double *__restrict__ ap,
                                                      const double *__restrict__ w,
                                                      const double *__restrict__ mag,
                                                      const double *__restrict__ delta,
                                                      const double *__restrict__ vol,
                                                      double *__restrict__ lower,
                                                      double *__restrict__ upper) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf) return;
  // The face arrays are read once: nontemporal loads keep the gathered A_p
  // and volumes cached.
  const int o = ld<true>(own + f), q = ld<true>(nei + f);
  const double c = (inv_face(ap, ld<true>(w + f), o, q) * ld<true>(mag + f)) * ld<true>(delta + f);
  lower[f] = (-c) / vol[o];
  upper[f] = (-c) / vol[q];
}

// Per row: diag = sum(c/V_P over owner faces) + sum(c/V_N over neighbour
// faces) (c/V == -lower / -upper exactly) (:131-135); source =
// ((0 + sum(-phi, owner faces)) + sum(phi, neighbour faces)
//   + sum(-phi, boundary faces)) / V (:137-150).
__global__ __launch_bounds__(kBlock) void k_pe_rows(int n, int nf, int nslices,
                                                     const SliceInfo *__restrict__ sl,
                                                     const int32_t *__restrict__ map,
                                                     const int32_t *__restrict__ own_start,
                                                     const double *__restrict__ lower,
                                                     const double *__restrict__ upper,
                                                     const double *__restrict__ phi,
                                                     const int32_t *__restrict__ bptr,
                                                     const int32_t *__restrict__ bface,
                                                     const double *__restrict__ vol,
                                                     double *__restrict__ diag,
                                                     double *__restrict__ source) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  double d1 = 0.0, d2 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
  int base = si.ptr + lane;
  if (si.wl <= 4 && si.wu <= 4) {
    // Every face index of the row loaded at once, then every gather, then the
    // sums in slot order: two dependent round trips instead of one per slot.
    constexpr int U = 4;
    const int bu = base + si.wl * kSlice;
    int fl[U], fu[U];
    // owner-sorted faces: the row's owner faces follow from own_start, no map
    const int o0 = own_start && row < n ? own_start[row] : 0;
    const int o1 = own_start && row < n ? own_start[row + 1] : 0;
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (own_start)
        fl[j] = o0 + j < o1 ? o0 + j : -1;
      else
        fl[j] = j < si.wl ? map[base + j * kSlice] : -1;
      fu[j] = j < si.wu ? map[bu + j * kSlice] : -1;
    }
    double lv[U], pl[U], uv[U], pu[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      lv[j] = fl[j] >= 0 ? lower[fl[j]] : 0.0;
      pl[j] = fl[j] >= 0 ? phi[fl[j]] : 0.0;
      uv[j] = fu[j] >= 0 ? upper[fu[j]] : 0.0;
      pu[j] = fu[j] >= 0 ? phi[fu[j]] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (fl[j] >= 0) {
        d1 += -lv[j];
        t1 += -pl[j];
      }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (fu[j] >= 0) {
        d2 += -uv[j];
        t2 += pu[j];
      }
  } else {
    for (int j = 0; j < si.wl; ++j, base += kSlice) {
      const int f = map[base];
      if (f >= 0) {
        d1 += -lower[f];
        t1 += -phi[f];
      }
    }
    for (int j = 0; j < si.wu; ++j, base += kSlice) {
      const int f = map[base];
      if (f >= 0) {
        d2 += -upper[f];
        t2 += phi[f];
      }
    }
  }
  if (row < n) {
    for (int k = bptr[row]; k < bptr[row + 1]; ++k) t3 += -phi[nf + bface[k]];
    diag[row] = d1 + d2;
    source[row] = (((0.0 + t1) + t2) + t3) / vol[row];
  }
}

// phi_out = phi + ((((1/A_p)_f * (p_P - p_N)) * |S_f|) * delta_f) on internal
// faces, phi on boundary faces (:305-330).
__global__ __launch_bounds__(kBlock) void k_cff(int nf_all, int nf,
                                                 const int32_t *__restrict__ own,
                                                 const int32_t *__restrict__ nei,
                                                 const double *__restrict__ phi,
                                                 const double *__restrict__ p,
                                                 const double *__restrict__ ap,
                                                 const double *__restrict__ w,
                                                 const double *__restrict__ mag,
                                                 const double *__restrict__ delta,
                                                 double *__restrict__ out) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf_all) return;
  if (f >= nf) {
    out[f] = phi[f];
    return;
  }
  const int o = ld<true>(own + f), q = ld<true>(nei + f);  // face arrays: read once
  const double dp = p[o] - p[q];
  out[f] = ld<true>(phi + f) +
           (((inv_face(ap, ld<true>(w + f), o, q) * dp) * ld<true>(mag + f)) * ld<true>(delta + f));
}

// Gauss gradient with linearly interpolated face pressure (boundary faces
// take the owner value), one running sum per cell in index_add_ order; then
// U = HbyA - (alpha/A_p) grad p, limited to |U| <= 1e4 (:228-265).  A slot of
// row r holds a face of r and the face's other cell in its column: on the
// owner side p[own] = p[r] and p[nei] = p[col], on the neighbour side the
// reverse, so the face's two pressures come from the row and the slot column
// (no face-indexed owner/neighbour gathers).
//
// Latency: the row's own operands are loaded first, then every slot's face
// index and column, then every gather, then the sums in slot order (two
// dependent round trips for rows of at most 4 + 4 couplings).  kPair reads
// the columns from the paired 16-bit offsets.  kOwnRun (faces oriented and
// sorted by owner): a row's owner-side faces are consecutive from its entry
// of own_start, loaded with the row's operands.  HbyA and U, read and written
// once, go nontemporally.  At 128^3: 158 us with the per-group round trips,
// 140 us with the owner-side run from the first lower slot's face index and
// plain HbyA and U, 131 us in this form.
template <bool kPair, bool kOwnRun>
__global__ __launch_bounds__(kBlock) void k_cvel(int n, int nf, int nslices,
                                                  const SliceInfo *__restrict__ sl,
                                                  const int32_t *__restrict__ map,
                                                  const int32_t *__restrict__ col,
                                                  const uint32_t *__restrict__ cp,
                                                  const int32_t *__restrict__ own_start,
                                                  const double *__restrict__ w,
                                                  const double *__restrict__ p,
                                                  const double *__restrict__ sf,
                                                  const int32_t *__restrict__ bptr,
                                                  const int32_t *__restrict__ bface,
                                                  const double *__restrict__ vol,
                                                  const double *__restrict__ ap,
                                                  const double *__restrict__ hbya, double alpha,
                                                  double *__restrict__ u) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  const bool in = row < n;
  const double prow = in ? p[row] : 0.0;
  const double v0 = in ? vol[row] : 1.0;
  const double apr = in ? ap[row] : 1.0;
  const double h0 = in ? ld<true>(hbya + 3 * row) : 0.0;
  const double h1 = in ? ld<true>(hbya + 3 * row + 1) : 0.0;
  const double h2 = in ? ld<true>(hbya + 3 * row + 2) : 0.0;
  const int b0 = in ? bptr[row] : 0, b1 = in ? bptr[row + 1] : 0;
  const int f0 = kOwnRun && in ? own_start[row] : 0;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  // One slot's contribution, in slot order (owner side: + p_f S_f; neighbour
  // side: + (-p_f S_f)).
  auto add = [&](bool owner_side, double wf, double pc, double sx, double sy, double sz) {
    const double p_own = owner_side ? prow : pc, p_nei = owner_side ? pc : prow;
    const double pf = wf * p_own + (1.0 - wf) * p_nei;
    const double cx = pf * sx, cy = pf * sy, cz = pf * sz;
    if (owner_side) {
      gx += cx;
      gy += cy;
      gz += cz;
    } else {
      gx += -cx;
      gy += -cy;
      gz += -cz;
    }
  };
  const int wl = si.wl, wu = si.wu;
  if (wl <= 4 && wu <= 4) {
    constexpr int U = 8;  // slots 0 .. wl-1 owner side, wl .. wl+wu-1 neighbour side
    const int nsl = wl + wu;
    const int b = si.ptr + lane;
    int f[U], c[U];
    if constexpr (kPair) {
      const int pb = pair_base(si, s);
      const bool ori = pair_oriented(cp);
      uint32_t wd[U / 2];
#pragma unroll
      for (int j = 0; j < U / 2; ++j) wd[j] = 2 * j < nsl ? cp[pb + j * kSlice + lane] : 0u;
#pragma unroll
      for (int j = 0; j < U; ++j) c[j] = j < nsl ? pair_col(wd[j >> 1], j & 1, row, j < wl, ori) : -1;
    } else {
#pragma unroll
      for (int j = 0; j < U; ++j) c[j] = j < nsl ? col[b + j * kSlice] : -1;
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      if (kOwnRun && j < wl)
        f[j] = c[j] >= 0 ? f0 + j : -1;
      else
        f[j] = j < nsl ? map[b + j * kSlice] : -1;
    }
    double wf[U], pc[U], sx[U], sy[U], sz[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const bool ok = f[j] >= 0;
      wf[j] = ok ? w[f[j]] : 0.0;
      pc[j] = ok ? p[c[j]] : 0.0;
      sx[j] = ok ? sf[3 * f[j]] : 0.0;
      sy[j] = ok ? sf[3 * f[j] + 1] : 0.0;
      sz[j] = ok ? sf[3 * f[j] + 2] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (f[j] >= 0) add(j < wl, wf[j], pc[j], sx[j], sy[j], sz[j]);
  } else {
    int base = si.ptr + lane;
    for (int j = 0; j < wl + wu; ++j, base += kSlice) {
      const int f = map[base];
      if (f < 0) continue;
      add(j < wl, w[f], p[col[base]], sf[3 * f], sf[3 * f + 1], sf[3 * f + 2]);
    }
  }
  if (!in) return;
  for (int k = b0; k < b1; ++k) {
    const int64_t bf = (int64_t)nf + bface[k];
    gx += prow * sf[3 * bf];
    gy += prow * sf[3 * bf + 1];
    gz += prow * sf[3 * bf + 2];
  }
  const double v = clamp_min(v0, 1e-30);
  gx = gx / v;
  gy = gy / v;
  gz = gz / v;
  const double a = alpha * inv_ap(apr);
  double ux = h0 - a * gx;
  double uy = h1 - a * gy;
  double uz = h2 - a * gz;
  const double m = clamp_min(sqrt(fma(uz, uz, fma(uy, uy, ux * ux))), 1e-30);
  if (m > 1e4) {
    const double sc = 1e4 / m;
    ux = ux * sc;
    uy = uy * sc;
    uz = uz * sc;
  }
  __builtin_nontemporal_store(ux, u + 3 * row);
  __builtin_nontemporal_store(uy, u + 3 * row + 1);
  __builtin_nontemporal_store(uz, u + 3 * row + 2);
}

constexpr int kAdjThreads = 1024;

// Net boundary flux, one workgroup: fixed strided partial sums, then a fixed
// tree (deterministic; the reference leaves the order to torch's sum).
__global__ __launch_bounds__(kAdjThreads) void k_adj_sum(int nb, const double *__restrict__ phib,
                                                          double *__restrict__ net) {
  __shared__ double red[kAdjThreads];
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += kAdjThreads) s += phib[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kAdjThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) net[0] = red[0];
}

// phi_b -= net / n_b when |net| >= 1e-30 (:366-375).
__global__ __launch_bounds__(kBlock) void k_adj_apply(int nb, double *__restrict__ phib,
                                                       const double *__restrict__ net) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  const double t = net[0];
  if (k >= nb || fabs(t) < 1e-30) return;
  phib[k] = phib[k] - t / (double)nb;
}

int check_bnd(const pfm_ldu_s *h, int32_t n_faces, const int32_t *bptr, const int32_t *bface) {
  PFM_REQUIRE(n_faces >= h->nf, "n_faces (%d) < internal faces of the handle (%d)", n_faces,
              h->nf);
  PFM_REQUIRE(h->n == 0 || bptr, "null boundary pointer array");
  PFM_REQUIRE(n_faces == h->nf || bface, "null boundary face list");
  return PFM_OK;
}

}  // namespace

extern "C" {

int pfm_pressure_assemble(pfm_ldu h, int32_t n_faces, const int32_t *bnd_ptr,
                          const int32_t *bnd_face, const double *phi_hbya, const double *a_p,
                          const double *weights, const double *mag_sf, const double *delta,
                          const double *vol, double *diag, double *lower, double *upper,
                          double *source, void *stream) {
  PFM_NONNULL(h);
  int rc = check_bnd(h, n_faces, bnd_ptr, bnd_face);
  if (rc != PFM_OK) return rc;
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(phi_hbya);
  PFM_NONNULL(a_p);
  PFM_NONNULL(vol);
  PFM_NONNULL(diag);
  PFM_NONNULL(source);
  if (h->nf > 0) {
    PFM_NONNULL(weights);
    PFM_NONNULL(mag_sf);
    PFM_NONNULL(delta);
    PFM_NONNULL(lower);
    PFM_NONNULL(upper);
  }
  hipStream_t s = (hipStream_t)stream;
  if (h->nf > 0)
    k_pe_faces<<<grid1d(h->nf), kBlock, 0, s>>>(h->nf, h->d_own, h->d_nei, a_p, weights, mag_sf,
                                                 delta, vol, lower, upper);
  k_pe_rows<<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
      h->n, h->nf, h->nslices, h->d_slices, h->d_map,
      h->oriented_sorted ? h->d_own_start : nullptr, lower, upper, phi_hbya, bnd_ptr, bnd_face,
      vol, diag, source);
  PFM_CHECK_HIP(hipGetLastError());
  return PFM_OK;
}

int pfm_correct_face_flux(pfm_ldu h, int32_t n_faces, const double *phi, const double *p,
                          const double *a_p, const double *weights, const double *mag_sf,
                          const double *delta, double *phi_out, void *stream) {
  PFM_NONNULL(h);
  PFM_REQUIRE(n_faces >= h->nf, "n_faces (%d) < internal faces of the handle (%d)", n_faces,
              h->nf);
  if (n_faces == 0) return PFM_OK;
  PFM_NONNULL(phi);
  PFM_NONNULL(phi_out);
  if (h->nf > 0) {
    PFM_NONNULL(p);
    PFM_NONNULL(a_p);
    PFM_NONNULL(weights);
    PFM_NONNULL(mag_sf);
    PFM_NONNULL(delta);
  }
  hipStream_t s = (hipStream_t)stream;
  k_cff<<<grid1d(n_faces), kBlock, 0, s>>>(n_faces, h->nf, h->d_own, h->d_nei, phi, p, a_p,
                                            weights, mag_sf, delta, phi_out);
  PFM_CHECK_HIP(hipGetLastError());
  return PFM_OK;
}

int pfm_correct_velocity(pfm_ldu h, int32_t n_faces, const int32_t *bnd_ptr,
                         const int32_t *bnd_face, const double *hbya, const double *p,
                         const double *a_p, const double *weights, const double *face_areas,
                         const double *vol, double alpha, double *u_out, void *stream) {
  PFM_NONNULL(h);
  int rc = check_bnd(h, n_faces, bnd_ptr, bnd_face);
  if (rc != PFM_OK) return rc;
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(hbya);
  PFM_NONNULL(p);
  PFM_NONNULL(a_p);
  PFM_NONNULL(vol);
  PFM_NONNULL(u_out);
  if (n_faces > 0) PFM_NONNULL(face_areas);
  if (h->nf > 0) PFM_NONNULL(weights);
  hipStream_t s = (hipStream_t)stream;
  const int G = blocks_for_slices(h->nslices);
#define PFM_CVEL_ARGS                                                                    \
  h->n, h->nf, h->nslices, h->d_slices, h->d_map, h->d_col, h->d_colp2, h->d_own_start, weights, \
      p, face_areas, \
      bnd_ptr, bnd_face, vol, a_p, hbya, alpha, u_out
  const bool run = h->oriented_sorted && h->d_own_start;
  if (h->d_colp2 && run)
    k_cvel<true, true><<<G, kBlock, 0, s>>>(PFM_CVEL_ARGS);
  else if (h->d_colp2)
    k_cvel<true, false><<<G, kBlock, 0, s>>>(PFM_CVEL_ARGS);
  else if (run)
    k_cvel<false, true><<<G, kBlock, 0, s>>>(PFM_CVEL_ARGS);
  else
    k_cvel<false, false><<<G, kBlock, 0, s>>>(PFM_CVEL_ARGS);
#undef PFM_CVEL_ARGS
  PFM_CHECK_HIP(hipGetLastError());
  return PFM_OK;
}

int pfm_adjust_phi(int32_t n_faces, int32_t n_internal, double *phi, int closed, double *work,
                   double *net_out_h, int32_t *adjusted_out_h, void *stream) {
  PFM_REQUIRE(n_faces >= 0 && n_internal >= 0 && n_internal <= n_faces,
              "bad sizes (faces %d, internal %d)", n_faces, n_internal);
  PFM_REQUIRE(adjusted_out_h, "null output");
  *adjusted_out_h = 0;
  if (net_out_h) *net_out_h = 0.0;
  const int nb = n_faces - n_internal;
  if (nb == 0 || !closed) return PFM_OK;
  PFM_NONNULL(phi);
  PFM_NONNULL(work);
  hipStream_t s = (hipStream_t)stream;
  k_adj_sum<<<1, kAdjThreads, 0, s>>>(nb, phi + n_internal, work);
  k_adj_apply<<<grid1d(nb), kBlock, 0, s>>>(nb, phi + n_internal, work);
  PFM_CHECK_HIP(hipGetLastError());
  double net = 0.0;
  PFM_CHECK_HIP(hipMemcpyAsync(&net, work, sizeof(double), hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  if (net_out_h) *net_out_h = net;
  *adjusted_out_h = fabs(net) < 1e-30 ? 0 : 1;
  return PFM_OK;
}

}  // extern "C"
