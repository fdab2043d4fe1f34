// LDU handle, coefficient packing, SpMV, DIC/DILU factorisation, Laplacian
// assembly and adjoint gradient kernels.
//
// All kernels are HBM-bandwidth bound; none uses MFMA (nothing here is a dense
// contraction).  The file is compiled with -ffp-contract=off so every product
// and sum is rounded on its own, as in the reference's op-by-op torch code:
// the DIC/DILU factors and the SpMV are bit-identical to the sequential
// reference recurrences.

#include <algorithm>
#include <cmath>

#include <cstdlib>
#include <cstring>

#include "device_common.h"

using namespace pfm;

namespace {

// ---------------------------------------------------------------------------
// Packing: face arrays -> row-slice order.  Slot-major: each thread writes one
// slot (coalesced), reading the face coefficient through the slot->face map.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack(int64_t nslots, int nslices,
                                                  const SliceInfo *__restrict__ sl,
                                                  const int32_t *__restrict__ map,
                                                  const double *__restrict__ lo,
                                                  const double *__restrict__ up,
                                                  T *__restrict__ packed) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  int base = si.ptr + lane;
  for (int j = 0; j < si.wl; ++j, base += kSlice) {
    const int f = map[base];
    packed[base] = f >= 0 ? (T)lo[f] : T(0);
  }
  for (int j = 0; j < si.wu; ++j, base += kSlice) {
    const int f = map[base];
    packed[base] = f >= 0 ? (T)up[f] : T(0);
  }
}

// One wave per slice.  kNT: the streams read or written once per product
// (coefficients, columns, diagonal, y) are nontemporal, so they pass through
// L2 and the MALL without displacing x, whose entries every row gathers.
template <typename T, bool kNT>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, int nslices,
                                                  const SliceInfo *__restrict__ sl,
                                                  const int32_t *__restrict__ col,
                                                  const T *__restrict__ val,
                                                  const T *__restrict__ diag,
                                                  const T *__restrict__ x,
                                                  T *__restrict__ y) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  const T dx = row < n ? ld<kNT>(diag + row) * x[row] : T(0);
  auto gx = [x](int c) { return x[c]; };
  const T v = row_spmv_g<T, decltype(gx), false, ColAbsT<kNT>, kNT>(si, lane, dx,
                                                                      ColAbsT<kNT>{col}, val, gx);
  if (row < n) {
    if constexpr (kNT)
      __builtin_nontemporal_store(v, y + row);
    else
      y[row] = v;
  }
}

// k_spmv over the handle's paired int16 column offsets (row_spmv_pair): the
// same products and sums, 2 index bytes per slot instead of 4.
template <typename T, bool kNT>
__global__ __launch_bounds__(kBlock) void k_spmv_p2(int n, int nslices,
                                                     const SliceInfo *__restrict__ sl,
                                                     const uint32_t *__restrict__ cp,
                                                     const T *__restrict__ val,
                                                     const T *__restrict__ diag,
                                                     const T *__restrict__ x,
                                                     T *__restrict__ y) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  const T dx = row < n ? ld<kNT>(diag + row) * x[row] : T(0);
  auto gx = [x](int c) { return x[c]; };
  const T v = row_spmv_pair<T, decltype(gx), false, kNT>(si, s, lane, row, dx, cp, val, gx);
  if (row < n) {
    if constexpr (kNT)
      __builtin_nontemporal_store(v, y + row);
    else
      y[row] = v;
  }
}

// Y = A X for k right-hand sides stored row-major (X[i*k + j]): one wave per
// slice, the row's coefficients and columns loaded once and reused for every
// column j, each column summed in the association of the single product.
__global__ __launch_bounds__(kBlock) void k_spmv_multi(int n, int nslices, int k,
                                                        const SliceInfo *__restrict__ sl,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ val,
                                                        const double *__restrict__ diag,
                                                        const double *__restrict__ X,
                                                        double *__restrict__ Y) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  const double dg = row < n ? diag[row] : 0.0;
  constexpr int U = 4;
  const int wl = si.wl, wu = si.wu;
  if (wl <= U && wu <= U) {
    const int bl = si.ptr + lane, bu = si.ptr + wl * kSlice + lane;
    int c[2 * U];
    double v[2 * U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      c[j] = j < wl ? col[bl + j * kSlice] : -1;
      v[j] = j < wl ? val[bl + j * kSlice] : 0.0;
      c[U + j] = j < wu ? col[bu + j * kSlice] : -1;
      v[U + j] = j < wu ? val[bu + j * kSlice] : 0.0;
    }
    for (int q = 0; q < k; ++q) {
      double accL = 0.0, accU = 0.0;
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (c[j] >= 0) accL += v[j] * X[(int64_t)c[j] * k + q];
#pragma unroll
      for (int j = 0; j < U; ++j)
        if (c[U + j] >= 0) accU += v[U + j] * X[(int64_t)c[U + j] * k + q];
      if (row < n) Y[(int64_t)row * k + q] = (dg * X[(int64_t)row * k + q] + accL) + accU;
    }
    return;
  }
  for (int q = 0; q < k; ++q) {  // wide rows: the single-product row sum per column
    auto gx = [X, k, q](int cc) { return X[(int64_t)cc * k + q]; };
    const double dx = row < n ? dg * X[(int64_t)row * k + q] : 0.0;
    const double y = row_spmv_g<double>(si, lane, dx, ColAbs{col}, val, gx);
    if (row < n) Y[(int64_t)row * k + q] = y;
  }
}

// FvMatrix.relax (core/fv_matrix.py:192-235): d' = d / alpha,
// s' = s + ((1 - alpha) * d') * phi_old, rounded as the reference's ops.
// (element-wise: outputs may alias inputs, so no restrict qualifiers)
__global__ __launch_bounds__(kBlock) void k_relax(int n, double alpha, const double *d,
                                                   const double *src, const double *phi,
                                                   double *d_out, double *s_out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double dn = d[i] / alpha;
  d_out[i] = dn;
  s_out[i] = src[i] + ((1.0 - alpha) * dn) * phi[i];
}

// Achievable-bandwidth probes: a nontemporal 16-byte-per-lane read of a
// buffer (one partial per wave, so the loads cannot be dropped) and a copy.
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(kBlock) void k_bw_read(const dvec2 *__restrict__ a, int64_t n2,
                                                     double *__restrict__ out) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n2;
       i += (int64_t)gridDim.x * kBlock) {
    const dvec2 v = __builtin_nontemporal_load(a + i);
    s += v.x + v.y;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) out[blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)] = s;
}

__global__ __launch_bounds__(kBlock) void k_bw_copy(const dvec2 *__restrict__ a,
                                                     dvec2 *__restrict__ b, int64_t n2) {
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n2;
       i += (int64_t)gridDim.x * kBlock)
    b[i] = __builtin_nontemporal_load(a + i);
}

constexpr int kBwGrid = 2048;

// Bytes one product streams (coefficients and columns per slot; diagonal, x
// and y per row) from which pfm_spmv streams nontemporally: half the 256 MB
// MALL.  A product that large is evicted between calls by any other work of
// its size; a smaller one repeated alone stays MALL-resident, where plain
// loads are faster (128^3, 200 MB: 28.4 us resident with plain loads, 33.7 us
// with nontemporal ones; from HBM 39.5 us plain, 34.5 us nontemporal).
constexpr int64_t kStreamBytes = int64_t(128) << 20;

// PFM_SPMV_COLS=int32 makes pfm_spmv read the int32 columns even where the
// paired int16 offsets exist (for comparing the two layouts).
bool spmv_int32_columns() {
  static const bool v = [] {
    const char *e = std::getenv("PFM_SPMV_COLS");
    return e && std::strcmp(e, "int32") == 0;
  }();
  return v;
}

template <typename T>
int launch_spmv(const pfm_ldu_s *h, const T *diag, const T *packed, const T *x, T *y,
                hipStream_t s) {
  const int64_t bytes = h->nslots * (int64_t)(sizeof(T) + 4) + 3 * (int64_t)h->n * sizeof(T);
  if (h->d_colp2 && !spmv_int32_columns()) {
    if (bytes >= kStreamBytes)
      k_spmv_p2<T, true><<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
          h->n, h->nslices, h->d_slices, h->d_colp2, packed, diag, x, y);
    else
      k_spmv_p2<T, false><<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
          h->n, h->nslices, h->d_slices, h->d_colp2, packed, diag, x, y);
  } else if (bytes >= kStreamBytes)
    k_spmv<T, true><<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
        h->n, h->nslices, h->d_slices, h->d_col, packed, diag, x, y);
  else
    k_spmv<T, false><<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
        h->n, h->nslices, h->d_slices, h->d_col, packed, diag, x, y);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

// r = b - A x
__global__ __launch_bounds__(kBlock) void k_residual(int n, int nslices,
                                                      const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ b,
                                                      double *__restrict__ r) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  const double dx = row < n ? diag[row] * x[row] : 0.0;
  const double v = row_spmv<double>(si, lane, dx, col, val, x);
  if (row < n) r[row] = b[row] - v;
}

// y = A x or A^T x straight from face arrays (no packing).  A^T follows the
// reference Ax_T association: (diag*x + sum over neighbour-faces of lower*x[own])
// + sum over owner-faces of upper*x[nei].
template <bool kTranspose>
__global__ __launch_bounds__(kBlock) void k_spmv_ldu(int n, int nslices,
                                                      const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ col,
                                                      const int32_t *__restrict__ map,
                                                      const double *__restrict__ diag,
                                                      const double *__restrict__ lower,
                                                      const double *__restrict__ upper,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ y) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  double accO = 0.0, accN = 0.0;  // owner-side slots, neighbour-side slots
  int base = si.ptr + lane;
  const double *co = kTranspose ? upper : lower;
  const double *cn = kTranspose ? lower : upper;
  for (int j = 0; j < si.wl; ++j, base += kSlice) {
    const int c = col[base];
    if (c >= 0) accO += co[map[base]] * x[c];
  }
  for (int j = 0; j < si.wu; ++j, base += kSlice) {
    const int c = col[base];
    if (c >= 0) accN += cn[map[base]] * x[c];
  }
  if (row < n) {
    const double dx = diag[row] * x[row];
    y[row] = kTranspose ? (dx + accN) + accO : (dx + accO) + accN;
  }
}

// ---------------------------------------------------------------------------
// DIC factor, one launch per dependency level.  Cell i walks its
// neighbour-faces in face order (the packed upper part of its row); owners
// j < i are already final, owners j > i still hold their original diagonal.
__global__ __launch_bounds__(kBlock) void k_dic_level(const int32_t *__restrict__ cells, int count,
                                                       const SliceInfo *__restrict__ sl,
                                                       const int32_t *__restrict__ col,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ diag,
                                                       double *__restrict__ d) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int i = cells[t];
  const int2 raw = reinterpret_cast<const int2 *>(sl)[i >> 6];
  const int wl = raw.y & 0xffff, wu = (unsigned)raw.y >> 16;
  int base = raw.x + wl * kSlice + (i & 63);
  double di = diag[i];
  for (int j = 0; j < wu; ++j, base += kSlice) {
    const int c = col[base];
    if (c < 0) break;
    const double a = val[base];
    const double dj = c < i ? d[c] : diag[c];
    if (fabs(dj) > 1e-30) di = di - (a * a) / dj;
  }
  d[i] = di;
}

// rD = |d|.clamp(min=1e-30), NaN propagates like torch.clamp.
__global__ __launch_bounds__(kBlock) void k_abs_clamp(int n, double *__restrict__ d) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double a = fabs(d[i]);
  d[i] = a < 1e-30 ? 1e-30 : a;
}

// DILU factor.  The face recurrence of the reference has a dependency depth
// of the order of the largest RCM level (tens of thousands of levels at 128^3,
// a few hundred faces each), far too many for one launch per level.  One
// persistent workgroup walks the levels instead, with a barrier per level; a
// parallel gather first lays the faces out in level order as (owner,
// neighbour, lower*upper) so each level is one contiguous, coalesced read.
// lower*upper == upper*lower exactly, so both updates use the same product.
__global__ __launch_bounds__(kBlock) void k_dilu_gather(int nf, const int32_t *__restrict__ faces,
                                                         const int32_t *__restrict__ own,
                                                         const int32_t *__restrict__ nei,
                                                         const double *__restrict__ lower,
                                                         const double *__restrict__ upper,
                                                         int2 *__restrict__ pq,
                                                         double *__restrict__ lu) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nf) return;
  const int f = faces[k];
  pq[k] = make_int2(own[f], nei[f]);
  lu[k] = lower[f] * upper[f];
}

// Gauss-Seidel schedule entries (GsEnt), one thread per scheduled cell: the
// cell, its lower-slot base, group widths and its columns merged into face
// order (the valid prefixes of both groups by face index), so a sweep reads
// the entry, then the coefficients and x values at once.
__global__ __launch_bounds__(kBlock) void k_gs_entries(int count, const int32_t *__restrict__ cells,
                                                        const SliceInfo *__restrict__ sl,
                                                        const int32_t *__restrict__ col,
                                                        const int32_t *__restrict__ map,
                                                        GsEnt *__restrict__ out) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int c = cells[t];
  const SliceInfo si = load_slice_lane(sl, c / kSlice);
  const int bl = si.ptr + c % kSlice, bu = bl + si.wl * kSlice;
  int cols[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int n = 0, mask = 0, il = 0, iu = 0;
  bool wide = false;
  while (true) {
    const bool hl = il < si.wl && col[bl + il * kSlice] >= 0;
    const bool hu = iu < si.wu && col[bu + iu * kSlice] >= 0;
    if (!hl && !hu) break;
    if (n == 8) {
      wide = true;
      break;
    }
    const bool take_l = hl && (!hu || map[bl + il * kSlice] < map[bu + iu * kSlice]);
    if (take_l) {
      cols[n] = col[bl + il * kSlice];
      mask |= 1 << n;
      ++il;
    } else {
      cols[n] = col[bu + iu * kSlice];
      ++iu;
    }
    ++n;
  }
  GsEnt e;
  e.a = make_int4(c, bl, (int)(uint16_t)si.wl | ((int)si.wu << 16),
                  wide ? (255 << 8) : (mask | (n << 8)));
  e.b = make_int4(cols[0], cols[1], cols[2], cols[3]);
  e.c = make_int4(cols[4], cols[5], cols[6], cols[7]);
  out[t] = e;
}

constexpr int kDiluThreads = 1024;

// Only the d values depend on the previous level: each thread loads its
// first face of level l + 1 (and that level's end) while level l runs, so a
// level costs one round trip for d, the division and the barrier.
__device__ __forceinline__ void dilu_face(int2 f, double a, double *__restrict__ d) {
  const double dp = d[f.x], dq = d[f.y];
  if (fabs(dq) > 1e-30) d[f.x] = dp - a / dq;
  if (fabs(dp) > 1e-30) d[f.y] = dq - a / dp;
}

__global__ __launch_bounds__(kDiluThreads) void k_dilu_sweep(const int32_t *__restrict__ level_ptr,
                                                              int nlevels,
                                                              const int2 *__restrict__ pq,
                                                              const double *__restrict__ lu,
                                                              double *__restrict__ d) {
  if (nlevels <= 0) return;
  const int t = (int)threadIdx.x;
  int b = level_ptr[0], e = level_ptr[1];
  bool has = b + t < e;
  int2 f = has ? pq[b + t] : make_int2(0, 0);
  double a = has ? lu[b + t] : 0.0;
  for (int l = 0; l < nlevels; ++l) {
    // level l + 1: [e, en), this thread's first face of it
    const int en = l + 1 < nlevels ? level_ptr[l + 2] : e;
    const bool hn = e + t < en;
    const int2 fn = hn ? pq[e + t] : make_int2(0, 0);
    const double an = hn ? lu[e + t] : 0.0;
    if (has) dilu_face(f, a, d);
    for (int k = b + t + kDiluThreads; k < e; k += kDiluThreads) dilu_face(pq[k], lu[k], d);
    __syncthreads();
    b = e;
    e = en;
    has = hn;
    f = fn;
    a = an;
  }
}

// The same sweep with the diagonal in LDS.  Each face level's faces touch
// cells spread over the whole mesh, but a cell is touched only within a
// window of levels (first to last face holding it): 16,449 cells are live at
// once at 128^3 (1,041 at 32^3, ~n^2).  The host gives each cell an LDS slot
// for its window.  Level l is one phase between barriers: its faces (the
// arithmetic of dilu_face, on LDS), the write-back to d of the cells whose
// window ended at level l - 1, and the load of the cells whose window starts
// at level l + 1.  The three touch different slots: a slot written back in
// phase l is given to another cell from phase l + 1 on.  After the last
// level, the cells whose window ended there are written back.  Each of the
// kT threads holds at most one face and one phase entry per level (the host
// picks kT = 512 or 1024 to fit), fetched kDiluAhead levels before use into a
// register ring; the level pointers come a chunk of kDiluChunk levels at a
// time into LDS.  The fetches are unconditional (clamped indices, validity
// kept beside the raw values) so that the compiler's waits before a ring
// entry is used cover the loads of earlier levels only.  At 128^3 (33,147
// levels, at most 254 faces and 130 phase entries per level) the sweep takes
// 14.2 ms with 512 threads (16.6 ms with 1024, 15.0 ms with 256, at two turns
// of the ring per loop iteration); on the global diagonal it took 34.7-36.8 ms.
constexpr int kDiluSlots = 19000;   // 152,000 B of the 160 KB LDS
constexpr int kDiluAhead = 8;
constexpr int kDiluUnroll = 4;
constexpr int kDiluChunk = 256;

__device__ __forceinline__ void dilu_face_lds(uint32_t w, double a, double *dl) {
  const int sp = (int)(w & 0xffffu), sq = (int)(w >> 16);
  const double dp = dl[sp], dq = dl[sq];
  if (fabs(dq) > 1e-30) dl[sp] = dp - a / dq;
  if (fabs(dp) > 1e-30) dl[sq] = dq - a / dp;
}

// Phase lists: iptr[0] .. iptr[1] the cells loaded before level 0, iptr[l + 1]
// .. iptr[l + 2] those of phase l (0 <= l <= nl; phase nl holds write-backs
// only).  Entries (cell, slot) load, (cell, -(slot + 1)) write back.  nf,
// nio >= 1.
template <int kT>
__global__ __launch_bounds__(kT) void k_dilu_lds(int nl, int nf, int nio,
                                                  const int32_t *__restrict__ fptr,
                                                  const uint32_t *__restrict__ fslot,
                                                  const double *__restrict__ lu,
                                                  const int32_t *__restrict__ iptr,
                                                  const int2 *__restrict__ io,
                                                  const double *__restrict__ din,
                                                  double *__restrict__ d) {
  __shared__ double dl[kDiluSlots];
  __shared__ int32_t pb[2][2][kDiluChunk + 1];  // chunk buffer, {face, phase} level offsets
  const int t = (int)threadIdx.x;
  if (nl <= 0) return;
  // Level offsets of chunk c (levels c*C .. c*C + C, clamped to nl) into
  // buffer c & 1.
  auto load_chunk = [&](int c) {
    const int l0 = c * kDiluChunk;
    for (int i = t; i < 2 * (kDiluChunk + 1); i += kT) {
      const int which = i / (kDiluChunk + 1), j = i - which * (kDiluChunk + 1);
      const int l = min(l0 + j, nl);
      pb[c & 1][which][j] = which == 0 ? fptr[l] : iptr[l + 1];
    }
  };
  for (int k = iptr[0] + t; k < iptr[1]; k += kT) dl[io[k].y] = din[k];
  load_chunk(0);
  load_chunk(1);
  __syncthreads();
  uint32_t rw[kDiluAhead];
  double ra[kDiluAhead], rv[kDiluAhead];
  int2 re[kDiluAhead];
  int rok[kDiluAhead];  // bit 0: a face, bit 1: a phase entry
  // This thread's face and phase entry of level l (levels past the end read
  // the empty range of level nl) into ring slot j.
  auto fetch = [&](int j, int l) {
    const int lc = min(l, nl);
    const int c = lc / kDiluChunk, o = lc - c * kDiluChunk;
    const int32_t *pf = pb[c & 1][0], *pi = pb[c & 1][1];
    const int fb = pf[o] + t, ib = pi[o] + t;
    rok[j] = (fb < pf[o + 1] ? 1 : 0) | (ib < pi[o + 1] ? 2 : 0);
    const int fi = min(fb, nf - 1), ii = min(ib, nio - 1);
    rw[j] = fslot[fi];
    ra[j] = lu[fi];
    re[j] = io[ii];
    rv[j] = din[ii];
  };
#pragma unroll
  for (int j = 0; j < kDiluAhead; ++j) fetch(j, j);
  // kDiluUnroll turns of the ring per loop iteration: the compiler's waits are
  // exact inside the unrolled body and drain the outstanding loads only where
  // the loop closes (14.8 / 14.5 / 14.2 / 14.2 ms for 1 / 2 / 4 / 8 turns).
  for (int l0 = 0; l0 < nl; l0 += kDiluUnroll * kDiluAhead) {
#pragma unroll
    for (int jj = 0; jj < kDiluUnroll * kDiluAhead; ++jj) {
      const int j = jj % kDiluAhead, l = l0 + jj;
      if (l >= nl) break;
      const int c = l / kDiluChunk, o = l - c * kDiluChunk;
      if (o == 0 && c > 0) load_chunk(c + 1);  // chunk c - 1 is no longer read
      const int ok = rok[j];
      const uint32_t w = rw[j];
      const double a = ra[j], v = rv[j];
      const int2 e = re[j];
      fetch(j, l + kDiluAhead);
      if (ok & 1) dilu_face_lds(w, a, dl);
      if (ok & 2) {
        if (e.y >= 0)
          dl[e.y] = v;
        else
          d[e.x] = dl[-e.y - 1];
      }
      __syncthreads();
    }
  }
  for (int k = iptr[nl + 1] + t; k < iptr[nl + 2]; k += kT) {  // phase nl
    const int2 e = io[k];
    d[e.x] = dl[-e.y - 1];
  }
}

// The initial value of every cell the DILU LDS sweep loads (din[k] = d[cell]
// for a load entry), read before the sweep so that its loads need no index.
__global__ __launch_bounds__(kBlock) void k_dilu_din(int nio, const int2 *__restrict__ io,
                                                      const double *__restrict__ d,
                                                      double *__restrict__ din) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nio) return;
  const int2 e = io[k];
  din[k] = e.y >= 0 ? d[e.x] : 0.0;
}

// Small schedules: one launch per face level (used when the level count is
// low enough that launches are cheaper than the single-workgroup sweep).
__global__ __launch_bounds__(kBlock) void k_dilu_level(const int32_t *__restrict__ faces, int count,
                                                        const int32_t *__restrict__ own,
                                                        const int32_t *__restrict__ nei,
                                                        const double *__restrict__ lower,
                                                        const double *__restrict__ upper,
                                                        double *__restrict__ d) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  const int f = faces[t];
  const int p = own[f], q = nei[f];
  const double l = lower[f], u = upper[f];
  const double dp = d[p], dq = d[q];
  if (fabs(dq) > 1e-30) d[p] = dp - (l * u) / dq;
  if (fabs(dp) > 1e-30) d[q] = dq - (u * l) / dp;
}

// ---------------------------------------------------------------------------
// Laplacian assembly.
__global__ __launch_bounds__(kBlock) void k_lap_faces(int nf, const double *__restrict__ gamma,
                                                       const double *__restrict__ mag,
                                                       const double *__restrict__ delta,
                                                       const double *__restrict__ vol,
                                                       const int32_t *__restrict__ own,
                                                       const int32_t *__restrict__ nei,
                                                       double *__restrict__ lower,
                                                       double *__restrict__ upper) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf) return;
  // The face arrays are read once: nontemporal loads keep the gathered
  // volumes cached (128^3 assembly 94 -> 85 us).
  const double c = (ld<true>(gamma + f) * ld<true>(mag + f)) * ld<true>(delta + f);
  lower[f] = (-c) / vol[ld<true>(own + f)];
  upper[f] = (-c) / vol[ld<true>(nei + f)];
}

// diag_i = (sum over owner-faces of c/V_i) + (sum over neighbour-faces of c/V_i);
// c/V_P == -lower_f and c/V_N == -upper_f exactly, so the sums read the
// freshly written face coefficients through the slot map.
// own_start (owner-sorted faces, else null): a row's owner faces are
// own_start[row] .. own_start[row + 1] - 1, so the lower group needs no map.
__global__ __launch_bounds__(kBlock) void k_lap_diag(int n, int nslices,
                                                      const SliceInfo *__restrict__ sl,
                                                      const int32_t *__restrict__ map,
                                                      const int32_t *__restrict__ own_start,
                                                      const double *__restrict__ lower,
                                                      const double *__restrict__ upper,
                                                      double *__restrict__ diag) {
  const int s = xcd_remap(blockIdx.x, gridDim.x) * kWavesPerBlock + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const SliceInfo si = load_slice(sl, s);
  const int lane = threadIdx.x & 63;
  const int row = s * kSlice + lane;
  double s1 = 0.0, s2 = 0.0;
  int base = si.ptr + lane;
  if (own_start) {
    if (row < n)
      for (int f = own_start[row]; f < own_start[row + 1]; ++f) s1 += -lower[f];
    base += si.wl * kSlice;
  } else {
    for (int j = 0; j < si.wl; ++j, base += kSlice) {
      const int f = map[base];
      if (f >= 0) s1 += -lower[f];
    }
  }
  for (int j = 0; j < si.wu; ++j, base += kSlice) {
    const int f = map[base];
    if (f >= 0) s2 += -upper[f];
  }
  if (row < n) diag[row] = s1 + s2;
}

// Dirichlet faces (sorted by owner cell): the first face of each cell's run
// sums the run's gamma_b*|S_b|*delta_b/V in face order and adds it.
// One thread per distinct Dirichlet cell (group g of the plan).  The group's
// entries are in the caller's order (patch, then face); each patch's terms
// are summed from zero as its scatter_add does, and each patch's sum is added
// to the diagonal in patch order.
__global__ __launch_bounds__(kBlock) void k_lap_dirichlet(int ngroups,
                                                           const int32_t *__restrict__ gptr,
                                                           const int32_t *__restrict__ gcell,
                                                           const int32_t *__restrict__ order,
                                                           const int32_t *__restrict__ patch,
                                                           const double *__restrict__ g,
                                                           const double *__restrict__ mag,
                                                           const double *__restrict__ dlt,
                                                           const double *__restrict__ vol,
                                                           double *__restrict__ diag) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= ngroups) return;
  const int c = gcell[k];
  const double v = vol[c];
  double d = diag[c], s = 0.0;
  const int j0 = gptr[k], j1 = gptr[k + 1];
  int cur = patch[j0];
  for (int j = j0; j < j1; ++j) {
    if (patch[j] != cur) {
      d = d + s;
      s = 0.0;
      cur = patch[j];
    }
    const int e = order[j];
    s += ((g[e] * mag[e]) * dlt[e]) / v;
  }
  diag[c] = d + s;
}

// ---------------------------------------------------------------------------
// Adjoint gradients.
__global__ __launch_bounds__(kBlock) void k_adj_diag(int n, const double *__restrict__ lam,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ g) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) g[i] = (-lam[i]) * x[i];
}

__global__ __launch_bounds__(kBlock) void k_adj_faces(int nf, const int32_t *__restrict__ own,
                                                       const int32_t *__restrict__ nei,
                                                       const double *__restrict__ lam,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ gl,
                                                       double *__restrict__ gu) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf) return;
  const int p = own[f], q = nei[f];
  if (gl) gl[f] = (-lam[p]) * x[q];
  if (gu) gu[f] = (-lam[q]) * x[p];
}

__global__ __launch_bounds__(kBlock) void k_lap_backward(int nf, const int32_t *__restrict__ own,
                                                          const int32_t *__restrict__ nei,
                                                          const double *__restrict__ gd,
                                                          const double *__restrict__ gl,
                                                          const double *__restrict__ gu,
                                                          const double *__restrict__ mag,
                                                          const double *__restrict__ delta,
                                                          const double *__restrict__ vol,
                                                          double *__restrict__ gg) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf) return;
  const int p = own[f], q = nei[f];
  const double sd = mag[f] * delta[f];
  gg[f] = sd * ((gd[p] - gl[f]) / vol[p] + (gd[q] - gu[f]) / vol[q]);
}

__global__ __launch_bounds__(kBlock) void k_adj_gamma(int nf, const int32_t *__restrict__ own,
                                                       const int32_t *__restrict__ nei,
                                                       const double *__restrict__ lam,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ mag,
                                                       const double *__restrict__ delta,
                                                       const double *__restrict__ vol,
                                                       double *__restrict__ gg) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= nf) return;
  const int p = ld<true>(own + f), q = ld<true>(nei + f);  // face arrays: read once
  const double sd = ld<true>(mag + f) * ld<true>(delta + f);
  gg[f] = -sd * ((x[p] - x[q]) * (lam[p] / vol[p] - lam[q] / vol[q]));
}

// Dirichlet-face gradients: g_b = |S_b| delta_b * w[P] / V_P with
// w = g_diag (chain rule) or w = -lam*x (fused adjoint).
__global__ __launch_bounds__(kBlock) void k_dir_grad(int nd, const int32_t *__restrict__ cell,
                                                      const double *__restrict__ w,
                                                      const double *__restrict__ lam,
                                                      const double *__restrict__ x,
                                                      const double *__restrict__ mag,
                                                      const double *__restrict__ dlt,
                                                      const double *__restrict__ vol,
                                                      double *__restrict__ g) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nd) return;
  const int c = cell[k];
  const double wc = w ? w[c] : -(lam[c] * x[c]);
  g[k] = (mag[k] * dlt[k]) * wc / vol[c];
}

constexpr int kDiluLaunchLevels = 64;

// PFM_DILU_LDS=0 keeps the DILU sweep on the global-memory diagonal (for
// comparing the two forms; read at every factorisation).
bool dilu_lds_enabled() {
  const char *e = std::getenv("PFM_DILU_LDS");
  return !(e && std::strcmp(e, "0") == 0);
}

inline int grid1d(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

template <typename T>
int upload(T **dst, const std::vector<T> &src) {
  *dst = nullptr;
  if (src.empty()) return PFM_OK;
  PFM_CHECK_HIP(hipMalloc((void **)dst, src.size() * sizeof(T)));
  PFM_CHECK_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return PFM_OK;
}

}  // namespace

namespace pfm {
// The handle's Gauss-Seidel schedule entries, built on first use (48 bytes per
// cell and direction: only solves that sweep Gauss-Seidel need them).  The
// build finishes before the entries are published, so a solve on another
// stream never reads them half-built.
int ensure_gs_entries(pfm_ldu_s *h, hipStream_t s) {
  std::lock_guard<std::mutex> lock(h->gs_ent_mu);
  if (h->d_gs_fwd_ent || h->n == 0) return PFM_OK;
  const int nf = (int)h->gs_fwd_ptr.back(), nb = (int)h->gs_bwd_ptr.back();
  GsEnt *fwd = nullptr, *bwd = nullptr;
  PFM_CHECK_HIP(hipMalloc((void **)&fwd, sizeof(GsEnt) * (size_t)std::max(nf, 1)));
  if (hipMalloc((void **)&bwd, sizeof(GsEnt) * (size_t)std::max(nb, 1)) != hipSuccess) {
    (void)hipFree(fwd);
    set_error("hipMalloc of the Gauss-Seidel schedule entries failed");
    return PFM_ERR_HIP;
  }
  if (nf) k_gs_entries<<<grid1d(nf), kBlock, 0, s>>>(nf, h->d_gs_fwd_cells, h->d_slices, h->d_col,
                                                     h->d_map, fwd);
  if (nb) k_gs_entries<<<grid1d(nb), kBlock, 0, s>>>(nb, h->d_gs_bwd_cells, h->d_slices, h->d_col,
                                                     h->d_map, bwd);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFree(fwd);
    (void)hipFree(bwd);
    set_error("building the Gauss-Seidel schedule entries failed");
    return PFM_ERR_HIP;
  }
  h->d_gs_bwd_ent = bwd;
  h->d_gs_fwd_ent = fwd;
  return PFM_OK;
}

int ensure_dilu_slots(pfm_ldu_s *h, hipStream_t s) {
  std::lock_guard<std::mutex> lock(h->dilu_slots_mu);
  if (h->dilu_slots_state != 0) return PFM_OK;
  const int nf = h->nf, n = h->n;
  const int nl = (int)h->dilu_level_ptr.size() - 1;
  if (nf == 0 || nl <= 0) {
    h->dilu_slots_state = -1;
    return PFM_OK;
  }
  std::vector<int32_t> faces(nf), own(nf), nei(nf);
  PFM_CHECK_HIP(hipMemcpyAsync(faces.data(), h->d_dilu_faces, sizeof(int32_t) * nf,
                               hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipMemcpyAsync(own.data(), h->d_own, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipMemcpyAsync(nei.data(), h->d_nei, sizeof(int32_t) * nf, hipMemcpyDeviceToHost, s));
  PFM_CHECK_HIP(hipStreamSynchronize(s));
  const std::vector<int32_t> &lp = h->dilu_level_ptr;
  // each cell's window of levels
  std::vector<int32_t> first(n, INT32_MAX), last(n, -1);
  for (int l = 0; l < nl; ++l)
    for (int k = lp[l]; k < lp[l + 1]; ++k) {
      const int f = faces[k];
      for (const int c : {own[f], nei[f]}) {
        first[c] = std::min(first[c], l);
        last[c] = std::max(last[c], l);
      }
    }
  // cells by first and by last level (counting sorts)
  std::vector<int32_t> fptr(nl + 2, 0), lptr(nl + 2, 0);
  for (int c = 0; c < n; ++c)
    if (last[c] >= 0) {
      ++fptr[first[c] + 1];
      ++lptr[last[c] + 1];
    }
  for (int l = 0; l <= nl; ++l) {
    fptr[l + 1] += fptr[l];
    lptr[l + 1] += lptr[l];
  }
  std::vector<int32_t> byfirst(fptr[nl + 1]), bylast(lptr[nl + 1]);
  {
    std::vector<int32_t> fp(fptr.begin(), fptr.end() - 1), lq(lptr.begin(), lptr.end() - 1);
    for (int c = 0; c < n; ++c)
      if (last[c] >= 0) {
        byfirst[fp[first[c]]++] = c;
        bylast[lq[last[c]]++] = c;
      }
  }
  // slots; phase l: write-backs (window ended at l - 1), loads (window starts
  // at l + 1); a slot written back in phase l is reused from phase l + 1 on
  std::vector<int32_t> slot(n, -1), freelist, released;
  std::vector<int2> io;
  std::vector<int32_t> iptr(nl + 3, 0);
  io.reserve(2 * (size_t)fptr[nl + 1]);
  int nslots = 0;
  auto take = [&](int c) {
    int v;
    if (!freelist.empty()) {
      v = freelist.back();
      freelist.pop_back();
    } else {
      v = nslots++;
    }
    slot[c] = v;
    io.push_back(make_int2(c, v));
  };
  auto write_back = [&](int l) {  // cells whose window ended at level l
    for (int k = lptr[l]; k < lptr[l + 1]; ++k) {
      const int c = bylast[k];
      io.push_back(make_int2(c, -(slot[c] + 1)));
      released.push_back(slot[c]);
    }
  };
  for (int k = fptr[0]; k < fptr[1]; ++k) take(byfirst[k]);  // before level 0
  iptr[1] = (int32_t)io.size();
  for (int l = 0; l < nl && nslots <= kDiluSlots; ++l) {
    if (l > 0) write_back(l - 1);
    if (l + 1 < nl)
      for (int k = fptr[l + 1]; k < fptr[l + 2]; ++k) take(byfirst[k]);
    freelist.insert(freelist.end(), released.begin(), released.end());
    released.clear();
    iptr[l + 2] = (int32_t)io.size();
  }
  if (nslots > kDiluSlots) {  // more cells live at once than LDS holds
    h->dilu_slots_state = -1;
    return PFM_OK;
  }
  write_back(nl - 1);  // phase nl
  iptr[nl + 2] = (int32_t)io.size();
  int widest = 0;  // faces or phase entries of the fullest level
  for (int l = 0; l < nl; ++l)
    widest = std::max(widest, std::max(lp[l + 1] - lp[l], iptr[l + 2] - iptr[l + 1]));
  if (nslots > kDiluSlots || io.empty() || widest > kDiluThreads) {
    h->dilu_slots_state = -1;
    return PFM_OK;
  }
  h->dilu_lds_threads = widest <= 512 ? 512 : kDiluThreads;
  std::vector<uint32_t> fslot(nf);
  for (int k = 0; k < nf; ++k) {
    const int f = faces[k];
    fslot[k] = (uint32_t)slot[own[f]] | ((uint32_t)slot[nei[f]] << 16);
  }
  uint32_t *d_fslot = nullptr;
  int2 *d_io = nullptr;
  int32_t *d_iptr = nullptr;
  double *d_din = nullptr;
  const size_t nio = io.size();
  auto fail = [&] {
    (void)hipFree(d_fslot);
    (void)hipFree(d_io);
    (void)hipFree(d_iptr);
    (void)hipFree(d_din);
    set_error("building the DILU slot schedule failed");
    return PFM_ERR_HIP;
  };
  if (hipMalloc((void **)&d_fslot, sizeof(uint32_t) * nf) != hipSuccess ||
      hipMalloc((void **)&d_io, sizeof(int2) * std::max<size_t>(nio, 1)) != hipSuccess ||
      hipMalloc((void **)&d_iptr, sizeof(int32_t) * (nl + 3)) != hipSuccess ||
      hipMalloc((void **)&d_din, sizeof(double) * nio) != hipSuccess)
    return fail();
  if (hipMemcpyAsync(d_fslot, fslot.data(), sizeof(uint32_t) * nf, hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      (nio && hipMemcpyAsync(d_io, io.data(), sizeof(int2) * nio, hipMemcpyHostToDevice, s) !=
                  hipSuccess) ||
      hipMemcpyAsync(d_iptr, iptr.data(), sizeof(int32_t) * (nl + 3), hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail();
  h->d_dilu_fslot = d_fslot;
  h->d_dilu_io = d_io;
  h->d_dilu_iptr = d_iptr;
  h->d_dilu_din = d_din;
  h->n_dilu_io = (int32_t)nio;
  h->dilu_slots_state = 1;
  return PFM_OK;
}
}  // namespace pfm

extern "C" {

int pfm_ldu_destroy(pfm_ldu h) {
  if (!h) return PFM_OK;
  int dev_prev = 0;
  (void)hipGetDevice(&dev_prev);
  (void)hipSetDevice(h->device);
  (void)hipFree(h->d_slices);
  (void)hipFree(h->d_col);
  (void)hipFree(h->d_col16);
  (void)hipFree(h->d_colp2_alloc);
  (void)hipFree(h->d_map);
  (void)hipFree(h->d_own);
  (void)hipFree(h->d_nei);
  (void)hipFree(h->d_own_start);
  (void)hipFree(h->d_dic_cells);
  (void)hipFree(h->d_dilu_faces);
  (void)hipFree(h->d_dilu_level_ptr);
  (void)hipFree(h->d_dilu_pq);
  (void)hipFree(h->d_dilu_lu);
  (void)hipFree(h->d_dilu_fslot);
  (void)hipFree(h->d_dilu_io);
  (void)hipFree(h->d_dilu_iptr);
  (void)hipFree(h->d_dilu_din);
  (void)hipFree(h->d_gs_fwd_ptr);
  (void)hipFree(h->d_gs_fwd_cells);
  (void)hipFree(h->d_gs_bwd_ptr);
  (void)hipFree(h->d_gs_bwd_cells);
  (void)hipFree(h->d_gs_fwd_ent);
  (void)hipFree(h->d_gs_bwd_ent);
  poll_pool_free(h->poll);
  (void)hipSetDevice(dev_prev);
  delete h;
  return PFM_OK;
}

int pfm_ldu_create(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei, pfm_ldu *out) {
  PFM_REQUIRE(out, "null output handle");
  *out = nullptr;
  HostLayout L;
  int st = build_host_layout(n, nf, own, nei, L);
  if (st != PFM_OK) return st;
  pfm_ldu h = new pfm_ldu_s();
  h->n = n;
  h->nf = nf;
  h->nslices = L.nslices;
  h->nslots = L.nslots;
  h->max_lower = L.max_lower;
  h->max_upper = L.max_upper;
  h->dic_level_ptr = L.dic_level_ptr;
  h->dilu_level_ptr = L.dilu_level_ptr;
  h->gs_fwd_ptr = L.gs_fwd_ptr;
  h->gs_bwd_ptr = L.gs_bwd_ptr;
  h->oriented_sorted = true;
  for (int32_t f = 0; f < nf && h->oriented_sorted; ++f)
    h->oriented_sorted = own[f] < nei[f] && (f == 0 || own[f - 1] <= own[f]);
  hipError_t e = hipGetDevice(&h->device);
  if (e != hipSuccess) {
    delete h;
    set_error("no HIP device: %s", hipGetErrorString(e));
    return PFM_ERR_HIP;
  }
  std::vector<int32_t> o(own, own + nf), q(nei, nei + nf);
  if ((st = upload(&h->d_slices, L.slices)) || (st = upload(&h->d_col, L.col)) ||
      (st = upload(&h->d_map, L.map)) || (st = upload(&h->d_own, o)) ||
      (st = upload(&h->d_nei, q)) || (st = upload(&h->d_dic_cells, L.dic_cells)) ||
      (st = upload(&h->d_dilu_faces, L.dilu_faces)) ||
      (st = upload(&h->d_dilu_level_ptr, L.dilu_level_ptr)) ||
      (st = upload(&h->d_gs_fwd_ptr, L.gs_fwd_ptr)) ||
      (st = upload(&h->d_gs_fwd_cells, L.gs_fwd_cells)) ||
      (st = upload(&h->d_gs_bwd_ptr, L.gs_bwd_ptr)) ||
      (st = upload(&h->d_gs_bwd_cells, L.gs_bwd_cells))) {
    pfm_ldu_destroy(h);
    return st;
  }
  if (h->oriented_sorted) {  // first owner-face of each row (faces sorted by owner)
    std::vector<int32_t> os(n + 1, nf);
    for (int32_t f = nf - 1; f >= 0; --f) os[own[f]] = f;
    for (int32_t r = n - 1; r >= 0; --r) os[r] = std::min(os[r], os[r + 1]);
    if ((st = upload(&h->d_own_start, os))) {
      pfm_ldu_destroy(h);
      return st;
    }
  }
  {
    // Column offsets from the row in int16 when every one fits: the PCG's
    // grid-stride SpMV then reads 2 index bytes per slot instead of 4 (3 %
    // faster per iteration at 128^3).  PFM_COL16=0 keeps the int32 columns.
    std::vector<int16_t> c16(L.col.size(), (int16_t)kColPad16);
    const char *env = std::getenv("PFM_COL16");
    bool fits = !(env && env[0] == '0');
    for (size_t s = 0; s < L.slices.size() && fits; ++s) {
      const SliceInfo si = L.slices[s];
      const int width = si.wl + si.wu;
      for (int j = 0; j < width && fits; ++j)
        for (int lane = 0; lane < kSlice; ++lane) {
          const int64_t k = (int64_t)si.ptr + (int64_t)j * kSlice + lane;
          const int32_t c = L.col[k];
          if (c < 0) continue;
          const int64_t d = (int64_t)c - ((int64_t)s * kSlice + lane);
          if (d <= kColPad16 || d > 32767) {
            fits = false;
            break;
          }
          c16[k] = (int16_t)d;
        }
    }
    if (fits && (st = upload(&h->d_col16, c16))) {
      pfm_ldu_destroy(h);
      return st;
    }
    // The offsets two slots per 4-byte word (pair_base layout: slice s owns
    // 32 (wl + wu + 1) words from ptr/2 + 32 s), in one of two encodings,
    // recorded in the word before the array (kPairHeader words precede it):
    // on an oriented topology (every owner < its neighbour) whose columns lie
    // within 65535 of their rows, unsigned magnitudes (a lower-group column is
    // row + u, an upper-group one row - u, 0 = padding: twice the signed
    // reach); otherwise the signed int16 offsets above when they fit.
    bool ori = !(env && env[0] == '0');
    for (int32_t f = 0; f < nf && ori; ++f) ori = own[f] < nei[f] && nei[f] - own[f] <= 65535;
    if (ori || fits) {
      const uint32_t pad = ori ? 0u : 0x8000u;
      const int64_t nwords = L.nslots / 2 + (int64_t)32 * L.nslices;
      std::vector<uint32_t> p2((size_t)(kPairHeader + nwords), pad | (pad << 16));
      p2[kPairHeader - 1] = ori ? kPairOriented : kPairSigned;
      for (size_t s = 0; s < L.slices.size(); ++s) {
        const SliceInfo si = L.slices[s];
        const int64_t pb = kPairHeader + si.ptr / 2 + (int64_t)32 * s;
        const int width = si.wl + si.wu;
        for (int j = 0; j < width; ++j)
          for (int lane = 0; lane < kSlice; ++lane) {
            const int32_t c = L.col[(int64_t)si.ptr + (int64_t)j * kSlice + lane];
            const int64_t d = (int64_t)c - ((int64_t)s * kSlice + lane);
            uint32_t half = pad;
            if (c >= 0) half = ori ? (uint32_t)(j < si.wl ? d : -d) : (uint32_t)(uint16_t)(int16_t)d;
            uint32_t &w = p2[pb + (int64_t)(j >> 1) * kSlice + lane];
            w = (j & 1) ? ((w & 0xffffu) | (half << 16)) : ((w & 0xffff0000u) | half);
          }
      }
      if ((st = upload(&h->d_colp2_alloc, p2))) {
        pfm_ldu_destroy(h);
        return st;
      }
      h->d_colp2 = h->d_colp2_alloc + kPairHeader;
    }
  }
  if ((st = poll_pool_init(h->poll)) != PFM_OK) {
    pfm_ldu_destroy(h);
    return st;
  }
  *out = h;
  return PFM_OK;
}

int pfm_ldu_set_progress(pfm_ldu h, pfm_progress_fn fn, void *user, double interval_s) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(interval_s >= 0.0, "negative progress interval");
  h->progress = fn;
  h->progress_user = user;
  h->progress_interval = interval_s;
  return PFM_OK;
}

int pfm_ldu_get_info(pfm_ldu h, pfm_ldu_info *out) {
  PFM_REQUIRE(h && out, "null handle or output");
  out->n_cells = h->n;
  out->n_faces = h->nf;
  out->n_slices = h->nslices;
  out->n_dic_levels = (int32_t)h->dic_level_ptr.size() - 1;
  out->n_dilu_levels = (int32_t)h->dilu_level_ptr.size() - 1;
  out->max_lower = h->max_lower;
  out->max_upper = h->max_upper;
  out->device = h->device;
  out->n_slots = h->nslots;
  return PFM_OK;
}

int pfm_ldu_pack(pfm_ldu h, int transpose, const double *lower, const double *upper,
                 double *packed, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->nslices == 0 || h->nslots == 0) return PFM_OK;
  PFM_NONNULL(packed);
  if (h->nf > 0) {
    PFM_NONNULL(lower);
    PFM_NONNULL(upper);
  }
  const double *lo = transpose ? upper : lower, *up = transpose ? lower : upper;
  k_pack<double><<<blocks_for_slices(h->nslices), kBlock, 0, (hipStream_t)stream>>>(
      h->nslots, h->nslices, h->d_slices, h->d_map, lo, up, packed);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_ldu_pack_f32(pfm_ldu h, int transpose, const double *lower, const double *upper,
                     float *packed, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->nslices == 0 || h->nslots == 0) return PFM_OK;
  PFM_NONNULL(packed);
  if (h->nf > 0) {
    PFM_NONNULL(lower);
    PFM_NONNULL(upper);
  }
  const double *lo = transpose ? upper : lower, *up = transpose ? lower : upper;
  k_pack<float><<<blocks_for_slices(h->nslices), kBlock, 0, (hipStream_t)stream>>>(
      h->nslots, h->nslices, h->d_slices, h->d_map, lo, up, packed);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_spmv(pfm_ldu h, const double *diag, const double *packed, const double *x, double *y,
             void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(x);
  PFM_NONNULL(y);
  if (h->nslots) PFM_NONNULL(packed);
  // One slice per wave with the int32 columns: the fastest form standalone
  // (the int16 offsets pay off in the solvers' grid-stride kernels only).
  return launch_spmv<double>(h, diag, packed, x, y, (hipStream_t)stream);
}

int pfm_spmv_multi(pfm_ldu h, const double *diag, const double *packed, int32_t k,
                   const double *x, double *y, void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(k >= 1, "k must be >= 1");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(x);
  PFM_NONNULL(y);
  if (h->nslots) PFM_NONNULL(packed);
  k_spmv_multi<<<blocks_for_slices(h->nslices), kBlock, 0, (hipStream_t)stream>>>(
      h->n, h->nslices, k, h->d_slices, h->d_col, packed, diag, x, y);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_relax(int32_t n, double alpha, const double *diag, const double *source,
              const double *field_old, double *diag_out, double *source_out, void *stream) {
  PFM_REQUIRE(n >= 0, "negative n");
  PFM_REQUIRE(alpha > 0.0 && alpha <= 1.0, "under-relaxation factor must be in (0, 1], got %g",
              alpha);
  if (n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(source);
  PFM_NONNULL(field_old);
  PFM_NONNULL(diag_out);
  PFM_NONNULL(source_out);
  k_relax<<<grid1d(n), kBlock, 0, (hipStream_t)stream>>>(n, alpha, diag, source, field_old,
                                                         diag_out, source_out);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_bandwidth_probe(int mode, const void *src, void *dst, int64_t nbytes, void *stream) {
  PFM_REQUIRE(mode == 0 || mode == 1, "unknown probe mode %d", mode);
  PFM_REQUIRE(nbytes >= 0 && nbytes % 16 == 0, "nbytes must be a non-negative multiple of 16");
  PFM_REQUIRE(((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & 15) == 0,
              "buffers must be 16-byte aligned");
  PFM_NONNULL(src);
  PFM_NONNULL(dst);
  hipStream_t s = (hipStream_t)stream;
  if (mode == 0)
    k_bw_read<<<kBwGrid, kBlock, 0, s>>>((const dvec2 *)src, nbytes / 16, (double *)dst);
  else
    k_bw_copy<<<kBwGrid, kBlock, 0, s>>>((const dvec2 *)src, (dvec2 *)dst, nbytes / 16);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int64_t pfm_bandwidth_probe_out_bytes(void) {
  return (int64_t)kBwGrid * kWavesPerBlock * sizeof(double);
}

int pfm_spmv_f32(pfm_ldu h, const float *diag, const float *packed, const float *x, float *y,
                 void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(x);
  PFM_NONNULL(y);
  if (h->nslots) PFM_NONNULL(packed);
  return launch_spmv<float>(h, diag, packed, x, y, (hipStream_t)stream);
}

int pfm_spmv_ldu(pfm_ldu h, int transpose, const double *diag, const double *lower,
                 const double *upper, const double *x, double *y, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(x);
  PFM_NONNULL(y);
  if (h->nf) {
    PFM_NONNULL(lower);
    PFM_NONNULL(upper);
  }
  const int G = blocks_for_slices(h->nslices);
  if (transpose)
    k_spmv_ldu<true><<<G, kBlock, 0, (hipStream_t)stream>>>(h->n, h->nslices, h->d_slices, h->d_col,
                                                           h->d_map, diag, lower, upper, x, y);
  else
    k_spmv_ldu<false><<<G, kBlock, 0, (hipStream_t)stream>>>(h->n, h->nslices, h->d_slices, h->d_col,
                                                            h->d_map, diag, lower, upper, x, y);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_residual(pfm_ldu h, const double *diag, const double *packed, const double *x,
                 const double *b, double *r, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(x);
  PFM_NONNULL(b);
  PFM_NONNULL(r);
  if (h->nslots) PFM_NONNULL(packed);
  k_residual<<<blocks_for_slices(h->nslices), kBlock, 0, (hipStream_t)stream>>>(
      h->n, h->nslices, h->d_slices, h->d_col, packed, diag, x, b, r);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_dic_factor(pfm_ldu h, const double *diag, const double *packed, double *rD, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(rD);
  hipStream_t s = (hipStream_t)stream;
  if (h->nf == 0) {
    PFM_CHECK_HIP(hipMemcpyAsync(rD, diag, sizeof(double) * h->n, hipMemcpyDeviceToDevice, s));
  } else {
    PFM_NONNULL(packed);
    const int nl = (int)h->dic_level_ptr.size() - 1;
    for (int l = 0; l < nl; ++l) {
      const int b = h->dic_level_ptr[l], cnt = h->dic_level_ptr[l + 1] - b;
      k_dic_level<<<grid1d(cnt), kBlock, 0, s>>>(h->d_dic_cells + b, cnt, h->d_slices, h->d_col,
                                                 packed, diag, rD);
    }
    PFM_LAUNCH_CHECK();
  }
  k_abs_clamp<<<grid1d(h->n), kBlock, 0, s>>>(h->n, rD);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_dilu_factor(pfm_ldu h, int transpose, const double *diag, const double *lower,
                    const double *upper, double *rD, void *stream) {
  PFM_REQUIRE(h, "null handle");
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(diag);
  PFM_NONNULL(rD);
  hipStream_t s = (hipStream_t)stream;
  PFM_CHECK_HIP(hipMemcpyAsync(rD, diag, sizeof(double) * h->n, hipMemcpyDeviceToDevice, s));
  if (h->nf) {
    PFM_NONNULL(lower);
    PFM_NONNULL(upper);
    const double *lo = transpose ? upper : lower, *up = transpose ? lower : upper;
    const int nl = (int)h->dilu_level_ptr.size() - 1;
    if (nl <= kDiluLaunchLevels) {
      for (int l = 0; l < nl; ++l) {
        const int b = h->dilu_level_ptr[l], cnt = h->dilu_level_ptr[l + 1] - b;
        k_dilu_level<<<grid1d(cnt), kBlock, 0, s>>>(h->d_dilu_faces + b, cnt, h->d_own, h->d_nei,
                                                    lo, up, rD);
      }
    } else {
      if (!h->d_dilu_pq) {
        PFM_CHECK_HIP(hipMalloc((void **)&h->d_dilu_pq, sizeof(int2) * (size_t)h->nf));
        PFM_CHECK_HIP(hipMalloc((void **)&h->d_dilu_lu, sizeof(double) * (size_t)h->nf));
      }
      if (dilu_lds_enabled()) {
        const int rc = ensure_dilu_slots(h, s);
        if (rc != PFM_OK) return rc;
      }
      k_dilu_gather<<<grid1d(h->nf), kBlock, 0, s>>>(h->nf, h->d_dilu_faces, h->d_own, h->d_nei,
                                                     lo, up, (int2 *)h->d_dilu_pq, h->d_dilu_lu);
      if (dilu_lds_enabled() && h->dilu_slots_state == 1) {
        if (h->n_dilu_io)
          k_dilu_din<<<grid1d(h->n_dilu_io), kBlock, 0, s>>>(h->n_dilu_io, h->d_dilu_io, rD,
                                                             h->d_dilu_din);
        if (h->dilu_lds_threads == 512)
          k_dilu_lds<512><<<1, 512, 0, s>>>(nl, h->nf, h->n_dilu_io, h->d_dilu_level_ptr,
                                            h->d_dilu_fslot, h->d_dilu_lu, h->d_dilu_iptr,
                                            h->d_dilu_io, h->d_dilu_din, rD);
        else
          k_dilu_lds<kDiluThreads><<<1, kDiluThreads, 0, s>>>(
              nl, h->nf, h->n_dilu_io, h->d_dilu_level_ptr, h->d_dilu_fslot, h->d_dilu_lu,
              h->d_dilu_iptr, h->d_dilu_io, h->d_dilu_din, rD);
      } else {
        k_dilu_sweep<<<1, kDiluThreads, 0, s>>>(h->d_dilu_level_ptr, nl,
                                                 (const int2 *)h->d_dilu_pq, h->d_dilu_lu, rD);
      }
    }
    PFM_LAUNCH_CHECK();
  }
  k_abs_clamp<<<grid1d(h->n), kBlock, 0, s>>>(h->n, rD);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_dirichlet_create(pfm_ldu h, int32_t n_dir, const int32_t *dir_cell_h, int32_t n_patches,
                         const int32_t *patch_ptr_h, pfm_dirichlet *out) {
  PFM_REQUIRE(out, "null output handle");
  *out = nullptr;
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(n_dir >= 0 && n_patches >= 0, "negative size");
  PFM_REQUIRE(n_dir == 0 || (dir_cell_h && patch_ptr_h && n_patches >= 1),
              "Dirichlet faces need cells and patch offsets");
  std::vector<int32_t> patch(n_dir);
  if (n_dir) {
    PFM_REQUIRE(patch_ptr_h[0] == 0 && patch_ptr_h[n_patches] == n_dir,
                "patch offsets must run from 0 to n_dir");
    for (int32_t p = 0; p < n_patches; ++p) {
      PFM_REQUIRE(patch_ptr_h[p] <= patch_ptr_h[p + 1], "patch offsets must not decrease");
      for (int32_t j = patch_ptr_h[p]; j < patch_ptr_h[p + 1]; ++j) patch[j] = p;
    }
    for (int32_t j = 0; j < n_dir; ++j)
      PFM_REQUIRE(dir_cell_h[j] >= 0 && dir_cell_h[j] < h->n,
                  "Dirichlet cell %d of entry %d outside [0, %d)", dir_cell_h[j], j, h->n);
  }
  // entries grouped by cell; within a cell in the caller's (patch, face) order
  std::vector<int32_t> order(n_dir);
  for (int32_t j = 0; j < n_dir; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t b) { return dir_cell_h[a] < dir_cell_h[b]; });
  std::vector<int32_t> gptr, gcell, spatch(n_dir);
  for (int32_t j = 0; j < n_dir; ++j) {
    const int32_t c = dir_cell_h[order[j]];
    if (j == 0 || c != gcell.back()) {
      gptr.push_back(j);
      gcell.push_back(c);
    }
    spatch[j] = patch[order[j]];
  }
  gptr.push_back(n_dir);
  pfm_dirichlet d = new pfm_dirichlet_s();
  d->n_dir = n_dir;
  d->n_groups = (int32_t)gcell.size();
  d->ldu = h;
  int st;
  if ((st = upload(&d->d_gptr, gptr)) || (st = upload(&d->d_gcell, gcell)) ||
      (st = upload(&d->d_order, order)) || (st = upload(&d->d_patch, spatch))) {
    pfm_dirichlet_destroy(d);
    return st;
  }
  *out = d;
  return PFM_OK;
}

int pfm_dirichlet_destroy(pfm_dirichlet d) {
  if (!d) return PFM_OK;
  (void)hipFree(d->d_gptr);
  (void)hipFree(d->d_gcell);
  (void)hipFree(d->d_order);
  (void)hipFree(d->d_patch);
  delete d;
  return PFM_OK;
}

int pfm_laplacian_assemble(pfm_ldu h, const double *gamma, const double *mag_sf,
                           const double *delta, const double *vol, pfm_dirichlet dir,
                           const double *dir_gamma, const double *dir_mag_sf,
                           const double *dir_delta, double *diag, double *lower, double *upper,
                           void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(!dir || dir->ldu == h, "Dirichlet plan made for another handle");
  const int32_t n_dir = dir ? dir->n_dir : 0;
  hipStream_t s = (hipStream_t)stream;
  if (h->n == 0) return PFM_OK;
  PFM_NONNULL(vol);
  PFM_NONNULL(diag);
  if (h->nf) {
    PFM_NONNULL(gamma);
    PFM_NONNULL(mag_sf);
    PFM_NONNULL(delta);
    PFM_NONNULL(lower);
    PFM_NONNULL(upper);
    k_lap_faces<<<grid1d(h->nf), kBlock, 0, s>>>(h->nf, gamma, mag_sf, delta, vol, h->d_own,
                                                 h->d_nei, lower, upper);
  }
  k_lap_diag<<<blocks_for_slices(h->nslices), kBlock, 0, s>>>(
      h->n, h->nslices, h->d_slices, h->d_map, h->oriented_sorted ? h->d_own_start : nullptr,
      lower, upper, diag);
  if (n_dir > 0) {
    PFM_NONNULL(dir_gamma);
    PFM_NONNULL(dir_mag_sf);
    PFM_NONNULL(dir_delta);
    k_lap_dirichlet<<<grid1d(dir->n_groups), kBlock, 0, s>>>(
        dir->n_groups, dir->d_gptr, dir->d_gcell, dir->d_order, dir->d_patch, dir_gamma,
        dir_mag_sf, dir_delta, vol, diag);
  }
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_adjoint_ldu_grads(pfm_ldu h, const double *lam, const double *x, double *g_diag,
                          double *g_lower, double *g_upper, void *stream) {
  PFM_REQUIRE(h, "null handle");
  hipStream_t s = (hipStream_t)stream;
  PFM_NONNULL(lam);
  PFM_NONNULL(x);
  if (g_diag && h->n) k_adj_diag<<<grid1d(h->n), kBlock, 0, s>>>(h->n, lam, x, g_diag);
  if ((g_lower || g_upper) && h->nf)
    k_adj_faces<<<grid1d(h->nf), kBlock, 0, s>>>(h->nf, h->d_own, h->d_nei, lam, x, g_lower,
                                                 g_upper);
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_laplacian_backward(pfm_ldu h, const double *g_diag, const double *g_lower,
                           const double *g_upper, const double *mag_sf, const double *delta,
                           const double *vol, int32_t n_dir, const int32_t *dir_cell,
                           const double *dir_mag_sf, const double *dir_delta, double *g_gamma,
                           double *g_dir_gamma, void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(n_dir >= 0, "negative n_dir");
  hipStream_t s = (hipStream_t)stream;
  PFM_NONNULL(g_diag);
  PFM_NONNULL(vol);
  if (h->nf) {
    PFM_NONNULL(g_lower);
    PFM_NONNULL(g_upper);
    PFM_NONNULL(mag_sf);
    PFM_NONNULL(delta);
    PFM_NONNULL(g_gamma);
    k_lap_backward<<<grid1d(h->nf), kBlock, 0, s>>>(h->nf, h->d_own, h->d_nei, g_diag, g_lower,
                                                    g_upper, mag_sf, delta, vol, g_gamma);
  }
  if (n_dir > 0 && g_dir_gamma) {
    PFM_NONNULL(dir_cell);
    PFM_NONNULL(dir_mag_sf);
    PFM_NONNULL(dir_delta);
    k_dir_grad<<<grid1d(n_dir), kBlock, 0, s>>>(n_dir, dir_cell, g_diag, nullptr, nullptr,
                                                dir_mag_sf, dir_delta, vol, g_dir_gamma);
  }
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

int pfm_adjoint_gamma_grad(pfm_ldu h, const double *lam, const double *x, const double *mag_sf,
                           const double *delta, const double *vol, int32_t n_dir,
                           const int32_t *dir_cell, const double *dir_mag_sf,
                           const double *dir_delta, double *g_gamma, double *g_dir_gamma,
                           void *stream) {
  PFM_REQUIRE(h, "null handle");
  PFM_REQUIRE(n_dir >= 0, "negative n_dir");
  hipStream_t s = (hipStream_t)stream;
  PFM_NONNULL(lam);
  PFM_NONNULL(x);
  PFM_NONNULL(vol);
  if (h->nf) {
    PFM_NONNULL(mag_sf);
    PFM_NONNULL(delta);
    PFM_NONNULL(g_gamma);
    k_adj_gamma<<<grid1d(h->nf), kBlock, 0, s>>>(h->nf, h->d_own, h->d_nei, lam, x, mag_sf, delta,
                                                 vol, g_gamma);
  }
  if (n_dir > 0 && g_dir_gamma) {
    PFM_NONNULL(dir_cell);
    PFM_NONNULL(dir_mag_sf);
    PFM_NONNULL(dir_delta);
    k_dir_grad<<<grid1d(n_dir), kBlock, 0, s>>>(n_dir, dir_cell, nullptr, lam, x, dir_mag_sf,
                                                dir_delta, vol, g_dir_gamma);
  }
  PFM_LAUNCH_CHECK();
  return PFM_OK;
}

}  // extern "C"
