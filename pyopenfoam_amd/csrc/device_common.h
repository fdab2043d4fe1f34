// Device-side helpers shared by the libpfm kernels.
#pragma once

#include <cstdlib>
#include <cstring>

#include "pfm_internal.h"

namespace pfm {

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Remapping gives each XCD one contiguous range of logical blocks, so rows
// that share x[] cache lines (RCM keeps neighbours within a narrow band) are
// served from the same XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
  const int per = G / kXcds;
  const int full = per * kXcds;
  if (b >= full) return b;
  return (b % kXcds) * per + (b / kXcds);
}

__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic block sum (fixed shuffle pattern, waves added in order).
// Every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double *lds /*[kWavesPerBlock]*/) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double s = lds[0];
#pragma unroll
  for (int i = 1; i < kWavesPerBlock; ++i) s += lds[i];
  __syncthreads();
  return s;
}

// Sum of partials[0..np) in a fixed order, result on every thread of the
// (single) finalize workgroup.
__device__ __forceinline__ double sum_partials(const double *partials, int np, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += kBlock) s += partials[i];
  return block_sum(s, lds);
}

// Producer kernels use a fixed grid of at most kMaxRedBlocks workgroups so
// that their per-block partial sums (reduced by a one-workgroup finalize
// kernel in a fixed order) stay few; each workgroup owns one contiguous range
// of slices, its four waves interleaved over it.
#ifndef PFM_MAX_RED_BLOCKS
#define PFM_MAX_RED_BLOCKS 2048
#endif
constexpr int kMaxRedBlocks = PFM_MAX_RED_BLOCKS;

struct SliceRange {
  int begin, end;
};

__device__ __forceinline__ SliceRange block_slices(int nslices) {
  const int G = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, G);
  const int per = (nslices + G - 1) / G;
  const int b = lb * per;
  SliceRange r;
  r.begin = b < nslices ? b : nslices;
  r.end = b + per < nslices ? b + per : nslices;
  return r;
}

// Load of an element that one pass reads once (kNT: a nontemporal load, so the
// stream does not displace data that is read again from L2 and the MALL).
template <bool kNT, typename T>
__device__ __forceinline__ T ld(const T *p) {
  if constexpr (kNT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// Column of slot k of a row: the int32 column (-1 = padding slot), or an
// int16 offset from the row (kColPad16 = padding slot), the compact form the
// handle keeps when every offset fits (half the index bytes per slot).
constexpr int kColPad16 = -32768;
template <bool kNT = false>
struct ColAbsT {
  const int32_t *__restrict__ c;
  __device__ __forceinline__ int operator()(int k) const { return ld<kNT>(c + k); }
};
template <bool kNT = false>
struct ColRelT {
  const int16_t *__restrict__ d;
  int row;
  __device__ __forceinline__ int operator()(int k) const {
    const int v = ld<kNT>(d + k);
    return v == kColPad16 ? -1 : row + v;
  }
};
using ColAbs = ColAbsT<false>;
using ColRel = ColRelT<false>;

// One row of y = A x on the packed row-slice layout.  Returns
// (diag*x + sum_lower) + sum_upper, each sum in face order from 0, which is
// the association of the reference's Ax (diag*x, then scatter of the owner
// side, then scatter of the neighbour side).
// `gx(c)` supplies the x value of column c (a plain load, or a value formed
// where it is read).  kUpperFirst returns (diag*x + sum_upper) + sum_lower,
// the association of the reference's Ax_T (ldu_matrix.py:248-257: the
// neighbour-side scatter first) when the packed values are the transposed
// ones.  kNT loads the coefficients nontemporally (pass a ColF of the same
// kind for the columns).
template <typename T, typename Gather, bool kUpperFirst = false, typename ColF = ColAbs,
          bool kNT = false>
__device__ __forceinline__ T row_spmv_g(const SliceInfo si, int lane, T dx, ColF col,
                                        const T *__restrict__ val, Gather gx) {
  constexpr int U = 4;
  T accL = T(0), accU = T(0);
  const int wl = si.wl, wu = si.wu;
  const int bl = si.ptr + lane, bu = si.ptr + wl * kSlice + lane;
  if (wl <= U && wu <= U) {
    // Common case (hexahedral and most polyhedral rows): all column and
    // coefficient loads of both groups in flight at once, then the gathers,
    // then the two sums in face order.
    int c[2 * U];
    T v[2 * U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      c[j] = j < wl ? col(bl + j * kSlice) : -1;
      v[j] = j < wl ? ld<kNT>(val + bl + j * kSlice) : T(0);
      c[U + j] = j < wu ? col(bu + j * kSlice) : -1;
      v[U + j] = j < wu ? ld<kNT>(val + bu + j * kSlice) : T(0);
    }
    T xv[2 * U];
#pragma unroll
    for (int j = 0; j < 2 * U; ++j) xv[j] = c[j] >= 0 ? gx(c[j]) : T(0);
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (c[j] >= 0) accL += v[j] * xv[j];
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (c[U + j] >= 0) accU += v[U + j] * xv[U + j];
    return kUpperFirst ? (dx + accU) + accL : (dx + accL) + accU;
  }
  int base = bl;
  for (int j0 = 0; j0 < wl; j0 += U) {
    int c[U];
    T v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      c[j] = (j0 + j < wl) ? col(base + j * kSlice) : -1;
      v[j] = (j0 + j < wl) ? ld<kNT>(val + base + j * kSlice) : T(0);
    }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (c[j] >= 0) accL += v[j] * gx(c[j]);
    base += U * kSlice;
  }
  base = bu;
  for (int j0 = 0; j0 < wu; j0 += U) {
    int c[U];
    T v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      c[j] = (j0 + j < wu) ? col(base + j * kSlice) : -1;
      v[j] = (j0 + j < wu) ? ld<kNT>(val + base + j * kSlice) : T(0);
    }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (c[j] >= 0) accU += v[j] * gx(c[j]);
    base += U * kSlice;
  }
  return kUpperFirst ? (dx + accU) + accL : (dx + accL) + accU;
}
template <typename T>
__device__ __forceinline__ T row_spmv(const SliceInfo si, int lane, T dx,
                                      const int32_t *__restrict__ col,
                                      const T *__restrict__ val,
                                      const T *__restrict__ x) {
  return row_spmv_g<T>(si, lane, dx, ColAbs{col}, val, [x](int c) { return x[c]; });
}

// Column forms of a handle: int32 columns, int16 offsets from the row, the
// int16 offsets paired two slots per 4-byte word (row_spmv_pair).
enum ColForm { kCols32 = 0, kCols16 = 1, kColsPair = 2 };

// Column form of the Krylov solvers' SpMV kernels: the paired int16 offsets
// where the handle has them, else int32 columns.  PFM_PCG_COLS=int16 / int32
// selects the unpaired offsets or the int32 columns (for comparing layouts).
inline int krylov_col_form(const pfm_ldu_s *h) {
  static const int forced = [] {
    const char *e = std::getenv("PFM_PCG_COLS");
    if (!e) return -1;
    if (std::strcmp(e, "int32") == 0) return (int)kCols32;
    if (std::strcmp(e, "int16") == 0) return (int)kCols16;
    return -1;
  }();
  if (forced == kCols32) return kCols32;
  if (forced == kCols16 || !h->d_colp2) return h->d_col16 ? kCols16 : kCols32;
  return kColsPair;
}

// Paired 16-bit column offsets.  Word j of lane l in slice s, at
// pair_base(s) + 64 j + l, holds the offsets from the row of the slice's slots
// 2j (low half) and 2j+1 (high half), slots counted over the lower group then
// the upper group.  The word before the array names the encoding: signed
// (col = row + int16, kColPad16 = padding) or oriented (col = row + u in the
// lower group, row - u in the upper one, u = 0 = padding).  A 4-byte load
// per lane then fetches two slots' columns: half the index bytes of int32
// columns with as many bytes per load instruction (a 2-byte load per slot
// moves half as much per instruction and was slower than int32 columns).
// Slice s owns 32 (wl + wu + 1) words from pair_base(s) = ptr/2 + 32 s, enough
// for 64 ceil((wl + wu) / 2); the words past them are never read.
__device__ __forceinline__ int pair_base(const SliceInfo si, int s) { return si.ptr / 2 + 32 * s; }

__device__ __forceinline__ bool pair_oriented(const uint32_t *__restrict__ cp) {
  return cp[-1] == kPairOriented;
}

__device__ __forceinline__ int pair_col(uint32_t w, int hi, int row, bool lower, bool ori) {
  const uint32_t u = hi ? (w >> 16) : (w & 0xffffu);
  if (ori) return u == 0 ? -1 : (lower ? row + (int)u : row - (int)u);
  const int v = (int)(int16_t)u;
  return v == kColPad16 ? -1 : row + v;
}

// The common case of row_spmv_pair for a slice with WL lower slots (known at
// compile time, so every slot's word and half are too) and at most 4 upper
// slots: every load of the row in flight at once, then the gathers, then the
// two sums in face order.
template <int WL, typename T, typename Gather, bool kUpperFirst, bool kNT>
__device__ __forceinline__ T row_pair_fixed(int wu, int ptr, int pb, int lane, int row, T dx,
                                            const uint32_t *__restrict__ cp, bool ori,
                                            const T *__restrict__ val, Gather gx) {
  constexpr int U = 4;
  constexpr int NS = WL + U;
  constexpr int NW = (NS + 1) / 2;
  const int nsl = WL + wu;
  uint32_t w[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j)
    w[j] = 2 * j < nsl ? ld<kNT>(cp + pb + j * kSlice + lane) : 0x80008000u;
  T v[NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) v[j] = j < nsl ? ld<kNT>(val + ptr + j * kSlice + lane) : T(0);
  int c[NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) c[j] = j < nsl ? pair_col(w[j >> 1], j & 1, row, j < WL, ori) : -1;
  T xv[NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) xv[j] = c[j] >= 0 ? gx(c[j]) : T(0);
  T accL = T(0), accU = T(0);
#pragma unroll
  for (int j = 0; j < WL; ++j)
    if (c[j] >= 0) accL += v[j] * xv[j];
#pragma unroll
  for (int j = WL; j < NS; ++j)
    if (c[j] >= 0) accU += v[j] * xv[j];
  return kUpperFirst ? (dx + accU) + accL : (dx + accL) + accU;
}

// row_spmv_g over the paired int16 columns (same association and order).
template <typename T, typename Gather, bool kUpperFirst = false, bool kNT = false>
__device__ __forceinline__ T row_spmv_pair(const SliceInfo si, int s, int lane, int row, T dx,
                                           const uint32_t *__restrict__ cp,
                                           const T *__restrict__ val, Gather gx) {
  const int wl = si.wl, wu = si.wu;
  const int pb = pair_base(si, s);
  const bool ori = pair_oriented(cp);  // one scalar load per wave
  if (wu <= 4) {
    switch (wl) {  // wave-uniform
      case 0: return row_pair_fixed<0, T, Gather, kUpperFirst, kNT>(wu, si.ptr, pb, lane, row, dx, cp, ori, val, gx);
      case 1: return row_pair_fixed<1, T, Gather, kUpperFirst, kNT>(wu, si.ptr, pb, lane, row, dx, cp, ori, val, gx);
      case 2: return row_pair_fixed<2, T, Gather, kUpperFirst, kNT>(wu, si.ptr, pb, lane, row, dx, cp, ori, val, gx);
      case 3: return row_pair_fixed<3, T, Gather, kUpperFirst, kNT>(wu, si.ptr, pb, lane, row, dx, cp, ori, val, gx);
      case 4: return row_pair_fixed<4, T, Gather, kUpperFirst, kNT>(wu, si.ptr, pb, lane, row, dx, cp, ori, val, gx);
      default: break;
    }
  }
  // wide rows: one slot at a time, each sum in face order
  T accL = T(0), accU = T(0);
  for (int j = 0; j < wl + wu; ++j) {
    const uint32_t wd = ld<kNT>(cp + pb + (j >> 1) * kSlice + lane);
    const int c = pair_col(wd, j & 1, row, j < wl, ori);
    if (c < 0) continue;
    const T p = ld<kNT>(val + si.ptr + j * kSlice + lane) * gx(c);
    if (j < wl)
      accL += p;
    else
      accU += p;
  }
  return kUpperFirst ? (dx + accU) + accL : (dx + accL) + accU;
}

// Slice descriptor for a slice index that is the same on every lane of the
// wave (the row-slice kernels: one wave per slice); the value is moved to
// scalar registers.
__device__ __forceinline__ SliceInfo load_slice(const SliceInfo *sl, int s) {
  const int2 raw = reinterpret_cast<const int2 *>(sl)[s];
  SliceInfo si;
  si.ptr = uniform(raw.x);
  const int w = uniform(raw.y);
  si.wl = (int16_t)(w & 0xffff);
  si.wu = (int16_t)((unsigned)w >> 16);
  return si;
}

// Slice descriptor when lanes of a wave may address different slices.
__device__ __forceinline__ SliceInfo load_slice_lane(const SliceInfo *sl, int s) {
  const int2 raw = reinterpret_cast<const int2 *>(sl)[s];
  SliceInfo si;
  si.ptr = raw.x;
  si.wl = (int16_t)(raw.y & 0xffff);
  si.wu = (int16_t)((unsigned)raw.y >> 16);
  return si;
}

// Gauss-Seidel cell updates (solvers/smoothers.py:113-149).  A row's
// couplings sit in two groups of its slice (owner-faces, column = neighbour,
// then neighbour-faces, column = owner), each in face order; merged by face
// index they follow the reference's face order.  kAssign: x[c] = new value
// (SmoothSolver's sweep, smooth_solver.py:170-180) instead of
// x[c] + omega (new - x[c]) (GaussSeidelSmoother).
//
// x[c] from the row sum (diag x_c + couplings in face order).
template <typename T, bool kAssign>
__device__ __forceinline__ void gs_store(int c, T d, T xc, T row, const T *__restrict__ b, T *x,
                                         T omega, bool live) {
  const T off = row - d * xc;
  if (live && fabs(d) > T(1e-30)) {
    const T nv = (b[c] - off) / d;
    if (kAssign)
      x[c] = nv;
    else
      x[c] = xc + omega * (nv - xc);
  }
}

// The update of cell c whose lower slots start at bl (slot ptr + lane) in a
// slice of group widths wl, wu, for rows of more than 8 couplings: the two
// groups, each in face order, merged by face index slot by slot (one chain of
// dependent loads per coupling); `live` false computes without storing (a
// sweep enqueued after its solve stopped).
template <typename T, bool kAssign = false>
__device__ __forceinline__ void gs_core(int c, int bl, int wl, int wu,
                                        const int32_t *__restrict__ col,
                                        const int32_t *__restrict__ map,
                                        const T *__restrict__ val,
                                        const T *__restrict__ diag,
                                        const T *__restrict__ b, T *x, T omega, bool live) {
  const int bu = bl + wl * kSlice;
  const T d = diag[c];
  const T xc = x[c];
  T row = d * xc;
  int jl = 0, ju = 0;
  int cl = jl < wl ? col[bl] : -1;
  int cu = ju < wu ? col[bu] : -1;
  while (cl >= 0 || cu >= 0) {
    bool take_l = cl >= 0;
    if (cl >= 0 && cu >= 0) take_l = map[bl + jl * kSlice] < map[bu + ju * kSlice];
    if (take_l) {
      row += val[bl + jl * kSlice] * x[cl];
      ++jl;
      cl = jl < wl ? col[bl + jl * kSlice] : -1;
    } else {
      row += val[bu + ju * kSlice] * x[cu];
      ++ju;
      cu = ju < wu ? col[bu + ju * kSlice] : -1;
    }
  }
  gs_store<T, kAssign>(c, d, xc, row, b, x, omega, live);
}

// The update of a schedule entry's cell (GsEnt): the columns already in face
// order, so the coefficients, the gathered x values, diag, b and x_c are all
// loaded at once after the entry; rows of more than 8 couplings go through
// gs_core's merge by the slot map.
template <typename T, bool kAssign = false>
__device__ __forceinline__ void gs_entry(const GsEnt &e, const int32_t *__restrict__ col,
                                         const int32_t *__restrict__ map,
                                         const T *__restrict__ val,
                                         const T *__restrict__ diag,
                                         const T *__restrict__ b, T *x, T omega, bool live) {
  const int c = e.a.x, bl = e.a.y, wl = e.a.z & 0xffff, wu = (int)((unsigned)e.a.z >> 16);
  const int mask = e.a.w & 0xff, n = (e.a.w >> 8) & 0xff;
  if (n > 8) {
    gs_core<T, kAssign>(c, bl, wl, wu, col, map, val, diag, b, x, omega, live);
    return;
  }
  const int cols[8] = {e.b.x, e.b.y, e.b.z, e.b.w, e.c.x, e.c.y, e.c.z, e.c.w};
  const int bu = bl + wl * kSlice;
  const T d = diag[c];
  const T xc = x[c];
  T v[8], xv[8];
  int il = 0, iu = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool lo = (mask >> k) & 1;
    const int slot = lo ? bl + il * kSlice : bu + iu * kSlice;
    il += lo ? 1 : 0;
    iu += lo ? 0 : 1;
    v[k] = k < n ? val[slot] : T(0);
    xv[k] = k < n ? x[cols[k]] : T(0);
  }
  T row = d * xc;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < n) row += v[k] * xv[k];
  gs_store<T, kAssign>(c, d, xc, row, b, x, omega, live);
}

}  // namespace pfm
