// Finite-volume mesh geometry on the device (mesh/mesh_geometry.py:35-395).
//
// Polygonal faces in CSR form (face_ptr / face_verts, vertex loops whose
// right-hand normal points from owner to neighbour); cells as CSR lists of
// face entries from pfm_cell_faces (owned faces ascending, then neighboured
// internal faces ascending, a neighbour-side entry stored as -(f + 1)).
// Every quantity follows the reference's operation order, one rounding per
// operation (the file is built with -ffp-contract=off):
//   face centre     sum of the vertices from 0, then / n_v          (:35-63)
//   area vector     0.5 * sum_j v_j x v_{j+1}; zero for n_v < 3     (:66-108)
//   |S_f|           sqrt(fma(z,z,fma(y,y,x*x))), the order of the torch
//                   norm the reference applies to S_f (delta :384,
//                   fvm.laplacian operators.py:490)
//   cell reference  sum of the cell's distinct vertices in ascending
//                   vertex order, / their count                     (:199-211)
//   volume, centre  per face (in list order) and per edge j, the tet
//                   (ref, fc, v_j, v_{j+1}) with the loop reversed on
//                   the neighbour side: t = (fc-ref).((v_j-ref)x(v_{j+1}-ref))/6,
//                   skipped when |t| < 1e-30, t = |t|;
//                   V += t; W += t * (((ref + fc) + v_j) + v_{j+1}) / 4;
//                   C = W / V, or the mean of the face centres when
//                   |V| <= 1e-30 (V = 0)                            (:213-254)
//   delta           1 / |(C_N - C_P) . S/|S||, internal faces only,
//                   both divisors guarded at 1e-30                  (:330-395)
// One thread per face or per cell: the work is a few hundred flops per
// element over gathered points, bound by the point gathers.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pfm.h"
#include "pfm_internal.h"

namespace {

constexpr int kGeoBlock = 256;
constexpr int kMaxCellVerts = 64;  // vertex references per cell (hex: 24)

struct V3 {
  double x, y, z;
};

__device__ inline V3 ld3(const double *__restrict__ a, int64_t i) {
  return {a[3 * i], a[3 * i + 1], a[3 * i + 2]};
}
__device__ inline void st3(double *__restrict__ a, int64_t i, V3 v) {
  a[3 * i] = v.x;
  a[3 * i + 1] = v.y;
  a[3 * i + 2] = v.z;
}
__device__ inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline V3 divs(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// |v| as torch's CPU norm over a row of 3 evaluates it: sqrt(fma(z,z,fma(y,y,x*x))).
__device__ inline double norm_fma(V3 v) { return sqrt(fma(v.z, v.z, fma(v.y, v.y, v.x * v.x))); }

__global__ __launch_bounds__(kGeoBlock) void k_face_geom(int32_t nf,
                                                          const int32_t *__restrict__ fptr,
                                                          const int32_t *__restrict__ fv,
                                                          const double *__restrict__ pts,
                                                          double *__restrict__ fc,
                                                          double *__restrict__ sf,
                                                          double *__restrict__ mag) {
  const int32_t f = blockIdx.x * kGeoBlock + threadIdx.x;
  if (f >= nf) return;
  const int32_t b = fptr[f], nv = fptr[f + 1] - b;
  V3 c = {0.0, 0.0, 0.0};
  V3 a = {0.0, 0.0, 0.0};
  if (nv > 0) {
    V3 v0 = ld3(pts, fv[b]);
    V3 vj = v0;
    for (int32_t j = 0; j < nv; ++j) {
      c = add(c, vj);
      const V3 vn = j + 1 < nv ? ld3(pts, fv[b + j + 1]) : v0;
      if (nv >= 3) a = add(a, cross(vj, vn));
      vj = vn;
    }
    c = divs(c, (double)nv);
    a = scale(a, 0.5);
  }
  st3(fc, f, c);
  st3(sf, f, a);
  mag[f] = norm_fma(a);
}

__global__ __launch_bounds__(kGeoBlock) void k_cell_geom(int32_t nc,
                                                          const int32_t *__restrict__ cptr,
                                                          const int32_t *__restrict__ cf,
                                                          const int32_t *__restrict__ fptr,
                                                          const int32_t *__restrict__ fv,
                                                          const double *__restrict__ pts,
                                                          const double *__restrict__ fc,
                                                          double *__restrict__ vol,
                                                          double *__restrict__ cc) {
  const int32_t c = blockIdx.x * kGeoBlock + threadIdx.x;
  if (c >= nc) return;
  const int32_t e0 = cptr[c], e1 = cptr[c + 1];
  if (e1 == e0) {  // no faces: volume 0, centre 0 (the reference skips the cell)
    vol[c] = 0.0;
    st3(cc, c, {0.0, 0.0, 0.0});
    return;
  }
  // Distinct vertices in ascending order (insertion sort; the host checked
  // that no cell references more than kMaxCellVerts vertices).
  int32_t ids[kMaxCellVerts];
  int n = 0;
  for (int32_t e = e0; e < e1; ++e) {
    const int32_t ent = cf[e];
    const int32_t f = ent >= 0 ? ent : -ent - 1;
    for (int32_t k = fptr[f]; k < fptr[f + 1] && n < kMaxCellVerts; ++k) {
      const int32_t v = fv[k];
      int i = n++;
      while (i > 0 && ids[i - 1] > v) {
        ids[i] = ids[i - 1];
        --i;
      }
      ids[i] = v;
    }
  }
  V3 ref = {0.0, 0.0, 0.0};
  int cnt = 0;
  for (int i = 0; i < n; ++i)
    if (i == 0 || ids[i] != ids[i - 1]) {
      ref = add(ref, ld3(pts, ids[i]));
      ++cnt;
    }
  ref = divs(ref, (double)cnt);

  double tv = 0.0;
  V3 wc = {0.0, 0.0, 0.0};
  for (int32_t e = e0; e < e1; ++e) {
    const int32_t ent = cf[e];
    const bool flip = ent < 0;
    const int32_t f = flip ? -ent - 1 : ent;
    const int32_t b = fptr[f], nv = fptr[f + 1] - b;
    if (nv < 3) continue;
    const V3 fcf = ld3(fc, f);
    const V3 d = sub(fcf, ref);
    const V3 rf = add(ref, fcf);
    // loop position j -> vertex slot: j (owner side) or nv-1-j (reversed)
    V3 va = ld3(pts, fv[b + (flip ? nv - 1 : 0)]);
    const V3 vfirst = va;
    for (int32_t j = 0; j < nv; ++j) {
      const int32_t jn = j + 1;
      const V3 vb = jn < nv ? ld3(pts, fv[b + (flip ? nv - 1 - jn : jn)]) : vfirst;
      double t = dot(d, cross(sub(va, ref), sub(vb, ref))) / 6.0;
      if (!(fabs(t) < 1e-30)) {
        t = fabs(t);
        const V3 cen = divs(add(add(rf, va), vb), 4.0);
        tv = tv + t;
        wc = add(wc, scale(cen, t));
      }
      va = vb;
    }
  }
  if (fabs(tv) > 1e-30) {
    vol[c] = fabs(tv);
    st3(cc, c, divs(wc, tv));
  } else {
    V3 s = {0.0, 0.0, 0.0};
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t ent = cf[e];
      s = add(s, ld3(fc, ent >= 0 ? ent : -ent - 1));
    }
    vol[c] = 0.0;
    st3(cc, c, divs(s, (double)(e1 - e0)));
  }
}

__global__ __launch_bounds__(kGeoBlock) void k_face_delta(int32_t nf, int32_t nfi,
                                                           const int32_t *__restrict__ own,
                                                           const int32_t *__restrict__ nei,
                                                           const double *__restrict__ sf,
                                                           const double *__restrict__ mag,
                                                           const double *__restrict__ cc,
                                                           double *__restrict__ delta) {
  const int32_t f = blockIdx.x * kGeoBlock + threadIdx.x;
  if (f >= nf) return;
  if (f >= nfi) {
    delta[f] = 0.0;
    return;
  }
  const V3 d = sub(ld3(cc, nei[f]), ld3(cc, own[f]));
  const double m = mag[f];
  const V3 nh = divs(ld3(sf, f), m > 1e-30 ? m : 1.0);
  const double dn = fabs(dot(d, nh));
  delta[f] = 1.0 / (dn > 1e-30 ? dn : 1.0);
}

// w = |x_f - C_N| / (|x_f - C_P| + |x_f - C_N|) on internal faces (norms in
// torch's order), 1 on boundary faces (:262-327).
__global__ __launch_bounds__(kGeoBlock) void k_face_weights(int32_t nf, int32_t nfi,
                                                             const int32_t *__restrict__ own,
                                                             const int32_t *__restrict__ nei,
                                                             const double *__restrict__ fc,
                                                             const double *__restrict__ cc,
                                                             double *__restrict__ w) {
  const int32_t f = blockIdx.x * kGeoBlock + threadIdx.x;
  if (f >= nf) return;
  if (f >= nfi) {
    w[f] = 1.0;
    return;
  }
  const V3 x = ld3(fc, f);
  const double dp = norm_fma(sub(x, ld3(cc, own[f])));
  const double dn = norm_fma(sub(x, ld3(cc, nei[f])));
  const double den = dp + dn;
  w[f] = dn / (den > 1e-30 ? den : 1.0);
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kGeoBlock - 1) / kGeoBlock); }

}  // namespace

extern "C" {

int pfm_mesh_geometry(int32_t n_points, const double *points, int32_t n_faces,
                      const int32_t *face_ptr, const int32_t *face_verts, int32_t n_internal,
                      const int32_t *owner, const int32_t *neighbour, int32_t n_cells,
                      const int32_t *cell_ptr, const int32_t *cell_faces,
                      int32_t max_cell_verts, double *face_centres, double *face_areas,
                      double *mag_sf, double *cell_volumes, double *cell_centres,
                      double *delta, void *stream) {
  PFM_REQUIRE(n_points >= 0 && n_faces >= 0 && n_cells >= 0 && n_internal >= 0 &&
                  n_internal <= n_faces,
              "bad sizes (points %d, faces %d, internal %d, cells %d)", n_points, n_faces,
              n_internal, n_cells);
  PFM_REQUIRE(max_cell_verts >= 0 && max_cell_verts <= kMaxCellVerts,
              "a cell references %d vertices; at most %d are supported", max_cell_verts,
              kMaxCellVerts);
  PFM_NONNULL(face_ptr);
  PFM_NONNULL(cell_ptr);
  if (n_faces > 0) {
    PFM_NONNULL(points);
    PFM_NONNULL(face_verts);
    PFM_NONNULL(owner);
    PFM_NONNULL(face_centres);
    PFM_NONNULL(face_areas);
    PFM_NONNULL(mag_sf);
    PFM_NONNULL(delta);
  }
  if (n_internal > 0) PFM_NONNULL(neighbour);
  if (n_cells > 0) {
    PFM_NONNULL(cell_faces);
    PFM_NONNULL(cell_volumes);
    PFM_NONNULL(cell_centres);
  }
  hipStream_t s = (hipStream_t)stream;
  if (n_faces > 0)
    k_face_geom<<<blocks(n_faces), kGeoBlock, 0, s>>>(n_faces, face_ptr, face_verts, points,
                                                       face_centres, face_areas, mag_sf);
  if (n_cells > 0)
    k_cell_geom<<<blocks(n_cells), kGeoBlock, 0, s>>>(n_cells, cell_ptr, cell_faces, face_ptr,
                                                       face_verts, points, face_centres,
                                                       cell_volumes, cell_centres);
  if (n_faces > 0)
    k_face_delta<<<blocks(n_faces), kGeoBlock, 0, s>>>(n_faces, n_internal, owner, neighbour,
                                                        face_areas, mag_sf, cell_centres, delta);
  PFM_CHECK_HIP(hipGetLastError());
  return PFM_OK;
}

int pfm_face_weights(int32_t n_faces, int32_t n_internal, const int32_t *owner,
                     const int32_t *neighbour, const double *face_centres,
                     const double *cell_centres, double *weights, void *stream) {
  PFM_REQUIRE(n_faces >= 0 && n_internal >= 0 && n_internal <= n_faces,
              "bad sizes (faces %d, internal %d)", n_faces, n_internal);
  if (n_faces == 0) return PFM_OK;
  PFM_NONNULL(weights);
  if (n_internal > 0) {
    PFM_NONNULL(owner);
    PFM_NONNULL(neighbour);
    PFM_NONNULL(face_centres);
    PFM_NONNULL(cell_centres);
  }
  k_face_weights<<<blocks(n_faces), kGeoBlock, 0, (hipStream_t)stream>>>(
      n_faces, n_internal, owner, neighbour, face_centres, cell_centres, weights);
  PFM_CHECK_HIP(hipGetLastError());
  return PFM_OK;
}

}  // extern "C"
