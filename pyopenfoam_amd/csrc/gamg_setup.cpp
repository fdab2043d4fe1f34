// GAMG hierarchy setup on the host (topology only; values are coarsened on
// the device every solve).
//
// Aggregation follows GAMGSolver._build_aggregates (solvers/gamg.py:164-239):
// cells in index order; an unaggregated cell seeds a new aggregate and takes
// every unaggregated neighbour in face-append order; once the aggregate count
// reaches the target, every remaining cell joins its first aggregated
// neighbour (or aggregate 0).  Coarse faces follow _build_coarse_matrix
// (:283-324): fine faces whose cells fall in different aggregates, keyed by
// the unordered aggregate pair in first-seen order, oriented like the first
// fine face of the pair; faces inside an aggregate are dropped.
//
// Pair agglomeration follows PairGamgAgglomeration._pair_cells
// (fv_agglomeration/pair_agglomeration.py:90-131), computed in O(F) per level
// with the same result as the reference's O(N*F) relabelling.

#include <algorithm>
#include <numeric>
#include <unordered_map>

#include "gamg_internal.h"

namespace pfm {

int gamg_aggregate(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei, int32_t target,
                   std::vector<int32_t> &f2c, int32_t &nagg) {
  std::vector<int32_t> ptr(n + 1, 0), adj(2 * (size_t)nf);
  for (int32_t f = 0; f < nf; ++f) {
    ptr[own[f] + 1]++;
    ptr[nei[f] + 1]++;
  }
  for (int32_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  {
    std::vector<int32_t> c(ptr.begin(), ptr.end() - 1);
    for (int32_t f = 0; f < nf; ++f) {
      adj[c[own[f]]++] = nei[f];
      adj[c[nei[f]]++] = own[f];
    }
  }
  f2c.assign(n, -1);
  nagg = 0;
  for (int32_t cell = 0; cell < n; ++cell) {
    if (f2c[cell] >= 0) continue;
    const int32_t id = nagg++;
    f2c[cell] = id;
    for (int32_t k = ptr[cell]; k < ptr[cell + 1]; ++k)
      if (f2c[adj[k]] < 0) f2c[adj[k]] = id;
    if (nagg >= target) {
      for (int32_t rem = 0; rem < n; ++rem) {
        if (f2c[rem] >= 0) continue;
        int32_t a = 0;
        for (int32_t k = ptr[rem]; k < ptr[rem + 1]; ++k) {
          if (f2c[adj[k]] >= 0) {
            a = f2c[adj[k]];
            break;
          }
        }
        f2c[rem] = a;
      }
      break;
    }
  }
  for (auto &v : f2c)
    if (v < 0) v = 0;
  return PFM_OK;
}

int gamg_coarse_faces(int32_t nf, const int32_t *own, const int32_t *nei,
                      const std::vector<int32_t> &f2c, std::vector<int32_t> &c_own,
                      std::vector<int32_t> &c_nei, std::vector<int32_t> &face_map) {
  face_map.assign(nf, -1);
  c_own.clear();
  c_nei.clear();
  std::unordered_map<uint64_t, int32_t> seen;
  seen.reserve((size_t)nf);
  for (int32_t f = 0; f < nf; ++f) {
    const int32_t cp = f2c[own[f]], cn = f2c[nei[f]];
    if (cp == cn) continue;
    const uint64_t key = ((uint64_t)(uint32_t)std::min(cp, cn) << 32) | (uint32_t)std::max(cp, cn);
    auto it = seen.find(key);
    if (it != seen.end()) {
      face_map[f] = it->second;
    } else {
      const int32_t idx = (int32_t)c_own.size();
      seen.emplace(key, idx);
      c_own.push_back(cp);
      c_nei.push_back(cn);
      face_map[f] = idx;
    }
  }
  return PFM_OK;
}

int pair_level(int32_t n_fine, int32_t n_cur, int32_t nf, const int32_t *own, const int32_t *nei,
               const std::vector<int32_t> &cell_map, std::vector<int32_t> &out, int32_t &n_new) {
  std::vector<char> paired(n_cur, 0);
  std::vector<int32_t> redirect(n_cur);
  std::iota(redirect.begin(), redirect.end(), 0);
  bool any = false;
  for (int32_t f = 0; f < nf; ++f) {
    const int32_t c0 = cell_map[own[f]], c1 = cell_map[nei[f]];
    if (c0 == c1 || paired[c0] || paired[c1]) continue;
    redirect[c1] = c0;  // c1 unpaired: nothing maps to it yet; c0 is never redirected later
    paired[c0] = paired[c1] = 1;
    any = true;
  }
  n_new = 0;
  out.clear();
  if (!any) return PFM_OK;
  std::vector<int32_t> dense(n_cur, -1);
  std::vector<char> used(n_cur, 0);
  for (int32_t i = 0; i < n_fine; ++i) used[redirect[cell_map[i]]] = 1;
  for (int32_t v = 0; v < n_cur; ++v)
    if (used[v]) dense[v] = n_new++;
  out.resize(n_fine);
  for (int32_t i = 0; i < n_fine; ++i) out[i] = dense[redirect[cell_map[i]]];
  return PFM_OK;
}

}  // namespace pfm

using namespace pfm;

extern "C" {

int pfm_gamg_aggregate(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei,
                       int32_t target, int32_t *f2c_out, int32_t *n_coarse) {
  int st = validate_topology(n, nf, own, nei);
  if (st != PFM_OK) return st;
  PFM_REQUIRE(n == 0 || (f2c_out && n_coarse), "null output");
  std::vector<int32_t> f2c;
  int32_t nagg = 0;
  gamg_aggregate(n, nf, own, nei, target, f2c, nagg);
  std::copy(f2c.begin(), f2c.end(), f2c_out);
  if (n_coarse) *n_coarse = nagg;
  return PFM_OK;
}

int pfm_pair_agglomerate(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei,
                         int32_t n_levels, int32_t *maps_out, int32_t *n_coarse_out,
                         int32_t *n_levels_out) {
  int st = validate_topology(n, nf, own, nei);
  if (st != PFM_OK) return st;
  PFM_REQUIRE(n_levels >= 0 && n_levels_out, "bad arguments");
  std::vector<int32_t> map(n), next;
  for (int32_t i = 0; i < n; ++i) map[i] = i;
  int32_t cur = n, made = 0;
  for (int32_t level = 0; level < n_levels; ++level) {
    if (cur <= 4) break;
    int32_t k = 0;
    pair_level(n, cur, nf, own, nei, map, next, k);
    if (k == 0 || k > 0.9 * cur) break;
    if (maps_out) std::copy(next.begin(), next.end(), maps_out + (size_t)made * n);
    if (n_coarse_out) n_coarse_out[made] = k;
    map.swap(next);
    cur = k;
    ++made;
  }
  *n_levels_out = made;
  return PFM_OK;
}

}  // extern "C"
