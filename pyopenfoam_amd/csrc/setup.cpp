#include <chrono>
// Host-side setup of libpfm: error state, RCM ordering, dependency levels of
// the DIC/DILU recurrences and the row-slice layout used by every SpMV.
// Sequential graph algorithms stay on the host (they are setup, done once per
// topology); their results are uploaded by pfm_ldu_create.

#include <algorithm>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <numeric>
#include <vector>

#include "pfm_internal.h"

namespace pfm {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int validate_topology(int32_t n, int32_t nf, const int32_t *own,
                      const int32_t *nei) {
  PFM_REQUIRE(n >= 0 && nf >= 0, "negative size (n_cells=%d, n_faces=%d)", n, nf);
  PFM_REQUIRE(nf == 0 || (own && nei), "null owner/neighbour array");
  for (int32_t f = 0; f < nf; ++f) {
    PFM_REQUIRE(own[f] >= 0 && own[f] < n && nei[f] >= 0 && nei[f] < n,
                "face %d addresses a cell outside [0, %d)", f, n);
    PFM_REQUIRE(own[f] != nei[f], "face %d has owner == neighbour", f);
  }
  return PFM_OK;
}

// Row-slice layout ----------------------------------------------------------
// Row i's slots: first its owner-faces (lower coefficients, column =
// neighbour), then its neighbour-faces (upper coefficients, column = owner),
// each group in face order.  64 consecutive rows form one slice whose slots
// are stored slot-major, so lane l of a wavefront reads entry j of row
// 64*s + l at ptr + j*64 + l: every load of the SpMV is a contiguous 64-wide
// access.
int build_host_layout(int32_t n, int32_t nf, const int32_t *own,
                      const int32_t *nei, HostLayout &L) {
  int st = validate_topology(n, nf, own, nei);
  if (st != PFM_OK) return st;
  L = HostLayout();
  L.n = n;
  L.nf = nf;
  L.nslices = (n + kSlice - 1) / kSlice;

  std::vector<int32_t> lo_ptr(n + 1, 0), up_ptr(n + 1, 0);
  for (int32_t f = 0; f < nf; ++f) {
    lo_ptr[own[f] + 1]++;
    up_ptr[nei[f] + 1]++;
  }
  for (int32_t i = 0; i < n; ++i) {
    L.max_lower = std::max(L.max_lower, lo_ptr[i + 1]);
    L.max_upper = std::max(L.max_upper, up_ptr[i + 1]);
    lo_ptr[i + 1] += lo_ptr[i];
    up_ptr[i + 1] += up_ptr[i];
  }
  PFM_REQUIRE(L.max_lower < 32768 && L.max_upper < 32768, "cell degree too large");
  std::vector<int32_t> lo_face(nf), up_face(nf);
  {
    std::vector<int32_t> a(lo_ptr.begin(), lo_ptr.end() - 1), b(up_ptr.begin(), up_ptr.end() - 1);
    for (int32_t f = 0; f < nf; ++f) {  // face order preserved inside each row
      lo_face[a[own[f]]++] = f;
      up_face[b[nei[f]]++] = f;
    }
  }

  L.slices.resize(L.nslices);
  int64_t ptr = 0;
  for (int32_t s = 0; s < L.nslices; ++s) {
    int wl = 0, wu = 0;
    for (int32_t r = s * kSlice; r < std::min(n, (s + 1) * kSlice); ++r) {
      wl = std::max(wl, lo_ptr[r + 1] - lo_ptr[r]);
      wu = std::max(wu, up_ptr[r + 1] - up_ptr[r]);
    }
    PFM_REQUIRE(ptr <= INT32_MAX, "matrix too large for int32 slot offsets");
    L.slices[s] = SliceInfo{(int32_t)ptr, (int16_t)wl, (int16_t)wu};
    ptr += (int64_t)(wl + wu) * kSlice;
  }
  PFM_REQUIRE(ptr <= INT32_MAX, "matrix too large for int32 slot offsets");
  L.nslots = ptr;
  L.col.assign(ptr, -1);
  L.map.assign(ptr, -1);
  for (int32_t s = 0; s < L.nslices; ++s) {
    const SliceInfo si = L.slices[s];
    for (int32_t r = s * kSlice; r < std::min(n, (s + 1) * kSlice); ++r) {
      const int lane = r - s * kSlice;
      for (int32_t k = lo_ptr[r]; k < lo_ptr[r + 1]; ++k) {
        const int64_t slot = si.ptr + (int64_t)(k - lo_ptr[r]) * kSlice + lane;
        L.col[slot] = nei[lo_face[k]];
        L.map[slot] = lo_face[k];
      }
      for (int32_t k = up_ptr[r]; k < up_ptr[r + 1]; ++k) {
        const int64_t slot = si.ptr + (int64_t)(si.wl + k - up_ptr[r]) * kSlice + lane;
        L.col[slot] = own[up_face[k]];
        L.map[slot] = up_face[k];
      }
    }
  }

  // DIC cell levels, grouped.
  std::vector<int32_t> lev(n);
  int32_t nl = 0;
  if (pfm_dic_levels(n, nf, own, nei, lev.data(), &nl) != PFM_OK) return PFM_ERR_ARG;
  L.dic_level_ptr.assign(nl + 1, 0);
  for (int32_t i = 0; i < n; ++i) L.dic_level_ptr[lev[i] + 1]++;
  for (int32_t l = 0; l < nl; ++l) L.dic_level_ptr[l + 1] += L.dic_level_ptr[l];
  L.dic_cells.resize(n);
  {
    std::vector<int32_t> c(L.dic_level_ptr.begin(), L.dic_level_ptr.end() - 1);
    for (int32_t i = 0; i < n; ++i) L.dic_cells[c[lev[i]]++] = i;
  }
  // DILU face levels, grouped.
  std::vector<int32_t> flev(nf);
  int32_t nfl = 0;
  if (pfm_dilu_levels(n, nf, own, nei, flev.data(), &nfl) != PFM_OK) return PFM_ERR_ARG;
  L.dilu_level_ptr.assign(nfl + 1, 0);
  for (int32_t f = 0; f < nf; ++f) L.dilu_level_ptr[flev[f] + 1]++;
  for (int32_t l = 0; l < nfl; ++l) L.dilu_level_ptr[l + 1] += L.dilu_level_ptr[l];
  L.dilu_faces.resize(nf);
  {
    std::vector<int32_t> c(L.dilu_level_ptr.begin(), L.dilu_level_ptr.end() - 1);
    for (int32_t f = 0; f < nf; ++f) L.dilu_faces[c[flev[f]]++] = f;
  }
  // Gauss-Seidel forward / backward sweep levels, grouped.
  std::vector<int32_t> gf(n), gb(n);
  int32_t ngf = 0, ngb = 0;
  if (pfm_gs_levels(n, nf, own, nei, gf.data(), gb.data(), &ngf, &ngb) != PFM_OK)
    return PFM_ERR_ARG;
  auto group = [n](const std::vector<int32_t> &lev, int32_t nl, std::vector<int32_t> &ptr,
                   std::vector<int32_t> &cells) {
    ptr.assign(nl + 1, 0);
    for (int32_t i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int32_t l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    cells.resize(n);
    std::vector<int32_t> c(ptr.begin(), ptr.end() - 1);
    for (int32_t i = 0; i < n; ++i) cells[c[lev[i]]++] = i;
  };
  group(gf, ngf, L.gs_fwd_ptr, L.gs_fwd_cells);
  group(gb, ngb, L.gs_bwd_ptr, L.gs_bwd_cells);
  return PFM_OK;
}

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void report_progress(const pfm_ldu_s *h, int32_t iterations, double &last) {
  if (!h->progress) return;
  const double t = now_seconds();
  if (t - last < h->progress_interval) return;
  last = t;
  h->progress(iterations, h->progress_user);
}

}  // namespace pfm

using namespace pfm;

extern "C" {

const char *pfm_last_error(void) { return g_err; }

int pfm_abi_version(void) { return 2; }

int pfm_rcm(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei,
            int32_t *perm_out) {
  PFM_REQUIRE(n >= 0 && nf >= 0, "negative size");
  PFM_REQUIRE(n == 0 || perm_out, "null output");
  for (int32_t f = 0; f < nf; ++f)
    PFM_REQUIRE(own[f] >= 0 && own[f] < n && nei[f] >= 0 && nei[f] < n,
                "face %d addresses a cell outside [0, %d)", f, n);
  if (n == 0) return PFM_OK;
  // Symmetric adjacency, coalesced: each row's columns ascending, unique.
  std::vector<int32_t> ptr(n + 1, 0);
  for (int32_t f = 0; f < nf; ++f) {
    ptr[own[f] + 1]++;
    ptr[nei[f] + 1]++;
  }
  for (int32_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  std::vector<int32_t> adj(ptr[n]);
  {
    std::vector<int32_t> c(ptr.begin(), ptr.end() - 1);
    for (int32_t f = 0; f < nf; ++f) {
      adj[c[own[f]]++] = nei[f];
      adj[c[nei[f]]++] = own[f];
    }
  }
  std::vector<int32_t> deg(n, 0), len(n, 0);
  for (int32_t i = 0; i < n; ++i) {
    auto b = adj.begin() + ptr[i], e = adj.begin() + ptr[i + 1];
    std::sort(b, e);
    len[i] = (int32_t)(std::unique(b, e) - b);
    deg[i] = len[i];  // degree counts coalesced entries, self-loops included
  }
  // Neighbour lists without self-loops, stably sorted by degree.
  for (int32_t i = 0; i < n; ++i) {
    auto b = adj.begin() + ptr[i];
    auto e = std::remove(b, b + len[i], i);
    len[i] = (int32_t)(e - b);
    std::stable_sort(b, e, [&](int32_t a, int32_t c) { return deg[a] < deg[c]; });
  }
  // Candidate start cells in (degree, index) order.
  std::vector<int32_t> by_deg(n);
  std::iota(by_deg.begin(), by_deg.end(), 0);
  std::stable_sort(by_deg.begin(), by_deg.end(),
                   [&](int32_t a, int32_t c) { return deg[a] < deg[c]; });
  std::vector<char> seen(n, 0);
  std::vector<int32_t> order;
  order.reserve(n);
  size_t cand = 0;
  while ((int32_t)order.size() < n) {
    while (seen[by_deg[cand]]) ++cand;
    size_t head = order.size();
    int32_t s = by_deg[cand];
    seen[s] = 1;
    order.push_back(s);
    while (head < order.size()) {
      int32_t v = order[head++];
      for (int32_t k = ptr[v]; k < ptr[v] + len[v]; ++k) {
        int32_t w = adj[k];
        if (!seen[w]) {
          seen[w] = 1;
          order.push_back(w);
        }
      }
    }
  }
  for (int32_t i = 0; i < n; ++i) perm_out[i] = order[n - 1 - i];
  return PFM_OK;
}

int pfm_dic_levels(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei,
                   int32_t *level, int32_t *n_levels) {
  int st = validate_topology(n, nf, own, nei);
  if (st != PFM_OK) return st;
  PFM_REQUIRE(n == 0 || level, "null output");
  // faces grouped by neighbour
  std::vector<int32_t> ptr(n + 1, 0), fl(nf);
  for (int32_t f = 0; f < nf; ++f) ptr[nei[f] + 1]++;
  for (int32_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  {
    std::vector<int32_t> c(ptr.begin(), ptr.end() - 1);
    for (int32_t f = 0; f < nf; ++f) fl[c[nei[f]]++] = f;
  }
  int32_t nl = 0;
  for (int32_t i = 0; i < n; ++i) {
    int32_t l = 0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      int32_t j = own[fl[k]];
      if (j < i) l = std::max(l, level[j] + 1);
    }
    level[i] = l;
    nl = std::max(nl, l + 1);
  }
  if (n_levels) *n_levels = nl;
  return PFM_OK;
}

int pfm_gs_levels(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei,
                  int32_t *fwd_level, int32_t *bwd_level, int32_t *n_fwd, int32_t *n_bwd) {
  int st = validate_topology(n, nf, own, nei);
  if (st != PFM_OK) return st;
  PFM_REQUIRE(n == 0 || (fwd_level && bwd_level), "null output");
  // cell adjacency from both sides of every face
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
  int32_t nlf = 0, nlb = 0;
  for (int32_t i = 0; i < n; ++i) {  // forward sweep: i reads the new x of every j < i
    int32_t l = 0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
      if (adj[k] < i) l = std::max(l, fwd_level[adj[k]] + 1);
    fwd_level[i] = l;
    nlf = std::max(nlf, l + 1);
  }
  for (int32_t i = n - 1; i >= 0; --i) {  // backward sweep: i reads the new x of every j > i
    int32_t l = 0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
      if (adj[k] > i) l = std::max(l, bwd_level[adj[k]] + 1);
    bwd_level[i] = l;
    nlb = std::max(nlb, l + 1);
  }
  if (n_fwd) *n_fwd = nlf;
  if (n_bwd) *n_bwd = nlb;
  return PFM_OK;
}

int pfm_dilu_levels(int32_t n, int32_t nf, const int32_t *own, const int32_t *nei,
                    int32_t *level, int32_t *n_levels) {
  int st = validate_topology(n, nf, own, nei);
  if (st != PFM_OK) return st;
  PFM_REQUIRE(nf == 0 || level, "null output");
  std::vector<int32_t> last(n, -1);
  int32_t nl = 0;
  for (int32_t f = 0; f < nf; ++f) {
    int32_t l = std::max(last[own[f]], last[nei[f]]) + 1;
    level[f] = l;
    last[own[f]] = last[nei[f]] = l;
    nl = std::max(nl, l + 1);
  }
  if (n_levels) *n_levels = nl;
  return PFM_OK;
}

int pfm_cell_faces(int32_t n_cells, int32_t n_faces, int32_t n_internal, const int32_t *owner,
                   const int32_t *neighbour, const int32_t *face_ptr, int32_t *cell_ptr_out,
                   int32_t *cell_faces_out, int32_t *max_cell_verts_out) {
  PFM_REQUIRE(n_cells >= 0 && n_faces >= 0 && n_internal >= 0 && n_internal <= n_faces,
              "bad sizes (cells %d, faces %d, internal %d)", n_cells, n_faces, n_internal);
  PFM_REQUIRE(cell_ptr_out && max_cell_verts_out, "null output");
  PFM_REQUIRE(n_faces == 0 || (owner && face_ptr && cell_faces_out), "null array");
  PFM_REQUIRE(n_internal == 0 || neighbour, "null neighbour array");
  for (int32_t f = 0; f < n_faces; ++f) {
    PFM_REQUIRE(owner[f] >= 0 && owner[f] < n_cells, "face %d: owner %d outside [0, %d)", f,
                owner[f], n_cells);
    PFM_REQUIRE(face_ptr[f + 1] >= face_ptr[f], "face_ptr decreases at face %d", f);
  }
  for (int32_t f = 0; f < n_internal; ++f)
    PFM_REQUIRE(neighbour[f] >= 0 && neighbour[f] < n_cells && neighbour[f] != owner[f],
                "face %d: neighbour %d invalid", f, neighbour[f]);
  // Counting sort: owned faces in face order, then neighboured faces in face
  // order (mesh_geometry.py:188-194).
  std::fill(cell_ptr_out, cell_ptr_out + n_cells + 1, 0);
  for (int32_t f = 0; f < n_faces; ++f) cell_ptr_out[owner[f] + 1]++;
  for (int32_t f = 0; f < n_internal; ++f) cell_ptr_out[neighbour[f] + 1]++;
  for (int32_t c = 0; c < n_cells; ++c) cell_ptr_out[c + 1] += cell_ptr_out[c];
  std::vector<int32_t> fill(cell_ptr_out, cell_ptr_out + n_cells);
  for (int32_t f = 0; f < n_faces; ++f) cell_faces_out[fill[owner[f]]++] = f;
  for (int32_t f = 0; f < n_internal; ++f) cell_faces_out[fill[neighbour[f]]++] = -f - 1;
  int32_t mx = 0;
  for (int32_t c = 0; c < n_cells; ++c) {
    int64_t nv = 0;
    for (int32_t e = cell_ptr_out[c]; e < cell_ptr_out[c + 1]; ++e) {
      const int32_t ent = cell_faces_out[e];
      const int32_t f = ent >= 0 ? ent : -ent - 1;
      nv += face_ptr[f + 1] - face_ptr[f];
    }
    if (nv > mx) mx = nv > INT32_MAX ? INT32_MAX : (int32_t)nv;
  }
  *max_cell_verts_out = mx;
  return PFM_OK;
}

int pfm_relabel_faces(int32_t n_cells, int32_t n_internal, const int32_t *owner,
                      const int32_t *neighbour, const int32_t *new_of_old, int32_t *order_out,
                      uint8_t *flip_out, int32_t *owner_out, int32_t *neighbour_out) {
  PFM_REQUIRE(n_cells >= 0 && n_internal >= 0, "negative size");
  PFM_REQUIRE(n_cells == 0 || new_of_old, "null permutation");
  PFM_REQUIRE(n_internal == 0 || (owner && neighbour && order_out && flip_out && owner_out &&
                                  neighbour_out),
              "null array");
  {
    std::vector<uint8_t> seen(n_cells, 0);
    for (int32_t c = 0; c < n_cells; ++c) {
      const int32_t q = new_of_old[c];
      PFM_REQUIRE(q >= 0 && q < n_cells && !seen[q], "new_of_old is not a permutation (%d)", c);
      seen[q] = 1;
    }
  }
  std::vector<int32_t> lo(n_internal), hi(n_internal);
  for (int32_t f = 0; f < n_internal; ++f) {
    PFM_REQUIRE(owner[f] >= 0 && owner[f] < n_cells && neighbour[f] >= 0 &&
                    neighbour[f] < n_cells,
                "face %d addresses a cell outside [0, %d)", f, n_cells);
    const int32_t a = new_of_old[owner[f]], b = new_of_old[neighbour[f]];
    flip_out[f] = a > b;
    lo[f] = a > b ? b : a;
    hi[f] = a > b ? a : b;
  }
  // Two stable counting sorts (by hi, then by lo) give the (lo, hi) order
  // with ties in face order.
  std::vector<int32_t> cnt(n_cells + 1), tmp(n_internal);
  auto pass = [&](const std::vector<int32_t> &key, const int32_t *in, int32_t *out) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int32_t k = 0; k < n_internal; ++k) cnt[key[in ? in[k] : k] + 1]++;
    for (int32_t c = 0; c < n_cells; ++c) cnt[c + 1] += cnt[c];
    for (int32_t k = 0; k < n_internal; ++k) {
      const int32_t f = in ? in[k] : k;
      out[cnt[key[f]]++] = f;
    }
  };
  pass(hi, nullptr, tmp.data());
  pass(lo, tmp.data(), order_out);
  std::vector<uint8_t> fl(flip_out, flip_out + n_internal);
  for (int32_t k = 0; k < n_internal; ++k) {
    const int32_t f = order_out[k];
    owner_out[k] = lo[f];
    neighbour_out[k] = hi[f];
    flip_out[k] = fl[f];
  }
  return PFM_OK;
}

}  // extern "C"
