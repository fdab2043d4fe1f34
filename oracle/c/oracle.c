/*
 * CPU oracle, C part.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke
 * test and bench.py's cpu_baseline leg as the results compared against; never
 * by the product.
 *
 * Plain sequential restatements of the reference's per-element Python loops,
 * which are too slow to run at the benchmark sizes.  Each function cites the
 * loop it restates (paths relative to src/pyfoam/).  Built with
 * -ffp-contract=off so each product/quotient/sum is rounded like the
 * reference's scalar float64 arithmetic.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* solvers/preconditioners.py:102-133 (DICPreconditioner._build).
 * lower_adj[n] lists (owner, upper) of every face with neighbour n in face
 * order; cells are processed in ascending order with the current d. */
int orc_dic_build(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                  const double *diag, const double *upper, double *out) {
  int64_t *ptr = calloc(n + 1, sizeof(int64_t));
  int64_t *lst = malloc(sizeof(int64_t) * (nf ? nf : 1));
  if (!ptr || !lst) return 1;
  for (int64_t f = 0; f < nf; ++f) ptr[nei[f] + 1]++;
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  int64_t *c = malloc(sizeof(int64_t) * (n ? n : 1));
  memcpy(c, ptr, sizeof(int64_t) * n);
  for (int64_t f = 0; f < nf; ++f) lst[c[nei[f]]++] = f;
  free(c);
  memcpy(out, diag, sizeof(double) * n);
  if (nf > 0) {
    for (int64_t i = 0; i < n; ++i) {
      for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int64_t f = lst[k];
        const double a = upper[f];
        const double dj = out[own[f]];
        if (fabs(dj) > 1e-30) out[i] = out[i] - (a * a) / dj;
      }
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    const double a = fabs(out[i]);
    out[i] = a < 1e-30 ? 1e-30 : a;
  }
  free(ptr);
  free(lst);
  return 0;
}

/* solvers/preconditioners.py:244-272 (DILUPreconditioner._build). */
int orc_dilu_build(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                   const double *diag, const double *lower, const double *upper, double *out) {
  memcpy(out, diag, sizeof(double) * n);
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t p = own[f], q = nei[f];
    const double l = lower[f], u = upper[f];
    const double dp = out[p], dq = out[q];
    if (fabs(dq) > 1e-30) out[p] = out[p] - (l * u) / dq;
    if (fabs(dp) > 1e-30) out[q] = out[q] - (u * l) / dp;
  }
  for (int64_t i = 0; i < n; ++i) {
    const double a = fabs(out[i]);
    out[i] = a < 1e-30 ? 1e-30 : a;
  }
  return 0;
}

/* DICPreconditioner.apply_full (preconditioners.py:153-201): forward sweep
 * over faces with p < n, scale by 1/d, reverse backward sweep. */
int orc_dic_apply_full(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                       const double *rD, const double *upper, const double *r, double *z) {
  memcpy(z, r, sizeof(double) * n);
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t p = own[f], q = nei[f];
    if (p < q) {
      const double dp = rD[p];
      if (fabs(dp) > 1e-30) z[q] = z[q] - upper[f] * z[p] / dp;
    }
  }
  for (int64_t i = 0; i < n; ++i) z[i] = z[i] / rD[i];
  for (int64_t f = nf - 1; f >= 0; --f) {
    const int64_t p = own[f], q = nei[f];
    if (p < q) {
      const double dq = rD[q];
      if (fabs(dq) > 1e-30) z[p] = z[p] - upper[f] * z[q] / dq;
    }
  }
  return 0;
}

/* DILUPreconditioner.apply_full (preconditioners.py:290-337). */
int orc_dilu_apply_full(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                        const double *rD, const double *lower, const double *upper,
                        const double *r, double *z) {
  memcpy(z, r, sizeof(double) * n);
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t p = own[f], q = nei[f];
    const double dp = rD[p];
    if (fabs(dp) > 1e-30) z[q] = z[q] - upper[f] * z[p] / dp;
  }
  for (int64_t i = 0; i < n; ++i) z[i] = z[i] / rD[i];
  for (int64_t f = nf - 1; f >= 0; --f) {
    const int64_t p = own[f], q = nei[f];
    const double dq = rD[q];
    if (fabs(dq) > 1e-30) z[p] = z[p] - lower[f] * z[q] / dq;
  }
  return 0;
}

/* GaussSeidelSmoother.smooth (solvers/smoothers.py:92-149).  The row sum of
 * cell c runs over all faces in face order using the latest x:
 * row_ax = diag*x + sum(lower*x[n] if p==c, upper*x[p] if n==c); the cell's
 * faces are visited in face order, which is what the O(N*F) reference loop
 * does.  off = row_ax - diag*x; x[c] += omega*((b - off)/d - x[c]). */
int orc_gauss_seidel(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                     const double *diag, const double *lower, const double *upper,
                     const double *b, double omega, int64_t sweeps, double *x) {
  int64_t *ptr = calloc(n + 1, sizeof(int64_t));
  int64_t *lst = malloc(sizeof(int64_t) * (2 * nf + 1));
  int64_t *c = malloc(sizeof(int64_t) * (n + 1));
  if (!ptr || !lst || !c) return 1;
  for (int64_t f = 0; f < nf; ++f) {
    ptr[own[f] + 1]++;
    ptr[nei[f] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  memcpy(c, ptr, sizeof(int64_t) * n);
  for (int64_t f = 0; f < nf; ++f) {  /* face order inside each cell list */
    lst[c[own[f]]++] = f;
    lst[c[nei[f]]++] = f;
  }
  for (int64_t s = 0; s < sweeps; ++s) {
    for (int pass = 0; pass < 2; ++pass) {
      for (int64_t t = 0; t < n; ++t) {
        const int64_t cc = pass == 0 ? t : n - 1 - t;
        const double d = diag[cc];
        double row = d * x[cc];
        for (int64_t k = ptr[cc]; k < ptr[cc + 1]; ++k) {
          const int64_t f = lst[k];
          if (own[f] == cc)
            row += lower[f] * x[nei[f]];
          else
            row += upper[f] * x[own[f]];
        }
        const double off = row - d * x[cc];
        if (fabs(d) > 1e-30) {
          const double nv = (b[cc] - off) / d;
          x[cc] = x[cc] + omega * (nv - x[cc]);
        }
      }
    }
  }
  free(ptr);
  free(lst);
  free(c);
  return 0;
}

/* SmoothSolver._solve_gauss_seidel sweep (solvers/smooth_solver.py:150-180):
 * cells 0..n-1 in order, row sum over the cell's faces in face order with the
 * latest x (diag*x first), off = row - diag*x, x[c] = (b - off)/diag when
 * |diag| > 1e-30 (direct assignment, forward only). */
int orc_smooth_gs(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                  const double *diag, const double *lower, const double *upper,
                  const double *b, int64_t sweeps, double *x) {
  int64_t *ptr = calloc(n + 1, sizeof(int64_t));
  int64_t *lst = malloc(sizeof(int64_t) * (2 * nf + 1));
  int64_t *c = malloc(sizeof(int64_t) * (n + 1));
  if (!ptr || !lst || !c) return 1;
  for (int64_t f = 0; f < nf; ++f) {
    ptr[own[f] + 1]++;
    ptr[nei[f] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  memcpy(c, ptr, sizeof(int64_t) * n);
  for (int64_t f = 0; f < nf; ++f) {
    lst[c[own[f]]++] = f;
    lst[c[nei[f]]++] = f;
  }
  for (int64_t s = 0; s < sweeps; ++s) {
    for (int64_t cc = 0; cc < n; ++cc) {
      const double d = diag[cc];
      double row = d * x[cc];
      for (int64_t k = ptr[cc]; k < ptr[cc + 1]; ++k) {
        const int64_t f = lst[k];
        if (own[f] == cc)
          row += lower[f] * x[nei[f]];
        else
          row += upper[f] * x[own[f]];
      }
      const double off = row - d * x[cc];
      if (fabs(d) > 1e-30) x[cc] = (b[cc] - off) / d;
    }
  }
  free(ptr);
  free(lst);
  free(c);
  return 0;
}

/* GAMGSolver._build_aggregates (solvers/gamg.py:164-239).  adj[c] lists the
 * other cell of every face touching c, in face order. */
int64_t orc_gamg_aggregate(int64_t n, int64_t nf, const int64_t *own, const int64_t *nei,
                           int64_t target, int64_t *f2c) {
  int64_t *ptr = calloc(n + 1, sizeof(int64_t));
  int64_t *adj = malloc(sizeof(int64_t) * (2 * nf + 1));
  int64_t *c = malloc(sizeof(int64_t) * (n + 1));
  if (!ptr || !adj || !c) return -1;
  for (int64_t f = 0; f < nf; ++f) {
    ptr[own[f] + 1]++;
    ptr[nei[f] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  memcpy(c, ptr, sizeof(int64_t) * n);
  for (int64_t f = 0; f < nf; ++f) {
    adj[c[own[f]]++] = nei[f];
    adj[c[nei[f]]++] = own[f];
  }
  for (int64_t i = 0; i < n; ++i) f2c[i] = -1;
  int64_t nagg = 0;
  for (int64_t cell = 0; cell < n; ++cell) {
    if (f2c[cell] >= 0) continue;
    const int64_t id = nagg++;
    f2c[cell] = id;
    for (int64_t k = ptr[cell]; k < ptr[cell + 1]; ++k)
      if (f2c[adj[k]] < 0) f2c[adj[k]] = id;
    if (nagg >= target) {
      for (int64_t rem = 0; rem < n; ++rem) {
        if (f2c[rem] >= 0) continue;
        int assigned = 0;
        for (int64_t k = ptr[rem]; k < ptr[rem + 1]; ++k) {
          if (f2c[adj[k]] >= 0) {
            f2c[rem] = f2c[adj[k]];
            assigned = 1;
            break;
          }
        }
        if (!assigned) f2c[rem] = 0;
      }
      break;
    }
  }
  for (int64_t i = 0; i < n; ++i)
    if (f2c[i] < 0) f2c[i] = 0;
  free(ptr);
  free(adj);
  free(c);
  return nagg;
}

/* GAMGSolver._build_coarse_matrix off-diagonal part (solvers/gamg.py:283-324):
 * faces whose cells fall in different aggregates, keyed by (min, max) of the
 * aggregate pair in first-seen order; the first face fixes the coarse
 * (owner, neighbour) orientation; later faces add lower->lower and
 * upper->upper.  Returns the number of coarse faces; face_map[f] = coarse face
 * or -1.  Uses an open-addressing hash on the aggregate pair. */
int64_t orc_gamg_coarse_faces(int64_t nf, const int64_t *own, const int64_t *nei,
                              const int64_t *f2c, const double *lower, const double *upper,
                              int64_t *c_own, int64_t *c_nei, double *c_lower, double *c_upper,
                              int64_t *face_map) {
  uint64_t cap = 16;
  while (cap < 2 * (uint64_t)nf + 16) cap <<= 1;
  uint64_t *keys = malloc(sizeof(uint64_t) * cap);
  int64_t *vals = malloc(sizeof(int64_t) * cap);
  if (!keys || !vals) return -1;
  for (uint64_t i = 0; i < cap; ++i) keys[i] = UINT64_MAX;
  int64_t ncf = 0;
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t cp = f2c[own[f]], cn = f2c[nei[f]];
    if (cp == cn) {
      face_map[f] = -1;
      continue;
    }
    const uint64_t a = (uint64_t)(cp < cn ? cp : cn), b = (uint64_t)(cp < cn ? cn : cp);
    const uint64_t key = (a << 32) | b;
    uint64_t hsh = (key * 0x9E3779B97F4A7C15ull) & (cap - 1);
    while (keys[hsh] != UINT64_MAX && keys[hsh] != key) hsh = (hsh + 1) & (cap - 1);
    if (keys[hsh] == key) {
      const int64_t idx = vals[hsh];
      c_lower[idx] = c_lower[idx] + lower[f];
      c_upper[idx] = c_upper[idx] + upper[f];
      face_map[f] = idx;
    } else {
      keys[hsh] = key;
      vals[hsh] = ncf;
      c_own[ncf] = cp;
      c_nei[ncf] = cn;
      c_lower[ncf] = lower[f];
      c_upper[ncf] = upper[f];
      face_map[f] = ncf;
      ++ncf;
    }
  }
  free(keys);
  free(vals);
  return ncf;
}

/* PairGamgAgglomeration._pair_cells (fv_agglomeration/pair_agglomeration.py:90-131).
 * cell_map maps finest cells to current-level cells; for every internal face
 * in order: c0 = map[own], c1 = map[nei]; skip if equal or either paired;
 * else every finest cell whose new_map equals c1 gets c0, both marked paired.
 * Then labels are renumbered by sorted unique value.  Returns the new coarse
 * count, or 0 when nothing paired (reference returns None).
 * The relabel "new_map[new_map == c1] = c0" is done with a union-style
 * redirect table on current-level labels, which gives the same result because
 * a label c1 is merged at most once (it is marked paired). */
int64_t orc_pair_level(int64_t n_fine, int64_t n_cur, int64_t nf, const int64_t *own,
                       const int64_t *nei, const int64_t *cell_map, int64_t *out) {
  char *paired = calloc(n_cur + 1, 1);
  int64_t *redirect = malloc(sizeof(int64_t) * (n_cur + 1));
  if (!paired || !redirect) return -1;
  for (int64_t i = 0; i < n_cur; ++i) redirect[i] = i;
  int64_t any = 0;
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t c0 = cell_map[own[f]], c1 = cell_map[nei[f]];
    if (c0 == c1) continue;
    if (paired[c0] || paired[c1]) continue;
    redirect[c1] = c0;
    paired[c0] = paired[c1] = 1;
    any = 1;
  }
  if (!any) {
    free(paired);
    free(redirect);
    return 0;
  }
  /* new label of current-level label v: redirect[v] (c0 is never redirected
   * itself because it is paired once), then dense renumber by sorted value. */
  char *used = calloc(n_cur + 1, 1);
  int64_t *dense = malloc(sizeof(int64_t) * (n_cur + 1));
  for (int64_t i = 0; i < n_fine; ++i) used[redirect[cell_map[i]]] = 1;
  int64_t k = 0;
  for (int64_t v = 0; v < n_cur; ++v) dense[v] = used[v] ? k++ : -1;
  for (int64_t i = 0; i < n_fine; ++i) out[i] = dense[redirect[cell_map[i]]];
  free(paired);
  free(redirect);
  free(used);
  free(dense);
  return k;
}

/* torch's CPU vector norm of 3-component rows (Tensor.norm(dim=1), used by
 * fvm.laplacian for |S_f|, discretisation/operators.py:490) accumulates the
 * squares with fused multiply-adds: sqrt(fma(z, z, fma(y, y, x*x))). */
int orc_norm3(int64_t n, const double *v, double *out) {
  for (int64_t i = 0; i < n; ++i) {
    const double x = v[3 * i], y = v[3 * i + 1], z = v[3 * i + 2];
    out[i] = sqrt(fma(z, z, fma(y, y, x * x)));
  }
  return 0;
}
