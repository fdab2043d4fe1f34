/*
 * pfm.h -- C ABI of libpfm, the MI355X (gfx950) pressure-equation path:
 * LDU assembly, LDU SpMV, PCG(DIC) / PBiCGStab(DILU) / GAMG solves and the
 * adjoint face-gradient scatter.
 *
 * Conventions
 *   - Every function returns an int status (PFM_OK = 0).  On failure the
 *     message is available from pfm_last_error() (thread-local).
 *   - Device arrays are raw pointers into caller-owned memory; the library
 *     never frees them.  Host arrays are marked _h.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     All device work is enqueued on it; solver entry points synchronise
 *     it once at the end to return their result struct.
 *   - Floating point: FP64 unless a name ends in _f32.  Indices: int32.
 *   - LDU storage (reference src/pyfoam/core/ldu_matrix.py:27-71):
 *       diag[n_cells]; lower[f] = A[owner[f], neighbour[f]];
 *       upper[f] = A[neighbour[f], owner[f]].
 *
 * Reference citations are relative to the pyOpenFOAM source tree
 * (src/pyfoam/...).
 */
#ifndef PFM_H
#define PFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFM_OK 0
#define PFM_ERR_ARG 1      /* bad argument: maps to ValueError / TypeError */
#define PFM_ERR_HIP 2      /* HIP runtime error */
#define PFM_ERR_STATE 3    /* handle not usable for this call */
#define PFM_ERR_NOMEM 4

/* Solver status codes (pfm_solve_result.status). */
#define PFM_SOLVE_RUNNING 0
#define PFM_SOLVE_CONVERGED 1
#define PFM_SOLVE_MAXITER 2
#define PFM_SOLVE_BREAKDOWN 3

#define PFM_PRECOND_NONE 0
#define PFM_PRECOND_DIAG 1 /* z = r / rD, rD from pfm_dic_factor / pfm_dilu_factor */

const char *pfm_last_error(void);
/* ABI version of this header (no reference counterpart: checked by the loader). */
int pfm_abi_version(void);

/* ------------------------------------------------------------------ */
/* Host-side setup algorithms (no GPU needed)                          */
/* ------------------------------------------------------------------ */

/* Reverse Cuthill-McKee ordering, perm_out[new] = old.
 * Replaces tools/renumber_mesh.py:163-239 (_reverse_cuthill_mckee) with the
 * same tie-breaks: neighbours in ascending index stably sorted by degree,
 * start = lowest-index unvisited cell of minimum degree, FIFO BFS, restart
 * on disconnected parts, whole list reversed. */
int pfm_rcm(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
            const int32_t *neighbour_h, int32_t *perm_out_h);

/* Dependency level of every cell in the DIC factor recurrence
 * (solvers/preconditioners.py:102-133): cell i waits for every owner j < i
 * of a face whose neighbour is i.  Returns the level count in *n_levels. */
int pfm_dic_levels(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
                   const int32_t *neighbour_h, int32_t *level_out_h,
                   int32_t *n_levels);

/* Dependency level of every face in the DILU factor recurrence
 * (solvers/preconditioners.py:244-272): face f waits for the previous face
 * (in face order) touching either of its cells. */
int pfm_dilu_levels(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
                    const int32_t *neighbour_h, int32_t *level_out_h,
                    int32_t *n_levels);

/* Cell -> face lists of a polyMesh, the order of mesh/mesh_geometry.py:188-194:
 * for each cell the faces it owns in ascending face order, then the internal
 * faces it neighbours in ascending order; a neighbour-side entry is stored as
 * -(f + 1).  face_ptr_h (n_faces + 1) delimits each face's vertex loop;
 * *max_cell_verts_out = the largest number of vertex references of one cell
 * (the bound pfm_mesh_geometry checks).  cell_ptr_out_h: n_cells + 1,
 * cell_faces_out_h: n_faces + n_internal. */
int pfm_cell_faces(int32_t n_cells, int32_t n_faces, int32_t n_internal,
                   const int32_t *owner_h, const int32_t *neighbour_h,
                   const int32_t *face_ptr_h, int32_t *cell_ptr_out_h,
                   int32_t *cell_faces_out_h, int32_t *max_cell_verts_out);

/* Applies a cell permutation to the internal faces, the step
 * tools/renumber_mesh.py:60-97 leaves to its caller: owner and neighbour are
 * relabelled (new_of_old[old] = new), a face whose new owner exceeds its new
 * neighbour is swapped (flip_out[k] = 1: reverse its vertex loop), and the
 * faces are sorted by (owner, neighbour), ties in face order, which gives the
 * owner < neighbour orientation checked by mesh/topology.py:81-87.
 * order_out[k] = old index of new face k. */
int pfm_relabel_faces(int32_t n_cells, int32_t n_internal, const int32_t *owner_h,
                      const int32_t *neighbour_h, const int32_t *new_of_old_h,
                      int32_t *order_out_h, uint8_t *flip_out_h, int32_t *owner_out_h,
                      int32_t *neighbour_out_h);

/* Parsers of the large ASCII lists of an OpenFOAM polyMesh, the inner loops
 * of io/mesh_io.py:154-369 (_parse_ascii_points / _parse_ascii_faces /
 * _parse_ascii_label_list).  `text` is what follows the list's opening
 * parenthesis (len bytes, not null-terminated; reading stops at the list's
 * closing parenthesis); n is the count the file states, and a different
 * count is an error, as in the reference.  Points:
 * one "(x y z)" group per point -> out[3n].  Faces: one "<k>(v ...)" group
 * per face (k unchecked, as the reference's regex leaves it) -> face_ptr[n+1]
 * and verts[verts_cap] (*n_verts = labels found; more than verts_cap is an
 * error).  Labels: whitespace-separated integers -> out[n]. */
int pfm_parse_ascii_points(const char *text, int64_t len, int64_t n,
                           double *out, int64_t *n_read);
int pfm_parse_ascii_faces(const char *text, int64_t len, int64_t n,
                          int64_t verts_cap, int32_t *face_ptr, int32_t *verts,
                          int64_t *n_read, int64_t *n_verts);
int pfm_parse_ascii_labels(const char *text, int64_t len, int64_t n,
                           int64_t *out, int64_t *n_read);

/* Sweep levels of the symmetric Gauss-Seidel smoother
 * (solvers/smoothers.py:92-149): in the forward sweep cell i reads the
 * updated x of every neighbour j < i, so fwd_level[i] = 1 + max over those;
 * in the backward sweep likewise for neighbours j > i.  Cells of one level
 * are independent, so level-by-level updates equal the sequential sweeps. */
int pfm_gs_levels(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
                  const int32_t *neighbour_h, int32_t *fwd_level_out_h,
                  int32_t *bwd_level_out_h, int32_t *n_fwd_levels, int32_t *n_bwd_levels);

/* Greedy aggregation of GAMGSolver._build_aggregates (solvers/gamg.py:164-239)
 * on one level: f2c_out[n_cells], *n_coarse = aggregate count. */
int pfm_gamg_aggregate(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
                       const int32_t *neighbour_h, int32_t target, int32_t *f2c_out_h,
                       int32_t *n_coarse);

/* PairGamgAgglomeration.build_levels (fv_agglomeration/pair_agglomeration.py:47-131):
 * cumulative finest->level maps, level k in maps_out[k*n_cells ...],
 * coarse counts in n_coarse_out[k]; *n_levels_out levels were made. */
int pfm_pair_agglomerate(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
                         const int32_t *neighbour_h, int32_t n_levels, int32_t *maps_out_h,
                         int32_t *n_coarse_out_h, int32_t *n_levels_out);

/* ------------------------------------------------------------------ */
/* LDU addressing handle                                              */
/* ------------------------------------------------------------------ */

typedef struct pfm_ldu_s *pfm_ldu;

typedef struct {
  int32_t n_cells;
  int32_t n_faces;
  int32_t n_slices;      /* row slices of 64 cells (one wavefront each) */
  int32_t n_dic_levels;
  int32_t n_dilu_levels;
  int32_t max_lower;     /* max owner-faces of one cell */
  int32_t max_upper;     /* max neighbour-faces of one cell */
  int32_t device;
  int64_t n_slots;       /* length of a packed coefficient array */
} pfm_ldu_info;

/* Build the device addressing of one LDU topology (owner/neighbour of the
 * internal faces, reference LduMatrix.__init__ ldu_matrix.py:74-100):
 * row slices for the SpMV, DIC cell levels and DILU face levels.
 * Topology arrays are host int32; the current HIP device is used. */
int pfm_ldu_create(int32_t n_cells, int32_t n_faces, const int32_t *owner_h,
                   const int32_t *neighbour_h, pfm_ldu *out);
int pfm_ldu_destroy(pfm_ldu h);
int pfm_ldu_get_info(pfm_ldu h, pfm_ldu_info *out);

/* Progress reports of long solves (no reference counterpart: the reference
 * logs residuals per iteration on the host).  When set, the host loops that
 * drive pfm_pcg_solve, pfm_pbicgstab_solve, pfm_pcg_mixed_solve and
 * pfm_gamg_solve (hierarchies built on this handle) call fn(iterations_enqueued, user) at most every interval_s seconds
 * (from the calling thread, between polls of the device stop flag).  fn =
 * NULL turns it off. */
typedef void (*pfm_progress_fn)(int32_t iterations, void *user);
int pfm_ldu_set_progress(pfm_ldu h, pfm_progress_fn fn, void *user, double interval_s);

/* Pack face coefficients into row-slice order: packed[n_slots].
 * transpose = 0 packs A, transpose = 1 packs A^T (lower and upper swapped,
 * as LduMatrix.Ax_T ldu_matrix.py:219-258 and the adjoint matrix of
 * differentiable/linear_solver.py:127-131). */
int pfm_ldu_pack(pfm_ldu h, int transpose, const double *lower,
                 const double *upper, double *packed, void *stream);
int pfm_ldu_pack_f32(pfm_ldu h, int transpose, const double *lower,
                     const double *upper, float *packed, void *stream);

/* y = A x on a packed matrix.  Row sums follow the reference association
 * y = (diag*x + sum_owner-faces lower*x[nei]) + sum_neighbour-faces upper*x[own]
 * with each sum in face order, i.e. LduMatrix.Ax ldu_matrix.py:200-217.
 * A product that streams at least 128 MiB (half the MALL) loads its
 * coefficients, columns and diagonal and stores y nontemporally. */
int pfm_spmv(pfm_ldu h, const double *diag, const double *packed,
             const double *x, double *y, void *stream);
int pfm_spmv_f32(pfm_ldu h, const float *diag, const float *packed,
                 const float *x, float *y, void *stream);

/* Y = A X for k right-hand sides: LduMatrix.Ax_batched / Ax_sparse on an
 * (n_cells, k) input (ldu_matrix.py:345-381), x and y row-major
 * (x[i*k + j]).  Each column is summed in the association of pfm_spmv
 * (LduMatrix.Ax), so column j equals the single product of x[:, j] bit for
 * bit; the reference's CSR product rounds in its own order. */
int pfm_spmv_multi(pfm_ldu h, const double *diag, const double *packed, int32_t k,
                   const double *x, double *y, void *stream);

/* FvMatrix.relax (core/fv_matrix.py:192-235): diag_out = diag / alpha,
 * source_out = source + ((1 - alpha) * diag_out) * field_old; alpha in
 * (0, 1] (PFM_ERR_ARG otherwise).  Outputs may alias the inputs. */
int pfm_relax(int32_t n, double alpha, const double *diag, const double *source,
              const double *field_old, double *diag_out, double *source_out,
              void *stream);

/* Achievable-bandwidth probes (no reference counterpart: measurement of the
 * ceiling the bandwidth-bound kernels are compared with).  mode 0 reads
 * nbytes of src with nontemporal 16-byte loads and writes one partial sum per
 * wave to dst (pfm_bandwidth_probe_out_bytes() bytes); mode 1 copies nbytes
 * from src to dst.  nbytes: a multiple of 16; buffers 16-byte aligned. */
int pfm_bandwidth_probe(int mode, const void *src, void *dst, int64_t nbytes,
                        void *stream);
int64_t pfm_bandwidth_probe_out_bytes(void);
/* y = A x (transpose = 0) or A^T x (transpose = 1) straight from the face
 * arrays, no packing: LduMatrix.Ax / Ax_T (ldu_matrix.py:172-258). */
int pfm_spmv_ldu(pfm_ldu h, int transpose, const double *diag,
                 const double *lower, const double *upper, const double *x,
                 double *y, void *stream);
/* r = b - A x (packed A). */
int pfm_residual(pfm_ldu h, const double *diag, const double *packed,
                 const double *x, const double *b, double *r, void *stream);

/* DIC factor (preconditioners.py:88-133): rD = max(|d|, 1e-30) with the
 * sequential recurrence d[i] -= upper^2 / d[owner] evaluated level by level.
 * `packed` must come from pfm_ldu_pack with the same transpose flag as the
 * matrix being solved. Bit-exact with the sequential recurrence. */
int pfm_dic_factor(pfm_ldu h, const double *diag, const double *packed,
                   double *rD, void *stream);
/* DILU factor (preconditioners.py:228-272), face levels, bit-exact. */
int pfm_dilu_factor(pfm_ldu h, int transpose, const double *diag,
                    const double *lower, const double *upper, double *rD,
                    void *stream);

/* ------------------------------------------------------------------ */
/* Krylov solvers                                                     */
/* ------------------------------------------------------------------ */

typedef struct {
  double tolerance; /* absolute: converged when ||r|| < tolerance */
  double rel_tol;   /* relative: ||r|| / ||r0|| < rel_tol (r0 > 0) */
  int32_t max_iter;
  int32_t min_iter;
  int32_t precond;  /* PFM_PRECOND_* */
  int32_t reserved;
} pfm_solve_params;

typedef struct {
  int32_t iterations;  /* length of the residual history (reference
                          ConvergenceInfo.iterations, residual.py:162) */
  int32_t converged;
  int32_t status;      /* PFM_SOLVE_* */
  int32_t reserved;
  double initial_residual;
  double final_residual;
} pfm_solve_result;

/* Device workspace (bytes) for one solve with at most max_iter iterations.
 * No reference counterpart: PCGSolver._solve / PBiCGSTABSolver._solve
 * (pcg.py:97-187, pbicgstab.py:103-239) allocate their vectors per call; here
 * the caller allocates them once and passes them in. */
int64_t pfm_pcg_workspace_bytes(pfm_ldu h, int32_t max_iter);
int64_t pfm_pbicgstab_workspace_bytes(pfm_ldu h, int32_t max_iter);

/* PCG (solvers/pcg.py:97-187) with the stopping rule of
 * solvers/residual.py:112-145; x holds x0 on entry and the solution on exit.
 * history_h (host, may be NULL) receives result.iterations residual norms. */
int pfm_pcg_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag,
                  const double *packed, const double *rD, const double *b,
                  double *x, void *work, double *history_h,
                  pfm_solve_result *res, void *stream);
/* pfm_pcg_solve without a caller workspace and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace lives in a file-scope device array of
 * pfm_pcg_noalloc_max_bytes() bytes (68 MiB: 2^21 cells with max_iter up to
 * 250,000) that each GPU the library runs on holds until the process ends;
 * a solve needing more (pfm_pcg_workspace_bytes) is rejected with
 * PFM_ERR_ARG before anything is enqueued.  The handle's device must be the
 * current device.  Calls on one GPU share that array: calls that overlap in
 * time on one GPU race, and each call starts by writing every word it later
 * reads (nothing is cleared).  The host loop polls through the handle's
 * pooled pinned slots; with all of them leased by concurrent solves it
 * returns PFM_ERR_STATE instead of allocating. */
int64_t pfm_pcg_noalloc_max_bytes(void);
int pfm_pcg_solve_noalloc(pfm_ldu h, const pfm_solve_params *prm,
                          const double *diag, const double *packed,
                          const double *rD, const double *b, double *x,
                          double *history_h, pfm_solve_result *res,
                          void *stream);
/* PBiCGStab (solvers/pbicgstab.py:103-239), same conventions. */
int pfm_pbicgstab_solve(pfm_ldu h, const pfm_solve_params *prm,
                        const double *diag, const double *packed,
                        const double *rD, const double *b, double *x,
                        void *work, double *history_h, pfm_solve_result *res,
                        void *stream);

/* PBiCGSolver._solve (solvers/pbicg.py:107-220): BiCG with the shadow system
 * on A^T and the same diagonal preconditioner (the reference's M^-T = M^-1);
 * packed_t = pfm_ldu_pack(h, 1, lower, upper) (A^T).  Breakdown when
 * |rho| or |z~.v| < 1e-30 (not converged). */
int64_t pfm_pbicg_workspace_bytes(pfm_ldu h, int32_t max_iter);
int pfm_pbicg_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag,
                    const double *packed, const double *packed_t, const double *rD,
                    const double *b, double *x, void *work, double *history_h,
                    pfm_solve_result *res, void *stream);

/* Mixed precision (extension of PCGSolver, solvers/pcg.py:97-187): FP32
 * coefficients (pfm_ldu_pack_f32), FP32 SpMV and FP32 diagonal DIC apply in the
 * inner PCG loop (every dot product and norm accumulated in FP64), inside FP64
 * iterative refinement: r = b - A x in FP64, inner solve to
 * ||r_in|| < inner_rel_tol*||r||, x += e.  The reference's stopping rule is
 * applied to the FP64 true residual after each refinement step; prm->max_iter
 * bounds the total inner iterations; history_h receives the true residual
 * norms (result.iterations of them); *inner_iterations the inner total. */
int64_t pfm_pcg_mixed_workspace_bytes(pfm_ldu h, int32_t max_outer);
int pfm_pcg_mixed_solve(pfm_ldu h, const pfm_solve_params *prm, double inner_rel_tol,
                        int32_t max_outer, const double *diag, const double *packed,
                        const float *packed32, const double *rD, const double *b, double *x,
                        void *work, double *history_h, pfm_solve_result *res,
                        int32_t *inner_iterations, void *stream);

/* ------------------------------------------------------------------ */
/* Batched PCG                                                        */
/* ------------------------------------------------------------------ */

typedef struct pfm_batch_s *pfm_batch;

/* Many independent LDU systems (same solver, uneven sizes) as one batch.
 * System k has cells [cell_offsets[k], cell_offsets[k+1]) and faces
 * [face_offsets[k], face_offsets[k+1]) of the concatenated arrays; owner_h /
 * neighbour_h hold each system's LOCAL cell indices (0-based per system).
 * tile_slices = 0 autotunes the work tile (slices of 64 rows per workgroup) by
 * timing whole batched PCG iterations on this batch; no reference counterpart
 * (the reference solves systems one at a time through PCGSolver). */
int pfm_batch_create(int32_t n_systems, const int32_t *cell_offsets_h,
                     const int32_t *face_offsets_h, const int32_t *owner_h,
                     const int32_t *neighbour_h, int32_t tile_slices, pfm_batch *out);
int pfm_batch_destroy(pfm_batch b);
/* n_padded rows, tile count, chosen tile size, autotuner timings (us per
 * iteration for tile sizes 1, 2, 4, 8, 16, 32; 0 if not tuned). */
int pfm_batch_info(pfm_batch b, int32_t *n_padded, int32_t *n_tiles, int32_t *tile_slices,
                   float *tuned_us_h);
int64_t pfm_batch_workspace_bytes(pfm_batch b);
/* PCG (solvers/pcg.py:97-187, reference diagonal DIC) on every system of the
 * batch at once, each with its own monitor (residual.py:112-145) and stop.
 * Arrays are concatenated in the caller's (unpadded) layout; transpose = 1
 * solves A^T (the adjoint).  graph_chunk > 0 captures that many iterations in
 * a hipGraph and replays it; 0 launches eagerly.  Per-system iterations,
 * status and residuals go to the host arrays (any may be NULL). */
int pfm_batch_pcg_solve(pfm_batch b, const pfm_solve_params *prm, int transpose,
                        const double *diag, const double *lower, const double *upper,
                        const double *rhs, double *x, void *work, int32_t graph_chunk,
                        int32_t *iterations_h, int32_t *status_h, double *initial_res_h,
                        double *final_res_h, void *stream);

/* ------------------------------------------------------------------ */
/* Smoothers                                                          */
/* ------------------------------------------------------------------ */

/* GaussSeidelSmoother.smooth (solvers/smoothers.py:92-149): n_sweeps times a
 * forward sweep over cells 0..n-1 then a backward sweep n-1..0; each cell's
 * row sum runs over its faces in face order with the latest x (diag*x first),
 * off = row - diag*x, and x += omega*((b - off)/diag - x) where |diag| >
 * 1e-30.  x is updated in place (the reference returns a smoothed copy).
 * packed = pfm_ldu_pack(h, 0, lower, upper).  Level-scheduled: one launch
 * per sweep level, or one workgroup walking every level for small meshes. */
int pfm_gs_smooth(pfm_ldu h, const double *diag, const double *packed, const double *b,
                  double *x, int32_t n_sweeps, double omega, void *stream);

/* JacobiSmoother.smooth (solvers/smoothers.py:180-200): n_iter times
 * x = x + (omega * 1/max(|diag|, 1e-30)) * (b - A x), in place on x; work
 * holds n_cells doubles. */
int pfm_jacobi_smooth(pfm_ldu h, const double *diag, const double *packed, const double *b,
                      double *x, int32_t n_iter, double omega, double *work, void *stream);

/* DICPreconditioner.apply_full / DILUPreconditioner.apply_full
 * (solvers/preconditioners.py:153-201, :290-337): z = M^-1 r with the forward
 * face sweep z[n] -= (upper_f z[p]) / rD[p], z /= rD and the reverse face
 * sweep z[p] -= (c_f z[n]) / rD[n], c = upper (DIC, faces with owner <
 * neighbour only) or lower (DILU).  rD from pfm_dic_factor / pfm_dilu_factor.
 * Faces oriented owner < neighbour with non-decreasing owners run cell-level
 * schedules (one launch per level); any other face order one workgroup over
 * the face levels.  z may alias r. */
int pfm_precond_apply_full(pfm_ldu h, int dilu, const double *rD, const double *lower,
                           const double *upper, const double *r, double *z, void *stream);

/* DiagonalSolver._solve (solvers/diagonal.py:57-85): x = b / max(|diag|,
 * 1e-30); *residual_h = ||b - diag x||.  work holds n_cells/256 + 2 doubles. */
int pfm_diagonal_solve(pfm_ldu h, const double *diag, const double *b, double *x,
                       double *work, double *residual_h, void *stream);

#define PFM_SMOOTH_GAUSS_SEIDEL 0 /* forward sweeps, x[c] = (b - off) / diag */
#define PFM_SMOOTH_DIC 1          /* Richardson: x += omega DIC.apply_full(b - A x) */
#define PFM_SMOOTH_DILU 2         /* Richardson with DILU.apply_full */

/* SmoothSolver._solve (solvers/smooth_solver.py:96-232): monitor on b - A x0;
 * each iteration n_sweeps smoothing passes, then the monitor on b - A x.
 * packed = pfm_ldu_pack(h, 0, lower, upper); rD for DIC/DILU. */
int64_t pfm_smooth_workspace_bytes(pfm_ldu h, int32_t max_iter);
int pfm_smooth_solve(pfm_ldu h, const pfm_solve_params *prm, int smoother, int32_t n_sweeps,
                     double omega, const double *diag, const double *lower, const double *upper,
                     const double *packed, const double *rD, const double *b, double *x,
                     void *work, double *history_h, pfm_solve_result *res, void *stream);


/* ------------------------------------------------------------------ */
/* GAMG                                                               */
/* ------------------------------------------------------------------ */

typedef struct pfm_gamg_s *pfm_gamg;

#define PFM_AGGLOM_GREEDY 0   /* GAMGSolver._build_aggregates (solvers/gamg.py:164-239) */
#define PFM_AGGLOM_PAIR 1     /* PairGamgAgglomeration (fv_agglomeration/pair_agglomeration.py:47-131) */
#define PFM_COARSE_REFERENCE 0 /* _build_coarse_matrix (gamg.py:242-333): couplings
                                  inside an aggregate are dropped */
#define PFM_COARSE_GALERKIN 1  /* extension: they are added to the coarse diagonal,
                                  the Galerkin product R A P of piecewise-constant P */
#define PFM_SMOOTHER_JACOBI 0       /* GAMGSolver._smooth (gamg.py:385-429) */
#define PFM_SMOOTHER_GAUSS_SEIDEL 1 /* extension: every smoothing step on every level is a
                                       symmetric GaussSeidelSmoother sweep
                                       (smoothers.py:92-149), in place */

typedef struct {
  int32_t max_levels;       /* coarsening steps (reference default 10) */
  int32_t min_cells_coarse; /* stop at or below this many cells (default 10) */
  int32_t n_pre;            /* pre-smoothing sweeps (default 2) */
  int32_t n_post;           /* post-smoothing sweeps (default 2) */
  int32_t n_coarsest;       /* sweeps on the coarsest level (reference: 50) */
  int32_t agglomeration;    /* PFM_AGGLOM_* */
  int32_t coarse;           /* PFM_COARSE_* */
  int32_t smoother;         /* PFM_SMOOTHER_* */
} pfm_gamg_params;

/* Build the hierarchy of GAMGSolver._solve (gamg.py:541-570) from the fine
 * topology: aggregation maps and coarse face maps per level (host, topology
 * only) and an LDU handle per coarse level.  Coefficients are coarsened on the
 * device by every solve. `fine` must outlive the hierarchy. */
int pfm_gamg_create(pfm_ldu fine, const pfm_gamg_params *prm, pfm_gamg *out);
int pfm_gamg_destroy(pfm_gamg g);
int pfm_gamg_num_levels(pfm_gamg g);
int pfm_gamg_level_info(pfm_gamg g, int32_t level, int32_t *n_cells, int32_t *n_faces);
/* Maps of level -> level+1 (host arrays; any may be NULL): f2c[n_cells(level)],
 * face_map[n_faces(level)] (coarse face or -1), coarse owner/neighbour. */
int pfm_gamg_level_maps(pfm_gamg g, int32_t level, int32_t *f2c_h, int32_t *face_map_h,
                        int32_t *c_owner_h, int32_t *c_neighbour_h);

/* Workspace of pfm_gamg_solve (no reference counterpart, as for PCG above). */
int64_t pfm_gamg_workspace_bytes(pfm_gamg g, int32_t max_iter);
/* GAMGSolver._build_coarse_matrix (solvers/gamg.py:242-330) for every level:
 * coarsen the coefficients into the workspace (diag by aggregate sums in fine
 * cell order, off-diagonals by coarse-face sums in fine face order), pack every
 * level and compute the Jacobi inverse diagonals.  pfm_gamg_solve calls it. */
int pfm_gamg_setup_values(pfm_gamg g, const double *diag, const double *lower,
                          const double *upper, void *work, void *stream);
/* Copy the coarsened coefficients of `level` (>= 1) out of the workspace (the
 * reference's GAMGLevel diag / lower / upper, gamg.py:326-333). */
int pfm_gamg_level_values(pfm_gamg g, void *work, int32_t level, double *diag,
                          double *lower, double *upper, void *stream);
/* GAMGSolver._solve (gamg.py:499-585): V-cycles with undamped Jacobi
 * (1/|diag|) smoothing, sum restriction, injection prolongation and the
 * coarsest level smoothed n_coarsest times; the residual b - A x after every
 * cycle feeds the monitor (no initial residual is recorded). */
int pfm_gamg_solve(pfm_gamg g, const pfm_solve_params *prm, const double *diag,
                   const double *lower, const double *upper, const double *b,
                   double *x, void *work, double *history_h, pfm_solve_result *res,
                   void *stream);

/* Mixed-precision GAMG (no reference counterpart: the precision pattern on
 * GAMGSolver._solve, gamg.py:541-585).  FP64 iterative refinement: r = b - A x
 * in FP64 and the residual monitor on it (the initial residual included, as
 * PCG); each refinement step solves A e = r with V-cycles whose finest
 * fp32_levels levels run in FP32 (coefficients coarsened in FP64 in the
 * reference order, then cast; FP32 SpMV, smoothing, restriction and
 * prolongation) and whose coarser levels, the coarsest-level solve included,
 * run in FP64 (levels of at most 4096 cells are always FP64), until the inner
 * residual drops below inner_rel_tol times the outer one; then x += e in
 * FP64.  prm->max_iter bounds the total V-cycles, max_outer the refinement
 * steps; *inner_cycles returns the cycles run. */
int64_t pfm_gamg_mixed_workspace_bytes(pfm_gamg g, int32_t max_outer, int32_t fp32_levels);
int pfm_gamg_mixed_solve(pfm_gamg g, const pfm_solve_params *prm, double inner_rel_tol,
                         int32_t max_outer, int32_t fp32_levels, const double *diag,
                         const double *lower,
                         const double *upper, const double *b, double *x, void *work,
                         double *history_h, pfm_solve_result *res, int32_t *inner_cycles,
                         void *stream);

/* PCG (solvers/pcg.py:97-187, the reference's loop and stopping rule) with a
 * GAMG V-cycle as the preconditioner: z = M^-1 r is one cycle of this
 * hierarchy from zero on A z = r (no reference counterpart: the reference
 * PCG knows DIC, DILU and none; OpenFOAM's "PCG with GAMG preconditioner").
 * fp32_levels > 0 runs the cycle's finest levels in FP32 (mixed precision:
 * FP64 Krylov vectors, dot products, SpMV of the outer loop and coarse
 * levels).  pfm_gamg_pcg_workspace_bytes(g, max_iter, fp32_levels) bytes. */
int64_t pfm_gamg_pcg_workspace_bytes(pfm_gamg g, int32_t max_iter, int32_t fp32_levels);
int pfm_gamg_pcg_solve(pfm_gamg g, const pfm_solve_params *prm, int32_t fp32_levels,
                       const double *diag, const double *lower, const double *upper,
                       const double *b, double *x, void *work, double *history_h,
                       pfm_solve_result *res, void *stream);



/* ------------------------------------------------------------------ */
/* Assembly and adjoint                                               */
/* ------------------------------------------------------------------ */

/* Mesh geometry on the device, mesh/mesh_geometry.py:35-395: face centres
 * (n_faces x 3, vertex mean), area vectors (0.5 sum v_j x v_{j+1}; zero for
 * faces of fewer than 3 vertices) and |S_f|; cell volumes and centres by the
 * reference's tet decomposition around the mean of the cell's distinct
 * vertices (faces in pfm_cell_faces order, loops reversed on the neighbour
 * side, |t| < 1e-30 skipped, degenerate cells: V = 0 and the mean of their
 * face centres); delta = 1/|d . n_hat| on internal faces, 0 on boundary
 * faces (guards at 1e-30 as in :386-393).  Faces: CSR face_ptr /
 * face_verts (int32 vertex ids in [0, n_points)); points and all vector
 * outputs are row-major (n, 3) FP64.  All arrays are device arrays;
 * max_cell_verts (from pfm_cell_faces) must be <= 64. */
int pfm_mesh_geometry(int32_t n_points, const double *points, int32_t n_faces,
                      const int32_t *face_ptr, const int32_t *face_verts,
                      int32_t n_internal, const int32_t *owner,
                      const int32_t *neighbour, int32_t n_cells,
                      const int32_t *cell_ptr, const int32_t *cell_faces,
                      int32_t max_cell_verts, double *face_centres,
                      double *face_areas, double *mag_sf, double *cell_volumes,
                      double *cell_centres, double *delta, void *stream);

/* Linear-interpolation face weights, mesh/mesh_geometry.py:262-327:
 * w = |x_f - C_N| / (|x_f - C_P| + |x_f - C_N|) on internal faces (norms in
 * the order of torch's norm, the sum guarded at 1e-30), 1 on boundary faces.
 * Device arrays; face_centres / cell_centres row-major (n, 3). */
int pfm_face_weights(int32_t n_faces, int32_t n_internal, const int32_t *owner,
                     const int32_t *neighbour, const double *face_centres,
                     const double *cell_centres, double *weights, void *stream);

/* ------------------------------------------------------------------ */
/* The pressure-equation step around the solve                         */
/* (solvers/pressure_equation.py).  Face arrays hold the handle's       */
/* internal faces first, then n_faces - n_internal boundary faces;      */
/* bnd_ptr (n_cells + 1) / bnd_face list, per cell and in face order,   */
/* the boundary faces it owns (index b means face n_internal + b).      */
/* 1/A_p is 1 / clamp(|A_p|, 1e-30); face sums run in the order of the  */
/* reference's scatter_add / index_add_ (owner faces, neighbour faces,  */
/* boundary faces, each ascending).                                     */
/* ------------------------------------------------------------------ */

/* assemble_pressure_equation (pressure_equation.py:49-152): lower/upper =
 * -(1/A_p)_f |S_f| delta_f / V_{P,N} with (1/A_p)_f = w inv_P + (1-w) inv_N,
 * diag = the sum of their negatives per row, source = -div(phiHbyA) / V
 * (boundary faces included).  weights, mag_sf, delta: internal faces. */
int pfm_pressure_assemble(pfm_ldu h, int32_t n_faces, const int32_t *bnd_ptr,
                          const int32_t *bnd_face, const double *phi_hbya,
                          const double *a_p, const double *weights,
                          const double *mag_sf, const double *delta,
                          const double *vol, double *diag, double *lower,
                          double *upper, double *source, void *stream);

/* correct_face_flux (pressure_equation.py:270-332): phi_out = phi +
 * (1/A_p)_f (p_P - p_N) |S_f| delta_f on internal faces, phi elsewhere. */
int pfm_correct_face_flux(pfm_ldu h, int32_t n_faces, const double *phi,
                          const double *p, const double *a_p,
                          const double *weights, const double *mag_sf,
                          const double *delta, double *phi_out, void *stream);

/* correct_velocity (pressure_equation.py:187-267): Gauss gradient of p with
 * linearly interpolated face values (boundary faces: the owner value),
 * divided by clamp(V, 1e-30); u = hbya - (alpha/A_p) grad p, rescaled to
 * |u| = 1e4 where larger.  hbya, u_out, face_areas: row-major (n, 3). */
int pfm_correct_velocity(pfm_ldu h, int32_t n_faces, const int32_t *bnd_ptr,
                         const int32_t *bnd_face, const double *hbya,
                         const double *p, const double *a_p,
                         const double *weights, const double *face_areas,
                         const double *vol, double alpha, double *u_out,
                         void *stream);

/* adjust_phi (pressure_equation.py:335-377): when closed and the net
 * boundary flux is at least 1e-30 in magnitude, subtract net / n_boundary
 * from every boundary face of phi (in place).  work: one device double.
 * Synchronises the stream; *adjusted_out_h = 1 when phi changed,
 * *net_out_h (optional) = the net flux (summed in a fixed device order). */
int pfm_adjust_phi(int32_t n_faces, int32_t n_internal, double *phi,
                   int closed, double *work, double *net_out_h,
                   int32_t *adjusted_out_h, void *stream);

/* Dirichlet (fixedValue) patch faces of a mesh for pfm_laplacian_assemble,
 * the term of applications/laplacian_foam.py:318-366, which loops over the
 * patches and adds each one's scatter_add to the diagonal in turn.
 * dir_cell_h: the owner cell of each of n_dir faces (host array) in that
 * order -- patch after patch, patch p holding entries patch_ptr_h[p] ..
 * patch_ptr_h[p+1]-1, its faces in face order.  The library groups the
 * entries by cell once, on the host, so that the assembly kernel gives every
 * cell to one thread whatever the caller's order, and that thread adds the
 * cell's terms as the reference does: diag + s_1, then + s_2, ..., where s_p
 * is patch p's scatter_add sum over the cell's faces of that patch.  Cells
 * must lie in [0, n_cells of h).  Device state is allocated on the current
 * device; destroy with pfm_dirichlet_destroy. */
typedef struct pfm_dirichlet_s *pfm_dirichlet;
int pfm_dirichlet_create(pfm_ldu h, int32_t n_dir, const int32_t *dir_cell_h,
                         int32_t n_patches, const int32_t *patch_ptr_h,
                         pfm_dirichlet *out);
int pfm_dirichlet_destroy(pfm_dirichlet d);

/* fvm.laplacian (discretisation/operators.py:490-517) plus the
 * laplacianFoam Dirichlet diagonal term (applications/laplacian_foam.py:342-366):
 *   c_f = gamma_f*|S_f|*delta_f, lower_f = -c_f/V_P, upper_f = -c_f/V_N,
 *   diag_i = sum c/V_P over owner faces + sum c/V_N over neighbour faces
 *            + per patch, the sum over its faces b of cell i of
 *              gamma_b*|S_b|*delta_b/V_i.
 * Dirichlet faces (dir = NULL for none) are described by a plan from
 * pfm_dirichlet_create on the same handle; dir_gamma, dir_mag_sf (|S_b|) and
 * dir_delta (1/max(|(x_f - C_P).n_hat|, 1e-30)) hold one value per entry in
 * the plan's (the reference's) order.  Every product, quotient and sum is
 * rounded in the reference's order, so the result is bit-identical to the
 * reference's torch ops. */
int pfm_laplacian_assemble(pfm_ldu h, const double *gamma, const double *mag_sf,
                           const double *delta, const double *vol,
                           pfm_dirichlet dir, const double *dir_gamma,
                           const double *dir_mag_sf, const double *dir_delta,
                           double *diag, double *lower, double *upper,
                           void *stream);

/* Adjoint gradients of DifferentiableLinearSolve.backward
 * (differentiable/linear_solver.py:141-149):
 *   g_diag = -lam*x, g_lower = -lam[own]*x[nei], g_upper = -lam[nei]*x[own].
 * Any output pointer may be NULL. */
int pfm_adjoint_ldu_grads(pfm_ldu h, const double *lam, const double *x,
                          double *g_diag, double *g_lower, double *g_upper,
                          void *stream);

/* Chain rule of the assembly above (what torch autograd does through
 * fvm.laplacian, discretisation/operators.py:432-518, and the Dirichlet term,
 * applications/laplacian_foam.py:341-366): gradient w.r.t. gamma_f given
 * gradients w.r.t. (diag, lower, upper):
 *   g_gamma_f = |S_f| delta_f ((g_diag[P]-g_lower_f)/V_P + (g_diag[N]-g_upper_f)/V_N)
 * and for Dirichlet faces g_gamma_b = |S_b| delta_b g_diag[P] / V_P. */
int pfm_laplacian_backward(pfm_ldu h, const double *g_diag,
                           const double *g_lower, const double *g_upper,
                           const double *mag_sf, const double *delta,
                           const double *vol, int32_t n_dir,
                           const int32_t *dir_cell, const double *dir_mag_sf,
                           const double *dir_delta, double *g_gamma,
                           double *g_dir_gamma, void *stream);

/* Fused adjoint face-gradient scatter straight to dJ/dgamma: the two entry
 * points above composed (differentiable/linear_solver.py:141-149 followed by
 * autograd through the assembly) in one pass over the faces:
 *   dJ/dgamma_f = -|S_f| delta_f (x_P - x_N) (lam_P/V_P - lam_N/V_N)
 *   dJ/dgamma_b = -|S_b| delta_b lam_P x_P / V_P      (Dirichlet faces). */
int pfm_adjoint_gamma_grad(pfm_ldu h, const double *lam, const double *x,
                           const double *mag_sf, const double *delta,
                           const double *vol, int32_t n_dir,
                           const int32_t *dir_cell, const double *dir_mag_sf,
                           const double *dir_delta, double *g_gamma,
                           double *g_dir_gamma, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PFM_H */
