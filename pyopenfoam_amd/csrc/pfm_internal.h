// Internal declarations shared by the libpfm translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <cstdint>
#include <vector>

#include "../../include/pfm.h"

namespace pfm {

void set_error(const char *fmt, ...);

#define PFM_CHECK_HIP(call)                                                    \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::pfm::set_error("%s failed: %s", #call, hipGetErrorString(e_));         \
      return PFM_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

#define PFM_REQUIRE(cond, ...)                                                 \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::pfm::set_error(__VA_ARGS__);                                           \
      return PFM_ERR_ARG;                                                      \
    }                                                                          \
  } while (0)

// Entry points reject null device pointers with a named error.
#define PFM_NONNULL(p)                                                         \
  do {                                                                         \
    if (!(p)) {                                                                \
      ::pfm::set_error("null device pointer: %s", #p);                         \
      return PFM_ERR_ARG;                                                      \
    }                                                                          \
  } while (0)

#define PFM_LAUNCH_CHECK()                                                     \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      ::pfm::set_error("kernel launch failed: %s", hipGetErrorString(e_));     \
      return PFM_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

constexpr int kSlice = 64;     // rows per slice = one wavefront
constexpr int kBlock = 256;    // threads per workgroup
constexpr int kWavesPerBlock = kBlock / kSlice;
constexpr int kXcds = 8;       // MI355X: 8 XCDs, workgroups dealt round-robin

// Paired column offsets: the allocation holds kPairHeader words before the
// array (256-byte aligned), the last of which names the encoding.
constexpr int kPairHeader = 64;
constexpr uint32_t kPairSigned = 1;    // int16 col - row, -32768 = padding
constexpr uint32_t kPairOriented = 2;  // uint16 |col - row|, lower +, upper -, 0 = padding

// A Gauss-Seidel schedule entry (48 bytes): a = {cell, slot of its first
// lower coupling (ptr + lane), wl | wu << 16, group mask | n << 8}, then the
// row's n <= 8 columns in face order (b, c); bit k of the mask says whether
// the k-th coupling in face order is a lower one.  n = 255: the row has more
// than 8 couplings and is merged through the slot map instead.
struct GsEnt {
  int4 a, b, c;
};

// Per-slice layout: slot offset, lower and upper slot counts.
struct SliceInfo {
  int32_t ptr;
  int16_t wl;
  int16_t wu;
};

// Host-built layout of one LDU topology (pure C++, no device state).
struct HostLayout {
  int32_t n = 0, nf = 0, nslices = 0;
  int64_t nslots = 0;
  int32_t max_lower = 0, max_upper = 0;
  std::vector<SliceInfo> slices;
  std::vector<int32_t> col;   // column of each slot, -1 = padding
  std::vector<int32_t> map;   // face of each slot, -1 = padding
  std::vector<int32_t> dic_level_ptr, dic_cells;
  std::vector<int32_t> dilu_level_ptr, dilu_faces;
  std::vector<int32_t> gs_fwd_ptr, gs_fwd_cells;  // Gauss-Seidel sweep levels
  std::vector<int32_t> gs_bwd_ptr, gs_bwd_cells;
};

int build_host_layout(int32_t n, int32_t nf, const int32_t *own,
                      const int32_t *nei, HostLayout &L);
int validate_topology(int32_t n, int32_t nf, const int32_t *own,
                      const int32_t *nei);

// Pinned host words through which a solve polls its device stop flag, with
// the two events that order each read.  A handle keeps kPollSlots pairs and
// every solve in flight on it leases a pair of its own, so solves that run
// concurrently on one handle (threads sharing a topology) never read each
// other's flags.
constexpr int kPollSlots = 8;
struct PollPool {
  int32_t *flags = nullptr;  // pinned, two words per slot
  hipEvent_t ev[2 * kPollSlots] = {};
  std::atomic<uint32_t> busy{0};
};
int poll_pool_init(PollPool &p);  // allocates; called at handle creation
void poll_pool_free(PollPool &p);

class PollLease {
 public:
  PollLease() = default;
  PollLease(const PollLease &) = delete;
  PollLease &operator=(const PollLease &) = delete;
  // Waits for the last flag copy to land, then returns the pair.
  ~PollLease();
  // A free pair of `pool`.  With every pair leased: a pair allocated for this
  // solve when may_allocate, else PFM_ERR_STATE.
  int acquire(PollPool *pool, bool may_allocate);
  int32_t *flag = nullptr;
  hipEvent_t ev[2] = {};

 private:
  PollPool *pool_ = nullptr;
  int slot_ = -1;
};

}  // namespace pfm

struct pfm_batch_s;

// Dirichlet faces grouped by cell (pfm_dirichlet_create).
struct pfm_dirichlet_s {
  int32_t n_dir = 0, n_groups = 0;
  const struct pfm_ldu_s *ldu = nullptr;
  int32_t *d_gptr = nullptr;   // group g: sorted entries gptr[g] .. gptr[g+1]-1
  int32_t *d_gcell = nullptr;  // cell of group g
  int32_t *d_order = nullptr;  // caller's entry index of each sorted entry
  int32_t *d_patch = nullptr;  // patch of each sorted entry
};

// Device-resident addressing of one LDU topology.
struct pfm_ldu_s {
  int32_t n = 0, nf = 0, nslices = 0;
  int64_t nslots = 0;
  int32_t max_lower = 0, max_upper = 0;
  int device = 0;
  std::vector<int32_t> dic_level_ptr;   // host copies of level offsets
  std::vector<int32_t> dilu_level_ptr;
  pfm::SliceInfo *d_slices = nullptr;
  int32_t *d_col = nullptr;
  int16_t *d_col16 = nullptr;           // column - row per slot, when all fit (else null)
  uint32_t *d_colp2 = nullptr;          // column offsets paired per lane (pair_base), or null
  uint32_t *d_colp2_alloc = nullptr;    // its allocation: kPairHeader words, then d_colp2
  int32_t *d_map = nullptr;
  int32_t *d_own = nullptr;
  int32_t *d_nei = nullptr;
  int32_t *d_own_start = nullptr;       // n + 1: first owner-face of each row (oriented_sorted only)
  int32_t *d_dic_cells = nullptr;
  int32_t *d_dilu_faces = nullptr;
  int32_t *d_dilu_level_ptr = nullptr;
  void *d_dilu_pq = nullptr;            // level-ordered (owner, neighbour), lazily allocated
  double *d_dilu_lu = nullptr;          // level-ordered lower*upper
  // LDS-resident DILU sweep (ensure_dilu_slots): per face in level order the
  // LDS slots of its owner and neighbour (low / high 16 bits); per phase the
  // cells written back and loaded, as (cell, -(slot + 1)) or (cell, slot).
  int dilu_slots_state = 0;             // 0 not built, 1 built, -1 does not fit in LDS
  int dilu_lds_threads = 0;             // threads of the sweep (512 or 1024)
  std::mutex dilu_slots_mu;
  uint32_t *d_dilu_fslot = nullptr;
  int2 *d_dilu_io = nullptr;
  int32_t *d_dilu_iptr = nullptr;       // levels + 3: the loads before level 0, phases 0 .. levels
  double *d_dilu_din = nullptr;         // initial value of each loaded cell, per factor call
  int32_t n_dilu_io = 0;
  std::vector<int32_t> gs_fwd_ptr, gs_bwd_ptr;  // host copies of GS level offsets
  int32_t *d_gs_fwd_ptr = nullptr, *d_gs_fwd_cells = nullptr;
  int32_t *d_gs_bwd_ptr = nullptr, *d_gs_bwd_cells = nullptr;
  pfm::GsEnt *d_gs_fwd_ent = nullptr, *d_gs_bwd_ent = nullptr;  // ensure_gs_entries()
  std::mutex gs_ent_mu;
  pfm::PollPool poll;                   // pinned stop-flag slots of the solves
  bool oriented_sorted = false;         // every owner < neighbour, owners non-decreasing
  pfm_progress_fn progress = nullptr;   // optional host progress report
  void *progress_user = nullptr;
  double progress_interval = 0.0;
};

namespace pfm {
// The body of pfm_pcg_solve: `work` holds pfm_pcg_workspace_bytes(h,
// prm->max_iter) bytes of device memory; may_allocate = false restricts the
// host loop to the handle's pooled poll slots (no allocation at all).
int pcg_solve(pfm_ldu h, const pfm_solve_params *prm, const double *diag, const double *packed,
              const double *rD, const double *b, double *x, void *work, double *history_h,
              pfm_solve_result *res, void *stream, bool may_allocate);

// Builds the handle's Gauss-Seidel schedule entries if they are not built yet
// (allocates; synchronises `s` once).
int ensure_gs_entries(pfm_ldu_s *h, hipStream_t s);

// Builds the handle's LDS slot schedule of the DILU sweep if it is not built
// yet (host work over a download of the face levels; synchronises `s`), and
// records whether the cells live at once fit the workgroup's LDS.
int ensure_dilu_slots(pfm_ldu_s *h, hipStream_t s);

// Calls the handle's progress report when its interval has passed since `last`.
void report_progress(const pfm_ldu_s *h, int32_t iterations, double &last);
double now_seconds();
}  // namespace pfm

namespace pfm {
// Grid helpers --------------------------------------------------------------
inline int blocks_for_slices(int nslices) {
  return (nslices + kWavesPerBlock - 1) / kWavesPerBlock;
}
}  // namespace pfm
