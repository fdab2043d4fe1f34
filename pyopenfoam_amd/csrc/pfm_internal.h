// Internal declarations shared by the libpfm translation units.
#pragma once

#include <hip/hip_runtime.h>

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

}  // namespace pfm

struct pfm_batch_s;

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
  int32_t *d_map = nullptr;
  int32_t *d_own = nullptr;
  int32_t *d_nei = nullptr;
  int32_t *d_dic_cells = nullptr;
  int32_t *d_dilu_faces = nullptr;
  int32_t *d_dilu_level_ptr = nullptr;
  void *d_dilu_pq = nullptr;            // level-ordered (owner, neighbour), lazily allocated
  double *d_dilu_lu = nullptr;          // level-ordered lower*upper
  std::vector<int32_t> gs_fwd_ptr, gs_bwd_ptr;  // host copies of GS level offsets
  int32_t *d_gs_fwd_ptr = nullptr, *d_gs_fwd_cells = nullptr;
  int32_t *d_gs_bwd_ptr = nullptr, *d_gs_bwd_cells = nullptr;
  int32_t *h_flag = nullptr;            // pinned: solver progress polling
  bool oriented_sorted = false;         // every owner < neighbour, owners non-decreasing
  pfm_progress_fn progress = nullptr;   // optional host progress report
  void *progress_user = nullptr;
  double progress_interval = 0.0;
};

namespace pfm {
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
