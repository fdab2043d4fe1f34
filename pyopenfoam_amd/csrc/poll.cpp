// Leased poll slots (PollPool / PollLease): the pinned words and events
// through which the solver drivers read the device stop flag while the next
// chunk of iterations is already queued.

#include "pfm_internal.h"

namespace pfm {

int poll_pool_init(PollPool &p) {
  hipError_t e = hipHostMalloc((void **)&p.flags, 2 * kPollSlots * sizeof(int32_t),
                               hipHostMallocDefault);
  for (int i = 0; i < 2 * kPollSlots && e == hipSuccess; ++i)
    e = hipEventCreateWithFlags(&p.ev[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    poll_pool_free(p);
    set_error("poll slots: %s", hipGetErrorString(e));
    return PFM_ERR_HIP;
  }
  for (int i = 0; i < 2 * kPollSlots; ++i) p.flags[i] = 0;
  return PFM_OK;
}

void poll_pool_free(PollPool &p) {
  for (int i = 0; i < 2 * kPollSlots; ++i)
    if (p.ev[i]) {
      (void)hipEventDestroy(p.ev[i]);
      p.ev[i] = nullptr;
    }
  if (p.flags) (void)hipHostFree(p.flags);
  p.flags = nullptr;
}

int PollLease::acquire(PollPool *pool, bool may_allocate) {
  if (pool && pool->flags) {
    uint32_t cur = pool->busy.load(std::memory_order_relaxed);
    while (true) {
      int k = 0;
      while (k < kPollSlots && (cur >> k & 1u)) ++k;
      if (k == kPollSlots) break;
      if (pool->busy.compare_exchange_weak(cur, cur | (1u << k), std::memory_order_acquire)) {
        pool_ = pool;
        slot_ = k;
        flag = pool->flags + 2 * k;
        ev[0] = pool->ev[2 * k];
        ev[1] = pool->ev[2 * k + 1];
        flag[0] = flag[1] = 0;
        return PFM_OK;
      }
    }
  }
  if (!may_allocate) {
    set_error("all %d poll slots of the handle are in use by concurrent solves", kPollSlots);
    return PFM_ERR_STATE;
  }
  hipError_t e = hipHostMalloc((void **)&flag, 2 * sizeof(int32_t), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  if (e != hipSuccess) {
    set_error("poll slot: %s", hipGetErrorString(e));
    return PFM_ERR_HIP;
  }
  flag[0] = flag[1] = 0;
  return PFM_OK;
}

PollLease::~PollLease() {
  // An event never recorded counts as complete.
  if (ev[0]) (void)hipEventSynchronize(ev[0]);
  if (ev[1]) (void)hipEventSynchronize(ev[1]);
  if (pool_) {
    pool_->busy.fetch_and(~(1u << slot_), std::memory_order_release);
    return;
  }
  if (ev[0]) (void)hipEventDestroy(ev[0]);
  if (ev[1]) (void)hipEventDestroy(ev[1]);
  if (flag) (void)hipHostFree(flag);
}

}  // namespace pfm
