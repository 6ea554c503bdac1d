/*
 * sync.h -- inter-wavefront hand-off used by the level-scheduled ILU and
 * triangular solves.
 *
 * Work is handed out by an atomic ticket in level order, so a wavefront only
 * ever waits for rows owned by tickets issued before its own, i.e. by
 * wavefronts that are already running: no wait can deadlock, whatever the
 * dispatch order of workgroups.  Every wait is additionally bounded
 * (KF_SPIN_LIMIT polls); on expiry the wave records KF_ERR_SPIN_LIMIT in the
 * context's error word and moves on, so every wavefront reaches the end of
 * the kernel.
 *
 * Visibility across the 8 XCDs (each with its own L2): the values one row
 * publishes for others are written with device-scope (sc1) stores, the
 * producer waits for their completion (s_waitcnt vmcnt(0)) and only then
 * stores the row's completion epoch (sc1); consumers poll the epoch with
 * device-scope loads and read the published values with device-scope loads.
 * This avoids the full-L2 write-back / invalidate that a generic agent-scope
 * release/acquire fence costs on gfx950.
 */
#ifndef KF_SYNC_H
#define KF_SYNC_H

#include "kf_dev.h"

#define KF_SPIN_LIMIT (1u << 22)

template <typename T>
__device__ __forceinline__ T dev_load(const T *p)
{
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void dev_store(T *p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* all stores issued so far by this wave have completed */
__device__ __forceinline__ void stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

/* wait until rdone[row] == epoch; returns false if the poll limit expired */
__device__ __forceinline__ bool wait_row(const uint32_t *rdone, uint32_t row, uint32_t epoch, uint32_t *err)
{
  uint32_t n = 0;
  while (dev_load(rdone + row) != epoch) {
    if (++n > KF_SPIN_LIMIT) {
      atomicOr(err, KF_ERR_SPIN_LIMIT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

/* next work ticket of this wave (lane 0 draws, all lanes receive it) */
__device__ __forceinline__ uint32_t next_ticket(uint32_t *ticket, uint32_t base)
{
  uint32_t t = 0;
  if (threadIdx.x == 0) t = atomicAdd(ticket, 1u) - base;
  return __shfl(t, 0, 64);
}

#endif
