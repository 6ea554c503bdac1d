/*
 * sync.h -- inter-wavefront hand-off used by the level-scheduled ILU and
 * triangular solves.
 *
 * Work is handed out by an atomic ticket in level order, so a wavefront only
 * ever waits for rows owned by tickets issued before its own, i.e. by
 * wavefronts that are already running: no wait can deadlock, whatever the
 * dispatch order of workgroups.
 *
 * One atomic on a single counter serialises at the memory side (about
 * 11 ns each on MI355X), so the planned kernels take G consecutive tickets
 * per atomic for a workgroup of G wavefronts, one ticket each, and the
 * group's next G once all G are done.  Every ticket below the oldest
 * unfinished one is finished and every ticket of a running group is being
 * worked on, so the oldest unfinished ticket of the launch belongs to a
 * running wave whose dependences are all finished: no wait can deadlock.
 * After a launch of nb workgroups over ntask tickets the counter has
 * advanced by exactly kf_ticket_advance_groups(ntask, nb, G).  The
 * fallback ILU kernel draws one ticket per wave (next_ticket).
 *
 * The triangular solves and the planned ILU hand results over through
 * sentinel-preset slots (handoff.h).  The fallback ILU kernel (no plan) uses
 * the flags below: the producer stores its results, executes an
 * agent-scope release fence and then stores the row's completion epoch; a
 * consumer polls the epoch with agent-scope atomic loads and executes an
 * agent-scope acquire fence before reading the published values.
 *
 * Every wait is bounded.  A wave whose poll limit expires records
 * KF_ERR_SPIN_LIMIT in the context's error word and stops waiting for the
 * rest of the launch (`aborted`), so each wave can time out at most once and
 * every wavefront reaches the end of the kernel: a failure surfaces as an
 * error code, never as a hang.
 */
#ifndef KF_SYNC_H
#define KF_SYNC_H

#include "kf_dev.h"

#define KF_SPIN_LIMIT (1u << 20)

__device__ __forceinline__ uint32_t flag_load(const uint32_t *p)
{
  return __hip_atomic_load(const_cast<uint32_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void flag_store(uint32_t *p, uint32_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* all results this wave stored so far are visible agent-wide before any
 * flag stored after this call (wave-uniform) */
__device__ __forceinline__ void publish_fence() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

/* values published before the flags this wave has seen are visible to its
 * subsequent ordinary loads (wave-uniform) */
__device__ __forceinline__ void consume_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

/* wait until rdone[row] == epoch; false if this wave gave up */
__device__ __forceinline__ bool wait_row(const uint32_t *rdone, uint32_t row, uint32_t epoch, uint32_t *err,
                                         bool &aborted, uint32_t sleep)
{
  if (aborted) return false;
  uint32_t n = 0;
  while (flag_load(rdone + row) != epoch) {
    if (++n > KF_SPIN_LIMIT) {
      atomicOr(err, KF_ERR_SPIN_LIMIT);
      aborted = true;
      return false;
    }
    for (uint32_t z = 0; z < sleep; z++) __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

/* LDS written by some lanes of this wave is read by others after this point
 * (the wave's LDS operations execute in order; the fences keep the compiler
 * from moving accesses across) */
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* next work ticket of this wave (lane 0 draws, all lanes receive it) */
__device__ __forceinline__ uint32_t next_ticket(uint32_t *ticket, uint32_t base)
{
  uint32_t t = 0;
  if (threadIdx.x == 0) t = atomicAdd(ticket, 1u) - base;
  return __shfl(t, 0, 64);
}

/* counter advance of a launch of nb workgroups that take G tickets per
 * atomic, one task per wavefront, and draw the next G when all are done:
 * every batch that starts below ntask, plus the one each workgroup holds at
 * exit */
static inline uint32_t kf_ticket_advance_groups(uint32_t ntask, uint32_t nb, uint32_t G)
{
  return G * ((ntask + G - 1) / G + nb);
}

#endif
