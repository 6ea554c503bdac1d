/*
 * sync.h -- inter-wavefront hand-off used by the level-scheduled ILU and
 * triangular solves.
 *
 * Work is handed out by an atomic ticket in level order, so a wavefront only
 * ever waits for rows owned by tickets issued before its own, i.e. by
 * wavefronts that are already running: no wait can deadlock, whatever the
 * dispatch order of workgroups.
 *
 * Visibility across the 8 XCDs follows the agent-scope release/acquire of
 * the HIP memory model, whose gfx950 lowering keeps the per-XCD L2 caches
 * coherent: the producer stores its results with ordinary stores, executes
 * an agent-scope release fence (L2 write-back of dirty lines, then a wait for
 * completion) and only then stores the row's completion epoch; a consumer
 * polls the epoch with agent-scope atomic loads, executes an agent-scope
 * acquire fence (cache invalidate) once its rows are complete, and then reads
 * the published values with ordinary loads.
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

/* Tickets drawn one task ahead: draw_ticket issues the atomic (lane 0) and
 * returns at once; take_ticket, at the end of the current task, waits for the
 * value and gives it to every lane.  A wave works through its tickets in the
 * order drawn, so the oldest unfinished ticket always belongs to the task its
 * wave is working on and the hand-off stays deadlock-free. */
__device__ __forceinline__ uint32_t draw_ticket(uint32_t *ticket)
{
  uint32_t t = 0;
  if (threadIdx.x == 0) t = atomicAdd(ticket, 1u);
  return t;
}

__device__ __forceinline__ uint32_t take_ticket(uint32_t drawn, uint32_t base)
{
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)drawn) - base;
}

__device__ __forceinline__ uint32_t first_ticket(uint32_t *ticket, uint32_t base)
{
  return take_ticket(draw_ticket(ticket), base);
}

/* next work ticket of this wave (lane 0 draws, all lanes receive it) */
__device__ __forceinline__ uint32_t next_ticket(uint32_t *ticket, uint32_t base)
{
  uint32_t t = 0;
  if (threadIdx.x == 0) t = atomicAdd(ticket, 1u) - base;
  return __shfl(t, 0, 64);
}

#endif
