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
 * 11 ns each on MI355X), so tickets are handed out in batches of K
 * consecutive values: a wave works through its batch in order and draws the
 * next batch one ahead.  A wave then holds at most 2 K tickets and always
 * works on its oldest unfinished one, so the oldest unfinished ticket of
 * the launch belongs to a running wave whose dependences are all finished:
 * no wait can deadlock.  After a launch of nw waves over ntask tickets the
 * counter has advanced by exactly kf_ticket_advance(ntask, nw, K).
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

/* LDS written by some lanes of this wave is read by others after this point
 * (the wave's LDS operations execute in order; the fences keep the compiler
 * from moving accesses across) */
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* batches of K consecutive tickets (K a launch constant) */
__device__ __forceinline__ uint32_t draw_batch(uint32_t *ticket, uint32_t K)
{
  uint32_t t = 0;
  if (threadIdx.x == 0) t = atomicAdd(ticket, K);
  return t;
}

struct ticket_batch {
  uint32_t t;    /* the current ticket */
  uint32_t left; /* tickets left in the current batch, t included */
  uint32_t nb;   /* the next batch, drawn (value read when needed) */
};

__device__ __forceinline__ ticket_batch first_batch(uint32_t *ticket, uint32_t base, uint32_t K)
{
  ticket_batch b;
  b.t = take_ticket(draw_batch(ticket, K), base);
  b.left = K;
  b.nb = draw_batch(ticket, K);
  return b;
}

/* after finishing ticket b.t: the next one of the batch, or the first of the
 * next batch (and draw the one after) */
__device__ __forceinline__ void next_in_batch(ticket_batch &b, uint32_t *ticket, uint32_t base, uint32_t K)
{
  if (--b.left) {
    b.t++;
    return;
  }
  b.t = take_ticket(b.nb, base);
  b.left = K;
  b.nb = draw_batch(ticket, K);
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

/* counter advance of a launch of nw waves over ntask tickets in batches of
 * K: every batch that starts below ntask, plus two held at exit by every wave
 * except the one that exits inside the batch crossing ntask */
static inline uint32_t kf_ticket_advance(uint32_t ntask, uint32_t nw, uint32_t K)
{
  const uint32_t batches = (ntask + K - 1) / K;
  return K * (batches + 2 * nw - (ntask % K != 0 ? 1u : 0u));
}

#endif
