/*
 * plan.c -- host-side schedule of the level-scheduled triangular solves.
 *
 * The reference runs its solves by level sets of rows (matrix/init.cpp:
 * 194-274).  The device schedule works on contiguous ranges of R rows (R a
 * power of two up to 16): rows inside a range are resolved in-wave by a
 * local level, ranges are grouped into tasks of up to 16 / R ranges of one
 * level of the range DAG, and tasks are handed out in level order.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kf_host.h"

#define KF_TASK_ROWS 16   /* rows per task: one quad of lanes per row of a wavefront */

/* Local level of every row inside its contiguous range of R rows: forward
 * ranges [R g, R g + R), backward ranges counted from the last row down.  A
 * row's level is 0 when no column of its lower (forward) or upper (backward)
 * part falls inside its own range, else one more than the deepest such
 * column; depth[g] = deepest level + 1. */
static void range_levels(const kf_prep *p, int upper, uint32_t R, uint8_t *lvl, uint8_t *depth)
{
  const size_t N = p->N;
  const size_t ng = (N + R - 1) / R;
  for (size_t g = 0; g < ng; g++) {
    size_t lo, hi;   /* rows [lo, hi] of range g */
    if (!upper) { lo = g * R; hi = lo + R - 1 < N - 1 ? lo + R - 1 : N - 1; }
    else { hi = N - 1 - g * R; lo = hi + 1 >= R ? hi + 1 - R : 0; }
    uint8_t dmax = 0;
    for (size_t t = 0; t <= hi - lo; t++) {
      const size_t r = upper ? hi - t : lo + t;
      const uint32_t j0 = upper ? p->diag[r] + 1 : p->rowptr[r], j1 = upper ? p->rowptr[r + 1] : p->diag[r];
      uint8_t l = 0;
      for (uint32_t j = j0; j < j1; j++) {
        const uint32_t cidx = p->col[j];
        if (cidx >= lo && cidx <= hi && lvl[cidx] + 1 > l) l = (uint8_t)(lvl[cidx] + 1);
      }
      lvl[r] = l;
      if (l + 1 > dmax) dmax = (uint8_t)(l + 1);
    }
    depth[g] = dmax;
  }
}

/* The ranges by level of the range DAG (a range's level is one more than
 * the deepest range it reads from; ties by range index), cut into tasks of
 * at most S = 16 / R ranges of one level.  Ranges of one level are
 * independent, so a task never waits on itself, and tickets in task order
 * follow a topological order: deadlock-free, and the wavefronts in flight
 * hold the ranges whose inputs complete next. */
void kf_plan_trsv(const kf_prep *p, int upper, uint32_t R, kf_trsv_plan *out)
{
  const size_t N = p->N;
  if (R == 0 || R > KF_TASK_ROWS || (R & (R - 1)) || KF_TASK_ROWS % R) {
    fprintf(stderr, "kf: triangular-solve range of %u rows (a power of two up to %d)\n", R, KF_TASK_ROWS);
    abort();
  }
  const size_t ng = (N + R - 1) / R;
  const uint32_t S = KF_TASK_ROWS / R;
  uint8_t *lvl = (uint8_t *)kf_xmalloc(N + 1), *depth = (uint8_t *)kf_xmalloc(ng + 1);
  range_levels(p, upper, R, lvl, depth);
  uint32_t *lev = (uint32_t *)kf_xcalloc(ng, sizeof(uint32_t));
  uint32_t maxlev = 0;
  for (size_t g = 0; g < ng; g++) {
    uint32_t l = 0;
    for (size_t t = 0; t < R; t++) {
      if (g * R + t >= N) break;
      const size_t r = upper ? N - 1 - (g * R + t) : g * R + t;
      const uint32_t j0 = upper ? p->diag[r] + 1 : p->rowptr[r], j1 = upper ? p->rowptr[r + 1] : p->diag[r];
      for (uint32_t j = j0; j < j1; j++) {
        const size_t h = upper ? (N - 1 - p->col[j]) / R : p->col[j] / R;
        if (h != g && lev[h] + 1 > l) l = lev[h] + 1;
      }
    }
    lev[g] = l;
    if (l > maxlev) maxlev = l;
  }
  /* counting sort of the ranges by level; begin[l] .. begin[l + 1] */
  size_t *begin = (size_t *)kf_xcalloc((size_t)maxlev + 2, sizeof(size_t));
  for (size_t g = 0; g < ng; g++) begin[lev[g] + 1]++;
  size_t nt = 0;
  for (uint32_t l = 0; l <= maxlev; l++) nt += (begin[l + 1] + S - 1) / S;
  for (uint32_t l = 0; l <= maxlev; l++) begin[l + 1] += begin[l];
  uint32_t *ord = (uint32_t *)kf_xmalloc(ng * sizeof(uint32_t));
  size_t *fill = (size_t *)kf_xmalloc(((size_t)maxlev + 1) * sizeof(size_t));
  memcpy(fill, begin, ((size_t)maxlev + 1) * sizeof(size_t));
  for (size_t g = 0; g < ng; g++) ord[fill[lev[g]]++] = (uint32_t)g;
  free(fill);
  uint32_t *task = (uint32_t *)kf_xmalloc(nt * S * sizeof(uint32_t));
  uint8_t *tdep = (uint8_t *)kf_xcalloc(nt + 1, 1);
  size_t k = 0;
  for (uint32_t l = 0; l <= maxlev; l++) {
    const size_t e = begin[l + 1];
    for (size_t b = begin[l]; b < e; b += S, k++)
      for (uint32_t i = 0; i < S; i++) {
        const uint32_t g = b + i < e ? ord[b + i] : ~0u;
        task[k * S + i] = g;
        if (g != ~0u && depth[g] > tdep[k]) tdep[k] = depth[g];
      }
  }
  if (k != nt) {
    fprintf(stderr, "kf: triangular-solve plan wrote %zu of %zu tasks\n", k, nt);
    abort();
  }
  free(ord);
  free(begin);
  free(lev);
  free(depth);
  out->R = R;
  out->S = S;
  out->ntask = (uint32_t)nt;
  out->lvl = lvl;
  out->task = task;
  out->tdep = tdep;
}

void kf_free_trsv_plan(kf_trsv_plan *t)
{
  free(t->lvl);
  free(t->task);
  free(t->tdep);
  memset(t, 0, sizeof(*t));
}

/* exported for tests (include/kfun3d.h) */
size_t kf_trsv_plan_host(const GEOMETRY *g, int upper, unsigned R, unsigned *task, unsigned char *tdep,
                         unsigned char *lvl, size_t cap)
{
  kf_prep p;
  kf_build_prep(g, &p);
  kf_trsv_plan t;
  kf_plan_trsv(&p, upper, R, &t);
  const size_t nt = t.ntask;
  if (task && tdep && lvl && cap >= nt) {
    memcpy(task, t.task, nt * t.S * sizeof(uint32_t));
    memcpy(tdep, t.tdep, nt);
    memcpy(lvl, t.lvl, p.N);
  }
  kf_free_trsv_plan(&t);
  kf_free_prep(&p);
  return nt;
}
