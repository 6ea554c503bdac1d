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
  /* per task and row slot q (range slot q / R, row q % R of that range):
   * the row, its block range and its local level; ~0u / 0 / 0 / 0xff for a
   * slot without a row */
  uint32_t *trow = (uint32_t *)kf_xmalloc(nt * KF_TASK_ROWS * 3 * sizeof(uint32_t));
  uint8_t *tlvl = (uint8_t *)kf_xmalloc(nt * KF_TASK_ROWS + 1);
  for (size_t tt = 0; tt < nt; tt++)
    for (uint32_t q = 0; q < KF_TASK_ROWS; q++) {
      const uint32_t g = task[tt * S + q / R], kr = q % R;
      const size_t o = tt * KF_TASK_ROWS + q;
      uint32_t r = ~0u, ja = 0, jb = 0;
      uint8_t l = 0xff;
      if (g != ~0u && (size_t)g * R + kr < N) {
        r = upper ? (uint32_t)(N - 1 - ((size_t)g * R + kr)) : g * R + kr;
        ja = upper ? p->diag[r] : p->rowptr[r];
        jb = upper ? p->rowptr[r + 1] : p->diag[r];
        l = lvl[r];
      }
      trow[3 * o] = r; trow[3 * o + 1] = ja; trow[3 * o + 2] = jb;
      tlvl[o] = l;
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
  out->trow = trow;
  out->tlvl = tlvl;
}

void kf_free_trsv_plan(kf_trsv_plan *t)
{
  free(t->lvl);
  free(t->task);
  free(t->tdep);
  free(t->trow);
  free(t->tlvl);
  memset(t, 0, sizeof(*t));
}

/* exported for tests (include/kfun3d.h) */
size_t kf_trsv_plan_host(const GEOMETRY *g, int upper, unsigned R, unsigned *task, unsigned char *tdep,
                         unsigned char *lvl, unsigned *rowtab, size_t cap)
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
  if (rowtab && cap >= nt) memcpy(rowtab, t.trow, nt * KF_TASK_ROWS * 3 * sizeof(uint32_t));
  kf_free_trsv_plan(&t);
  kf_free_prep(&p);
  return nt;
}

/* ILU(0) rows in ticket (level) order with everything a row needs to know
 * about the rows it reads from, so that a wavefront fetches it in one round:
 * per ticket the record {i, j0, L, nlow, mbeg, mcnt}; per lower position j
 * of row i (block (i, m), m = col[j]) the diagonal position of row m and the
 * start of its matches; per match the block (m, k) of row m's upper part and
 * the position of column k in row i (matrix/ilu.cpp:49-97 walks the same
 * pairs by merging the two rows). */
/* position of column k in the sorted row [j0, j1), or ~0u */
static uint32_t find_in_row(const uint32_t *col, uint32_t j0, uint32_t j1, uint32_t k)
{
  uint32_t lo = j0, hi = j1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (col[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return (lo < j1 && col[lo] == k) ? lo - j0 : ~0u;
}

/* update pairs of row i (written when mjm is not NULL); returns their count */
static uint32_t ilu_row_pairs(const kf_prep *p, uint32_t i, uint32_t *ldiag, uint32_t *moff, uint32_t *mjm,
                              uint16_t *mlo)
{
  const uint32_t j0 = p->rowptr[i], j1 = p->rowptr[i + 1], jd = p->diag[i];
  uint32_t n = 0;
  for (uint32_t j = j0; j < jd; j++) {
    const uint32_t m = p->col[j];
    if (mjm) { ldiag[j] = p->diag[m]; moff[j] = n; }
    for (uint32_t jm = p->diag[m] + 1; jm < p->rowptr[m + 1]; jm++) {
      const uint32_t lo = find_in_row(p->col, j0, j1, p->col[jm]);
      if (lo == ~0u) continue;
      if (mjm) { mjm[n] = jm; mlo[n] = (uint16_t)lo; }
      n++;
    }
  }
  return n;
}

/* ILU(0) rows in ticket (level) order with everything a row needs to know
 * about the rows it reads from, so that a wavefront fetches it in one round:
 * per ticket the record {i, j0, L, nlow, mbeg, mcnt}; per lower position j
 * of row i (block (i, m), m = col[j]) the diagonal position of row m and the
 * start of its pairs; per pair the block (m, k) of row m's upper part and
 * the position of column k in row i (matrix/ilu.cpp:49-97 walks the same
 * pairs by merging the two rows).  Two passes over the rows (count, then
 * fill at the prefix offsets), each in parallel. */
void kf_plan_ilu(const kf_prep *p, kf_ilu_plan *out)
{
  const size_t N = p->N;
  uint32_t *rec = (uint32_t *)kf_xcalloc(8 * N + 8, sizeof(uint32_t));
  uint32_t *ldiag = (uint32_t *)kf_xcalloc(p->nnzb + 1, sizeof(uint32_t));
  uint32_t *moff = (uint32_t *)kf_xcalloc(p->nnzb + 1, sizeof(uint32_t));
  uint32_t maxrow = 0, maxlow = 0, maxmatch = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : maxrow, maxlow, maxmatch)
  for (size_t t = 0; t < N; t++) {
    const uint32_t i = p->lrows[t];
    const uint32_t L = p->rowptr[i + 1] - p->rowptr[i], nlow = p->diag[i] - p->rowptr[i];
    const uint32_t mc = ilu_row_pairs(p, i, NULL, NULL, NULL, NULL);
    uint32_t *r = rec + 8 * t;
    r[0] = i; r[1] = p->rowptr[i]; r[2] = L; r[3] = nlow; r[5] = mc;
    if (L > maxrow) maxrow = L;
    if (nlow > maxlow) maxlow = nlow;
    if (mc > maxmatch) maxmatch = mc;
  }
  size_t nm = 0;
  for (size_t t = 0; t < N; t++) {
    if (nm > 0xffffffffu - rec[8 * t + 5]) { fprintf(stderr, "kf: ILU pair table exceeds 32-bit offsets\n"); abort(); }
    rec[8 * t + 4] = (uint32_t)nm;
    nm += rec[8 * t + 5];
  }
  uint32_t *mjm = (uint32_t *)kf_xmalloc((nm + 1) * sizeof(uint32_t));
  uint16_t *mlo = (uint16_t *)kf_xmalloc((nm + 1) * sizeof(uint16_t));
#pragma omp parallel for schedule(dynamic, 1024)
  for (size_t t = 0; t < N; t++) {
    const uint32_t *r = rec + 8 * t;
    ilu_row_pairs(p, r[0], ldiag, moff, mjm + r[4], mlo + r[4]);
  }
  out->rec = rec;
  out->ldiag = ldiag;
  out->moff = moff;
  out->mjm = mjm;
  out->mlo = mlo;
  out->nmatch = nm;
  out->maxrow = maxrow;
  out->maxlow = maxlow;
  out->maxmatch = maxmatch;
}

void kf_free_ilu_plan(kf_ilu_plan *t)
{
  free(t->rec); free(t->ldiag); free(t->moff); free(t->mjm); free(t->mlo);
  memset(t, 0, sizeof(*t));
}

/* exported for tests (include/kfun3d.h) */
size_t kf_ilu_plan_host(const GEOMETRY *g, unsigned *rec, unsigned *ldiag, unsigned *moff, unsigned *mjm,
                        unsigned short *mlo, size_t cap)
{
  kf_prep p;
  kf_build_prep(g, &p);
  kf_ilu_plan t;
  kf_plan_ilu(&p, &t);
  const size_t nm = t.nmatch;
  if (rec && ldiag && moff && mjm && mlo && cap >= nm) {
    memcpy(rec, t.rec, 8 * p.N * sizeof(uint32_t));
    memcpy(ldiag, t.ldiag, p.nnzb * sizeof(uint32_t));
    memcpy(moff, t.moff, p.nnzb * sizeof(uint32_t));
    memcpy(mjm, t.mjm, nm * sizeof(uint32_t));
    memcpy(mlo, t.mlo, nm * sizeof(uint16_t));
  }
  kf_free_ilu_plan(&t);
  kf_free_prep(&p);
  return nm;
}
