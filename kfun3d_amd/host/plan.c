/*
 * plan.c -- host-side schedule of the level-scheduled triangular solves and
 * of the ILU(0) factorisation.
 *
 * The reference runs its solves by level sets of rows (matrix/init.cpp:
 * 194-274).  The device schedule works on contiguous ranges of R rows (R a
 * power of two up to 16), clustered into tasks of up to 16 / R ranges that
 * may depend on one another: rows of a task are resolved in-wave by a local
 * level, tasks are handed out in the order of the task DAG's levels.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kf_host.h"

#define KF_TASK_ROWS 16   /* rows per task: one quad of lanes per row of a wavefront */

/* row t of range g in sweep order (forward: ascending from g R; backward:
 * descending from N - 1 - g R), or ~0u past the end */
static inline uint32_t range_row(size_t N, int upper, uint32_t R, uint32_t g, uint32_t t)
{
  const size_t k = (size_t)g * R + t;
  if (k >= N) return ~0u;
  return upper ? (uint32_t)(N - 1 - k) : (uint32_t)k;
}

static inline uint32_t row_range(size_t N, int upper, uint32_t R, uint32_t r)
{
  return upper ? (uint32_t)((N - 1 - r) / R) : r / R;
}

/* the sweep's part of row r: forward the blocks left of the diagonal,
 * backward those right of it */
static inline void sweep_part(const kf_prep *p, int upper, uint32_t r, uint32_t *j0, uint32_t *j1)
{
  *j0 = upper ? p->diag[r] + 1 : p->rowptr[r];
  *j1 = upper ? p->rowptr[r + 1] : p->diag[r];
}

/* Local levels of the rows of range c if it joins task t: a row reads
 * members' rows (levels final in lvl) and earlier rows of c (levels in tl,
 * by position in sweep order).  Returns the task's depth with c in it. */
static uint32_t join_levels(const kf_prep *p, int upper, uint32_t R, uint32_t c, int32_t t, const int32_t *rtask,
                            const uint8_t *lvl, uint8_t tl[KF_TASK_ROWS], uint32_t depth)
{
  const size_t N = p->N;
  for (uint32_t q = 0; q < R; q++) {
    const uint32_t r = range_row(N, upper, R, c, q);
    if (r == ~0u) break;
    uint32_t j0, j1, l = 0;
    sweep_part(p, upper, r, &j0, &j1);
    for (uint32_t j = j0; j < j1; j++) {
      const uint32_t cc = p->col[j];
      uint32_t lc = ~0u;
      if (rtask[cc] == t) lc = lvl[cc];
      else if (row_range(N, upper, R, cc) == c)
        lc = tl[upper ? (uint32_t)((N - 1 - cc) - (size_t)c * R) : cc - c * R];
      if (lc != ~0u && lc + 1 > l) l = lc + 1;
    }
    tl[q] = (uint8_t)l;
    if (l + 1 > depth) depth = l + 1;
  }
  return depth;
}

/* The ranges clustered into tasks, greedily in the order of the range
 * DAG's levels.  A task starts from the lowest-level unassigned range (its
 * seed); its level L is one more than the deepest task the seed reads
 * from.  It then takes, one at a time, a range that reads from a member and
 * otherwise only from tasks of level < L -- of those the one that keeps the
 * task's in-wave depth (longest chain of its rows) smallest, the lowest
 * index on ties -- until it holds 16 / R ranges.  Every task reads only
 * from itself and from tasks of lower level, so tickets in level order
 * follow a topological order (no wait can deadlock); on a lattice the
 * tasks become 2 x 2 blocks of x-ranges across y and z, which shortens the
 * chain of levels from nx/R + ny + nz to about nx/R + ny/2 + nz/2.
 * With cluster = 0 a task takes instead the next ranges of the seed's level
 * of the range DAG in index order (ranges of one level are independent). */
void kf_plan_trsv(const kf_prep *p, int upper, uint32_t R, int cluster, kf_trsv_plan *out)
{
  const size_t N = p->N;
  if (R == 0 || R > KF_TASK_ROWS || (R & (R - 1)) || KF_TASK_ROWS % R) {
    fprintf(stderr, "kf: triangular-solve range of %u rows (a power of two up to %d)\n", R, KF_TASK_ROWS);
    abort();
  }
  const uint32_t ng = (uint32_t)((N + R - 1) / R);
  const uint32_t S = KF_TASK_ROWS / R;
  /* range DAG: deps (ranges a range reads from) and kids, as CSR */
  uint32_t *dptr = (uint32_t *)kf_xcalloc((size_t)ng + 1, sizeof(uint32_t));
  uint32_t *stamp = (uint32_t *)kf_xmalloc((size_t)ng * sizeof(uint32_t));
  for (uint32_t g = 0; g < ng; g++) stamp[g] = ~0u;
  size_t nd = 0;
  for (int pass = 0; pass < 2; pass++) {
    uint32_t *dep = pass ? (uint32_t *)kf_xmalloc((nd ? nd : 1) * sizeof(uint32_t)) : NULL;
    size_t k = 0;
    for (uint32_t g = 0; g < ng; g++) {
      if (pass) dptr[g] = (uint32_t)k;
      for (uint32_t t = 0; t < R; t++) {
        const uint32_t r = range_row(N, upper, R, g, t);
        if (r == ~0u) break;
        uint32_t j0, j1;
        sweep_part(p, upper, r, &j0, &j1);
        for (uint32_t j = j0; j < j1; j++) {
          const uint32_t h = row_range(N, upper, R, p->col[j]);
          if (h == g || stamp[h] == 2 * g + pass) continue;
          stamp[h] = 2 * g + pass;
          if (pass) dep[k] = h;
          k++;
        }
      }
    }
    if (!pass) {
      nd = k;
    } else {
      dptr[ng] = (uint32_t)k;
      out->task = dep; /* parked; moved below */
    }
  }
  uint32_t *dep = out->task;
  out->task = NULL;
  uint32_t *kptr = (uint32_t *)kf_xcalloc((size_t)ng + 1, sizeof(uint32_t));
  for (size_t k = 0; k < nd; k++) kptr[dep[k] + 1]++;
  for (uint32_t g = 0; g < ng; g++) kptr[g + 1] += kptr[g];
  uint32_t *kid = (uint32_t *)kf_xmalloc((nd ? nd : 1) * sizeof(uint32_t));
  {
    uint32_t *fill = (uint32_t *)kf_xmalloc(((size_t)ng + 1) * sizeof(uint32_t));
    memcpy(fill, kptr, ((size_t)ng + 1) * sizeof(uint32_t));
    for (uint32_t g = 0; g < ng; g++)
      for (uint32_t k = dptr[g]; k < dptr[g + 1]; k++) kid[fill[dep[k]]++] = g;
    free(fill);
  }
  /* range levels (deps of a range have smaller indices in sweep order) and
   * the seed order: by level, then index */
  uint32_t *lev = (uint32_t *)kf_xcalloc(ng, sizeof(uint32_t));
  uint32_t maxlev = 0;
  for (uint32_t g = 0; g < ng; g++) {
    uint32_t l = 0;
    for (uint32_t k = dptr[g]; k < dptr[g + 1]; k++)
      if (lev[dep[k]] + 1 > l) l = lev[dep[k]] + 1;
    lev[g] = l;
    if (l > maxlev) maxlev = l;
  }
  uint32_t *begin = (uint32_t *)kf_xcalloc((size_t)maxlev + 2, sizeof(uint32_t));
  for (uint32_t g = 0; g < ng; g++) begin[lev[g] + 1]++;
  for (uint32_t l = 0; l <= maxlev; l++) begin[l + 1] += begin[l];
  uint32_t *ord = (uint32_t *)kf_xmalloc(((size_t)ng + 1) * sizeof(uint32_t));
  {
    uint32_t *fill = (uint32_t *)kf_xmalloc(((size_t)maxlev + 1) * sizeof(uint32_t));
    memcpy(fill, begin, ((size_t)maxlev + 1) * sizeof(uint32_t));
    for (uint32_t g = 0; g < ng; g++) ord[fill[lev[g]]++] = g;
    free(fill);
  }
  /* greedy clustering */
  int32_t *asg = (int32_t *)kf_xmalloc((size_t)ng * sizeof(int32_t));
  for (uint32_t g = 0; g < ng; g++) asg[g] = -1;
  uint32_t *mem = (uint32_t *)kf_xmalloc(((size_t)ng * S + 1) * sizeof(uint32_t)); /* members, S per task */
  uint32_t *tlev = (uint32_t *)kf_xmalloc(((size_t)ng + 1) * sizeof(uint32_t));
  uint8_t *tdep0 = (uint8_t *)kf_xcalloc((size_t)ng + 1, 1);
  uint8_t *lvl = (uint8_t *)kf_xcalloc(N + 1, 1);          /* local level of each row in its task */
  int32_t *rtask = (int32_t *)kf_xmalloc((N + 1) * sizeof(int32_t));
  for (size_t r = 0; r < N; r++) rtask[r] = -1;
  for (uint32_t g = 0; g < ng; g++) stamp[g] = ~0u;
  uint32_t nt = 0, epoch = 0, cand[256];
  uint8_t tl[KF_TASK_ROWS], best_l[KF_TASK_ROWS];
  for (uint32_t oi = 0; oi < ng; oi++) {
    const uint32_t g0 = ord[oi];
    if (asg[g0] >= 0) continue;
    const uint32_t t = nt++;
    uint32_t L = 0;
    for (uint32_t k = dptr[g0]; k < dptr[g0 + 1]; k++) {
      const int32_t a = asg[dep[k]];
      if (a < 0) {
        fprintf(stderr, "kf: triangular-solve plan: range %u read before it is scheduled\n", dep[k]);
        abort();
      }
      if (tlev[a] + 1 > L) L = tlev[a] + 1;
    }
    tlev[t] = L;
    uint32_t nm = 0, depth = 0, nx = oi + 1;
    for (;;) {
      uint32_t pick = ~0u, pdepth = ~0u;
      if (nm == 0) {
        pick = g0;
        pdepth = join_levels(p, upper, R, g0, (int32_t)t, rtask, lvl, best_l, 0);
      } else if (!cluster) {
        while (nx < ng && asg[ord[nx]] >= 0) nx++;
        if (nx == ng || lev[ord[nx]] != lev[g0]) break;
        pick = ord[nx];
        pdepth = join_levels(p, upper, R, pick, (int32_t)t, rtask, lvl, best_l, depth);
      } else {
        /* candidates: unassigned ranges that read from a member */
        uint32_t nc = 0;
        epoch++;
        for (uint32_t m = 0; m < nm; m++) {
          const uint32_t gm = mem[(size_t)t * S + m];
          for (uint32_t k = kptr[gm]; k < kptr[gm + 1]; k++) {
            const uint32_t c = kid[k];
            if (asg[c] >= 0 || stamp[c] == epoch || nc == 256) continue;
            stamp[c] = epoch;
            cand[nc++] = c;
          }
        }
        for (uint32_t ci = 0; ci < nc; ci++) {
          const uint32_t c = cand[ci];
          int ok = 1;
          for (uint32_t k = dptr[c]; k < dptr[c + 1] && ok; k++) {
            const int32_t a = asg[dep[k]];
            ok = a == (int32_t)t || (a >= 0 && tlev[a] < L);
          }
          if (!ok) continue;
          const uint32_t dmax = join_levels(p, upper, R, c, (int32_t)t, rtask, lvl, tl, depth);
          if (dmax < pdepth || (dmax == pdepth && c < pick)) {
            pick = c;
            pdepth = dmax;
            memcpy(best_l, tl, sizeof(tl));
          }
        }
        if (pick == ~0u) break;
      }
      /* take it */
      asg[pick] = (int32_t)t;
      mem[(size_t)t * S + nm++] = pick;
      for (uint32_t q = 0; q < R; q++) {
        const uint32_t r = range_row(N, upper, R, pick, q);
        if (r == ~0u) break;
        rtask[r] = (int32_t)t;
        lvl[r] = best_l[q];
      }
      depth = pdepth;
      if (nm == S) break;
    }
    for (uint32_t m = nm; m < S; m++) mem[(size_t)t * S + m] = ~0u;
    tdep0[t] = (uint8_t)depth;
  }
  /* tickets: tasks by level, then by seed order */
  uint32_t tmax = 0;
  for (uint32_t t = 0; t < nt; t++)
    if (tlev[t] > tmax) tmax = tlev[t];
  uint32_t *tb = (uint32_t *)kf_xcalloc((size_t)tmax + 2, sizeof(uint32_t));
  for (uint32_t t = 0; t < nt; t++) tb[tlev[t] + 1]++;
  for (uint32_t l = 0; l <= tmax; l++) tb[l + 1] += tb[l];
  uint32_t *tord = (uint32_t *)kf_xmalloc(((size_t)nt + 1) * sizeof(uint32_t)); /* ticket -> task */
  uint32_t *tpos = (uint32_t *)kf_xmalloc(((size_t)nt + 1) * sizeof(uint32_t)); /* task -> ticket */
  for (uint32_t t = 0; t < nt; t++) {
    tpos[t] = tb[tlev[t]]++;
    tord[tpos[t]] = t;
  }
  uint32_t *task = (uint32_t *)kf_xmalloc(((size_t)nt * S + 1) * sizeof(uint32_t));
  uint8_t *tdep = (uint8_t *)kf_xcalloc((size_t)nt + 1, 1);
  for (uint32_t k = 0; k < nt; k++) {
    memcpy(task + (size_t)k * S, mem + (size_t)tord[k] * S, S * sizeof(uint32_t));
    tdep[k] = tdep0[tord[k]];
  }
  /* per task (ticket order) and row slot q (member q / R, row q % R of it):
   * the row, its block range and its local level; ~0u / 0 / 0 / 0xff for a
   * slot without a row.  Per BSR position of a row's sweep part: the slot
   * of the column's row when it is in the row's task, else 0xff. */
  uint32_t *trow = (uint32_t *)kf_xmalloc(((size_t)nt * KF_TASK_ROWS * 3 + 1) * sizeof(uint32_t));
  uint8_t *tlvl = (uint8_t *)kf_xmalloc((size_t)nt * KF_TASK_ROWS + 1);
  uint8_t *eslot = (uint8_t *)kf_xmalloc(p->nnzb + 1);
  memset(eslot, 0xff, p->nnzb + 1);
  uint8_t *rslot = (uint8_t *)kf_xmalloc(N + 1);
  for (uint32_t k = 0; k < nt; k++)
    for (uint32_t q = 0; q < KF_TASK_ROWS; q++) {
      const uint32_t g = task[(size_t)k * S + q / R];
      const uint32_t r = g == ~0u ? ~0u : range_row(N, upper, R, g, q % R);
      if (r != ~0u) rslot[r] = (uint8_t)q;
    }
  for (uint32_t k = 0; k < nt; k++)
    for (uint32_t q = 0; q < KF_TASK_ROWS; q++) {
      const uint32_t g = task[(size_t)k * S + q / R];
      const uint32_t r = g == ~0u ? ~0u : range_row(N, upper, R, g, q % R);
      const size_t o = (size_t)k * KF_TASK_ROWS + q;
      uint32_t ja = 0, jb = 0;
      uint8_t l = 0xff;
      if (r != ~0u) {
        ja = upper ? p->diag[r] : p->rowptr[r];
        jb = upper ? p->rowptr[r + 1] : p->diag[r];
        l = lvl[r];
        uint32_t j0, j1;
        sweep_part(p, upper, r, &j0, &j1);
        for (uint32_t j = j0; j < j1; j++)
          if (rtask[p->col[j]] == rtask[r]) eslot[j] = rslot[p->col[j]];
      }
      trow[3 * o] = r; trow[3 * o + 1] = ja; trow[3 * o + 2] = jb;
      tlvl[o] = l;
    }
  free(rslot);
  free(tb); free(tord); free(tpos); free(tdep0); free(tlev); free(mem); free(asg); free(rtask);
  free(ord); free(begin); free(lev); free(kid); free(kptr); free(dep); free(dptr); free(stamp);
  out->R = R;
  out->S = S;
  out->ntask = nt;
  out->nlevel = nt ? tmax + 1 : 0;
  out->lvl = lvl;
  out->task = task;
  out->tdep = tdep;
  out->trow = trow;
  out->tlvl = tlvl;
  out->eslot = eslot;
}

/* the forward and the backward schedule, built side by side */
void kf_plan_trsv_both(const kf_prep *p, uint32_t R, int cluster, kf_trsv_plan *lower, kf_trsv_plan *upper)
{
#pragma omp parallel sections num_threads(2)
  {
#pragma omp section
    kf_plan_trsv(p, 0, R, cluster, lower);
#pragma omp section
    kf_plan_trsv(p, 1, R, cluster, upper);
  }
}

void kf_free_trsv_plan(kf_trsv_plan *t)
{
  free(t->lvl);
  free(t->task);
  free(t->tdep);
  free(t->trow);
  free(t->tlvl);
  free(t->eslot);
  memset(t, 0, sizeof(*t));
}

/* exported for tests (include/kfun3d.h) */
size_t kf_trsv_plan_host(const GEOMETRY *g, int upper, unsigned R, int cluster, unsigned *task, unsigned char *tdep,
                         unsigned char *lvl, unsigned *rowtab, unsigned char *eslot, size_t cap)
{
  kf_prep p;
  kf_build_prep(g, &p);
  kf_trsv_plan t;
  kf_plan_trsv(&p, upper, R, cluster, &t);
  const size_t nt = t.ntask;
  if (task && tdep && lvl && cap >= nt) {
    memcpy(task, t.task, nt * t.S * sizeof(uint32_t));
    memcpy(tdep, t.tdep, nt);
    memcpy(lvl, t.lvl, p.N);
  }
  if (rowtab && cap >= nt) memcpy(rowtab, t.trow, nt * KF_TASK_ROWS * 3 * sizeof(uint32_t));
  if (eslot && cap >= nt) memcpy(eslot, t.eslot, p.nnzb);
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
