/*
 * aggregate.c -- the coarse level of the two-level preconditioner
 * (kf_options.coarse): aggregates of vertices, the pattern of the Galerkin
 * product P^T A P, and the index lists the device kernels gather with.
 *
 * P is piecewise constant: component c of fine vertex v maps to component c
 * of its aggregate.  So block (i, J) of A P is the sum of the fine blocks
 * (i, j) of row i with j in aggregate J, a coarse block (I, J) of P^T A P is
 * the sum of the blocks (i, J) of A P with i in aggregate I, restriction sums
 * an aggregate's members, and prolongation copies the aggregate's value to
 * each member.  The reference has no counterpart; the fine level is its
 * block ILU(0) (matrix/ilu.cpp), unchanged.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kf_host.h"

/* Greedy aggregation in vertex order over the BSR pattern's graph: an
 * unaggregated vertex with at least one unaggregated neighbour starts an
 * aggregate with all of them; one whose neighbours are all taken joins its
 * lowest neighbour's aggregate (an isolated vertex forms its own).  On a
 * Kuhn lattice in lexicographic order this gives the 2 x 2 x 2 cubes. */
static uint32_t aggregate(const kf_prep *p, uint32_t *agg)
{
  const size_t N = p->N;
  const uint32_t none = ~0u;
  for (size_t v = 0; v < N; v++) agg[v] = none;
  uint32_t nc = 0;
  for (size_t v = 0; v < N; v++) {
    if (agg[v] != none) continue;
    uint32_t first = none;
    int any_free = 0;
    for (uint32_t j = p->rowptr[v]; j < p->rowptr[v + 1]; j++) {
      const uint32_t u = p->col[j];
      if (u == v) continue;
      if (first == none) first = u;
      if (agg[u] == none) any_free = 1;
    }
    if (!any_free && first != none) {
      agg[v] = agg[first];
      continue;
    }
    agg[v] = nc;
    for (uint32_t j = p->rowptr[v]; j < p->rowptr[v + 1]; j++)
      if (agg[p->col[j]] == none) agg[p->col[j]] = nc;
    nc++;
  }
  return nc;
}

static int cmp_u32(const void *a, const void *b)
{
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

void kf_build_coarse(const kf_prep *p, kf_coarse_host *out)
{
  memset(out, 0, sizeof(*out));
  const size_t N = p->N, nnzb = p->nnzb;
  uint32_t *agg = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  const uint32_t nc = aggregate(p, agg);

  /* members of each aggregate, ascending */
  uint32_t *mptr = (uint32_t *)kf_xcalloc((size_t)nc + 1, sizeof(uint32_t));
  uint32_t *mlist = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  for (size_t v = 0; v < N; v++) mptr[agg[v] + 1]++;
  for (uint32_t I = 0; I < nc; I++) mptr[I + 1] += mptr[I];
  {
    uint32_t *fill = (uint32_t *)kf_xmalloc(((size_t)nc + 1) * sizeof(uint32_t));
    memcpy(fill, mptr, ((size_t)nc + 1) * sizeof(uint32_t));
    for (size_t v = 0; v < N; v++) mlist[fill[agg[v]]++] = (uint32_t)v;
    free(fill);
  }

  /* coarse pattern: the aggregates each coarse row's members reach, sorted */
  uint32_t *mark = (uint32_t *)kf_xmalloc(((size_t)nc + 1) * sizeof(uint32_t));
  for (uint32_t I = 0; I < nc; I++) mark[I] = ~0u;
  uint32_t *crp = (uint32_t *)kf_xcalloc((size_t)nc + 1, sizeof(uint32_t));
  size_t cap = nnzb + 1, ncol = 0;
  uint32_t *ccol = (uint32_t *)kf_xmalloc(cap * sizeof(uint32_t));
  for (uint32_t I = 0; I < nc; I++) {
    const size_t start = ncol;
    for (uint32_t m = mptr[I]; m < mptr[I + 1]; m++) {
      const uint32_t i = mlist[m];
      for (uint32_t j = p->rowptr[i]; j < p->rowptr[i + 1]; j++) {
        const uint32_t J = agg[p->col[j]];
        if (mark[J] == I) continue;
        mark[J] = I;
        ccol[ncol++] = J;   /* at most one per fine block: ncol <= nnzb */
      }
    }
    qsort(ccol + start, ncol - start, sizeof(uint32_t), cmp_u32);
    crp[I + 1] = (uint32_t)ncol;
  }
  const size_t cnnzb = ncol;
  uint32_t *cdiag = (uint32_t *)kf_xmalloc(((size_t)nc + 1) * sizeof(uint32_t));
  for (uint32_t I = 0; I < nc; I++) {
    cdiag[I] = ~0u;
    for (uint32_t j = crp[I]; j < crp[I + 1]; j++)
      if (ccol[j] == I) cdiag[I] = j;
    if (cdiag[I] == ~0u) {
      fprintf(stderr, "kf: coarse row %u has no diagonal block\n", I);
      abort();
    }
  }

  /* A P: per fine row, the aggregates its columns fall in, ascending; each
   * fine block's slot there; the fine blocks of each A P block ascending */
  uint32_t *aprp = (uint32_t *)kf_xcalloc(N + 1, sizeof(uint32_t));
  uint32_t *apcol = (uint32_t *)kf_xmalloc((nnzb + 1) * sizeof(uint32_t));
  uint32_t *apos = (uint32_t *)kf_xmalloc((nnzb + 1) * sizeof(uint32_t));
  size_t nap = 0;
  for (size_t i = 0; i < N; i++) {
    const size_t start = nap;
    for (uint32_t j = p->rowptr[i]; j < p->rowptr[i + 1]; j++) {
      const uint32_t J = agg[p->col[j]];
      size_t k = start;
      while (k < nap && apcol[k] != J) k++;   /* a row holds few aggregates */
      if (k == nap) apcol[nap++] = J;
    }
    qsort(apcol + start, nap - start, sizeof(uint32_t), cmp_u32);
    aprp[i + 1] = (uint32_t)nap;
    for (uint32_t j = p->rowptr[i]; j < p->rowptr[i + 1]; j++) {
      const uint32_t J = agg[p->col[j]];
      size_t k = start;
      while (apcol[k] != J) k++;
      apos[j] = (uint32_t)k;
    }
  }
  uint32_t *apgptr = (uint32_t *)kf_xcalloc(nap + 1, sizeof(uint32_t));
  uint32_t *apglist = (uint32_t *)kf_xmalloc((nnzb + 1) * sizeof(uint32_t));
  for (size_t j = 0; j < nnzb; j++) apgptr[apos[j] + 1]++;
  for (size_t b = 0; b < nap; b++) apgptr[b + 1] += apgptr[b];
  {
    uint32_t *fill = (uint32_t *)kf_xmalloc((nap + 1) * sizeof(uint32_t));
    memcpy(fill, apgptr, (nap + 1) * sizeof(uint32_t));
    for (size_t j = 0; j < nnzb; j++) apglist[fill[apos[j]]++] = (uint32_t)j;
    free(fill);
  }
  free(apos);

  /* the A P blocks of each coarse block: block (i, J) of A P goes to
   * (agg[i], J); ascending A P position, i.e. ascending member */
  uint32_t *gpos = (uint32_t *)kf_xmalloc((nap + 1) * sizeof(uint32_t));
  for (size_t i = 0; i < N; i++) {
    const uint32_t I = agg[i];
    for (uint32_t k = aprp[i]; k < aprp[i + 1]; k++) {
      const uint32_t J = apcol[k];
      uint32_t lo = crp[I], hi = crp[I + 1];
      while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (ccol[mid] < J) lo = mid + 1;
        else hi = mid;
      }
      gpos[k] = lo;
    }
  }
  uint32_t *gptr = (uint32_t *)kf_xcalloc(cnnzb + 1, sizeof(uint32_t));
  uint32_t *glist = (uint32_t *)kf_xmalloc((nap + 1) * sizeof(uint32_t));
  for (size_t k = 0; k < nap; k++) gptr[gpos[k] + 1]++;
  for (size_t b = 0; b < cnnzb; b++) gptr[b + 1] += gptr[b];
  {
    uint32_t *fill = (uint32_t *)kf_xmalloc((cnnzb + 1) * sizeof(uint32_t));
    memcpy(fill, gptr, (cnnzb + 1) * sizeof(uint32_t));
    for (size_t k = 0; k < nap; k++) glist[fill[gpos[k]]++] = (uint32_t)k;
    free(fill);
  }
  free(gpos);
  free(mark);

  kf_prep *c = &out->prep;
  c->N = nc;
  c->nnzb = cnnzb;
  c->rowptr = crp;
  c->col = ccol;
  c->diag = cdiag;
  kf_prep_levels(c);
  out->agg = agg;
  out->mptr = mptr;
  out->mlist = mlist;
  out->nap = nap;
  out->aprp = aprp;
  out->apcol = apcol;
  out->apgptr = apgptr;
  out->apglist = apglist;
  out->gptr = gptr;
  out->glist = glist;
}

void kf_free_coarse(kf_coarse_host *h)
{
  kf_free_prep(&h->prep);
  free(h->agg); free(h->mptr); free(h->mlist);
  free(h->aprp); free(h->apcol); free(h->apgptr); free(h->apglist);
  free(h->gptr); free(h->glist);
  memset(h, 0, sizeof(*h));
}

/* exported for tests (include/kfun3d.h) */
size_t kf_coarse_plan_host(const GEOMETRY *g, unsigned *agg, size_t *cnnzb, unsigned *crowptr, unsigned *ccol)
{
  kf_prep p;
  kf_build_prep(g, &p);
  kf_coarse_host h;
  kf_build_coarse(&p, &h);
  const size_t nc = h.prep.N;
  if (cnnzb) *cnnzb = h.prep.nnzb;
  if (agg) memcpy(agg, h.agg, p.N * sizeof(uint32_t));
  if (crowptr) memcpy(crowptr, h.prep.rowptr, (nc + 1) * sizeof(uint32_t));
  if (ccol) memcpy(ccol, h.prep.col, h.prep.nnzb * sizeof(uint32_t));
  kf_free_coarse(&h);
  kf_free_prep(&p);
  return nc;
}
