/*
 * setup.c -- host preprocessing of the mesh, restated from the reference's
 * setup path (src/mesh.c:113-153): one owner part (the GPU replaces the
 * OpenMP owner-computes partition of src/subdomains.c, whose one-thread
 * result is part = 0 and ie = {0, E}), solid-face colouring, LSQ weights,
 * and the derived arrays the device context is built from.
 *
 * The floating-point expressions of kf_lsq_weights keep the reference's
 * operation order exactly (this file is compiled with -ffp-contract=off):
 * the LSQ terms feed every residual and are compared bit for bit with the
 * reference's own setup.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kf_host.h"

/* ---- src/index.c:8-112: merge sort with exchange-sorted leaves of <= 500 -- */

#define KF_SORT_LEAF 500

static void leaf_sort(const uint32_t *key, uint32_t *l, uint32_t *h)
{
  for (uint32_t *i = l; i < h; i++) {
    uint32_t ki = key[*i];
    for (uint32_t *j = i + 1; j < h; j++)
      if (ki > key[*j]) {
        uint32_t t = *i; *i = *j; *j = t;
        ki = key[*i];
      }
  }
}

static void merge_runs(const uint32_t *key, const uint32_t *l1, const uint32_t *h1, const uint32_t *l2,
                       const uint32_t *h2, uint32_t *out)
{
  while (l1 != h1 && l2 != h2) {
    if (key[*l2] < key[*l1]) *out++ = *l2++;
    else *out++ = *l1++;
  }
  while (l1 != h1) *out++ = *l1++;
  while (l2 != h2) *out++ = *l2++;
}

/* sorts [l,h); the result lands in l when in_place, else in tmp */
static void index_sort_rec(const uint32_t *key, uint32_t *l, uint32_t *h, uint32_t *tmp, int in_place)
{
  if (h - l <= KF_SORT_LEAF) {
    leaf_sort(key, l, h);
    if (!in_place) memcpy(tmp, l, (size_t)(h - l) * sizeof(uint32_t));
    return;
  }
  uint32_t *m = l + (h - l) / 2;
  uint32_t *mt = tmp + (m - l), *ht = tmp + (h - l);
  /* the two halves are independent: large ones run as tasks (the merge
   * tree, and so the result, is that of the sequential recursion) */
#pragma omp task if (h - l > (1 << 16))
  index_sort_rec(key, l, m, tmp, !in_place);
#pragma omp task if (h - l > (1 << 16))
  index_sort_rec(key, m, h, mt, !in_place);
#pragma omp taskwait
  if (in_place) merge_runs(key, tmp, mt, mt, ht, l);
  else merge_runs(key, l, m, m, h, tmp);
}

void kf_index_sort(size_t sz, const uint32_t *key, uint32_t *perm)
{
  int sorted = 1;
#pragma omp parallel for schedule(static) reduction(&& : sorted)
  for (size_t i = 0; i < sz; i++) {
    perm[i] = (uint32_t)i;
    if (i && key[i - 1] > key[i]) sorted = 0;
  }
  /* on keys already in order the sort is the identity: the leaf exchange
   * sort swaps only on a strictly greater key, the merge takes from the
   * right run only on a strictly smaller one */
  if (sz == 0 || sorted) return;
  uint32_t *tmp = (uint32_t *)kf_xmalloc(sz * sizeof(uint32_t));
#pragma omp parallel
#pragma omp single
  index_sort_rec(key, perm, perm + sz, tmp, 1);
  free(tmp);
}

/* ---- src/sbface_coloring.c:11-93: greedy node-disjoint colour classes ---- */

void kf_color_solid_faces(GEOMETRY *g)
{
  const size_t N = g->n->sz, S = g->b->fc->sz;
  uint32_t *n0 = g->b->fc->fptr->n0, *n1 = g->b->fc->fptr->n1, *n2 = g->b->fc->fptr->n2;
  uint32_t *per_color = (uint32_t *)kf_xcalloc(S + 1, sizeof(uint32_t));
  uint8_t *used = (uint8_t *)kf_xmalloc(N + 1);
  uint32_t k = 0, c = 0;
  for (uint32_t i = 0; i < S;) {
    memset(used, 0, N + 1);
    for (uint32_t j = i; j < S; j++) {
      if (used[n0[j]] | used[n1[j]] | used[n2[j]]) continue;
      used[n0[j]] = used[n1[j]] = used[n2[j]] = 1;
      uint32_t a = n0[j], b = n1[j], d = n2[j];
      n0[j] = n0[k]; n1[j] = n1[k]; n2[j] = n2[k];
      n0[k] = a; n1[k] = b; n2[k] = d;
      k++;
      per_color[c]++;
    }
    c++;
    i = k;
  }
  free(used);
  uint32_t *ic = (uint32_t *)kf_xmalloc((c + 1) * sizeof(uint32_t));
  uint32_t sum = 0;
  for (uint32_t i = 0; i < c; i++) { ic[i] = sum; sum += per_color[i]; }
  ic[c] = sum;
  free(per_color);
  g->t->i = ic;
  g->t->sz = c;
}

/* vertex -> incident edges in ascending edge id, ent = edge << 1 | side
 * (side 1: the vertex is n1): counts and slots drawn with atomics, then
 * each vertex's short list sorted, so the result is deterministic */
void kf_incidence(size_t N, size_t E, const uint32_t *n0, const uint32_t *n1, uint32_t **vptr, uint32_t **vent)
{
  uint32_t *pt = (uint32_t *)kf_xcalloc(N + 1, sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < E; i++) {
#pragma omp atomic
    pt[n0[i] + 1]++;
#pragma omp atomic
    pt[n1[i] + 1]++;
  }
  for (size_t i = 0; i < N; i++) pt[i + 1] += pt[i];
  uint32_t *fillp = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  memcpy(fillp, pt, (N + 1) * sizeof(uint32_t));
  uint32_t *en = (uint32_t *)kf_xmalloc((2 * E + 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < E; i++) {
    uint32_t s0, s1;
#pragma omp atomic capture
    s0 = fillp[n0[i]]++;
#pragma omp atomic capture
    s1 = fillp[n1[i]]++;
    en[s0] = (uint32_t)(i << 1);
    en[s1] = (uint32_t)(i << 1) | 1u;
  }
  free(fillp);
#pragma omp parallel for schedule(static)
  for (size_t v = 0; v < N; v++)
    for (uint32_t s = pt[v] + 1; s < pt[v + 1]; s++) {   /* insertion sort: ~14 entries */
      const uint32_t x = en[s];
      uint32_t t = s;
      for (; t > pt[v] && en[t - 1] > x; t--) en[t] = en[t - 1];
      en[t] = x;
    }
  *vptr = pt;
  *vent = en;
}

/* ---- src/weights.c:9-398, one owner: every edge updates both ends -------- */

/* The reference scatters each edge's terms to both ends, edge by edge, in
 * three passes (src/weights.c:103-297), each pass reading only the end's own
 * sums of the earlier passes.  So a vertex's sums see its incident edges in
 * ascending edge id, pass after pass, whatever the other vertices do: here
 * each vertex gathers them in that order (the same operations, the same
 * bits), all vertices in parallel. */
void kf_lsq_weights(GEOMETRY *g)
{
  const size_t N = g->n->sz, E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  const double *x0 = g->n->xyz->x0, *x1 = g->n->xyz->x1, *x2 = g->n->xyz->x2;
  double *w = (double *)kf_xcalloc(7 * N, sizeof(double));
  uint32_t *vptr, *vent;
  kf_incidence(N, E, n0, n1, &vptr, &vent);

#pragma omp parallel for schedule(static)
  for (size_t v = 0; v < N; v++) {
    double *wv = w + v * 7;
    /* w00, w01, w02: src/weights.c:237-297 */
    for (uint32_t s = vptr[v]; s < vptr[v + 1]; s++) {
      const uint32_t i = vent[s] >> 1, a = n0[i], b = n1[i];
      if (!(vent[s] & 1u)) {
        const double rx = x0[b] - x0[a], ry = x1[b] - x1[a], rz = x2[b] - x2[a];
        wv[0] += rx * rx; wv[1] += rx * ry; wv[2] += rx * rz;
      } else {
        const double rx = x0[a] - x0[b], ry = x1[a] - x1[b], rz = x2[a] - x2[b];
        wv[0] += rx * rx; wv[1] += rx * ry; wv[2] += rx * rz;
      }
    }
    wv[0] = sqrt(wv[0]);
    wv[1] /= wv[0];
    wv[2] /= wv[0];
    /* w11, w12: src/weights.c:173-232 */
    for (uint32_t s = vptr[v]; s < vptr[v + 1]; s++) {
      const uint32_t i = vent[s] >> 1, a = n0[i], b = n1[i];
      const double dx = x0[b] - x0[a], dy = x1[b] - x1[a], dz = x2[b] - x2[a];
      const double d = wv[1] / wv[0];
      if (!(vent[s] & 1u)) {
        const double d_ = dy - dx * d;
        wv[3] += d_ * d_;
        wv[4] += d_ * dz;
      } else {
        const double d_ = -dy + dx * d;
        wv[3] += d_ * d_;
        wv[4] -= d_ * dz;
      }
    }
    wv[3] = sqrt(wv[3]);
    wv[4] /= wv[3];
    /* w22: src/weights.c:103-170 */
    for (uint32_t s = vptr[v]; s < vptr[v + 1]; s++) {
      const uint32_t i = vent[s] >> 1, a = n0[i], b = n1[i];
      const double dx = x0[b] - x0[a], dy = x1[b] - x1[a], dz = x2[b] - x2[a];
      const double d0 = wv[1] / wv[0];
      const double d1 = wv[2] / wv[0];
      const double d2 = wv[4] / wv[3];
      if (!(vent[s] & 1u)) {
        const double d0_ = dy - dx * d0;
        const double d1_ = dz - dx * d1;
        const double d2_ = d1_ - d2 * d0_;
        wv[5] += d2_ * d2_;
      } else {
        const double d0_ = -dy + dx * d0;
        const double d1_ = -dz + dx * d1;
        const double d2_ = d1_ - d2 * d0_;
        wv[5] += d2_ * d2_;
      }
    }
  }
  free(vptr);
  free(vent);
  /* magnitudes: src/weights.c:313-366 */
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; i++) {
    double *v = w + i * 7;
    v[5] = sqrt(v[5]);
    const double sw00 = v[0] * v[0], sw11 = v[3] * v[3], sw22 = v[5] * v[5];
    const double w00 = 1.0 / sw00, w11 = 1.0 / sw11, w22 = 1.0 / sw22;
    const double w01 = v[1] / v[0], w02 = v[2] / v[0], w12 = v[4] / v[3];
    double m0 = v[1] * v[4];
    m0 -= v[2] * v[3];
    const double m1 = v[0] * v[3] * sw22;
    const double w33 = m0 / m1;
    v[0] = w00; v[3] = w11; v[5] = w22; v[1] = w01; v[2] = w02; v[4] = w12; v[6] = w33;
  }
  /* per-edge terms: src/weights.c:9-100 */
  struct weights *wt = (struct weights *)kf_xmalloc(sizeof(struct weights));
  wt->w0 = (struct xyz *)kf_xmalloc(sizeof(struct xyz));
  wt->w1 = (struct xyz *)kf_xmalloc(sizeof(struct xyz));
  double **wa[6] = {&wt->w0->x0, &wt->w0->x1, &wt->w0->x2, &wt->w1->x0, &wt->w1->x1, &wt->w1->x2};
  for (int k = 0; k < 6; k++) {   /* every entry but the spare last one is written below */
    *wa[k] = (double *)kf_xmalloc((E + 1) * sizeof(double));
    (*wa[k])[E] = 0.0;
  }
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < E; i++) {
    const uint32_t a = n0[i], b = n1[i];
    const double dx = x0[b] - x0[a], dy = x1[b] - x1[a], dz = x2[b] - x2[a];
    double c0, c1, tx, ty, tz;
    const double *v = w + (size_t)a * 7;
    c0 = -dx * v[1] + dy;
    c1 = -dx * v[2] + dz;
    c1 = -v[4] * c0 + c1;
    tx = v[3] * v[1] * c0;
    tx = dx * v[0] - tx;
    tx += v[6] * c1;
    ty = v[4] * v[5] * c1;
    ty = v[3] * c0 - ty;
    tz = v[5] * c1;
    wt->w0->x0[i] = tx; wt->w0->x1[i] = ty; wt->w0->x2[i] = tz;
    v = w + (size_t)b * 7;
    c0 = dx * v[1] - dy;
    c1 = dx * v[2] - dz;
    c1 = -v[4] * c0 + c1;
    tx = v[3] * v[1] * c0;
    tx = -dx * v[0] - tx;
    tx += v[6] * c1;
    ty = v[4] * v[5] * c1;
    ty = v[3] * c0 - ty;
    tz = v[5] * c1;
    wt->w1->x0[i] = tx; wt->w1->x1[i] = ty; wt->w1->x2[i] = tz;
  }
  free(w);
  g->e->w = wt;
}

/* inc/phy.h:8-9: U = cos(alpha / (180 / M_PI)), V = sin(...) */
void kf_free_stream(double alpha_deg, double *U, double *V)
{
  const double rad = alpha_deg / (180 / M_PI);
  *U = cos(rad);
  *V = sin(rad);
}

void kf_host_setup(GEOMETRY *g)
{
  const size_t N = g->n->sz;
  g->c = (struct ctbl *)kf_xcalloc(1, sizeof(struct ctbl));
  g->c->b = 4;
  g->c->b2 = 16;
  g->c->sz = 4 * N;
  g->n->part = (uint32_t *)kf_xcalloc(N + 1, sizeof(uint32_t));
  g->s = (struct stbl *)kf_xcalloc(1, sizeof(struct stbl));
  g->s->sz = 1;
  g->s->i = (uint32_t *)kf_xmalloc(2 * sizeof(uint32_t));
  g->s->i[0] = 0;
  g->s->i[1] = (uint32_t)g->e->sz;
  g->t = (struct stbl *)kf_xcalloc(1, sizeof(struct stbl));
  kf_color_solid_faces(g);
  kf_phase("face colouring");
  kf_lsq_weights(g);
  kf_phase("LSQ weights");
  g->q = (struct qtbl *)kf_xcalloc(1, sizeof(struct qtbl));
}

/* ---- derived arrays ------------------------------------------------------- */

static uint32_t find_block(const kf_prep *p, uint32_t row, uint32_t col)
{
  uint32_t lo = p->rowptr[row], hi = p->rowptr[row + 1];
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (p->col[mid] < col) lo = mid + 1;
    else hi = mid;
  }
  if (lo >= p->rowptr[row + 1] || p->col[lo] != col) {
    fprintf(stderr, "kf: block (%u,%u) missing from the pattern\n", row, col);
    abort();
  }
  return lo;
}

/* a list per owner: counts -> offsets -> fill in the order of the loop */
static void bucket(size_t n, size_t m, const uint32_t *owner, const uint32_t *payload, uint32_t **ptr, uint32_t **ent)
{
  uint32_t *pt = (uint32_t *)kf_xcalloc(n + 1, sizeof(uint32_t));
  for (size_t i = 0; i < m; i++) pt[owner[i] + 1]++;
  for (size_t i = 0; i < n; i++) pt[i + 1] += pt[i];
  uint32_t *en = (uint32_t *)kf_xmalloc((m + 1) * sizeof(uint32_t));
  uint32_t *fillp = (uint32_t *)kf_xmalloc((n + 1) * sizeof(uint32_t));
  memcpy(fillp, pt, (n + 1) * sizeof(uint32_t));
  for (size_t i = 0; i < m; i++) en[fillp[owner[i]]++] = payload[i];
  free(fillp);
  *ptr = pt;
  *ent = en;
}

static void levels(const kf_prep *p, int upper, uint32_t *nlev, uint32_t **lptr, uint32_t **lrows)
{
  const size_t N = p->N;
  uint32_t *depth = (uint32_t *)kf_xcalloc(N + 1, sizeof(uint32_t));
  uint32_t maxd = 1;
  for (size_t t = 0; t < N; t++) {
    const size_t i = upper ? N - 1 - t : t;
    uint32_t d = 1;
    const uint32_t j0 = upper ? p->diag[i] + 1 : p->rowptr[i];
    const uint32_t j1 = upper ? p->rowptr[i + 1] : p->diag[i];
    for (uint32_t j = j0; j < j1; j++)
      if (depth[p->col[j]] + 1 > d) d = depth[p->col[j]] + 1;
    depth[i] = d;
    if (d > maxd) maxd = d;
  }
  uint32_t *lvl = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  uint32_t *rows = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  for (size_t i = 0; i < N; i++) { lvl[i] = depth[i] - 1; rows[i] = (uint32_t)i; }
  bucket(maxd, N, lvl, rows, lptr, lrows);
  *nlev = maxd;
  free(depth); free(lvl); free(rows);
}

void kf_prep_levels(kf_prep *p)
{
  levels(p, 0, &p->nlev_l, &p->lptr, &p->lrows);
  levels(p, 1, &p->nlev_u, &p->uptr, &p->urows);
}

void kf_build_prep(const GEOMETRY *g, kf_prep *p)
{
  memset(p, 0, sizeof(*p));
  const size_t N = g->n->sz, E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  p->N = N;
  p->E = E;
  for (size_t i = 0; i < E; i++)
    if (n0[i] >= N || n1[i] >= N || n0[i] == n1[i]) {
      fprintf(stderr, "kf: edge %zu (%u,%u) is invalid for %zu vertices\n", i, n0[i], n1[i], N);
      abort();
    }

  /* vertex -> incident edges, ascending edge id */
  kf_incidence(N, E, n0, n1, &p->vptr, &p->vent);

  /* pattern: matrix/init.cpp:7-74 (self + both edge directions, sorted):
   * row i holds i and the other end of each incident edge */
  p->rowptr = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i <= N; i++) p->rowptr[i] = p->vptr[i] + (uint32_t)i;
  p->nnzb = p->rowptr[N];
  p->col = (uint32_t *)kf_xmalloc((p->nnzb + 1) * sizeof(uint32_t));
  p->diag = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; i++) {
    uint32_t *c = p->col + p->rowptr[i];
    const uint32_t len = p->rowptr[i + 1] - p->rowptr[i];
    c[0] = (uint32_t)i;
    for (uint32_t s = p->vptr[i]; s < p->vptr[i + 1]; s++) {
      const uint32_t e = p->vent[s] >> 1;
      c[1 + s - p->vptr[i]] = (p->vent[s] & 1) ? n0[e] : n1[e];
    }
    for (uint32_t a = 1; a < len; a++) {   /* insertion sort: ~15 entries */
      const uint32_t x = c[a];
      uint32_t b = a;
      for (; b > 0 && c[b - 1] > x; b--) c[b] = c[b - 1];
      c[b] = x;
    }
    for (uint32_t j = 1; j < len; j++)
      if (c[j] == c[j - 1]) {
        fprintf(stderr, "kf: duplicate edge (%zu,%u)\n", i, c[j]);
        abort();
      }
    p->diag[i] = find_block(p, (uint32_t)i, (uint32_t)i);
  }

  kf_phase("pattern");
  {
    p->vblk = (uint32_t *)kf_xmalloc((2 * E + 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
    for (size_t v = 0; v < N; v++)
      for (uint32_t s = p->vptr[v]; s < p->vptr[v + 1]; s++) {
        const uint32_t e = p->vent[s] >> 1;
        const uint32_t nb = (p->vent[s] & 1) ? n0[e] : n1[e];
        p->vblk[s] = find_block(p, (uint32_t)v, nb);
      }
  }

  kf_phase("incidences");
  levels(p, 0, &p->nlev_l, &p->lptr, &p->lrows);
  levels(p, 1, &p->nlev_u, &p->uptr, &p->urows);

  kf_phase("level sets");
  /* boundary incidences */
  {
    const size_t S = g->b->fc->sz;
    uint32_t *owner = (uint32_t *)kf_xmalloc((3 * S + 1) * sizeof(uint32_t));
    uint32_t *pay = (uint32_t *)kf_xmalloc((3 * S + 1) * sizeof(uint32_t));
    for (size_t f = 0; f < S; f++) {
      owner[3 * f + 0] = g->b->fc->fptr->n0[f]; pay[3 * f + 0] = (uint32_t)(f << 2) | 0u;
      owner[3 * f + 1] = g->b->fc->fptr->n1[f]; pay[3 * f + 1] = (uint32_t)(f << 2) | 1u;
      owner[3 * f + 2] = g->b->fc->fptr->n2[f]; pay[3 * f + 2] = (uint32_t)(f << 2) | 2u;
    }
    bucket(N, 3 * S, owner, pay, &p->sfptr, &p->sfent);
    free(owner); free(pay);
  }
  {
    const size_t ns = g->b->s->sz, nf = g->b->f->sz;
    uint32_t *idx = (uint32_t *)kf_xmalloc(((ns > nf ? ns : nf) + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < ns; i++) idx[i] = (uint32_t)i;
    bucket(N, ns, g->b->s->nptr, idx, &p->snptr, &p->snent);
    for (size_t i = 0; i < nf; i++) idx[i] = (uint32_t)i;
    bucket(N, nf, g->b->f->nptr, idx, &p->fnptr, &p->fnent);
    free(idx);
  }
  kf_phase("boundary incidences");
}

/* the device's per-edge records: en[2e..2e+1] = (n0, n1), egeo[4e..4e+3] =
 * (unit normal, dual-face area) */
void kf_edge_records(const GEOMETRY *g, uint32_t *en, double *egeo)
{
  const size_t E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  const struct xyzn *xn = g->e->xyzn;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < E; i++) {
    en[2 * i] = n0[i];
    en[2 * i + 1] = n1[i];
    egeo[4 * i] = xn->x0[i];
    egeo[4 * i + 1] = xn->x1[i];
    egeo[4 * i + 2] = xn->x2[i];
    egeo[4 * i + 3] = xn->x3[i];
  }
}

void kf_build_sell(const GEOMETRY *g, const kf_prep *p, kf_sell *out)
{
  const size_t N = p->N;
  const uint32_t nslice = (uint32_t)((N + 63) / 64);
  uint32_t *sptr = (uint32_t *)kf_xmalloc((nslice + 1) * sizeof(uint32_t));
  uint32_t *deg = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  uint32_t *len = (uint32_t *)kf_xmalloc((nslice + 1) * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (uint32_t s = 0; s < nslice; s++) {
    uint32_t l = 0;
    for (size_t v = (size_t)s * 64; v < N && v < (size_t)s * 64 + 64; v++) {
      deg[v] = p->vptr[v + 1] - p->vptr[v];
      if (deg[v] > l) l = deg[v];
    }
    len[s] = l;
  }
  size_t rows = 0;
  for (uint32_t s = 0; s < nslice; s++) {
    sptr[s] = (uint32_t)rows;
    rows += len[s];
  }
  sptr[nslice] = (uint32_t)rows;
  free(len);
  const size_t ne = rows * 64;
  uint32_t *nbr = (uint32_t *)kf_xcalloc(ne + 1, sizeof(uint32_t));
  uint32_t *ent = (uint32_t *)kf_xcalloc(ne + 1, sizeof(uint32_t));
  uint32_t *blk = (uint32_t *)kf_xcalloc(ne + 1, sizeof(uint32_t));
  double *wx = (double *)kf_xcalloc(ne + 1, sizeof(double));
  double *wy = (double *)kf_xcalloc(ne + 1, sizeof(double));
  double *wz = (double *)kf_xcalloc(ne + 1, sizeof(double));
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  const struct weights *w = g->e->w;
#pragma omp parallel for schedule(static)
  for (size_t v = 0; v < N; v++) {
    const size_t base = (size_t)sptr[v >> 6] * 64 + (v & 63);
    for (uint32_t k = 0; k < deg[v]; k++) {
      const uint32_t s = p->vptr[v] + k, en = p->vent[s], e = en >> 1;
      const size_t i = base + (size_t)k * 64;
      ent[i] = en;
      blk[i] = p->vblk[s];
      if (en & 1u) {
        nbr[i] = n0[e];
        wx[i] = w->w1->x0[e]; wy[i] = w->w1->x1[e]; wz[i] = w->w1->x2[e];
      } else {
        nbr[i] = n1[e];
        wx[i] = w->w0->x0[e]; wy[i] = w->w0->x1[e]; wz[i] = w->w0->x2[e];
      }
    }
  }
  out->nslice = nslice;
  out->ne = ne;
  out->sptr = sptr; out->deg = deg; out->nbr = nbr; out->ent = ent; out->blk = blk;
  out->wx = wx; out->wy = wy; out->wz = wz;
}

void kf_free_prep(kf_prep *p)
{
  free(p->rowptr); free(p->col); free(p->diag); free(p->vptr); free(p->vent); free(p->vblk);
  free(p->lptr); free(p->lrows); free(p->uptr); free(p->urows);
  free(p->sfptr); free(p->sfent); free(p->snptr); free(p->snent); free(p->fnptr); free(p->fnent);
  memset(p, 0, sizeof(*p));
}

void kf_free_lsq(GEOMETRY *g)
{
  if (!g->e || !g->e->w) return;
  free(g->e->w->w0->x0); free(g->e->w->w0->x1); free(g->e->w->w0->x2);
  free(g->e->w->w1->x0); free(g->e->w->w1->x1); free(g->e->w->w1->x2);
  free(g->e->w->w0); free(g->e->w->w1); free(g->e->w);
  g->e->w = NULL;
}

static void free_bn(struct bntbl *b)
{
  if (!b) return;
  free(b->nptr);
  if (b->xyz) { free(b->xyz->x0); free(b->xyz->x1); free(b->xyz->x2); free(b->xyz); }
  free(b);
}

void kf_free_host_geometry(GEOMETRY *g)
{
  if (g->e) {
    free(g->e->eptr->n0); free(g->e->eptr->n1); free(g->e->eptr);
    free(g->e->xyzn->x0); free(g->e->xyzn->x1); free(g->e->xyzn->x2); free(g->e->xyzn->x3); free(g->e->xyzn);
    kf_free_lsq(g);
    free(g->e);
  }
  if (g->n) {
    free(g->n->xyz->x0); free(g->n->xyz->x1); free(g->n->xyz->x2); free(g->n->xyz);
    free(g->n->part);
    free(g->n);
  }
  if (g->b) {
    free_bn(g->b->s); free_bn(g->b->f);
    if (g->b->fc) {
      free(g->b->fc->fptr->n0); free(g->b->fc->fptr->n1); free(g->b->fc->fptr->n2);
      free(g->b->fc->fptr); free(g->b->fc);
    }
    free(g->b);
  }
  if (g->s) { free(g->s->i); free(g->s); }
  if (g->t) { free(g->t->i); free(g->t); }
  free(g->c);
  free(g->q);
  g->e = NULL; g->n = NULL; g->b = NULL; g->s = NULL; g->t = NULL; g->c = NULL; g->q = NULL;
}
