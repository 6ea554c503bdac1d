/*
 * oracle.c -- one-thread CPU restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP path,
 * never part of it.  Built with gcc -O2 -ffp-contract=off (no FMA, plain
 * IEEE double, as the reference's x86-64 build), so every expression below
 * rounds exactly as the reference's one-thread run does.  Each function
 * cites the reference lines whose operation order it keeps; the owner test
 * part[node] == t of the reference is always true with one thread and is
 * dropped.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define BETA 15.0       /* inc/phy.h:6, B */
#define PINF 1.0        /* inc/phy.h:7, P */
#define WINF 0.0        /* inc/phy.h:10, W */
#define MAG0 (0.5 / 3)  /* src/flux.c:11 */
#define MAG1 (-MAG0)

typedef struct { uint32_t *i, *j, *d; double *a; } or_mat;

struct or_ctx {
  const OR_GEOMETRY *g;
  or_opts o;
  double U, V;
  size_t N;
  double *q, *cdt;
  /* fun3d::Matrix restated: A, L, U (pattern), Ut, index maps, levels */
  or_mat A, L, Up, Ut;
  uint32_t *ut_index;
  uint32_t tl_n, *tl_i, *tl_j, tu_n, *tu_i, *tu_j;
  double *grad[3], *r, *w, *t, *x;
};

static void *xmalloc(size_t b)
{
  void *p = malloc(b ? b : 8);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}
static void *xcalloc(size_t n, size_t s)
{
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* ---- matrix/init.cpp:96-155 block primitives ------------------------------ */

static void dgemv4(const double *A, const double *x, double *y)
{
  for (int i = 0; i < 4; i++) {
    y[i] = 0.0;
    for (int j = 0; j < 4; j++) y[i] += A[i * 4 + j] * x[j];
  }
}

static void dgemm4(const double *A, const double *B, double *C)
{
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      C[i * 4 + j] = 0.0;
      for (int k = 0; k < 4; k++) C[i * 4 + j] += A[i * 4 + k] * B[k * 4 + j];
    }
}

static void axpy4(double a, const double *x, double *y)
{
  for (int i = 0; i < 4; i++) { const double ax = a * x[i]; y[i] = ax + y[i]; }
}

static void axpy16(double a, const double *x, double *y)
{
  for (int i = 0; i < 16; i++) { const double ax = a * x[i]; y[i] = ax + y[i]; }
}

/* ---- matrix/inverse.cpp:195-353: cofactor inverse, det = 1.0f/det ---------- */

void or_inv4(const double *s, double *d)
{
  d[0] = +s[5] * s[10] * s[15] - s[5] * s[11] * s[14] - s[9] * s[6] * s[15] + s[9] * s[7] * s[14] + s[13] * s[6] * s[11] - s[13] * s[7] * s[10];
  d[1] = -s[1] * s[10] * s[15] + s[1] * s[11] * s[14] + s[9] * s[2] * s[15] - s[9] * s[3] * s[14] - s[13] * s[2] * s[11] + s[13] * s[3] * s[10];
  d[2] = +s[1] * s[6] * s[15] - s[1] * s[7] * s[14] - s[5] * s[2] * s[15] + s[5] * s[3] * s[14] + s[13] * s[2] * s[7] - s[13] * s[3] * s[6];
  d[3] = -s[1] * s[6] * s[11] + s[1] * s[7] * s[10] + s[5] * s[2] * s[11] - s[5] * s[3] * s[10] - s[9] * s[2] * s[7] + s[9] * s[3] * s[6];
  d[4] = -s[4] * s[10] * s[15] + s[4] * s[11] * s[14] + s[8] * s[6] * s[15] - s[8] * s[7] * s[14] - s[12] * s[6] * s[11] + s[12] * s[7] * s[10];
  d[5] = +s[0] * s[10] * s[15] - s[0] * s[11] * s[14] - s[8] * s[2] * s[15] + s[8] * s[3] * s[14] + s[12] * s[2] * s[11] - s[12] * s[3] * s[10];
  d[6] = -s[0] * s[6] * s[15] + s[0] * s[7] * s[14] + s[4] * s[2] * s[15] - s[4] * s[3] * s[14] - s[12] * s[2] * s[7] + s[12] * s[3] * s[6];
  d[7] = +s[0] * s[6] * s[11] - s[0] * s[7] * s[10] - s[4] * s[2] * s[11] + s[4] * s[3] * s[10] + s[8] * s[2] * s[7] - s[8] * s[3] * s[6];
  d[8] = +s[4] * s[9] * s[15] - s[4] * s[11] * s[13] - s[8] * s[5] * s[15] + s[8] * s[7] * s[13] + s[12] * s[5] * s[11] - s[12] * s[7] * s[9];
  d[9] = -s[0] * s[9] * s[15] + s[0] * s[11] * s[13] + s[8] * s[1] * s[15] - s[8] * s[3] * s[13] - s[12] * s[1] * s[11] + s[12] * s[3] * s[9];
  d[10] = +s[0] * s[5] * s[15] - s[0] * s[7] * s[13] - s[4] * s[1] * s[15] + s[4] * s[3] * s[13] + s[12] * s[1] * s[7] - s[12] * s[3] * s[5];
  d[11] = -s[0] * s[5] * s[11] + s[0] * s[7] * s[9] + s[4] * s[1] * s[11] - s[4] * s[3] * s[9] - s[8] * s[1] * s[7] + s[8] * s[3] * s[5];
  d[12] = -s[4] * s[9] * s[14] + s[4] * s[10] * s[13] + s[8] * s[5] * s[14] - s[8] * s[6] * s[13] - s[12] * s[5] * s[10] + s[12] * s[6] * s[9];
  d[13] = +s[0] * s[9] * s[14] - s[0] * s[10] * s[13] - s[8] * s[1] * s[14] + s[8] * s[2] * s[13] + s[12] * s[1] * s[10] - s[12] * s[2] * s[9];
  d[14] = -s[0] * s[5] * s[14] + s[0] * s[6] * s[13] + s[4] * s[1] * s[14] - s[4] * s[2] * s[13] - s[12] * s[1] * s[6] + s[12] * s[2] * s[5];
  d[15] = +s[0] * s[5] * s[10] - s[0] * s[6] * s[9] - s[4] * s[1] * s[10] + s[4] * s[2] * s[9] + s[8] * s[1] * s[6] - s[8] * s[2] * s[5];
  double det = +s[0] * d[0] + s[1] * d[4] + s[2] * d[8] + s[3] * d[12];
  det = 1.0 / det;
  for (int i = 0; i < 16; i++) d[i] *= det;
}

/* ---- matrix/init.cpp:7-74,194-274 and matrix/edmond.cpp:11-148 ------------- */

static void build_matrix(or_ctx *c)
{
  const OR_GEOMETRY *g = c->g;
  const size_t N = c->N, E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  or_mat *A = &c->A;
  A->i = xcalloc(N + 1, 4);
  A->d = xcalloc(N + 1, 4);
  for (size_t i = 0; i < E; i++) { A->i[n0[i]]++; A->i[n1[i]]++; }
  for (size_t i = 0; i <= N; i++) A->i[i]++;
  uint32_t sum = 0;
  for (size_t i = 0; i <= N; i++) { uint32_t t = A->i[i]; A->i[i] = sum; sum += t; }
  const size_t nnz = A->i[N];
  A->j = xcalloc(nnz + 1, 4);
  uint32_t *iw = xcalloc(N + 1, 4);
  for (size_t i = 0; i < N; i++) { A->j[A->i[i]] = (uint32_t)i; iw[i] = 1; }
  for (size_t i = 0; i < E; i++) {
    A->j[A->i[n0[i]] + iw[n0[i]]++] = n1[i];
    A->j[A->i[n1[i]] + iw[n1[i]]++] = n0[i];
  }
  for (size_t i = 0; i < N; i++) {
    /* insertion sort of the short row (std::sort gives the same order) */
    uint32_t *r = A->j + A->i[i];
    const size_t len = A->i[i + 1] - A->i[i];
    for (size_t a = 1; a < len; a++) {
      uint32_t key = r[a];
      size_t b = a;
      while (b > 0 && r[b - 1] > key) { r[b] = r[b - 1]; b--; }
      r[b] = key;
    }
    for (uint32_t j = A->i[i]; j < A->i[i + 1]; j++)
      if (A->j[j] == i) { A->d[i] = j; break; }
  }
  A->a = xcalloc(nnz * 16, sizeof(double));

  /* split: L keeps cols <= i with an identity diagonal, U keeps cols >= i */
  or_mat *L = &c->L, *U = &c->Up, *Ut = &c->Ut;
  L->i = xcalloc(N + 1, 4); L->d = xcalloc(N + 1, 4);
  U->i = xcalloc(N + 1, 4); U->d = xcalloc(N + 1, 4);
  Ut->i = xcalloc(N + 1, 4);
  for (size_t i = 0; i < N; i++)
    for (uint32_t j = A->i[i]; j < A->i[i + 1]; j++) {
      if (A->j[j] <= i) L->i[i]++;
      if (A->j[j] >= i) U->i[i]++;
    }
  sum = 0;
  for (size_t i = 0; i <= N; i++) { uint32_t t = L->i[i]; L->i[i] = sum; sum += t; }
  sum = 0;
  for (size_t i = 0; i <= N; i++) { uint32_t t = U->i[i]; U->i[i] = sum; sum += t; }
  L->j = xcalloc(L->i[N] + 1, 4);
  U->j = xcalloc(U->i[N] + 1, 4);
  L->a = xcalloc((size_t)L->i[N] * 16 + 1, sizeof(double));
  for (size_t i = 0; i < N; i++) {
    uint32_t il = L->i[i], iu = U->i[i];
    for (uint32_t j = A->i[i]; j < A->i[i + 1]; j++) {
      if (A->j[j] <= i) {
        L->j[il] = A->j[j];
        if (A->j[j] == i) {
          L->d[i] = il;
          double *b = L->a + (size_t)il * 16;
          for (int k = 0; k < 16; k++) b[k] = (k % 5 == 0) ? 1.0 : 0.0;
        }
        il++;
      }
      if (A->j[j] >= i) {
        U->j[iu] = A->j[j];
        if (A->j[j] == i) U->d[i] = iu;
        iu++;
      }
    }
  }
  /* csr2csc of U -> Ut (rows of Ut = columns of U, ascending) */
  uint32_t *offs = xcalloc(U->i[N] + 1, 4);
  for (size_t i = 0; i < N; i++)
    for (uint32_t j = U->i[i]; j < U->i[i + 1]; j++) offs[j] = Ut->i[U->j[j]]++;
  sum = 0;
  for (size_t i = 0; i <= N; i++) { uint32_t t = Ut->i[i]; Ut->i[i] = sum; sum += t; }
  Ut->j = xcalloc(Ut->i[N] + 1, 4);
  Ut->a = xcalloc((size_t)Ut->i[N] * 16 + 1, sizeof(double));
  for (size_t i = 0; i < N; i++)
    for (uint32_t j = U->i[i]; j < U->i[i + 1]; j++) Ut->j[Ut->i[U->j[j]] + offs[j]] = (uint32_t)i;
  free(offs);
  /* setUtIndex */
  c->ut_index = xcalloc(nnz + 1, 4);
  memset(iw, 0, (N + 1) * 4);
  for (size_t i = 0; i < N; i++)
    for (uint32_t j = A->i[i]; j < A->i[i + 1]; j++)
      if (A->j[j] >= i) c->ut_index[j] = Ut->i[A->j[j]] + iw[A->j[j]]++;
  free(iw);

  /* task_graph: matrix/init.cpp:194-274 */
  uint32_t *row = xcalloc(N + 1, 4), *top = xcalloc(N + 2, 4);
  c->tl_i = xcalloc(N + 2, 4);
  c->tl_n = 1;
  for (size_t i = 0; i < N; i++) {
    uint32_t depth = 1;
    for (uint32_t j = A->i[i]; j < A->d[i]; j++) if (row[A->j[j]] + 1 > depth) depth = row[A->j[j]] + 1;
    if (depth > c->tl_n) c->tl_n = depth;
    row[i] = depth;
    ++c->tl_i[depth];
  }
  c->tl_j = xmalloc((N + 1) * 4);
  for (uint32_t i = 1; i < c->tl_n + 1; i++) { c->tl_i[i] += c->tl_i[i - 1]; top[i] = c->tl_i[i]; }
  for (size_t i = 0; i < N; i++) { c->tl_j[top[row[i] - 1]] = (uint32_t)i; ++top[row[i] - 1]; }
  memset(row, 0, (N + 1) * 4);
  c->tu_i = xcalloc(N + 2, 4);
  c->tu_n = 1;
  for (long i = (long)N - 1; i >= 0; i--) {
    uint32_t depth = 1;
    for (uint32_t j = A->d[i] + 1; j < A->i[i + 1]; j++) if (row[A->j[j]] + 1 > depth) depth = row[A->j[j]] + 1;
    if (depth > c->tu_n) c->tu_n = depth;
    row[i] = depth;
    ++c->tu_i[depth];
  }
  memset(top, 0, (N + 2) * 4);
  c->tu_j = xmalloc((N + 1) * 4);
  for (uint32_t i = 1; i < c->tu_n + 1; i++) { c->tu_i[i] += c->tu_i[i - 1]; top[i] = c->tu_i[i]; }
  for (size_t i = 0; i < N; i++) { c->tu_j[top[row[i] - 1]] = (uint32_t)i; ++top[row[i] - 1]; }
  free(row); free(top);
}

/* ---- matrix/matrix.cpp:137-207: fill with the L / Ut copy side effects ----- */

static void mat_copy_after_fill(or_ctx *c, uint32_t row, uint32_t i)
{
  const double *blk = c->A.a + (size_t)i * 16;
  if (c->A.j[i] < row) memcpy(c->L.a + (size_t)c->L.i[row] * 16, blk, 16 * sizeof(double));
  if (c->A.j[i] >= row) memcpy(c->Ut.a + (size_t)c->ut_index[i] * 16, blk, 16 * sizeof(double));
}

static uint32_t mat_find(const or_ctx *c, uint32_t row, uint32_t col)
{
  for (uint32_t i = c->A.i[row]; i < c->A.i[row + 1]; i++) if (c->A.j[i] == col) return i;
  return UINT32_MAX;
}

static void fill_alpha(or_ctx *c, uint32_t row, uint32_t col, double a, const double *v)
{
  uint32_t i = mat_find(c, row, col);
  if (i == UINT32_MAX) return;
  for (int j = 0; j < 16; j++) c->A.a[(size_t)i * 16 + j] += a * v[j];
  mat_copy_after_fill(c, row, i);
}

static void fill_block(or_ctx *c, uint32_t row, uint32_t col, const double *v)
{
  uint32_t i = mat_find(c, row, col);
  if (i == UINT32_MAX) return;
  for (int j = 0; j < 16; j++) c->A.a[(size_t)i * 16 + j] += v[j];
  mat_copy_after_fill(c, row, i);
}

static void fill_diag(or_ctx *c, uint32_t row, uint32_t col, double v)
{
  uint32_t i = mat_find(c, row, col);
  if (i == UINT32_MAX) return;
  double *b = c->A.a + (size_t)i * 16;
  b[0] += v; b[5] += v; b[10] += v; b[15] += v;
  mat_copy_after_fill(c, row, i);
}

static void fill_boundary(or_ctx *c, uint32_t row, uint32_t col, uint32_t n, const double *v)
{
  uint32_t i = mat_find(c, row, col);
  if (i == UINT32_MAX) return;
  for (uint32_t j = 1; j <= n; j++) c->A.a[(size_t)i * 16 + (j * n + j)] += v[j - 1];
  mat_copy_after_fill(c, row, i);
}

/* ---- context -------------------------------------------------------------- */

or_ctx *or_create(const OR_GEOMETRY *g, const or_opts *o)
{
  or_ctx *c = xcalloc(1, sizeof(or_ctx));
  c->g = g;
  c->o = *o;
  c->N = g->n->sz;
  const double rad = o->alpha_deg / (180 / M_PI);
  c->U = cos(rad);
  c->V = sin(rad);
  const size_t sz = 4 * c->N;
  c->q = xmalloc(sz * sizeof(double));
  c->cdt = xcalloc(c->N + 1, sizeof(double));
  for (size_t i = 0; i < c->N; i++) {   /* src/guess.c:8-20 */
    c->q[4 * i + 0] = PINF; c->q[4 * i + 1] = c->U; c->q[4 * i + 2] = c->V; c->q[4 * i + 3] = WINF;
  }
  for (int k = 0; k < 3; k++) c->grad[k] = xcalloc(sz, sizeof(double));
  c->r = xcalloc(sz, sizeof(double));
  c->w = xcalloc(sz, sizeof(double));
  c->t = xcalloc(sz, sizeof(double));
  c->x = xcalloc(sz, sizeof(double));
  build_matrix(c);
  return c;
}

void or_destroy(or_ctx *c)
{
  if (!c) return;
  free(c->q); free(c->cdt);
  for (int k = 0; k < 3; k++) free(c->grad[k]);
  free(c->r); free(c->w); free(c->t); free(c->x);
  or_mat *m[4] = {&c->A, &c->L, &c->Up, &c->Ut};
  for (int k = 0; k < 4; k++) { free(m[k]->i); free(m[k]->j); free(m[k]->d); free(m[k]->a); }
  free(c->ut_index); free(c->tl_i); free(c->tl_j); free(c->tu_i); free(c->tu_j);
  free(c);
}

double *or_state(or_ctx *c) { return c->q; }
double *or_cdt(or_ctx *c) { return c->cdt; }
size_t or_nnzb(const or_ctx *c) { return c->A.i[c->N]; }
const double *or_blocks(const or_ctx *c) { return c->A.a; }
void or_pattern(const or_ctx *c, unsigned int *rowptr, unsigned int *col)
{
  memcpy(rowptr, c->A.i, (c->N + 1) * 4);
  memcpy(col, c->A.j, (size_t)c->A.i[c->N] * 4);
}

/* ---- src/flux.c:15-780 ------------------------------------------------------ */

void or_flux(or_ctx *c, const double *q, double *gradx0, double *gradx1, double *gradx2, double *r)
{
  const OR_GEOMETRY *g = c->g;
  const size_t dofs = 4 * c->N, E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  const double *x0 = g->e->xyzn->x0, *x1 = g->e->xyzn->x1, *x2 = g->e->xyzn->x2, *x3 = g->e->xyzn->x3;
  const double *xyz0 = g->n->xyz->x0, *xyz1 = g->n->xyz->x1, *xyz2 = g->n->xyz->x2;
  const double U = c->U, V = c->V, W = WINF, B = BETA, P = PINF;
  memset(gradx0, 0, dofs * sizeof(double));
  memset(gradx1, 0, dofs * sizeof(double));
  memset(gradx2, 0, dofs * sizeof(double));
  memset(r, 0, dofs * sizeof(double));

  /* LSQ gradient, src/flux.c:60-145 */
  const double *w0x = g->e->w->w0->x0, *w0y = g->e->w->w0->x1, *w0z = g->e->w->w0->x2;
  const double *w1x = g->e->w->w1->x0, *w1y = g->e->w->w1->x1, *w1z = g->e->w->w1->x2;
  for (size_t i = 0; i < E; i++) {
    const uint32_t idx0 = 4 * n0[i], idx1 = 4 * n1[i];
    double tx = w0x[i], ty = w0y[i], tz = w0z[i];
    for (int k = 0; k < 4; k++) {
      const double dq = q[idx1 + k] - q[idx0 + k];
      gradx0[idx0 + k] += tx * dq;
      gradx1[idx0 + k] += ty * dq;
      gradx2[idx0 + k] += tz * dq;
    }
    tx = w1x[i]; ty = w1y[i]; tz = w1z[i];
    for (int k = 0; k < 4; k++) {
      const double dq = q[idx0 + k] - q[idx1 + k];
      gradx0[idx1 + k] += tx * dq;
      gradx1[idx1 + k] += ty * dq;
      gradx2[idx1 + k] += tz * dq;
    }
  }
  if (c->o.first_order) {
    memset(gradx0, 0, dofs * sizeof(double));
    memset(gradx1, 0, dofs * sizeof(double));
    memset(gradx2, 0, dofs * sizeof(double));
  }

  /* Roe-type edge flux, src/flux.c:151-465 */
  for (size_t i = 0; i < E; i++) {
    const uint32_t node0 = n0[i], node1 = n1[i];
    const double xn = x0[i], yn = x1[i], zn = x2[i], ln = x3[i];
    const double xmean = 0.5 * (xyz0[node0] + xyz0[node1]);
    const double ymean = 0.5 * (xyz1[node0] + xyz1[node1]);
    const double zmean = 0.5 * (xyz2[node0] + xyz2[node1]);
    double X1, Y1, Z1, dot = xn;
    if (fabs(dot) < (double)0.95f) { X1 = 1.0 - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn; }
    else { dot = yn; X1 = -dot * xn; Y1 = 1.0 - dot * yn; Z1 = -dot * zn; }
    double size = X1 * X1;
    size += Y1 * Y1;
    size += Z1 * Z1;
    size = sqrt(size);
    X1 /= size; Y1 /= size; Z1 /= size;
    double X2 = yn * Z1; X2 -= zn * Y1;
    double Y2 = zn * X1; Y2 -= xn * Z1;
    double Z2 = xn * Y1; Z2 -= yn * X1;
    double rx = xmean - xyz0[node0], ry = ymean - xyz1[node0], rz = zmean - xyz2[node0];
    const uint32_t idx0 = 4 * node0, idx1 = 4 * node1;
    double pL = q[idx0 + 0] + gradx0[idx0 + 0] * rx; pL += gradx1[idx0 + 0] * ry; pL += gradx2[idx0 + 0] * rz;
    double uL = q[idx0 + 1] + gradx0[idx0 + 1] * rx; uL += gradx1[idx0 + 1] * ry; uL += gradx2[idx0 + 1] * rz;
    double vL = q[idx0 + 2] + gradx0[idx0 + 2] * rx; vL += gradx1[idx0 + 2] * ry; vL += gradx2[idx0 + 2] * rz;
    double wL = q[idx0 + 3] + gradx0[idx0 + 3] * rx; wL += gradx1[idx0 + 3] * ry; wL += gradx2[idx0 + 3] * rz;
    double ubarL = xn * uL; ubarL += yn * vL; ubarL += zn * wL;
    rx = xmean - xyz0[node1]; ry = ymean - xyz1[node1]; rz = zmean - xyz2[node1];
    double pR = q[idx1 + 0] + gradx0[idx1 + 0] * rx; pR += gradx1[idx1 + 0] * ry; pR += gradx2[idx1 + 0] * rz;
    double uR = q[idx1 + 1] + gradx0[idx1 + 1] * rx; uR += gradx1[idx1 + 1] * ry; uR += gradx2[idx1 + 1] * rz;
    double vR = q[idx1 + 2] + gradx0[idx1 + 2] * rx; vR += gradx1[idx1 + 2] * ry; vR += gradx2[idx1 + 2] * rz;
    double wR = q[idx1 + 3] + gradx0[idx1 + 3] * rx; wR += gradx1[idx1 + 3] * ry; wR += gradx2[idx1 + 3] * rz;
    double ubarR = xn * uR; ubarR += yn * vR; ubarR += zn * wR;
    const double u = 0.5 * (uL + uR), v = 0.5 * (vL + vR), w = 0.5 * (wL + wR);
    double ubar = xn * u; ubar += yn * v; ubar += zn * w;
    double phi1 = xn * B; phi1 += u * ubar;
    double phi2 = yn * B; phi2 += v * ubar;
    double phi3 = zn * B; phi3 += w * ubar;
    double phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
    double phi5 = Z2 * phi1; phi5 -= X2 * phi3;
    double phi6 = X2 * phi2; phi6 -= Y2 * phi1;
    double phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
    double phi8 = X1 * phi3; phi8 -= Z1 * phi1;
    double phi9 = Y1 * phi1; phi9 -= X1 * phi2;
    const double c2 = ubar * ubar + B;
    const double cc = sqrt(c2);
    const double eig1 = fabs(ubar), eig2 = fabs(ubar), eig3 = fabs(ubar + cc), eig4 = fabs(ubar - cc);
    const double dp = pR - pL, du = uR - uL, dv = vR - vL, dw = wR - wL;
    double ti11 = u * phi4; ti11 += v * phi5; ti11 += w * phi6; ti11 = -ti11 / B;
    double ti21 = u * phi7; ti21 += v * phi8; ti21 += w * phi9; ti21 = -ti21 / B;
    double ti31 = 0.5 * (cc - ubar); ti31 /= B;
    double ti41 = -0.5 * (cc + ubar); ti41 /= B;
    double dv1 = ti11 * dp; dv1 += phi4 * du; dv1 += phi5 * dv; dv1 += phi6 * dw; dv1 /= c2;
    double dv2 = ti21 * dp; dv2 += phi7 * du; dv2 += phi8 * dv; dv2 += phi9 * dw; dv2 /= c2;
    double dv3 = 2.0 * ti31 * dp; dv3 += xn * du; dv3 += yn * dv; dv3 += zn * dw; dv3 *= 0.5 / c2;
    double dv4 = 2.0 * ti41 * dp; dv4 += xn * du; dv4 += yn * dv; dv4 += zn * dw; dv4 *= 0.5 / c2;
    const double r13 = cc * B;
    double r23 = u * (ubar + cc); r23 += xn * B;
    double r33 = v * (ubar + cc); r33 += yn * B;
    double r43 = w * (ubar + cc); r43 += zn * B;
    const double r14 = -cc * B;
    double r24 = u * (ubar - cc); r24 += xn * B;
    double r34 = v * (ubar - cc); r34 += yn * B;
    double r44 = w * (ubar - cc); r44 += zn * B;
    const double t1 = eig3 * r13 * dv3 + eig4 * r14 * dv4;
    double t2 = eig1 * X1 * dv1 + eig2 * X2 * dv2; t2 += eig3 * r23 * dv3 + eig4 * r24 * dv4;
    double t3 = eig1 * Y1 * dv1 + eig2 * Y2 * dv2; t3 += eig3 * r33 * dv3 + eig4 * r34 * dv4;
    double t4 = eig1 * Z1 * dv1 + eig2 * Z2 * dv2; t4 += eig3 * r43 * dv3 + eig4 * r44 * dv4;
    const double fluxp1 = ln * B * ubarL;
    const double fluxp2 = ln * (uL * ubarL + xn * pL);
    const double fluxp3 = ln * (vL * ubarL + yn * pL);
    const double fluxp4 = ln * (wL * ubarL + zn * pL);
    const double fluxm1 = ln * B * ubarR;
    const double fluxm2 = ln * (uR * ubarR + xn * pR);
    const double fluxm3 = ln * (vR * ubarR + yn * pR);
    const double fluxm4 = ln * (wR * ubarR + zn * pR);
    const double res1 = 0.5 * (fluxp1 + fluxm1 - ln * t1);
    const double res2 = 0.5 * (fluxp2 + fluxm2 - ln * t2);
    const double res3 = 0.5 * (fluxp3 + fluxm3 - ln * t3);
    const double res4 = 0.5 * (fluxp4 + fluxm4 - ln * t4);
    r[idx0 + 0] = r[idx0 + 0] + res1; r[idx0 + 1] = r[idx0 + 1] + res2;
    r[idx0 + 2] = r[idx0 + 2] + res3; r[idx0 + 3] = r[idx0 + 3] + res4;
    r[idx1 + 0] = r[idx1 + 0] - res1; r[idx1 + 1] = r[idx1 + 1] - res2;
    r[idx1 + 2] = r[idx1 + 2] - res3; r[idx1 + 3] = r[idx1 + 3] - res4;
  }

  /* solid wall faces by colour, src/flux.c:467-539 */
  const uint32_t *sn0 = g->b->fc->fptr->n0, *sn1 = g->b->fc->fptr->n1, *sn2 = g->b->fc->fptr->n2;
  for (size_t cl = 0; cl < g->t->sz; cl++)
    for (uint32_t j = g->t->i[cl]; j < g->t->i[cl + 1]; j++) {
      const uint32_t node0 = sn0[j], node1 = sn1[j], node2 = sn2[j];
      const double p1 = q[4 * node0], p2 = q[4 * node1], p3 = q[4 * node2];
      const double ax = xyz0[node1] - xyz0[node0], ay = xyz1[node1] - xyz1[node0], az = xyz2[node1] - xyz2[node0];
      const double bx = xyz0[node2] - xyz0[node0], by = xyz1[node2] - xyz1[node0], bz = xyz2[node2] - xyz2[node0];
      double xn = ay * bz; xn -= az * by; xn *= MAG1;
      double yn = ax * bz; yn -= az * bx; yn *= MAG0;
      double zn = ax * by; zn -= ay * bx; zn *= MAG1;
      double pa = 0.125 * (p2 + p3); pa += 0.75 * p1;
      double pb = 0.125 * (p3 + p1); pb += 0.75 * p2;
      double pc = 0.125 * (p1 + p2); pc += 0.75 * p3;
      uint32_t idx = 4 * node0;
      r[idx + 1] += xn * pa; r[idx + 2] += yn * pa; r[idx + 3] += zn * pa;
      idx = 4 * node1;
      r[idx + 1] += xn * pb; r[idx + 2] += yn * pb; r[idx + 3] += zn * pb;
      idx = 4 * node2;
      r[idx + 1] += xn * pc; r[idx + 2] += yn * pc; r[idx + 3] += zn * pc;
    }

  /* free-stream nodes, src/flux.c:541-779 */
  const struct or_bntbl *fb = g->b->f;
  for (size_t i = 0; i < fb->sz; i++) {
    const uint32_t n = fb->nptr[i];
    double xn = fb->xyz->x0[i], yn = fb->xyz->x1[i], zn = fb->xyz->x2[i];
    double area = xn * xn; area += yn * yn; area += zn * zn; area = sqrt(area);
    xn /= area; yn /= area; zn /= area;
    double X1, Y1, Z1, dot = xn;
    if (fabs(dot) < (double)0.95f) { X1 = 1.0 - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn; }
    else { dot = yn; X1 = -dot * xn; Y1 = 1.0 - dot * yn; Z1 = -dot * zn; }
    double size = X1 * X1; size += Y1 * Y1; size += Z1 * Z1; size = sqrt(size);
    X1 /= size; Y1 /= size; Z1 /= size;
    double X2 = yn * Z1; X2 -= zn * Y1;
    double Y2 = zn * X1; Y2 -= xn * Z1;
    double Z2 = xn * Y1; Z2 -= yn * X1;
    double ubar0 = xn * U; ubar0 += yn * V; ubar0 += zn * W;
    const double c20 = ubar0 * ubar0 + B;
    const double c0 = sqrt(c20);
    double phi1 = xn * B; phi1 += U * ubar0;
    double phi2 = yn * B; phi2 += V * ubar0;
    double phi3 = zn * B; phi3 += W * ubar0;
    double phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
    double phi5 = Z2 * phi1; phi5 -= X2 * phi3;
    double phi6 = X2 * phi2; phi6 -= Y2 * phi1;
    double phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
    double phi8 = X1 * phi3; phi8 -= Z1 * phi1;
    double phi9 = Y1 * phi1; phi9 -= X1 * phi2;
    const double t13 = c0 * B;
    double t23 = U * (ubar0 + c0); t23 += xn * B;
    double t33 = V * (ubar0 + c0); t33 += yn * B;
    double t43 = W * (ubar0 + c0); t43 += zn * B;
    const double t14 = -c0 * B;
    double t24 = U * (ubar0 - c0); t24 += xn * B;
    double t34 = V * (ubar0 - c0); t34 += yn * B;
    double t44 = W * (ubar0 - c0); t44 += zn * B;
    double ti11 = U * phi4; ti11 += V * phi5; ti11 += W * phi6; ti11 = -ti11 / B;
    double ti21 = U * phi7; ti21 += V * phi8; ti21 += W * phi9; ti21 = -ti21 / B;
    double ti31 = 0.5 * (c0 - ubar0); ti31 /= B;
    double ti41 = -0.5 * (c0 + ubar0); ti41 /= B;
    const double pi = q[4 * n + 0], ui = q[4 * n + 1], vi = q[4 * n + 2], wi = q[4 * n + 3];
    double un = xn * ui; un += yn * vi; un += zn * wi;
    double pr, ur, vr, wr;
    if (un > 0.0) { pr = pi; ur = ui; vr = vi; wr = wi; }
    else { pr = P; ur = U; vr = V; wr = W; }
    double rhs1 = ti11 * pr; rhs1 += phi4 * ur; rhs1 += phi5 * vr; rhs1 += phi6 * wr; rhs1 /= c20;
    double rhs2 = ti21 * pr; rhs2 += phi7 * ur; rhs2 += phi8 * vr; rhs2 += phi9 * wr; rhs2 /= c20;
    double rhs3 = 2.0 * ti31 * pi; rhs3 += xn * ui; rhs3 += yn * vi; rhs3 += zn * wi; rhs3 = 0.5 * rhs3 / c20;
    double rhs4 = 2.0 * ti41 * P; rhs4 += xn * U; rhs4 += yn * V; rhs4 += zn * W; rhs4 = 0.5 * rhs4 / c20;
    double pb = t13 * rhs3; pb += t14 * rhs4;
    double ub = X1 * rhs1; ub += X2 * rhs2; ub += t23 * rhs3; ub += t24 * rhs4;
    double vb = Y1 * rhs1; vb += Y2 * rhs2; vb += t33 * rhs3; vb += t34 * rhs4;
    double wb = Z1 * rhs1; wb += Z2 * rhs2; wb += t43 * rhs3; wb += t44 * rhs4;
    double ubar = xn * ub; ubar += yn * vb; ubar += zn * wb;
    const uint32_t idx = 4 * n;
    r[idx + 0] += area * B * ubar;
    r[idx + 1] += area * (ub * ubar + xn * pb);
    r[idx + 2] += area * (vb * ubar + yn * pb);
    r[idx + 3] += area * (wb * ubar + zn * pb);
  }
}

/* ---- src/timestep.c:21-154 ------------------------------------------------- */

void or_timestep(or_ctx *c, double cfl)
{
  const OR_GEOMETRY *g = c->g;
  const size_t N = c->N, E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  const double *x0 = g->e->xyzn->x0, *x1 = g->e->xyzn->x1, *x2 = g->e->xyzn->x2, *x3 = g->e->xyzn->x3;
  const double *q = c->q, B = BETA;
  double *cdt = c->cdt;
  memset(cdt, 0, N * sizeof(double));
  for (size_t i = 0; i < E; i++) {
    const double xn = x0[i], yn = x1[i], zn = x2[i], ln = x3[i];
    const double xnorm = xn * ln, ynorm = yn * ln, znorm = zn * ln;
    const uint32_t idx0 = 4 * n0[i], idx1 = 4 * n1[i];
    const double u = 0.5 * (q[idx0 + 1] + q[idx1 + 1]);
    const double v = 0.5 * (q[idx0 + 2] + q[idx1 + 2]);
    const double w = 0.5 * (q[idx0 + 3] + q[idx1 + 3]);
    const double ubar = xn * u + yn * v + zn * w;
    const double cc = sqrt(ubar * ubar + B);
    double term = u * xnorm; term += v * ynorm; term += w * znorm;
    term = fabs(term) + cc * ln;
    cdt[n0[i]] = cdt[n0[i]] + term;
    cdt[n1[i]] = cdt[n1[i]] + term;
  }
  for (int sec = 0; sec < 2; sec++) {
    const struct or_bntbl *bn = sec ? g->b->f : g->b->s;
    for (size_t i = 0; i < bn->sz; i++) {
      const uint32_t n = bn->nptr[i];
      const double xn = bn->xyz->x0[i], yn = bn->xyz->x1[i], zn = bn->xyz->x2[i];
      const double ln = sqrt(xn * xn + yn * yn + zn * zn);
      const double u = q[4 * n + 1], v = q[4 * n + 2], w = q[4 * n + 3];
      const double ubar = u * xn + v * yn + w * zn;
      const double ubar_ = ubar / ln;
      const double cc = sqrt(ubar_ * ubar_ + B);
      const double Vn = fabs(ubar) + cc * ln;
      cdt[n] += Vn;
    }
  }
  for (size_t i = 0; i < N; i++) cdt[i] /= cfl;
}

/* ---- src/fill.c:115-879 ---------------------------------------------------- */

void or_fill(or_ctx *c)
{
  const OR_GEOMETRY *g = c->g;
  const size_t N = c->N, E = g->e->sz;
  const uint32_t *n0 = g->e->eptr->n0, *n1 = g->e->eptr->n1;
  const double *x0 = g->e->xyzn->x0, *x1 = g->e->xyzn->x1, *x2 = g->e->xyzn->x2, *x3 = g->e->xyzn->x3;
  const double *q = c->q, *cdt = c->cdt;
  const double U = c->U, V = c->V, W = WINF, B = BETA, P = PINF;
  memset(c->A.a, 0, (size_t)c->A.i[N] * 16 * sizeof(double));
  for (size_t i = 0; i < N; i++) fill_diag(c, (uint32_t)i, (uint32_t)i, cdt[i]);
  for (size_t i = 0; i < E; i++) {
    const uint32_t node0 = n0[i], node1 = n1[i];
    const double xn = x0[i], yn = x1[i], zn = x2[i], ln = x3[i];
    double dot = xn, X1, Y1, Z1;
    if (fabs(dot) < (double)0.95f) { X1 = 1.0 - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn; }
    else { dot = yn; X1 = -dot * xn; Y1 = 1.0 - dot * yn; Z1 = -dot * zn; }
    double size = X1 * X1; size += Y1 * Y1; size += Z1 * Z1; size = sqrt(size);
    X1 /= size; Y1 /= size; Z1 /= size;
    double X2 = yn * Z1; X2 -= zn * Y1;
    double Y2 = zn * X1; Y2 -= xn * Z1;
    double Z2 = xn * Y1; Z2 -= yn * X1;
    const double uL = q[4 * node0 + 1], vL = q[4 * node0 + 2], wL = q[4 * node0 + 3];
    double ubarL = xn * uL; ubarL += yn * vL; ubarL += zn * wL;
    const double uR = q[4 * node1 + 1], vR = q[4 * node1 + 2], wR = q[4 * node1 + 3];
    double ubarR = xn * uR; ubarR += yn * vR; ubarR += zn * wR;
    const double u = 0.5 * (uL + uR), v = 0.5 * (vL + vR), w = 0.5 * (wL + wR);
    double ubar = xn * u; ubar += yn * v; ubar += zn * w;
    const double c2 = ubar * ubar + B;
    const double cc = sqrt(c2);
    const double eig1 = ln * fabs(ubar), eig2 = ln * fabs(ubar), eig3 = ln * fabs(ubar + cc), eig4 = ln * fabs(ubar - cc);
    double phi1 = xn * B; phi1 += u * ubar;
    double phi2 = yn * B; phi2 += v * ubar;
    double phi3 = zn * B; phi3 += w * ubar;
    double phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
    double phi5 = Z2 * phi1; phi5 -= X2 * phi3;
    double phi6 = X2 * phi2; phi6 -= Y2 * phi1;
    double phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
    double phi8 = X1 * phi3; phi8 -= Z1 * phi1;
    double phi9 = Y1 * phi1; phi9 -= X1 * phi2;
    const double c2inv = 1.0 / c2;
    double y11 = u * phi4; y11 += v * phi5; y11 += w * phi6; y11 = -c2inv * y11 / B;
    double y21 = u * phi7; y21 += v * phi8; y21 += w * phi9; y21 = -c2inv * y21 / B;
    double y31 = c2inv * (cc - ubar); y31 = 0.5 * y31 / B;
    double y41 = c2inv * (cc + ubar); y41 = -0.5 * y41 / B;
    const double y12 = c2inv * phi4, y22 = c2inv * phi7, y32 = c2inv * 0.5 * xn, y42 = c2inv * 0.5 * xn;
    const double y13 = c2inv * phi5, y23 = c2inv * phi8, y33 = c2inv * 0.5 * yn, y43 = c2inv * 0.5 * yn;
    const double y14 = c2inv * phi6, y24 = c2inv * phi9, y34 = c2inv * 0.5 * zn, y44 = c2inv * 0.5 * zn;
    const double t13 = cc * B;
    double t23 = u * (ubar + cc); t23 += xn * B;
    double t33 = v * (ubar + cc); t33 += yn * B;
    double t43 = w * (ubar + cc); t43 += zn * B;
    const double t14 = -cc * B;
    double t24 = u * (ubar - cc); t24 += xn * B;
    double t34 = v * (ubar - cc); t34 += yn * B;
    double t44 = w * (ubar - cc); t44 += zn * B;
    double a11 = eig3 * t13 * y31; a11 += eig4 * t14 * y41;
    double a12 = eig3 * t13 * y32; a12 += eig4 * t14 * y42;
    double a13 = eig3 * t13 * y33; a13 += eig4 * t14 * y43;
    double a14 = eig3 * t13 * y34; a14 += eig4 * t14 * y44;
    double a21 = eig1 * X1 * y11; a21 += eig2 * X2 * y21; a21 += eig3 * t23 * y31; a21 += eig4 * t24 * y41;
    double a22 = eig1 * X1 * y12; a22 += eig2 * X2 * y22; a22 += eig3 * t23 * y32; a22 += eig4 * t24 * y42;
    double a23 = eig1 * X1 * y13; a23 += eig2 * X2 * y23; a23 += eig3 * t23 * y33; a23 += eig4 * t24 * y43;
    double a24 = eig1 * X1 * y14; a24 += eig2 * X2 * y24; a24 += eig3 * t23 * y34; a24 += eig4 * t24 * y44;
    double a31 = eig1 * Y1 * y11; a31 += eig2 * Y2 * y21; a31 += eig3 * t33 * y31; a31 += eig4 * t34 * y41;
    double a32 = eig1 * Y1 * y12; a32 += eig2 * Y2 * y22; a32 += eig3 * t33 * y32; a32 += eig4 * t34 * y42;
    double a33 = eig1 * Y1 * y13; a33 += eig2 * Y2 * y23; a33 += eig3 * t33 * y33; a33 += eig4 * t34 * y43;
    double a34 = eig1 * Y1 * y14; a34 += eig2 * Y2 * y24; a34 += eig3 * t33 * y34; a34 += eig4 * t34 * y44;
    double a41 = eig1 * Z1 * y11; a41 += eig2 * Z2 * y21; a41 += eig3 * t43 * y31; a41 += eig4 * t44 * y41;
    double a42 = eig1 * Z1 * y12; a42 += eig2 * Z2 * y22; a42 += eig3 * t43 * y32; a42 += eig4 * t44 * y42;
    double a43 = eig1 * Z1 * y13; a43 += eig2 * Z2 * y23; a43 += eig3 * t43 * y33; a43 += eig4 * t44 * y43;
    double a44 = eig1 * Z1 * y14; a44 += eig2 * Z2 * y24; a44 += eig3 * t43 * y34; a44 += eig4 * t44 * y44;
    const double lb = ln * B, lx = ln * xn, ly = ln * yn, lz = ln * zn;
    double v0[16], v1[16];
    v0[0] = 0.5 * a11; v0[4] = 0.5 * (lx + a21); v0[8] = 0.5 * (ly + a31); v0[12] = 0.5 * (lz + a41);
    v0[1] = 0.5 * ((lb * xn) + a12); v0[5] = 0.5 * ((ln * (ubarL + xn * uL)) + a22);
    v0[9] = 0.5 * ((lx * vL) + a32); v0[13] = 0.5 * ((lx * wL) + a42);
    v0[2] = 0.5 * ((lb * yn) + a13); v0[6] = 0.5 * ((ly * uL) + a23);
    v0[10] = 0.5 * ((ln * (ubarL + yn * vL)) + a33); v0[14] = 0.5 * ((ly * wL) + a43);
    v0[3] = 0.5 * ((lb * zn) + a14); v0[7] = 0.5 * ((lz * uL) + a24);
    v0[11] = 0.5 * ((lz * vL) + a34); v0[15] = 0.5 * ((ln * (ubarL + zn * wL)) + a44);
    v1[0] = 0.5 * -a11; v1[4] = 0.5 * (lx - a21); v1[8] = 0.5 * (ly - a31); v1[12] = 0.5 * (lz - a41);
    v1[1] = 0.5 * ((lb * xn) - a12); v1[5] = 0.5 * ((ln * (ubarR + xn * uR)) - a22);
    v1[9] = 0.5 * ((lx * vR) - a32); v1[13] = 0.5 * ((lx * wR) - a42);
    v1[2] = 0.5 * ((lb * yn) - a13); v1[6] = 0.5 * ((ly * uR) - a23);
    v1[10] = 0.5 * ((ln * (ubarR + yn * vR)) - a33); v1[14] = 0.5 * ((ly * wR) - a43);
    v1[3] = 0.5 * ((lb * zn) - a14); v1[7] = 0.5 * ((lz * uR) - a24);
    v1[11] = 0.5 * ((lz * vR) - a34); v1[15] = 0.5 * ((ln * (ubarR + zn * wR)) - a44);
    fill_block(c, node0, node0, v0);
    fill_block(c, node0, node1, v1);
    fill_alpha(c, node1, node0, -1.0, v0);
    fill_alpha(c, node1, node1, -1.0, v1);
  }
  const struct or_bntbl *sb = g->b->s;
  for (size_t i = 0; i < sb->sz; i++) {
    const double v[3] = {sb->xyz->x0[i], sb->xyz->x1[i], sb->xyz->x2[i]};
    fill_boundary(c, sb->nptr[i], sb->nptr[i], 3, v);
  }
  const struct or_bntbl *fb = g->b->f;
  for (size_t i = 0; i < fb->sz; i++) {
    const uint32_t n = fb->nptr[i];
    double xn = fb->xyz->x0[i], yn = fb->xyz->x1[i], zn = fb->xyz->x2[i];
    const double ln = sqrt(xn * xn + yn * yn + zn * zn);
    xn /= ln; yn /= ln; zn /= ln;
    double dot = xn, X1, Y1, Z1;
    if (fabs(dot) < (double)0.95f) { X1 = 1.0 - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn; }
    else { dot = yn; X1 = -dot * xn; Y1 = 1.0 - dot * yn; Z1 = -dot * zn; }
    const double size = sqrt(X1 * X1 + Y1 * Y1 + Z1 * Z1);
    X1 /= size; Y1 /= size; Z1 /= size;
    const double X2 = yn * Z1 - zn * Y1, Y2 = zn * X1 - xn * Z1, Z2 = xn * Y1 - yn * X1;
    double ubar0 = xn * U; ubar0 += yn * V; ubar0 += zn * W;
    const double c20 = ubar0 * ubar0 + B;
    const double c0 = sqrt(c20);
    double phi1 = xn * B; phi1 += U * ubar0;
    double phi2 = yn * B; phi2 += V * ubar0;
    double phi3 = zn * B; phi3 += W * ubar0;
    double phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
    double phi5 = Z2 * phi1; phi5 -= X2 * phi3;
    double phi6 = X2 * phi2; phi6 -= Y2 * phi1;
    double phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
    double phi8 = X1 * phi3; phi8 -= Z1 * phi1;
    double phi9 = Y1 * phi1; phi9 -= X1 * phi2;
    const double t13 = c0 * B;
    double t23 = U * (ubar0 + c0); t23 += xn * B;
    double t33 = V * (ubar0 + c0); t33 += yn * B;
    double t43 = W * (ubar0 + c0); t43 += zn * B;
    const double t14 = -c0 * B;
    double t24 = U * (ubar0 - c0); t24 += xn * B;
    double t34 = V * (ubar0 - c0); t34 += yn * B;
    double t44 = W * (ubar0 - c0); t44 += zn * B;
    double ti11 = U * phi4; ti11 += V * phi5; ti11 += W * phi6; ti11 = -ti11 / B / c20;
    double ti21 = U * phi7; ti21 += V * phi8; ti21 += W * phi9; ti21 = -ti21 / B / c20;
    const double ti31 = (c0 - ubar0) / (2.0 * B * c20);
    const double ti41 = -(c0 + ubar0) / (2.0 * B * c20);
    const double ti12 = phi4 / c20, ti22 = phi7 / c20, ti32 = 0.5 * xn / c20, ti42 = 0.5 * xn / c20;
    const double ti13 = phi5 / c20, ti23 = phi8 / c20, ti33 = 0.5 * yn / c20, ti43 = 0.5 * yn / c20;
    const double ti14 = phi6 / c20, ti24 = phi9 / c20, ti34 = 0.5 * zn / c20, ti44 = 0.5 * zn / c20;
    const double pi = q[4 * n + 0], ui = q[4 * n + 1], vi = q[4 * n + 2], wi = q[4 * n + 3];
    double un = xn * ui; un += yn * vi; un += zn * wi;
    double pr, prp, ur, uru, vr, vrv, wr, wrw;
    if (un > 0.0) { pr = pi; prp = 1.0; ur = ui; uru = 1.0; vr = vi; vrv = 1.0; wr = wi; wrw = 1.0; }
    else { pr = P; prp = 0.0; ur = U; uru = 0.0; vr = V; vrv = 0.0; wr = W; wrw = 0.0; }
    double rhs1 = ti11 * pr; rhs1 += ti12 * ur; rhs1 += ti13 * vr; rhs1 += ti14 * wr;
    const double rhs1p = ti11 * prp, rhs1u = ti12 * uru, rhs1v = ti13 * vrv, rhs1w = ti14 * wrw;
    double rhs2 = ti21 * pr; rhs2 += ti22 * ur; rhs2 += ti23 * vr; rhs2 += ti24 * wr;
    const double rhs2p = ti21 * prp, rhs2u = ti22 * uru, rhs2v = ti23 * vrv, rhs2w = ti24 * wrw;
    double rhs3 = ti31 * pi; rhs3 += ti32 * ui; rhs3 += ti33 * vi; rhs3 += ti34 * wi;
    double rhs4 = ti41 * P; rhs4 += ti42 * U; rhs4 += ti43 * V; rhs4 += ti44 * W;
    double pb = t13 * rhs3; pb += t14 * rhs4;
    const double pbp = t13 * ti31, pbu = t13 * ti32, pbv = t13 * ti33, pbw = t13 * ti34;
    double ub = X1 * rhs1; ub += X2 * rhs2; ub += t23 * rhs3; ub += t24 * rhs4;
    double ubp = X1 * rhs1p; ubp += X2 * rhs2p; ubp += t23 * ti31;
    double ubu = X1 * rhs1u; ubu += X2 * rhs2u; ubu += t23 * ti32;
    double ubv = X1 * rhs1v; ubv += X2 * rhs2v; ubv += t23 * ti33;
    double ubw = X1 * rhs1w; ubw += X2 * rhs2w; ubw += t23 * ti34;
    double vb = Y1 * rhs1; vb += Y2 * rhs2; vb += t33 * rhs3; vb += t34 * rhs4;
    double vbp = Y1 * rhs1p; vbp += Y2 * rhs2p; vbp += t33 * ti31;
    double vbu = Y1 * rhs1u; vbu += Y2 * rhs2u; vbu += t33 * ti32;
    double vbv = Y1 * rhs1v; vbv += Y2 * rhs2v; vbv += t33 * ti33;
    double vbw = Y1 * rhs1w; vbw += Y2 * rhs2w; vbw += t33 * ti34;
    double wb = Z1 * rhs1; wb += Z2 * rhs2; wb += t43 * rhs3; wb += t44 * rhs4;
    double wbp = Z1 * rhs1p; wbp += Z2 * rhs2p; wbp += t43 * ti31;
    double wbu = Z1 * rhs1u; wbu += Z2 * rhs2u; wbu += t43 * ti32;
    double wbv = Z1 * rhs1v; wbv += Z2 * rhs2v; wbv += t43 * ti33;
    double wbw = Z1 * rhs1w; wbw += Z2 * rhs2w; wbw += t43 * ti34;
    double unb = xn * ub; unb += yn * vb; unb += zn * wb;
    double unbp = xn * ubp; unbp += yn * vbp; unbp += zn * wbp;
    double unbu = xn * ubu; unbu += yn * vbu; unbu += zn * wbu;
    double unbv = xn * ubv; unbv += yn * vbv; unbv += zn * wbv;
    double unbw = xn * ubw; unbw += yn * vbw; unbw += zn * wbw;
    double v[16];
    v[0] = ln * B * unbp;
    v[4] = ln * (ub * unbp + unb * ubp + xn * pbp);
    v[8] = ln * (vb * unbp + unb * vbp + yn * pbp);
    v[12] = ln * (wb * unbp + unb * wbp + zn * pbp);
    v[1] = ln * B * unbu;
    v[5] = ln * (ub * unbu + unb * ubu + xn * pbu);
    v[9] = ln * (vb * unbu + unb * vbu + yn * pbu);
    v[13] = ln * (wb * unbu + unb * wbu + zn * pbu);
    v[2] = ln * B * unbv;
    v[6] = ln * (ub * unbv + unb * ubv + xn * pbv);
    v[10] = ln * (vb * unbv + unb * vbv + yn * pbv);
    v[14] = ln * (wb * unbv + unb * wbv + zn * pbv);
    v[3] = ln * B * unbw;
    v[7] = ln * (ub * unbw + unb * ubw + xn * pbw);
    v[11] = ln * (vb * unbw + unb * vbw + yn * pbw);
    v[15] = ln * (wb * unbw + unb * wbw + zn * pbw);
    fill_block(c, n, n, v);
  }
}

/* ---- matrix/ilu.cpp:49-138 (one thread) and a textbook variant -------------- */

void or_ilu(or_ctx *c)
{
  const size_t N = c->N;
  or_mat *A = &c->A, *L = &c->L, *Ut = &c->Ut;
  double a[16], d[16], T[16];
  if (c->o.ilu_textbook) {
    /* IKJ ILU(0) without the (a - d) + d rounding of the copy side effects */
    for (size_t i = 0; i < N; i++) {
      for (uint32_t j = A->i[i]; j < A->d[i]; j++) {
        const uint32_t k = A->j[j];
        double *lik = A->a + (size_t)j * 16;
        dgemm4(lik, A->a + (size_t)A->d[k] * 16, T);
        memcpy(lik, T, sizeof(T));
        for (uint32_t jj = A->d[k] + 1; jj < A->i[k + 1]; jj++) {
          const uint32_t col = A->j[jj];
          for (uint32_t t = j + 1; t < A->i[i + 1]; t++)
            if (A->j[t] == col) {
              dgemm4(T, A->a + (size_t)jj * 16, d);
              axpy16(-1.0, d, A->a + (size_t)t * 16);
              break;
            }
        }
      }
      or_inv4(A->a + (size_t)A->d[i] * 16, T);
      memcpy(A->a + (size_t)A->d[i] * 16, T, sizeof(T));
    }
    return;
  }
  for (size_t i = 0; i < N; i++)
    for (uint32_t j = A->i[i]; j < A->i[i + 1]; j++) {
      const uint32_t k = A->j[j];
      uint32_t il = L->i[i], iu = Ut->i[k];
      memcpy(a, A->a + (size_t)j * 16, sizeof(a));
      memset(d, 0, sizeof(d));
      while (il < L->i[i + 1] && iu < Ut->i[k + 1]) {
        const uint32_t jl = L->j[il], ju = Ut->j[iu];
        if (jl == ju) {
          dgemm4(L->a + (size_t)il * 16, Ut->a + (size_t)iu * 16, d);
          axpy16(-1.0, d, a);
        }
        if (jl <= ju) il++;
        if (jl >= ju) iu++;
      }
      axpy16(1.0, d, a);
      if (i > k) {
        or_inv4(Ut->a + (size_t)(Ut->i[k + 1] - 1) * 16, T);
        dgemm4(a, T, L->a + (size_t)(il - 1) * 16);
      } else {
        memcpy(Ut->a + (size_t)(iu - 1) * 16, a, sizeof(a));
      }
    }
  for (size_t i = 0; i < N; i++) {
    uint32_t idxL = L->i[i];
    for (uint32_t j = A->i[i]; j < A->i[i + 1]; j++) {
      if (A->j[j] < i) memcpy(A->a + (size_t)j * 16, L->a + (size_t)(idxL++) * 16, 16 * sizeof(double));
      if (A->j[j] >= i) memcpy(A->a + (size_t)j * 16, Ut->a + (size_t)c->ut_index[j] * 16, 16 * sizeof(double));
    }
    or_inv4(A->a + (size_t)A->d[i] * 16, T);
    memcpy(A->a + (size_t)A->d[i] * 16, T, sizeof(T));
  }
}

/* ---- matrix/sptrsv.cpp:34-73 ------------------------------------------------ */

void or_sptrsv(or_ctx *c, const double *b, double *x)
{
  const or_mat *A = &c->A;
  double t[4];
  for (uint32_t l = 0; l < c->tl_n; l++)
    for (uint32_t i = c->tl_i[l]; i < c->tl_i[l + 1]; i++) {
      const uint32_t row = c->tl_j[i];
      memcpy(x + 4 * row, b + 4 * row, 4 * sizeof(double));
      for (uint32_t j = A->i[row]; j < A->d[row]; j++) {
        dgemv4(A->a + (size_t)j * 16, x + 4 * A->j[j], t);
        axpy4(-1.0, t, x + 4 * row);
      }
    }
  for (uint32_t l = 0; l < c->tu_n; l++)
    for (uint32_t i = c->tu_i[l]; i < c->tu_i[l + 1]; i++) {
      const uint32_t row = c->tu_j[i];
      for (uint32_t j = A->d[row] + 1; j < A->i[row + 1]; j++) {
        dgemv4(A->a + (size_t)j * 16, x + 4 * A->j[j], t);
        axpy4(-1.0, t, x + 4 * row);
      }
      dgemv4(A->a + (size_t)A->d[row] * 16, x + 4 * row, t);
      memcpy(x + 4 * row, t, 4 * sizeof(double));
    }
}

/* fun3d::Matrix::sptrsv_Jacobi (matrix/sptrsv.cpp:75-132) on the in-place
 * factors: L = the strictly lower blocks, U = the strictly upper ones, D^-1
 * the inverted diagonal block.  (The reference's routine reads separate L
 * and U copies; the copy-back branch of its ILU that would fill U,
 * matrix/ilu.cpp:106-119, is compiled out, so its factors live in A as
 * here.)  `sweeps` >= 1 passes each way, as its do-while:
 *   x = b;  repeat: c_i = -1 * (sum_j dgemv(L_ij, x_j)) + b_i;  x = c
 *   x' = dgemv(D_i^-1, x_i);  repeat: c_i = dgemv(D_i^-1, -1 * (sum_j
 *   dgemv(U_ij, x'_j)) + x_i);  x' = c
 * every sum starting from zero and taking the row's blocks in ascending
 * column, each term added as axpy(1, t, sum) (matrix/init.cpp:96-142). */
void or_sptrsv_jacobi(or_ctx *c, const double *b, double *x, int sweeps)
{
  const or_mat *A = &c->A;
  const size_t n4 = 4 * c->N;
  double *xl = (double *)malloc(n4 * sizeof(double)), *cc = (double *)malloc(n4 * sizeof(double));
  if (!xl || !cc) abort();
  double t[4], s[4];
  memcpy(xl, b, n4 * sizeof(double));
  int sweep = 0;
  do {
    for (size_t i = 0; i < c->N; i++) {
      memset(s, 0, sizeof(s));
      for (uint32_t j = A->i[i]; j < A->d[i]; j++) {
        dgemv4(A->a + (size_t)j * 16, xl + 4 * (size_t)A->j[j], t);
        axpy4(1.0, t, s);
      }
      for (int k = 0; k < 4; k++) { const double ax = -1.0 * s[k]; cc[4 * i + k] = ax + b[4 * i + k]; }
    }
    memcpy(xl, cc, n4 * sizeof(double));
  } while (++sweep < sweeps);
  for (size_t i = 0; i < c->N; i++) dgemv4(A->a + (size_t)A->d[i] * 16, xl + 4 * i, x + 4 * i);
  sweep = 0;
  do {
    for (size_t i = 0; i < c->N; i++) {
      memset(s, 0, sizeof(s));
      for (uint32_t j = A->d[i] + 1; j < A->i[i + 1]; j++) {
        dgemv4(A->a + (size_t)j * 16, x + 4 * (size_t)A->j[j], t);
        axpy4(1.0, t, s);
      }
      double buf[4];
      for (int k = 0; k < 4; k++) { const double ax = -1.0 * s[k]; buf[k] = ax + xl[4 * i + k]; }
      dgemv4(A->a + (size_t)A->d[i] * 16, buf, cc + 4 * i);
    }
    memcpy(x, cc, n4 * sizeof(double));
  } while (++sweep < sweeps);
  free(xl);
  free(cc);
}

/* y = A x with the blocks as they stand (the fill's output before or_ilu).
 * No counterpart in the reference (its Krylov operator is matrix-free); the
 * order is the device product's: per row, the blocks in ascending column
 * (the CSR order of matrix/init.cpp:7-74), each component summed
 * a[c][0] x0 + a[c][1] x1 + a[c][2] x2 + a[c][3] x3 into one running sum. */
void or_spmv(const or_ctx *c, const double *x, double *y)
{
  const or_mat *A = &c->A;
  for (size_t row = 0; row < c->N; row++)
    for (int k = 0; k < 4; k++) {
      double acc = 0.0;
      for (uint32_t j = A->i[row]; j < A->i[row + 1]; j++) {
        const double *a = A->a + (size_t)j * 16 + 4 * k, *xc = x + 4 * (size_t)A->j[j];
        acc = acc + a[0] * xc[0];
        acc = acc + a[1] * xc[1];
        acc = acc + a[2] * xc[2];
        acc = acc + a[3] * xc[3];
      }
      y[4 * row + k] = acc;
    }
}

/* ---- src/forces.c:8-99 ------------------------------------------------------ */

void or_forces(or_ctx *c, double *forces)
{
  const OR_GEOMETRY *g = c->g;
  const double *X = g->n->xyz->x0, *Y = g->n->xyz->x1, *Z = g->n->xyz->x2, *q = c->q;
  const double U = c->U, V = c->V;
  double lift = 0.0, drag = 0.0, momn = 0.0;
  /* src/forces.c:33: reduction(+) per colour -- each colour sums into a
   * zero-initialised private copy that is then added to the total */
  for (size_t cl = 0; cl < g->t->sz; cl++) {
    const double lift_s = lift, drag_s = drag, momn_s = momn;
    lift = 0.0; drag = 0.0; momn = 0.0;
    for (uint32_t j = g->t->i[cl]; j < g->t->i[cl + 1]; j++) {
      const uint32_t n0 = g->b->fc->fptr->n0[j], n1 = g->b->fc->fptr->n1[j], n2 = g->b->fc->fptr->n2[j];
      const double x0 = X[n0], y0 = Y[n0], z0 = Z[n0];
      const double x1 = X[n1], y1 = Y[n1], z1 = Z[n1];
      const double x2 = X[n2], y2 = Y[n2], z2 = Z[n2];
      const double ax = x1 - x0, ay = y1 - y0, az = z1 - z0;
      const double bx = x2 - x0, by = y2 - y0, bz = z2 - z0;
      double xnorm = ay * bz; xnorm -= az * by; xnorm = -0.5 * xnorm;
      double ynorm = ax * bz; ynorm -= az * bx; ynorm = 0.5 * ynorm;
      const double p0 = q[4 * n0], p1 = q[4 * n1], p2 = q[4 * n2];
      const double press = (p0 + p1 + p2) / 3.0;
      const double cp = 2.0 * (press - 1.0);
      const double dcx = cp * xnorm, dcy = cp * ynorm;
      const double xmid = x0 + x1 + x2, ymid = y0 + y1 + y2;
      lift = lift - dcx * V + dcy * U;
      drag = drag + dcx * U + dcy * V;
      momn = momn + (xmid - 0.25) * dcy - ymid * dcx;
    }
    lift = lift_s + lift; drag = drag_s + drag; momn = momn_s + momn;
  }
  forces[0] = lift; forces[1] = drag; forces[2] = momn;
}

/* ---- src/utils.c:53-123 and src/kernel.c:24-45 ------------------------------- */

double or_norm(size_t n, const double *v)
{
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += v[i] * v[i];
  return sqrt(s);
}

/* the HIP path's reduction order: 1024 blocks x 256 threads, each thread a
 * strided sum over element pairs (2p, 2p + 1; the half pair of an odd
 * length last), an LDS tree per block, then a 256-thread strided sum of the
 * 1024 partials and one more tree (a check of the device, not the reference) */
static double tree256(double *sh)
{
  for (int s = 128; s > 0; s >>= 1)
    for (int t = 0; t < s; t++) sh[t] = sh[t] + sh[t + s];
  return sh[0];
}

static double dot_blocked(size_t n, const double *x, const double *y)
{
  static double part[1024];
  double sh[256];
  const size_t stride = 1024 * 256;
  for (int b = 0; b < 1024; b++) {
    for (int t = 0; t < 256; t++) {
      double acc = 0.0;
      for (size_t p = (size_t)b * 256 + t; 2 * p < n; p += stride) {
        acc += x[2 * p] * y[2 * p];
        if (2 * p + 1 < n) acc += x[2 * p + 1] * y[2 * p + 1];
      }
      sh[t] = acc;
    }
    part[b] = tree256(sh);
  }
  for (int t = 0; t < 256; t++) {
    double acc = 0.0;
    for (int i = t; i < 1024; i += 256) acc += part[i];
    sh[t] = acc;
  }
  return tree256(sh);
}

double or_dot_tree(size_t n, const double *x, const double *y) { return dot_blocked(n, x, y); }

static double ctx_dot(const or_ctx *c, size_t n, const double *x, const double *y)
{
  if (c->o.reduce_order) return dot_blocked(n, x, y);
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

static double ctx_norm(const or_ctx *c, size_t n, const double *v) { return sqrt(ctx_dot(c, n, v, v)); }

static void axpy_n(size_t n, double a, const double *x, double *y)
{
  for (size_t i = 0; i < n; i++) { const double ax = a * x[i]; y[i] = ax + y[i]; }
}

static void newaxpy_n(size_t n, double a, const double *x, const double *y, double *w)
{
  for (size_t i = 0; i < n; i++) { const double ax = a * x[i]; w[i] = ax + y[i]; }
}

static double normalize_n(const or_ctx *c, size_t n, double *x)
{
  const double nrm = ctx_norm(c, n, x);
  const double s = 1.0 / nrm;
  for (size_t i = 0; i < n; i++) x[i] *= s;
  return nrm;
}

#define FD_EPS 1.490116119384766e-08   /* src/kernel.c:12 */

void or_ijacvec(or_ctx *c, const double *x, const double *r, double norm, double *gx, double *gy, double *gz,
                double *w, double *y)
{
  const size_t sz = 4 * c->N;
  const double xnorm = ctx_norm(c, sz, x);
  const double h = FD_EPS * norm;
  newaxpy_n(sz, (h / xnorm), x, c->q, w);
  or_flux(c, w, gx, gy, gz, y);
  const double a = xnorm / h;
  for (size_t i = 0; i < c->N; i++)
    for (int k = 0; k < 4; k++) {
      const size_t idx = 4 * i + k;
      y[idx] = (y[idx] - r[idx] + c->cdt[i] * (w[idx] - c->q[idx])) * a;
    }
}

/* ---- src/kernel.c:47-217 ---------------------------------------------------- */

int or_kernel(or_ctx *c, or_step *rec, int max_rec)
{
  const size_t sz = 4 * c->N;
  const int restart = c->o.gmres_restart;
  const int nb = restart + 1;
  double *ghh = xcalloc((size_t)nb * nb + (size_t)nb, sizeof(double));
  double **vecs = xmalloc(nb * sizeof(double *));
  for (int v = 0; v < nb; v++) vecs[v] = xcalloc(sz, sizeof(double));
  double cc[64], ss[64], rs[64];
  double fnorm_init = 0.0, cfl = c->o.cfl_init;
  double *x = c->x, *r = c->r, *w = c->w, *t = c->t;
  int its = 0, nrec = 0;
  while (its < c->o.ts_max) {
    or_flux(c, c->q, c->grad[0], c->grad[1], c->grad[2], r);
    const double fnorm = ctx_norm(c, sz, r);
    if (its == 0) fnorm_init = fnorm;
    else {
      double cfl0 = 1.1 * c->o.cfl_init * fnorm_init;
      cfl0 /= fnorm;
      cfl = (cfl0 < c->o.cfl_max) ? cfl0 : c->o.cfl_max;
    }
    or_timestep(c, cfl);
    or_fill(c);
    const double qnorm = sqrt(1.0 + ctx_norm(c, sz, c->q));
    memset(x, 0, sz * sizeof(double));
    or_ilu(c);
    or_sptrsv(c, r, vecs[0]);
    double tol = 0.0;
    int total = 0, flag = 0;
    while (total <= c->o.gmres_max_it) {
      if (total > 0) {
        or_ijacvec(c, x, r, qnorm, c->grad[0], c->grad[1], c->grad[2], w, t);
        newaxpy_n(sz, -1.0, t, r, w);
        or_sptrsv(c, w, vecs[0]);
      }
      rs[0] = normalize_n(c, sz, vecs[0]);
      tol = (total == 0) ? c->o.gmres_rtol * rs[0] : tol;
      int it = 0;
      while (it < restart) {
        or_ijacvec(c, vecs[it], r, qnorm, c->grad[0], c->grad[1], c->grad[2], w, t);
        or_sptrsv(c, t, vecs[1 + it]);
        double *hh = ghh + (size_t)it * (nb + 1);
        for (int v = 0; v < it + 1; v++) {
          hh[v] = ctx_dot(c, sz, vecs[it + 1], vecs[v]);
          axpy_n(sz, -hh[v], vecs[v], vecs[it + 1]);
        }
        hh[it + 1] = normalize_n(c, sz, vecs[1 + it]);
        for (int i = 0; i < it; i++) {
          const double tt = *hh;
          *hh = cc[i] * *hh + ss[i] * *(hh + 1);
          hh++;
          *hh = cc[i] * *hh - ss[i] * tt;
        }
        const double tt = sqrt(*hh * *hh + *(hh + 1) * *(hh + 1));
        const double ct = *hh / tt, st = *(hh + 1) / tt;
        cc[it] = ct; ss[it] = st;
        rs[it + 1] = -(st * rs[it]);
        rs[it] = ct * rs[it];
        *hh = ct * *hh + st * *(hh + 1);
        const double res = fabs(rs[it + 1]);
        it++;
        if (res <= tol) { flag = 1; break; }
      }
      total += it;
      rs[it - 1] /= ghh[(size_t)(it - 1) * (nb + 1) + (it - 1)];
      for (int i = 1; i <= it - 1; i++) {
        const int k = (it - 1) - i;
        double tt = rs[k];
        for (int j = k + 1; j <= it - 1; j++) tt -= ghh[(size_t)j * (nb + 1) + k] * rs[j];
        rs[k] = tt / ghh[(size_t)k * (nb + 1) + k];
      }
      for (int v = 0; v < it; v++) axpy_n(sz, rs[v], vecs[v], x);
      if (flag) break;
    }
    axpy_n(sz, -1.0, x, c->q);
    double forces[3];
    or_forces(c, forces);
    if (nrec < max_rec && rec) {
      rec[nrec].step = its; rec[nrec].cfl = cfl; rec[nrec].fnorm = fnorm; rec[nrec].krylov = total;
      memcpy(rec[nrec].forces, forces, sizeof(forces));
    }
    nrec++;
    /* src/kernel.c:214 rounds the ratio; compared in double (no uint32 wrap) */
    if (round(fnorm_init / fnorm) >= c->o.fnorm_ratio) break;
    its++;
  }
  for (int v = 0; v < nb; v++) free(vecs[v]);
  free(vecs);
  free(ghh);
  return nrec;
}
