/*
 * ref_driver.c -- test harness linked against the reference's own path
 * sources (compiled from /root/reference by oracle/Makefile into
 * oracle/_ref/).  TEST INFRASTRUCTURE ONLY: it writes golden vectors for
 * the oracle and the HIP path, and times the reference's CPU path.
 *
 * The reference's Setup (src/mesh.c:22-154) is repeated here step by step
 * with one change: the owner partition of src/subdomains.c:117-190 comes
 * from contiguous vertex ranges (v * T / N) instead of METIS, so the METIS
 * call site is never linked (one thread: part == 0, exactly the reference's
 * one-thread partition).  Everything else runs the reference's functions.
 *
 *   kref setup   <mesh> <outdir>                 setup arrays
 *   kref kernels <mesh> <outdir> <seed> [zerow]  one call of each kernel
 *   kref solve   <mesh> <outdir> <threads>       Kernel() history (%.17g)
 *   kref time    <mesh> <threads> <reps>         per-kernel CPU times (JSON)
 */
#include <assert.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "allocator.h"
#include "core_kernel.h"
#include "geometry.h"
#include "mesh.h"
#include "mesh2geo.h"
#include "utils.h"

extern void *CXX_Matrix_Init(const size_t, const size_t, const unsigned int, const unsigned int,
                             const unsigned int *, const unsigned int *);
extern void CXX_Walk_Int(char *l, const char *h, const size_t sz, unsigned int *b);
extern void iJacVec(const GEOMETRY *, const double *, const double *, const double, GRADIENT *, double *, double *);
extern void Kernel(GEOMETRY *);
extern void fun3d_init_log(void);
/* oracle/ref_peek.cpp */
extern size_t kref_nnzb(void *m);
extern void kref_blocks(void *m, double *out);
extern void kref_pattern(void *m, unsigned int *rowptr, unsigned int *col);

/* ---- src/subdomains.c:117-190 with a contiguous partition ------------------ */

static void partition(GEOMETRY *g, uint32_t T)
{
  const size_t N = g->n->sz;
  uint32_t *part = (uint32_t *)fun3d_calloc(N, sizeof(uint32_t));
  for (size_t v = 0; v < N; v++) part[v] = (uint32_t)((v * (size_t)T) / N);
  uint32_t *ie = (uint32_t *)fun3d_calloc(T + 1, sizeof(uint32_t));
  for (size_t i = 0; i < g->e->sz; i++) {
    uint32_t t0 = part[g->e->eptr->n0[i]], t1 = part[g->e->eptr->n1[i]];
    ie[t0]++;
    if (t0 != t1) ie[t1]++;
  }
  uint32_t sum = 0;
  for (uint32_t i = 0; i < T; i++) { uint32_t t = ie[i]; ie[i] = sum; sum += t; }
  ie[T] = sum;
  g->s->i = ie;
  g->s->sz = T;
  g->n->part = part;
  /* replicate cut edges into both owners' ranges (emalloc_sub) */
  struct etbl *e = (struct etbl *)fun3d_malloc(1, sizeof(struct etbl));
  e->eptr = (struct edge *)fun3d_malloc(1, sizeof(struct edge));
  e->xyzn = (struct xyzn *)fun3d_malloc(1, sizeof(struct xyzn));
  e->sz = ie[T];
  e->eptr->n0 = (uint32_t *)fun3d_malloc(e->sz, sizeof(uint32_t));
  e->eptr->n1 = (uint32_t *)fun3d_malloc(e->sz, sizeof(uint32_t));
  double **dst[4] = {&e->xyzn->x0, &e->xyzn->x1, &e->xyzn->x2, &e->xyzn->x3};
  double *src[4] = {g->e->xyzn->x0, g->e->xyzn->x1, g->e->xyzn->x2, g->e->xyzn->x3};
  for (int k = 0; k < 4; k++) *dst[k] = (double *)fun3d_malloc(e->sz, sizeof(double));
  uint32_t *buf = (uint32_t *)fun3d_malloc(T, sizeof(uint32_t));
  memcpy(buf, ie, T * sizeof(uint32_t));
  for (size_t i = 0; i < g->e->sz; i++) {
    uint32_t t0 = part[g->e->eptr->n0[i]], t1 = part[g->e->eptr->n1[i]];
    uint32_t owners[2] = {t0, t1};
    for (int o = 0; o < (t0 != t1 ? 2 : 1); o++) {
      uint32_t at = buf[owners[o]]++;
      e->eptr->n0[at] = g->e->eptr->n0[i];
      e->eptr->n1[at] = g->e->eptr->n1[i];
      for (int k = 0; k < 4; k++) (*dst[k])[at] = src[k][i];
    }
  }
  fun3d_free(buf);
  g->e = e;   /* the unpartitioned table is left allocated (harness only) */
}

/* ---- src/mesh.c:22-154 ------------------------------------------------------ */

static void ref_setup(const char *fpath, GEOMETRY *g, uint32_t T)
{
  FILE *f = fopen(fpath, "rb");
  assert(f);
  fseek(f, 0, SEEK_END);
  size_t fsz = (size_t)ftell(f);
  rewind(f);
  char *fbuf = (char *)fun3d_malloc(fsz, 1);
  size_t got = fread(fbuf, 1, fsz, f);
  assert(got == fsz);
  fclose(f);
  uint32_t p[13];
  CXX_Walk_Int(fbuf, fbuf + 52, 13, p);
  size_t bytes = 52;
  struct etbl *e = (struct etbl *)fun3d_malloc(1, sizeof(struct etbl));
  struct edge *eptr = (struct edge *)fun3d_malloc(1, sizeof(struct edge));
  struct xyzn *xyzn = (struct xyzn *)fun3d_malloc(1, sizeof(struct xyzn));
  e->sz = p[2];
  eptr->n0 = (uint32_t *)fun3d_malloc(e->sz, sizeof(uint32_t));
  eptr->n1 = (uint32_t *)fun3d_malloc(e->sz, sizeof(uint32_t));
  xyzn->x0 = (double *)fun3d_malloc(e->sz, sizeof(double));
  xyzn->x1 = (double *)fun3d_malloc(e->sz, sizeof(double));
  xyzn->x2 = (double *)fun3d_malloc(e->sz, sizeof(double));
  xyzn->x3 = (double *)fun3d_malloc(e->sz, sizeof(double));
  e->eptr = eptr;
  e->xyzn = xyzn;
  bytes += emalloc(fbuf + bytes, e);
  g->e = e;
  struct ntbl *n = (struct ntbl *)fun3d_malloc(1, sizeof(struct ntbl));
  n->sz = p[1];
  n->xyz = (struct xyz *)fun3d_malloc(1, sizeof(struct xyz));
  n->xyz->x0 = (double *)fun3d_malloc(n->sz, sizeof(double));
  n->xyz->x1 = (double *)fun3d_malloc(n->sz, sizeof(double));
  n->xyz->x2 = (double *)fun3d_malloc(n->sz, sizeof(double));
  bytes += nmalloc(fbuf + bytes, n);
  g->n = n;
  struct btbl *b = (struct btbl *)fun3d_malloc(1, sizeof(struct btbl));
  bytes += bmalloc(p, fbuf + bytes, b);
  g->b = b;
  assert(bytes == fsz);
  fun3d_free(fbuf);
  g->c = (struct ctbl *)fun3d_malloc(1, sizeof(struct ctbl));
  m2csr(g);
  g->matrix = CXX_Matrix_Init(g->n->sz, g->e->sz, g->c->b, T, g->e->eptr->n0, g->e->eptr->n1);
  g->s = (struct stbl *)fun3d_malloc(1, sizeof(struct stbl));
  partition(g, T);
  g->t = (struct stbl *)fun3d_malloc(1, sizeof(struct stbl));
  ifcoloring(g);
  wmalloc(g);
  g->q = (struct qtbl *)fun3d_malloc(1, sizeof(struct qtbl));
  g->q->q = (double *)fun3d_malloc(g->c->sz, sizeof(double));
  iguess(g);
}

/* ---- dumps ------------------------------------------------------------------ */

static const char *g_out;

static void dump(const char *name, const void *p, size_t bytes)
{
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s.bin", g_out, name);
  FILE *f = fopen(path, "wb");
  assert(f);
  if (bytes) assert(fwrite(p, 1, bytes, f) == bytes);
  fclose(f);
}

static uint64_t rng_state;
static double urand(void)
{
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* ---- history capture around Kernel() (linker --wrap) ----------------------- */

static int g_log = 0;
static FILE *g_hist = NULL;
static const GEOMETRY *g_geo = NULL;

extern void __real_ComputeFlux(const GEOMETRY *, const double *, GRADIENT *, double *);
extern void __real_ComputeTimeStep(const double, GEOMETRY *);
extern void __real_ComputeForces(const GEOMETRY *, double *);

void __wrap_ComputeFlux(const GEOMETRY *g, const double *q, GRADIENT *grad, double *r)
{
  __real_ComputeFlux(g, q, grad, r);
  if (g_log && g_hist && q == g->q->q) {
    double s = 0.0;   /* the one-thread sum of src/utils.c:61-62 */
    for (size_t i = 0; i < g->c->sz; i++) s += r[i] * r[i];
    fprintf(g_hist, "fnorm %.17g\n", sqrt(s));
  }
}

void __wrap_ComputeTimeStep(const double cfl, GEOMETRY *g)
{
  if (g_log && g_hist) fprintf(g_hist, "cfl %.17g\n", cfl);
  __real_ComputeTimeStep(cfl, g);
}

void __wrap_ComputeForces(const GEOMETRY *g, double *forces)
{
  __real_ComputeForces(g, forces);
  if (g_log && g_hist) fprintf(g_hist, "forces %.17g %.17g %.17g\n", forces[0], forces[1], forces[2]);
}

static double now(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

#include "phy.h"  /* last: it defines the one-letter macros B P U V W */

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: kref setup|kernels|solve|time <mesh> ...\n");
    return 2;
  }
  const char *mode = argv[1], *mesh = argv[2];
  omp_set_dynamic(0);
  fun3d_init_log();   /* src/main.c:27: the timers every kernel logs into */
  if (!strcmp(mode, "time")) {
    const uint32_t T = (uint32_t)atoi(argv[3]);
    const int reps = atoi(argv[4]);
    omp_set_num_threads((int)T);
    GEOMETRY g;
    ref_setup(mesh, &g, T);
    omp_set_num_threads((int)T);
    const size_t sz = g.c->sz;
    GRADIENT grad;
    grad.x0 = (double *)fun3d_calloc(sz, sizeof(double));
    grad.x1 = (double *)fun3d_calloc(sz, sizeof(double));
    grad.x2 = (double *)fun3d_calloc(sz, sizeof(double));
    double *r = (double *)fun3d_calloc(sz, sizeof(double));
    double *x = (double *)fun3d_calloc(sz, sizeof(double));
    ComputeFlux(&g, g.q->q, &grad, r);
    double t0 = now();
    for (int k = 0; k < reps; k++) ComputeFlux(&g, g.q->q, &grad, r);
    double t_flux = (now() - t0) / reps;
    ComputeTimeStep(50.0, &g);
    t0 = now();
    ComputeA(&g);
    double t_fill = now() - t0;
    t0 = now();
    ComputeNumericalILU(&g);
    double t_ilu = now() - t0;
    t0 = now();
    for (int k = 0; k < reps; k++) ComputeSparseTriangularSolve(&g, r, x);
    double t_trsv = (now() - t0) / reps;
    /* one BLAS-1 pass over a state-length vector: ComputeAXPY (src/utils.c:69-87) */
    t0 = now();
    for (int k = 0; k < 10 * reps; k++) ComputeAXPY(sz, 1e-3, r, x);
    double t_blas = (now() - t0) / (10 * reps);
    printf("{\"threads\": %u, \"flux_s\": %.9g, \"fill_s\": %.9g, \"ilu_s\": %.9g, \"sptrsv_s\": %.9g, "
           "\"blas_pass_s\": %.9g, \"reps\": %d}\n", T, t_flux, t_fill, t_ilu, t_trsv, t_blas, reps);
    return 0;
  }
  g_out = argv[3];
  mkdir(g_out, 0755);
  if (!strcmp(mode, "setup")) {
    omp_set_num_threads(1);
    GEOMETRY g;
    ref_setup(mesh, &g, 1);
    const size_t E = g.e->sz, N = g.n->sz, S = g.b->fc->sz;
    dump("e_n0", g.e->eptr->n0, E * 4);
    dump("e_n1", g.e->eptr->n1, E * 4);
    dump("e_x0", g.e->xyzn->x0, E * 8);
    dump("e_x1", g.e->xyzn->x1, E * 8);
    dump("e_x2", g.e->xyzn->x2, E * 8);
    dump("e_x3", g.e->xyzn->x3, E * 8);
    dump("w0_x0", g.e->w->w0->x0, E * 8);
    dump("w0_x1", g.e->w->w0->x1, E * 8);
    dump("w0_x2", g.e->w->w0->x2, E * 8);
    dump("w1_x0", g.e->w->w1->x0, E * 8);
    dump("w1_x1", g.e->w->w1->x1, E * 8);
    dump("w1_x2", g.e->w->w1->x2, E * 8);
    dump("n_x0", g.n->xyz->x0, N * 8);
    dump("fc_n0", g.b->fc->fptr->n0, S * 4);
    dump("fc_n1", g.b->fc->fptr->n1, S * 4);
    dump("fc_n2", g.b->fc->fptr->n2, S * 4);
    dump("t_i", g.t->i, (g.t->sz + 1) * 4);
    dump("s_nptr", g.b->s->nptr, g.b->s->sz * 4);
    dump("s_x0", g.b->s->xyz->x0, g.b->s->sz * 8);
    dump("f_nptr", g.b->f->nptr, g.b->f->sz * 4);
    dump("f_x2", g.b->f->xyz->x2, g.b->f->sz * 8);
    const double uv[2] = {U, V};
    dump("phy_uv", uv, sizeof(uv));
    size_t nnzb = kref_nnzb(g.matrix);
    unsigned int *rp = (unsigned int *)malloc((N + 1) * 4), *col = (unsigned int *)malloc(nnzb * 4);
    kref_pattern(g.matrix, rp, col);
    dump("a_rowptr", rp, (N + 1) * 4);
    dump("a_col", col, nnzb * 4);
    return 0;
  }
  if (!strcmp(mode, "kernels")) {
    omp_set_num_threads(1);
    GEOMETRY g;
    ref_setup(mesh, &g, 1);
    rng_state = (uint64_t)strtoull(argv[4], NULL, 10);
    const int zerow = argc > 5 && !strcmp(argv[5], "zerow");
    const size_t sz = g.c->sz, E = g.e->sz, N = g.n->sz;
    if (zerow) {
      memset(g.e->w->w0->x0, 0, E * 8); memset(g.e->w->w0->x1, 0, E * 8); memset(g.e->w->w0->x2, 0, E * 8);
      memset(g.e->w->w1->x0, 0, E * 8); memset(g.e->w->w1->x1, 0, E * 8); memset(g.e->w->w1->x2, 0, E * 8);
    }
    /* state: free stream + 0.05 U(-1,1), then the same for a second state */
    for (size_t i = 0; i < sz; i++) g.q->q[i] += 0.05 * (2.0 * urand() - 1.0);
    dump("q", g.q->q, sz * 8);
    GRADIENT grad;
    grad.x0 = (double *)fun3d_calloc(sz, 8);
    grad.x1 = (double *)fun3d_calloc(sz, 8);
    grad.x2 = (double *)fun3d_calloc(sz, 8);
    double *r = (double *)fun3d_calloc(sz, 8), *x = (double *)fun3d_calloc(sz, 8);
    double *w = (double *)fun3d_calloc(sz, 8), *y = (double *)fun3d_calloc(sz, 8);
    double *v = (double *)fun3d_calloc(sz, 8);
    ComputeFlux(&g, g.q->q, &grad, r);
    dump("r", r, sz * 8);
    dump("grad_x0", grad.x0, sz * 8);
    dump("grad_x1", grad.x1, sz * 8);
    dump("grad_x2", grad.x2, sz * 8);
    ComputeTimeStep(50.0, &g);
    dump("cdt", g.n->cdt, N * 8);
    const size_t nnzb = kref_nnzb(g.matrix);
    double *blk = (double *)malloc(nnzb * 16 * 8);
    ComputeA(&g);
    kref_blocks(g.matrix, blk);
    dump("a_fill", blk, nnzb * 16 * 8);
    ComputeNumericalILU(&g);
    kref_blocks(g.matrix, blk);
    dump("a_ilu", blk, nnzb * 16 * 8);
    ComputeSparseTriangularSolve(&g, r, x);
    dump("trsv_x", x, sz * 8);
    for (size_t i = 0; i < sz; i++) v[i] = 2.0 * urand() - 1.0;
    dump("jv_x", v, sz * 8);
    const double qnorm = sqrt(1.f + Compute2ndNorm(sz, g.q->q));
    dump("jv_qnorm", &qnorm, 8);
    iJacVec(&g, v, r, qnorm, &grad, w, y);
    dump("jv_w", w, sz * 8);
    dump("jv_y", y, sz * 8);
    double forces[3];
    ComputeForces(&g, forces);
    dump("forces", forces, sizeof(forces));
    /* a second state, filled and factored again: the factorization now
     * starts from the L blocks the first one left behind */
    for (size_t i = 0; i < sz; i++) g.q->q[i] += 0.01 * (2.0 * urand() - 1.0);
    dump("q2", g.q->q, sz * 8);
    ComputeTimeStep(500.0, &g);
    dump("cdt2", g.n->cdt, N * 8);
    ComputeA(&g);
    kref_blocks(g.matrix, blk);
    dump("a_fill2", blk, nnzb * 16 * 8);
    ComputeNumericalILU(&g);
    kref_blocks(g.matrix, blk);
    dump("a_ilu2", blk, nnzb * 16 * 8);
    ComputeSparseTriangularSolve(&g, r, x);
    dump("trsv_x2", x, sz * 8);
    return 0;
  }
  if (!strcmp(mode, "solve")) {
    const uint32_t T = (uint32_t)atoi(argv[4]);
    omp_set_num_threads((int)T);
    GEOMETRY g;
    ref_setup(mesh, &g, T);
    omp_set_num_threads((int)T);
    char path[4096];
    snprintf(path, sizeof(path), "%s/history.txt", g_out);
    g_hist = fopen(path, "w");
    g_geo = &g;
    g_log = 1;
    double t0 = now();
    Kernel(&g);
    double t1 = now();
    g_log = 0;
    fprintf(g_hist, "seconds %.9g\n", t1 - t0);
    fclose(g_hist);
    return 0;
  }
  fprintf(stderr, "kref: unknown mode %s\n", mode);
  return 2;
}
