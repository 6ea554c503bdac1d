/*
 * msh.c -- reader of the reference's big-endian .msh mesh file.
 *
 * Layout and post-processing follow src/mesh.c:22-109 (header of 13 uint32,
 * then edges, nodes, boundaries; the byte count must equal the file size),
 * src/edges.c:22-87 (n0[E], n1[E] 1-based, then x0[E], x1[E], x2[E] unit
 * normal and x3[E] area; edges reordered by the index sort of src/index.c),
 * src/nodes.c:19-56 (x[N], y[N], z[N], vol[N]; vol ignored) and
 * src/boundaries.c:105-204 (per section: group table of 2*groups uint32
 * skipped, faces f0..f3 as four blocks, node ids, three normal blocks; the
 * solid faces are mapped to global node ids, the free faces dropped).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kf_host.h"

void *kf_xmalloc(size_t bytes)
{
  void *p = NULL;
  if (posix_memalign(&p, 64, bytes ? bytes : 64) != 0 || !p) {
    fprintf(stderr, "kf: host allocation of %zu bytes failed\n", bytes);
    abort();
  }
  return p;
}

void *kf_xcalloc(size_t n, size_t size)
{
  const size_t bytes = n * size, chunk = (size_t)8 << 20;
  void *p = kf_xmalloc(bytes);
  if (bytes <= 4 * chunk) {
    memset(p, 0, bytes);
    return p;
  }
  /* large: zeroed by all threads, which also spreads the first-touch page faults */
  const size_t nchunk = (bytes + chunk - 1) / chunk;
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < nchunk; c++) {
    const size_t off = c * chunk;
    memset((char *)p + off, 0, bytes - off < chunk ? bytes - off : chunk);
  }
  return p;
}

/* KF_SETUP_TIMES=1: the wall time of each setup phase on stderr */
void kf_phase(const char *name)
{
  static int on = -1;
  static double last = 0.0;
  if (on < 0) on = getenv("KF_SETUP_TIMES") && atoi(getenv("KF_SETUP_TIMES")) != 0;
  if (!on) return;
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  const double t = (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
  if (name) fprintf(stderr, "kf setup: %-24s %7.3f s\n", name, t - last);
  last = t;
}

static uint32_t rd_u32(const unsigned char *b)
{
  return ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | (uint32_t)b[3];
}

static double rd_f64(const unsigned char *b)
{
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = (v << 8) | b[i];
  double d;
  memcpy(&d, &v, 8);
  return d;
}

typedef struct { const unsigned char *buf; size_t pos, size; const char *path; } cursor;

static const unsigned char *take(cursor *c, size_t bytes)
{
  if (c->pos + bytes > c->size) {
    fprintf(stderr, "kf: %s: truncated mesh file (need %zu bytes at offset %zu of %zu)\n", c->path, bytes, c->pos, c->size);
    abort();
  }
  const unsigned char *p = c->buf + c->pos;
  c->pos += bytes;
  return p;
}

static uint32_t *take_u32(cursor *c, size_t n)
{
  const unsigned char *p = take(c, 4 * n);
  uint32_t *out = (uint32_t *)kf_xmalloc(n * sizeof(uint32_t));
#pragma omp parallel for schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; i++) out[i] = rd_u32(p + 4 * i);
  return out;
}

static double *take_f64(cursor *c, size_t n)
{
  const unsigned char *p = take(c, 8 * n);
  double *out = (double *)kf_xmalloc(n * sizeof(double));
#pragma omp parallel for schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; i++) out[i] = rd_f64(p + 8 * i);
  return out;
}

/* one boundary section: src/boundaries.c:105-141.  Every index is checked
 * against its range here (face corners against the section's node list,
 * node ids against the N vertices): the reference trusts the file, but an
 * out-of-range index would reach host arrays and device kernels alike. */
static void read_section(cursor *c, const char *what, size_t N, size_t nfaces, size_t nnodes, size_t ngroups,
                         uint32_t **faces, struct bntbl *bn)
{
  take(c, 2 * ngroups * sizeof(uint32_t));
  uint32_t *fb = take_u32(c, 4 * nfaces);
  *faces = (uint32_t *)kf_xmalloc((3 * nfaces + 1) * sizeof(uint32_t));
  for (size_t i = 0; i < nfaces; i++) {
    for (int k = 0; k < 3; k++) {
      const uint32_t v = fb[i + (size_t)k * nfaces];   /* 1-based in the file */
      if (v < 1 || v > nnodes) {
        fprintf(stderr, "kf: %s: %s face %zu corner %d is node-list entry %u, outside 1..%zu\n", c->path, what, i,
                k, v, nnodes);
        abort();
      }
      (*faces)[3 * i + k] = v - 1;
    }
  }
  free(fb);
  bn->sz = nnodes;
  bn->nptr = take_u32(c, nnodes);
  for (size_t i = 0; i < nnodes; i++) {
    if (bn->nptr[i] < 1 || bn->nptr[i] > N) {
      fprintf(stderr, "kf: %s: %s node %zu is vertex %u, outside 1..%zu\n", c->path, what, i, bn->nptr[i], N);
      abort();
    }
    bn->nptr[i]--;
  }
  double *nb = take_f64(c, 3 * nnodes);
  bn->xyz = (struct xyz *)kf_xmalloc(sizeof(struct xyz));
  bn->xyz->x0 = (double *)kf_xmalloc((nnodes + 1) * sizeof(double));
  bn->xyz->x1 = (double *)kf_xmalloc((nnodes + 1) * sizeof(double));
  bn->xyz->x2 = (double *)kf_xmalloc((nnodes + 1) * sizeof(double));
  for (size_t i = 0; i < nnodes; i++) {
    bn->xyz->x0[i] = nb[i];
    bn->xyz->x1[i] = nb[i + nnodes];
    bn->xyz->x2[i] = nb[i + 2 * nnodes];
  }
  free(nb);
}

void kf_read_msh(const char *fpath, GEOMETRY *g)
{
  FILE *f = fopen(fpath, "rb");
  if (!f) {
    fprintf(stderr, "kf: cannot open mesh file %s\n", fpath);
    abort();
  }
  fseek(f, 0, SEEK_END);
  size_t fsz = (size_t)ftell(f);
  rewind(f);
  unsigned char *buf = (unsigned char *)kf_xmalloc(fsz);
  if (fread(buf, 1, fsz, f) != fsz) {
    fprintf(stderr, "kf: short read on %s\n", fpath);
    abort();
  }
  fclose(f);
  kf_phase("read file");
  cursor c = {buf, 0, fsz, fpath};

  uint32_t p[13];
  const unsigned char *h = take(&c, 52);
  for (int i = 0; i < 13; i++) p[i] = rd_u32(h + 4 * i);

  /* edges: src/edges.c:22-87 */
  const size_t E = p[2], N = p[1];
  uint32_t *b0 = take_u32(&c, 2 * E);
  for (size_t i = 0; i < 2 * E; i++)
    if (b0[i] < 1 || b0[i] > N) {
      fprintf(stderr, "kf: %s: edge %zu end %zu is vertex %u, outside 1..%zu\n", fpath, i % E, i / E, b0[i], N);
      abort();
    }
  double *b1 = take_f64(&c, 4 * E);
  kf_phase("decode edges");
  uint32_t *perm = (uint32_t *)kf_xmalloc((E + 1) * sizeof(uint32_t));
  kf_index_sort(E, b0, perm);
  kf_phase("edge index sort");
  struct etbl *e = (struct etbl *)kf_xcalloc(1, sizeof(struct etbl));
  e->sz = E;
  e->eptr = (struct edge *)kf_xmalloc(sizeof(struct edge));
  e->xyzn = (struct xyzn *)kf_xmalloc(sizeof(struct xyzn));
  e->eptr->n0 = (uint32_t *)kf_xmalloc((E + 1) * sizeof(uint32_t));
  e->eptr->n1 = (uint32_t *)kf_xmalloc((E + 1) * sizeof(uint32_t));
  e->xyzn->x0 = (double *)kf_xmalloc((E + 1) * sizeof(double));
  e->xyzn->x1 = (double *)kf_xmalloc((E + 1) * sizeof(double));
  e->xyzn->x2 = (double *)kf_xmalloc((E + 1) * sizeof(double));
  e->xyzn->x3 = (double *)kf_xmalloc((E + 1) * sizeof(double));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < E; i++) {
    e->eptr->n0[i] = b0[perm[i]] - 1;
    e->eptr->n1[i] = b0[perm[i] + E] - 1;
    e->xyzn->x0[i] = b1[perm[i]];
    e->xyzn->x1[i] = b1[perm[i] + E];
    e->xyzn->x2[i] = b1[perm[i] + 2 * E];
    e->xyzn->x3[i] = b1[perm[i] + 3 * E];
  }
  free(b0); free(b1); free(perm);
  g->e = e;
  kf_phase("edge reorder");

  /* nodes: src/nodes.c:19-56 */
  double *nb = take_f64(&c, 4 * N);
  struct ntbl *n = (struct ntbl *)kf_xcalloc(1, sizeof(struct ntbl));
  n->sz = N;
  n->xyz = (struct xyz *)kf_xmalloc(sizeof(struct xyz));
  n->xyz->x0 = (double *)kf_xmalloc((N + 1) * sizeof(double));
  n->xyz->x1 = (double *)kf_xmalloc((N + 1) * sizeof(double));
  n->xyz->x2 = (double *)kf_xmalloc((N + 1) * sizeof(double));
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; i++) {
    n->xyz->x0[i] = nb[i];
    n->xyz->x1[i] = nb[i + N];
    n->xyz->x2[i] = nb[i + 2 * N];
  }
  free(nb);
  g->n = n;

  /* boundaries: src/boundaries.c:143-204 */
  struct btbl *b = (struct btbl *)kf_xcalloc(1, sizeof(struct btbl));
  b->s = (struct bntbl *)kf_xcalloc(1, sizeof(struct bntbl));
  b->f = (struct bntbl *)kf_xcalloc(1, sizeof(struct bntbl));
  b->fc = (struct bftbl *)kf_xcalloc(1, sizeof(struct bftbl));
  uint32_t *sfaces = NULL, *ffaces = NULL;
  read_section(&c, "solid", N, p[6], p[9], p[3], &sfaces, b->s);
  read_section(&c, "free-stream", N, p[8], p[11], p[5], &ffaces, b->f);
  free(ffaces);
  const size_t S = p[6];
  b->fc->sz = S;
  b->fc->fptr = (struct face *)kf_xmalloc(sizeof(struct face));
  b->fc->fptr->n0 = (uint32_t *)kf_xmalloc((S + 1) * sizeof(uint32_t));
  b->fc->fptr->n1 = (uint32_t *)kf_xmalloc((S + 1) * sizeof(uint32_t));
  b->fc->fptr->n2 = (uint32_t *)kf_xmalloc((S + 1) * sizeof(uint32_t));
  for (size_t i = 0; i < S; i++) {
    b->fc->fptr->n0[i] = b->s->nptr[sfaces[3 * i + 0]];
    b->fc->fptr->n1[i] = b->s->nptr[sfaces[3 * i + 1]];
    b->fc->fptr->n2[i] = b->s->nptr[sfaces[3 * i + 2]];
  }
  free(sfaces);
  g->b = b;

  if (c.pos != fsz) {
    fprintf(stderr, "kf: %s: %zu bytes parsed but the file holds %zu\n", fpath, c.pos, fsz);
    abort();
  }
  free(buf);
  kf_phase("nodes, boundaries");
}
