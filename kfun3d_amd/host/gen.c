/*
 * gen.c -- synthetic Kuhn-lattice tetrahedral mesh in the reference's
 * big-endian .msh layout (the layout read by src/mesh.c:22-154,
 * src/edges.c:22-87, src/nodes.c:19-56, src/boundaries.c:105-204).
 *
 * Lattice nx*ny*nz vertices on [0,3]x[0,1]x[0,1], vertex id
 * v = i + nx*(j + ny*k).  Every hex is split into the 6 Kuhn tetrahedra
 * around its (0,0,0)-(1,1,1) diagonal, so every vertex links to the offsets
 * +-x, +-y, +-z, +-(1,1,0), +-(1,0,1), +-(0,1,1), +-(1,1,1).  Interior
 * vertices are jittered by U(-0.15h, +0.15h) per axis (splitmix64 stream of
 * the seed, three draws per interior vertex in id order).  The bottom wall
 * follows z_b(x) = 0.1 sin^2(pi (x-1)) on 1 <= x <= 2 with z <- z_b + z(1-z_b).
 * The bottom (k = 0) is the solid (slip) section; the other five faces form
 * the free-stream section.
 *
 * Geometry written:
 *   edges     sorted by n0, n0 < n1 (so the reference's index sort is the
 *             identity), unit median-dual face normal pointing n0 -> n1 and
 *             its area;
 *   nodes     coordinates and median-dual volume;
 *   sections  triangles wound so that -(a x b)/6, a = x1-x0, b = x2-x0, points
 *             out of the domain (the convention of src/flux.c:498-508), and
 *             per-node boundary normals = sum of -(a x b)/6 over incident
 *             section triangles (the 1/3 rule).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kf_host.h"

static uint64_t splitmix64(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double unit_draw(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

/* the 7 forward link offsets in ascending id delta (= sorted n1 order) */
static const int OFF[7][3] = {
  {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

static void cross3(const double *a, const double *b, double *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static int link_type(int di, int dj, int dk)
{
  for (int t = 0; t < 7; t++)
    if (OFF[t][0] == di && OFF[t][1] == dj && OFF[t][2] == dk) return t;
  return -1;
}

typedef struct {
  int nx, ny, nz;
  size_t N, E, ncell;
  double *x, *y, *z, *vol;
  int64_t *eid;        /* eid[7 v + t] = edge index of (v, v+OFF[t]) or -1 */
  uint32_t *n0, *n1;
  double *en;          /* 3 E accumulated dual-face area vectors */
} lattice;

static inline size_t vid(const lattice *L, int i, int j, int k) { return (size_t)i + (size_t)L->nx * ((size_t)j + (size_t)L->ny * (size_t)k); }

/* side bit mask of a vertex: 1 x=0, 2 x=max, 4 y=0, 8 y=max, 16 z=0 (solid), 32 z=max */
static unsigned side_mask(const lattice *L, int i, int j, int k)
{
  unsigned m = 0;
  if (i == 0) m |= 1;
  if (i == L->nx - 1) m |= 2;
  if (j == 0) m |= 4;
  if (j == L->ny - 1) m |= 8;
  if (k == 0) m |= 16;
  if (k == L->nz - 1) m |= 32;
  return m;
}

typedef struct { uint32_t v[3]; uint32_t cell; } tri;
typedef struct { tri *t; size_t n, cap; } trilist;

static void tri_push(trilist *l, uint32_t a, uint32_t b, uint32_t c, uint32_t cell)
{
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 1024;
    l->t = (tri *)realloc(l->t, l->cap * sizeof(tri));
    if (!l->t) { fprintf(stderr, "kf: out of host memory\n"); abort(); }
  }
  l->t[l->n].v[0] = a; l->t[l->n].v[1] = b; l->t[l->n].v[2] = c; l->t[l->n].cell = cell;
  l->n++;
}

static void put_u32(FILE *f, uint32_t v)
{
  unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8), (unsigned char)v};
  fwrite(b, 1, 4, f);
}

static void put_f64(FILE *f, double d)
{
  uint64_t v;
  memcpy(&v, &d, 8);
  unsigned char b[8];
  for (int i = 0; i < 8; i++) b[i] = (unsigned char)(v >> (56 - 8 * i));
  fwrite(b, 1, 8, f);
}

/* n big-endian words v[i] + add (or add + i when v is NULL), encoded in
 * parallel into one buffer and written at once */
static void put_u32s(FILE *f, const uint32_t *v, size_t n, uint32_t add)
{
  unsigned char *b = (unsigned char *)kf_xmalloc(4 * n + 1);
#pragma omp parallel for schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; i++) {
    const uint32_t x = (v ? v[i] : (uint32_t)i) + add;
    b[4 * i] = (unsigned char)(x >> 24); b[4 * i + 1] = (unsigned char)(x >> 16);
    b[4 * i + 2] = (unsigned char)(x >> 8); b[4 * i + 3] = (unsigned char)x;
  }
  fwrite(b, 1, 4 * n, f);
  free(b);
}

/* n big-endian doubles v[i * stride] */
static void put_f64s(FILE *f, const double *v, size_t n, size_t stride)
{
  unsigned char *b = (unsigned char *)kf_xmalloc(8 * n + 1);
#pragma omp parallel for schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; i++) {
    uint64_t x;
    memcpy(&x, v + i * stride, 8);
    for (int k = 0; k < 8; k++) b[8 * i + k] = (unsigned char)(x >> (56 - 8 * k));
  }
  fwrite(b, 1, 8 * n, f);
  free(b);
}

/* writes one boundary section: group table, faces, node ids, node normals */
static void write_section(FILE *f, const lattice *L, const trilist *tl, const uint32_t *nodes, size_t nn,
                          const uint32_t *local, const double *nrm)
{
  put_u32(f, (uint32_t)tl->n);     /* group table: face counts ...   */
  put_u32(f, (uint32_t)nn);        /* ... then node counts (1 group) */
  for (int c = 0; c < 3; c++)
    for (size_t t = 0; t < tl->n; t++) put_u32(f, local[tl->t[t].v[c]] + 1);
  for (size_t t = 0; t < tl->n; t++) put_u32(f, tl->t[t].cell + 1);
  for (size_t i = 0; i < nn; i++) put_u32(f, nodes[i] + 1);
  for (int c = 0; c < 3; c++)
    for (size_t i = 0; i < nn; i++) put_f64(f, nrm[3 * (size_t)nodes[i] + c]);
  (void)L;
}

int kf_write_lattice_msh(const char *fpath, int nx, int ny, int nz, uint64_t seed)
{
  if (nx < 2 || ny < 2 || nz < 2) return -1;
  lattice L;
  memset(&L, 0, sizeof(L));
  L.nx = nx; L.ny = ny; L.nz = nz;
  L.N = (size_t)nx * ny * nz;
  L.ncell = 6 * (size_t)(nx - 1) * (ny - 1) * (nz - 1);
  L.x = (double *)kf_xmalloc(L.N * sizeof(double));
  L.y = (double *)kf_xmalloc(L.N * sizeof(double));
  L.z = (double *)kf_xmalloc(L.N * sizeof(double));
  L.vol = (double *)kf_xcalloc(L.N, sizeof(double));

  const double hx = 3.0 / (nx - 1), hy = 1.0 / (ny - 1), hz = 1.0 / (nz - 1);
  uint64_t rng = seed;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t v = vid(&L, i, j, k);
        double x = 3.0 * i / (nx - 1), y = (double)j / (ny - 1), z = (double)k / (nz - 1);
        if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1) {
          x += (2.0 * unit_draw(&rng) - 1.0) * 0.15 * hx;
          y += (2.0 * unit_draw(&rng) - 1.0) * 0.15 * hy;
          z += (2.0 * unit_draw(&rng) - 1.0) * 0.15 * hz;
        }
        double zb = 0.0;
        if (x >= 1.0 && x <= 2.0) {
          double s = sin(M_PI * (x - 1.0));
          zb = 0.1 * s * s;
        }
        L.x[v] = x; L.y[v] = y; L.z[v] = zb + z * (1.0 - zb);
      }

  /* edges in sorted (n0, n1) order */
  L.eid = (int64_t *)kf_xmalloc(7 * L.N * sizeof(int64_t));
  size_t E = 0;
  for (size_t v = 0; v < L.N; v++) {
    int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((size_t)nx * ny));
    for (int t = 0; t < 7; t++) {
      int a = i + OFF[t][0], b = j + OFF[t][1], c = k + OFF[t][2];
      L.eid[7 * v + t] = (a < nx && b < ny && c < nz) ? (int64_t)E++ : -1;
    }
  }
  L.E = E;
  L.n0 = (uint32_t *)kf_xmalloc(E * sizeof(uint32_t));
  L.n1 = (uint32_t *)kf_xmalloc(E * sizeof(uint32_t));
  L.en = (double *)kf_xcalloc(3 * E, sizeof(double));
  for (size_t v = 0; v < L.N; v++) {
    int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((size_t)nx * ny));
    for (int t = 0; t < 7; t++) {
      int64_t e = L.eid[7 * v + t];
      if (e < 0) continue;
      L.n0[e] = (uint32_t)v;
      L.n1[e] = (uint32_t)vid(&L, i + OFF[t][0], j + OFF[t][1], k + OFF[t][2]);
    }
  }

  trilist solid = {0}, freeb = {0};
  uint32_t cell = 0;
  static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int k = 0; k < nz - 1; k++)
    for (int j = 0; j < ny - 1; j++)
      for (int i = 0; i < nx - 1; i++)
        for (int p = 0; p < 6; p++, cell++) {
          int c[4][3];
          c[0][0] = i; c[0][1] = j; c[0][2] = k;
          for (int s = 1; s < 4; s++) {
            for (int d = 0; d < 3; d++) c[s][d] = c[s - 1][d];
            if (s < 3) c[s][PERM[p][s - 1]]++;
            else { c[s][0] = i + 1; c[s][1] = j + 1; c[s][2] = k + 1; }
          }
          size_t tv[4];
          double X[4][3];
          for (int s = 0; s < 4; s++) {
            tv[s] = vid(&L, c[s][0], c[s][1], c[s][2]);
            X[s][0] = L.x[tv[s]]; X[s][1] = L.y[tv[s]]; X[s][2] = L.z[tv[s]];
          }
          double g[3];
          for (int d = 0; d < 3; d++) g[d] = 0.25 * (X[0][d] + X[1][d] + X[2][d] + X[3][d]);
          /* tet volume -> median-dual volumes */
          double a[3], b[3], cc[3], axb[3];
          for (int d = 0; d < 3; d++) { a[d] = X[1][d] - X[0][d]; b[d] = X[2][d] - X[0][d]; cc[d] = X[3][d] - X[0][d]; }
          cross3(a, b, axb);
          double tvol = fabs(axb[0] * cc[0] + axb[1] * cc[1] + axb[2] * cc[2]) / 6.0;
          for (int s = 0; s < 4; s++) L.vol[tv[s]] += 0.25 * tvol;
          /* the 6 edges: dual-face pieces [m, c_pqr, g] + [m, g, c_pqs] */
          static const int EP[6][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2}, {1, 2, 0, 3}, {1, 3, 0, 2}, {2, 3, 0, 1}};
          for (int e = 0; e < 6; e++) {
            int P = EP[e][0], Q = EP[e][1], R = EP[e][2], S = EP[e][3];
            size_t vp = tv[P], vq = tv[Q];
            size_t lo = vp < vq ? vp : vq, hi = vp < vq ? vq : vp;
            int lo_s = vp < vq ? P : Q, hi_s = vp < vq ? Q : P;
            int li = (int)(lo % nx), lj = (int)((lo / nx) % ny), lk = (int)(lo / ((size_t)nx * ny));
            int hi_i = (int)(hi % nx), hj = (int)((hi / nx) % ny), hk = (int)(hi / ((size_t)nx * ny));
            int t = link_type(hi_i - li, hj - lj, hk - lk);
            int64_t eidx = (t >= 0) ? L.eid[7 * lo + t] : -1;
            if (eidx < 0) { fprintf(stderr, "kf: lattice edge lookup failed\n"); abort(); }
            double m[3], cr[3], cs[3], dir[3], u1[3], u2[3], w1[3], w2[3];
            for (int d = 0; d < 3; d++) {
              m[d] = 0.5 * (X[P][d] + X[Q][d]);
              cr[d] = (X[P][d] + X[Q][d] + X[R][d]) / 3.0;
              cs[d] = (X[P][d] + X[Q][d] + X[S][d]) / 3.0;
              dir[d] = X[hi_s][d] - X[lo_s][d];
            }
            for (int d = 0; d < 3; d++) { u1[d] = cr[d] - m[d]; u2[d] = g[d] - m[d]; }
            cross3(u1, u2, w1);
            for (int d = 0; d < 3; d++) { u1[d] = g[d] - m[d]; u2[d] = cs[d] - m[d]; }
            cross3(u1, u2, w2);
            double s1 = (w1[0] * dir[0] + w1[1] * dir[1] + w1[2] * dir[2]) < 0 ? -0.5 : 0.5;
            double s2 = (w2[0] * dir[0] + w2[1] * dir[1] + w2[2] * dir[2]) < 0 ? -0.5 : 0.5;
            for (int d = 0; d < 3; d++) L.en[3 * eidx + d] += s1 * w1[d] + s2 * w2[d];
          }
          /* boundary faces of this tet: all three vertices on one lattice side */
          static const int FV[4][4] = {{1, 2, 3, 0}, {0, 2, 3, 1}, {0, 1, 3, 2}, {0, 1, 2, 3}};
          for (int f = 0; f < 4; f++) {
            int A = FV[f][0], Bv = FV[f][1], C = FV[f][2], D = FV[f][3];
            unsigned mk = side_mask(&L, c[A][0], c[A][1], c[A][2]) & side_mask(&L, c[Bv][0], c[Bv][1], c[Bv][2]) &
                          side_mask(&L, c[C][0], c[C][1], c[C][2]);
            if (!mk) continue;
            double fa[3], fb[3], n[3], in[3];
            for (int d = 0; d < 3; d++) {
              fa[d] = X[Bv][d] - X[A][d];
              fb[d] = X[C][d] - X[A][d];
              in[d] = X[D][d] - (X[A][d] + X[Bv][d] + X[C][d]) / 3.0;
            }
            cross3(fa, fb, n);
            uint32_t t0 = (uint32_t)tv[A], t1 = (uint32_t)tv[Bv], t2 = (uint32_t)tv[C];
            if (n[0] * in[0] + n[1] * in[1] + n[2] * in[2] < 0) { uint32_t tmp = t1; t1 = t2; t2 = tmp; }
            /* a face on the bottom plane is solid; a face can lie on one side only */
            if (mk & 16) tri_push(&solid, t0, t1, t2, cell);
            else tri_push(&freeb, t0, t1, t2, cell);
          }
        }

  /* boundary node normals: sum of -(a x b)/6 over the section's triangles */
  double *sn = (double *)kf_xcalloc(3 * L.N, sizeof(double));
  double *fn = (double *)kf_xcalloc(3 * L.N, sizeof(double));
  uint8_t *ins = (uint8_t *)kf_xcalloc(L.N, 1), *inf = (uint8_t *)kf_xcalloc(L.N, 1);
  for (int sec = 0; sec < 2; sec++) {
    trilist *tl = sec ? &freeb : &solid;
    double *nn = sec ? fn : sn;
    uint8_t *in = sec ? inf : ins;
    for (size_t t = 0; t < tl->n; t++) {
      const uint32_t *tv = tl->t[t].v;
      double a[3] = {L.x[tv[1]] - L.x[tv[0]], L.y[tv[1]] - L.y[tv[0]], L.z[tv[1]] - L.z[tv[0]]};
      double b[3] = {L.x[tv[2]] - L.x[tv[0]], L.y[tv[2]] - L.y[tv[0]], L.z[tv[2]] - L.z[tv[0]]};
      double c[3];
      cross3(a, b, c);
      for (int s = 0; s < 3; s++) {
        in[tv[s]] = 1;
        for (int d = 0; d < 3; d++) nn[3 * (size_t)tv[s] + d] -= c[d] / 6.0;
      }
    }
  }
  size_t ns = 0, nf = 0;
  for (size_t v = 0; v < L.N; v++) { ns += ins[v]; nf += inf[v]; }
  uint32_t *snodes = (uint32_t *)kf_xmalloc((ns + 1) * sizeof(uint32_t));
  uint32_t *fnodes = (uint32_t *)kf_xmalloc((nf + 1) * sizeof(uint32_t));
  uint32_t *sloc = (uint32_t *)kf_xmalloc(L.N * sizeof(uint32_t));
  uint32_t *floc = (uint32_t *)kf_xmalloc(L.N * sizeof(uint32_t));
  ns = nf = 0;
  for (size_t v = 0; v < L.N; v++) {
    if (ins[v]) { sloc[v] = (uint32_t)ns; snodes[ns++] = (uint32_t)v; }
    if (inf[v]) { floc[v] = (uint32_t)nf; fnodes[nf++] = (uint32_t)v; }
  }

  FILE *f = fopen(fpath, "wb");
  if (!f) return -2;
  static char iobuf[1 << 20];
  setvbuf(f, iobuf, _IOFBF, sizeof(iobuf));
  uint32_t p[13] = {(uint32_t)L.ncell, (uint32_t)L.N, (uint32_t)L.E, 1, 0, 1, (uint32_t)solid.n, 0,
                    (uint32_t)freeb.n, (uint32_t)ns, 0, (uint32_t)nf, 0};
  for (int i = 0; i < 13; i++) put_u32(f, p[i]);
  put_u32s(f, L.n0, E, 1);
  put_u32s(f, L.n1, E, 1);
  double *area = (double *)kf_xmalloc(E * sizeof(double));
  double *unit = (double *)kf_xmalloc(E * sizeof(double));
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < E; e++) {
    const double *n = L.en + 3 * e;
    area[e] = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  }
  for (int d = 0; d < 3; d++) {
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < E; e++) unit[e] = L.en[3 * e + d] / area[e];
    put_f64s(f, unit, E, 1);
  }
  free(unit);
  put_f64s(f, area, E, 1);
  put_f64s(f, L.x, L.N, 1);
  put_f64s(f, L.y, L.N, 1);
  put_f64s(f, L.z, L.N, 1);
  put_f64s(f, L.vol, L.N, 1);
  write_section(f, &L, &solid, snodes, ns, sloc, sn);
  write_section(f, &L, &freeb, fnodes, nf, floc, fn);
  int rc = ferror(f) ? -3 : 0;
  fclose(f);

  free(area); free(sn); free(fn); free(ins); free(inf); free(snodes); free(fnodes); free(sloc); free(floc);
  free(solid.t); free(freeb.t);
  free(L.x); free(L.y); free(L.z); free(L.vol); free(L.eid); free(L.n0); free(L.n1); free(L.en);
  return rc;
}
