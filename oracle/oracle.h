/*
 * oracle.h -- CPU oracle of the KFUN3D pseudo-transient Newton-Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a one-thread C restatement of the
 * reference's algorithm, written to reproduce its floating-point operation
 * order (see the file:line citations in oracle.c).  It is used by tests/,
 * by __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the
 * checker; the product (kfun3d_amd) never links, loads or calls it.
 *
 * The GEOMETRY layout is the reference's inc/geometry.h:4-106 (host arrays).
 */
#ifndef KF_ORACLE_H
#define KF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct or_xyzn { double *x0, *x1, *x2, *x3; };
struct or_xyz { double *x0, *x1, *x2; };
struct or_edge { unsigned int *n0, *n1; };
struct or_face { unsigned int *n0, *n1, *n2; };
struct or_weights { struct or_xyz *w0, *w1; };
struct or_etbl { unsigned long int sz; struct or_edge *eptr; struct or_xyzn *xyzn; struct or_weights *w; };
struct or_ntbl { unsigned long int sz; struct or_xyz *xyz; unsigned int *part; double *cdt; };
struct or_stbl { unsigned long int sz; unsigned int *i; };
struct or_bntbl { unsigned long int sz; unsigned int *nptr; struct or_xyz *xyz; };
struct or_bftbl { unsigned long int sz; struct or_face *fptr; };
struct or_btbl { struct or_bntbl *s; struct or_bntbl *f; struct or_bftbl *fc; };
struct or_ctbl { unsigned int b; unsigned int b2; unsigned long int sz; };
struct or_qtbl { double *q; };
typedef struct or_geometry {
  struct or_etbl *e;
  struct or_ntbl *n;
  struct or_btbl *b;
  struct or_ctbl *c;
  struct or_stbl *s;
  struct or_stbl *t;
  struct or_qtbl *q;
  void *matrix;
} OR_GEOMETRY;

typedef struct or_ctx or_ctx;

typedef struct or_opts {
  double alpha_deg, cfl_init, cfl_max, fnorm_ratio, gmres_rtol;
  int ts_max, gmres_restart, gmres_max_it, first_order;
  int ilu_textbook;     /* 0: one-thread ilu_Edmond with its copy quirks */
  int reduce_order;     /* 0: one-thread sums (reference); 1: the HIP path's
                           fixed blocked tree (1024 x 256 strided, two trees) */
} or_opts;

typedef struct or_step {
  int step;
  double cfl, fnorm;
  int krylov;
  double forces[3];
} or_step;

/* geometry is borrowed (host arrays); q and cdt live in the context */
or_ctx *or_create(const OR_GEOMETRY *g, const or_opts *o);
void or_destroy(or_ctx *c);
double *or_state(or_ctx *c);      /* q[4N] */
double *or_cdt(or_ctx *c);        /* cdt[N] */
size_t or_nnzb(const or_ctx *c);
const double *or_blocks(const or_ctx *c);   /* A blocks, CSR order */
void or_pattern(const or_ctx *c, unsigned int *rowptr, unsigned int *col);

void or_flux(or_ctx *c, const double *q, double *gx, double *gy, double *gz, double *r);
void or_timestep(or_ctx *c, double cfl);
void or_fill(or_ctx *c);
void or_ilu(or_ctx *c);
void or_sptrsv(or_ctx *c, const double *b, double *x);
void or_sptrsv_jacobi(or_ctx *c, const double *b, double *x, int sweeps);
/* y = A x with the current blocks (after or_fill, before or_ilu) */
void or_spmv(const or_ctx *c, const double *x, double *y);
void or_forces(or_ctx *c, double *forces);
void or_ijacvec(or_ctx *c, const double *x, const double *r, double qnorm, double *gx, double *gy, double *gz,
                double *w, double *y);
double or_norm(size_t n, const double *v);
/* x . y summed in the HIP path's fixed tree (element pairs per thread,
 * 1024 x 256 blocks, two trees) */
double or_dot_tree(size_t n, const double *x, const double *y);
int or_kernel(or_ctx *c, or_step *rec, int max_rec);
void or_inv4(const double *src, double *dst);

#ifdef __cplusplus
}
#endif

#endif
