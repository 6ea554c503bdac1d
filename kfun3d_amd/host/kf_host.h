/*
 * kf_host.h -- host-side internals shared by the C preprocessing, the C
 * solver driver and the HIP device context.
 */
#ifndef KF_HOST_H
#define KF_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "../../include/kfun3d.h"

#ifdef __cplusplus
extern "C" {
#endif

void *kf_xmalloc(size_t bytes);
void *kf_xcalloc(size_t n, size_t size);
/* KF_SETUP_TIMES=1: prints the time since the previous call (NULL: resets) */
void kf_phase(const char *name);

/* Derived host arrays the device context is built from.  All orders are
 * the reference's: incidences ascend in edge (resp. face, node-list) index,
 * which is the order the reference's one-thread loops accumulate in. */
typedef struct kf_prep {
  size_t N, E, nnzb;
  /* BSR 4x4 pattern: CSR with sorted columns (matrix/init.cpp:7-74) */
  uint32_t *rowptr, *col, *diag;
  /* vertex -> incident edges in ascending edge id; ent = edge << 1 | side,
   * side 0: the vertex is n0 of the edge, side 1: it is n1 */
  uint32_t *vptr, *vent;
  /* BSR block of (vertex, neighbour) for each incidence entry */
  uint32_t *vblk;
  /* level sets of the lower / upper triangle (matrix/init.cpp:194-274) */
  uint32_t nlev_l, nlev_u;
  uint32_t *lptr, *lrows, *uptr, *urows;
  /* vertex -> (solid face << 2 | corner), solid-node entries, free-node
   * entries, each ascending */
  uint32_t *sfptr, *sfent, *snptr, *snent, *fnptr, *fnent;
  /* U and V of the free stream, computed as the reference's
   * cos/sin(alpha / (180 / M_PI)) */
  double U, V;
} kf_prep;

/* schedule of the level-scheduled triangular solves (plan.c) */
typedef struct kf_trsv_plan {
  uint32_t R, S;       /* rows per range, ranges per task (at most 16 / R) */
  uint32_t ntask;
  uint32_t nlevel;     /* levels of the task DAG (tickets are in level order) */
  uint8_t *lvl;        /* [N] local level of each row inside its task */
  uint32_t *task;      /* [ntask * S] ranges of each task, ~0u = none */
  uint8_t *tdep;       /* [ntask] local depth of each task */
  uint32_t *trow;      /* [ntask * 16 * 3] per task and row slot: row, then
                          forward: first block, diagonal position;
                          backward: diagonal position, row end (~0u, 0, 0
                          for an empty slot) */
  uint8_t *tlvl;       /* [ntask * 16] local level of the slot's row, 0xff empty */
  uint8_t *eslot;      /* [nnzb] per position of a row's sweep part: the slot of
                          the column's row in the row's task, 0xff outside */
} kf_trsv_plan;
void kf_plan_trsv(const kf_prep *p, int upper, uint32_t R, int cluster, kf_trsv_plan *out);
void kf_plan_trsv_both(const kf_prep *p, uint32_t R, int cluster, kf_trsv_plan *lower, kf_trsv_plan *upper);
void kf_free_trsv_plan(kf_trsv_plan *t);

/* ILU(0) rows in ticket (level) order (plan.c) */
typedef struct kf_ilu_plan {
  uint32_t *rec;       /* [8 N] per ticket: i, j0, L, nlow, mbeg, mcnt, 0, 0 */
  uint32_t *ldiag;     /* [nnzb] lower position j of row i: diag of row col[j] */
  uint32_t *moff;      /* [nnzb] lower position j: its first match, from mbeg */
  uint32_t *mjm;       /* [nmatch] block (m, k) of row m's upper part ... */
  uint16_t *mlo;       /* [nmatch] ... and the position of column k in row i */
  size_t nmatch;
  uint32_t maxrow, maxlow, maxmatch;
} kf_ilu_plan;
void kf_plan_ilu(const kf_prep *p, kf_ilu_plan *out);
void kf_free_ilu_plan(kf_ilu_plan *t);

/* the incidence lists as a sliced ELL (slice height 64: one wavefront):
 * vertex v of slice s = v / 64, lane v % 64, entry k at
 * (sptr[s] + k) * 64 + v % 64; ascending edge id; padding zero */
typedef struct kf_sell {
  uint32_t nslice;
  size_t ne;                        /* entries including padding */
  uint32_t *sptr, *deg;             /* [nslice + 1], [N] */
  uint32_t *nbr, *ent, *blk;        /* neighbour, edge << 1 | side, BSR block of (v, neighbour) */
  double *wx, *wy, *wz;             /* LSQ term of the edge seen from v */
} kf_sell;
void kf_build_sell(const GEOMETRY *g, const kf_prep *p, kf_sell *out);
/* the device's per-edge records: (n0, n1) pairs and (unit normal, area) */
void kf_edge_records(const GEOMETRY *g, uint32_t *en, double *egeo);

/* .msh reader (src/mesh.c:22-154 without the device part) */
void kf_read_msh(const char *fpath, GEOMETRY *g);
/* host preprocessing of src/mesh.c:113-153 with one owner */
void kf_host_setup(GEOMETRY *g);
/* derived arrays for the device */
void kf_build_prep(const GEOMETRY *g, kf_prep *p);
void kf_free_prep(kf_prep *p);
/* the level sets of a prep whose pattern (N, nnzb, rowptr, col, diag) is set */
void kf_prep_levels(kf_prep *p);

/* the coarse level of the two-level preconditioner (aggregate.c):
 * aggregates, the product A P (fine rows by aggregate columns), the Galerkin
 * pattern as a prep of its own, and the gather lists of both sums */
typedef struct kf_coarse_host {
  kf_prep prep;        /* N = aggregates; rowptr, col, diag, nnzb, level sets */
  uint32_t *agg;       /* [N fine] aggregate of each vertex */
  uint32_t *mptr, *mlist;   /* members of each aggregate, ascending */
  size_t nap;               /* blocks of A P */
  uint32_t *aprp, *apcol;   /* A P: per fine row, its aggregates ascending */
  uint32_t *apgptr, *apglist;   /* fine block positions of each A P block, ascending */
  uint32_t *gptr, *glist;   /* A P positions of each coarse block, ascending (= members ascending) */
} kf_coarse_host;
void kf_build_coarse(const kf_prep *p, kf_coarse_host *out);
void kf_free_coarse(kf_coarse_host *h);
void kf_free_host_geometry(GEOMETRY *g);
void kf_free_lsq(GEOMETRY *g);

/* restatements of the reference's host setup steps */
void kf_index_sort(size_t sz, const uint32_t *key, uint32_t *perm);       /* src/index.c:90-112 */
void kf_color_solid_faces(GEOMETRY *g);                                    /* src/sbface_coloring.c:11-93 */
void kf_lsq_weights(GEOMETRY *g);                                          /* src/weights.c:299-398 */
/* vertex -> incident edges in ascending edge id, ent = edge << 1 | side */
void kf_incidence(size_t N, size_t E, const uint32_t *n0, const uint32_t *n1, uint32_t **vptr, uint32_t **vent);
void kf_free_stream(double alpha_deg, double *U, double *V);               /* inc/phy.h:8-9 */

/* device context (HIP side) */
void kf_ctx_create(GEOMETRY *g, const kf_prep *p, const kf_options *o);
void kf_ctx_destroy(GEOMETRY *g);

/* Device vector operations for the C driver (blas.hip).  Reductions leave
 * their result in a device scalar slot; kfs_fetch copies slots to the host
 * (and synchronizes). */
enum {
  KFS_SLOT_API = 0,      /* Compute2ndNorm / Normalize */
  KFS_SLOT_XNORM = 1,    /* |x| of iJacVec */
  KFS_SLOT_H = 2,        /* Hessenberg column: slots 2 .. 2+restart+1 */
  KFS_SLOT_FNORM = 40,
  KFS_SLOT_QNORM = 41,
  KFS_SLOT_BETA = 42,
  KFS_SLOT_GA = 43,      /* learned correction guard: |ILU^-1 v|, |delta|, coefficient */
  KFS_SLOT_GB = 44,
  KFS_SLOT_GC = 45
};
void kfs_nrm2(GEOMETRY *g, size_t n, const double *x, int slot);
void kfs_dot(GEOMETRY *g, size_t n, const double *x, const double *y, int slot);
void kfs_fetch(GEOMETRY *g, int slot, int count, double *out);
void kfs_axpy(GEOMETRY *g, size_t n, double a, const double *x, double *y);
void kfs_newaxpy(GEOMETRY *g, size_t n, double a, const double *x, const double *y, double *w);
void kfs_scale_recip(GEOMETRY *g, size_t n, double *x, int slot);
void kfs_axpy_slot(GEOMETRY *g, size_t n, int slot, const double *x, double *y);
/* modified Gram-Schmidt of w against nv vectors vbase + k stride: h_k in
 * slots slot..slot+nv-1, |w| in slot+nv (src/kernel.c:129-141) */
/* x = U^-1 L^-1 y, y = ((b - r) + cdt (w - q)) |x| / h formed by the
 * forward sweep from b = R(w) (iJacVec's combination folded into the solve);
 * |x| in scal[slot_xnorm] */
void kfs_sptrsv_fd(GEOMETRY *g, const double *b, double *x, const double *r, const double *w, double h,
                   int slot_xnorm);
void kfs_mgs(GEOMETRY *g, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot);
/* kfs_mgs, then w scaled to unit length (x *= 1.0 / |w|, as kfs_scale_recip);
 * out (if not NULL) receives h_0 .. h_{nv-1}, |w|.  xslot >= 0: the norm of
 * the scaled w may be left in scal[xslot] as kfs_nrm2 would, and with wpert
 * also (fd_h / |w|) w + q there, as kfs_fd_perturb would; returns 2 if both
 * were, 1 if the norm only, 0 otherwise. */
int kfs_mgs_unit(GEOMETRY *g, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot,
                 double *out, int xslot, double fd_h, double *wpert);
/* resident launches that aborted (workgroups not co-resident) on this context */
unsigned long long kfs_mgs_aborts(GEOMETRY *g);
/* resident (one-launch) orthogonalisations so far on this context */
unsigned long long kfs_mgs_launches(GEOMETRY *g);
void kfs_fd_perturb(GEOMETRY *g, const double *x, double h, int slot, double *w);
void kfs_fd_combine(GEOMETRY *g, const double *r, const double *w, double h, int slot, double *y);
/* ComputeFlux(w) then kfs_fd_combine in one pass (y must not alias w or r) */
void kfs_flux_fd(GEOMETRY *g, const double *w, GRADIENT *grad, double *y, const double *r, double h, int slot);
/* whether GMRES's inner step has the residual form the combination */
int kfs_fd_residual(GEOMETRY *g);
void kfs_zero(GEOMETRY *g, size_t n, double *x);
void kfs_cgs_dots(GEOMETRY *g, size_t n, const double *vbase, int nv, const double *w, int slot);
void kfs_cgs_update(GEOMETRY *g, size_t n, const double *vbase, int nv, int slot, double *w);
/* residual of q together with its time-step sums, then cdt = sums / cfl
 * (src/timestep.c folded into the residual pass; q must not change between
 * the two calls) */
void kfs_flux_spectral(GEOMETRY *g, const double *q, GRADIENT *grad, double *r);
/* the forces of the state once the queued work is done, on a second stream
 * (overlapping what is queued next); collect them (the bits of
 * ComputeForces) before anything changes the state again */
void kfs_forces_async(GEOMETRY *g);
void kfs_forces_collect(GEOMETRY *g, double *forces);
void kfs_timestep_spectral(GEOMETRY *g, double cfl);
/* x[i] ~ N(0, 1) from a counter-based generator keyed by (seed, stream, i) */
void kfs_gaussian(GEOMETRY *g, size_t n, uint64_t seed, uint64_t stream, double *x);
/* slot_out = min(1, gamma slot_a / slot_b) */
void kfs_guard(GEOMETRY *g, int slot_a, int slot_b, double gamma, int slot_out);
/* the driver's learned-preconditioner state hangs off the device context */
void **kfs_learn_state(GEOMETRY *g);
void kfs_learn_detach(GEOMETRY *g);   /* solver.c: frees that state */
/* Kernel() work vectors: grad x/y/z, r, w, t, x, and nbasis Krylov vectors */
void kfs_work(GEOMETRY *g, int nbasis, double **grad3, double **r, double **w, double **t, double **x, double **vecs);

#ifdef __cplusplus
}
#endif

#endif /* KF_HOST_H */
