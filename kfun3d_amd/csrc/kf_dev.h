/*
 * kf_dev.h -- device context and launch helpers of the gfx950 path.
 *
 * HBM layout (one context per GEOMETRY, built once by kf_ctx_create):
 *   edges     SoA in edge order: n0[E], n1[E] (u32); nx, ny, nz, area[E] (f64)
 *   vertices  x[N], y[N], z[N] (f64); state q[4N] and every Krylov vector AoS
 *             (p,u,v,w) per vertex = one 32-byte segment per vertex
 *   incidence sliced-ELL with slice height 64 (one wavefront): vertex v of
 *             slice s = v / 64, lane l = v % 64, owns entries
 *             (sell_ptr[s] + k) * 64 + l for k < deg[v], ascending edge id
 *             (the order the reference's one-thread edge loops accumulate in):
 *               sell_nbr  neighbour vertex              (u32)
 *               sell_ent  edge << 1 | side              (u32, side 1: v is n1)
 *               sell_blk  BSR block of (v, neighbour)   (u32)
 *               sell_w*   LSQ term of the edge seen from v, 3 x f64 SoA
 *   BSR       rowptr[N+1], col[nnzb], diag[N]; blocks of 16 values row-major,
 *             FP64 (a64) or FP32 (a32); BF16 factor copy (abf)
 *   levels    rows in level order of the lower (L) / upper (U) triangle,
 *             cut into groups of at most 16 rows of one level
 */
#ifndef KF_DEV_H
#define KF_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../host/kf_host.h"

#define KF_CHECK(call)                                                                       \
  do {                                                                                       \
    hipError_t kf_e_ = (call);                                                               \
    if (kf_e_ != hipSuccess) {                                                               \
      fprintf(stderr, "kf: HIP error %s at %s:%d: %s\n", hipGetErrorName(kf_e_), __FILE__,  \
              __LINE__, hipGetErrorString(kf_e_));                                           \
      abort();                                                                               \
    }                                                                                        \
  } while (0)

#define KF_LAUNCH_CHECK() KF_CHECK(hipGetLastError())

#define KF_NRED 1024          /* partial sums of a deterministic reduction */
#define KF_NSCAL 256          /* device scalar slots */
#define KF_MGS_MAXV 64        /* basis vectors + 1 of one resident MGS launch, at most */
#define KF_MGS_ARRIVE 4096    /* polls a resident MGS workgroup waits for the others to arrive */
#define KF_GROUP 16           /* rows per triangular-solve group */
#define KF_SLOT_FASYNC (KF_NSCAL - 6)   /* scalar slots of the overlapped forces (3) */
#define KF_JACOBI_MAX 64      /* Jacobi triangular sweeps each way, at most */

/* device error word bits */
#define KF_ERR_NONFINITE_PIVOT 1u
#define KF_ERR_SPIN_LIMIT 2u

/* everything a kernel reads, passed by value */
/* an edge's (unit normal, area) record */
struct kf_egeo { double xn, yn, zn, ln; };
__device__ __forceinline__ kf_egeo kf_edge_geo(const double *__restrict__ egeo, uint32_t e)
{
  const double2 *p = reinterpret_cast<const double2 *>(egeo + 4 * (size_t)e);
  const double2 lo = p[0], hi = p[1];
  return {lo.x, lo.y, hi.x, hi.y};
}

struct kf_dv {
  uint32_t N, E, S, NS, NF, nslice;
  uint32_t spin_sleep;          /* s_sleep(1) repeats between two polls */
  double U, V;
  uint32_t *en;                 /* (n0, n1) per edge */
  double *egeo;                 /* (unit normal x, y, z, dual-face area) per edge: one 32-byte load */
  double *x, *y, *z;
  double *xyzw;                 /* (x, y, z, 0) per vertex: one 32-byte gather for the edge kernel */
  uint32_t *sell_ptr, *deg;
  uint32_t *sell_nbr, *sell_ent, *sell_blk;
  double *sell_wx, *sell_wy, *sell_wz;   /* the LSQ terms per incidence (NULL once lsq is verified) */
  double *lsq;                  /* 7 x N: per-vertex LSQ moments w00 w01 w02 w11 w12 w22 w33 (one array each) */
  uint32_t *sell_nbs;           /* neighbour | side << 31 per incidence (with lsq) */
  uint32_t *sf0, *sf1, *sf2;
  uint32_t *color, ncolor;    /* solid-face colour offsets */
  uint32_t *sfptr, *sfent;
  double *snx, *sny, *snz;
  uint32_t *snptr, *snent;
  double *fnx, *fny, *fnz;
  uint32_t *fnptr, *fnent;
  uint32_t *rowptr, *col, *diag;
  double *eres;
  double *red, *scal;
  uint32_t *err;
};

/* per-kernel HIP-event timing on the context stream (kf_profile) */
enum kf_prof_id { KP_GRAD, KP_EDGE, KP_RES, KP_TSTEP, KP_FILL, KP_ILU, KP_TRSV_L, KP_TRSV_U, KP_SPMV, KP_JAC_L, KP_JAC_U,
                  KP_COARSE_FACTOR, KP_COARSE, KP_N };
#define KF_PROF_POOL 512

struct kf_prof {
  int on;
  int used;
  hipEvent_t ev[2 * KF_PROF_POOL];
  int id[KF_PROF_POOL];
  double total_ms[KP_N];
  long count[KP_N];
};

struct kf_ctx {
  kf_dv d;
  kf_prof prof;
  hipStream_t stream;
  kf_options opt;
  size_t nnzb;
  uint32_t maxrow;          /* most blocks in one BSR row */
  /* ILU(0) schedule (kf_plan_ilu): record per ticket, per-position and
   * per-pair tables; the largest row, lower part and pair list */
  uint32_t *ilu_rec, *ilu_ldiag, *ilu_moff, *ilu_mjm;
  uint16_t *ilu_mlo;
  uint32_t ilu_maxrow, ilu_maxlow, ilu_maxmatch;
  void *ilu_apub, *ilu_udpub; /* published diagonal inverses and U blocks, U(i,i) */
  uint64_t icall;             /* ILU calls so far (sentinel numbering) */
  uint32_t trsv_waves, ilu_waves;   /* persistent wavefronts per launch (trsv: 0 = sized) */
  uint32_t trsv_auto;               /* triangular solves: waves for ~3.5 task levels in flight */
  uint32_t trsv_group, ilu_group;   /* wavefronts per workgroup of the triangular
                                       solves and the ILU: tickets per atomic */
  size_t lds_max;                   /* LDS bytes a workgroup may use */
  uint32_t nlev_l, nlev_u;
  /* the ILU kernel without a plan: rows in level order, per-row completion epoch */
  uint32_t *lrows;
  uint32_t *rdone;
  /* triangular solves: contiguous ranges of trsv_R rows, clustered into
   * tasks of up to trsv_S = 16 / R ranges (one wavefront each) */
  uint32_t trsv_R, trsv_S;
  int trsv_fold;              /* GMRES's combination in the forward sweep: -1 by regime, 0 no, 1 yes */
  int grad_mom;               /* 1: k_grad forms the LSQ terms from d.lsq, 0: reads them per incidence */
  int fd_res;                 /* 1: the residual forms iJacVec's combination where the preconditioner
                                 would run it on its own (KF_FD_RESIDUAL=0: never) */
  uint32_t trsv_pf;           /* off-diagonal entries per row held in registers (7 or 8):
                                7 when no row has more below or above the diagonal */
  uint32_t *trow_l, *trow_u; /* per task (ticket order) and row slot: row, block bounds */
  uint8_t *tlvl_l, *tlvl_u;  /* per task and row slot: local level inside the task */
  uint8_t *tdep_l, *tdep_u;  /* local depth of each task */
  uint8_t *eslot;            /* per block position: slot of the column's row in the
                                row's task (in the sweep that reads it), 0xff outside */
  uint32_t ntask_l, ntask_u;
  double *ty64;             /* forward-sweep output (FP64) */
  uint64_t tcall;           /* triangular solves so far (sentinel numbering) */
  /* modified Gram-Schmidt with w resident in registers (one launch per
   * orthogonalisation): per-step partial sums handed over through sentinel
   * slots, two tables (by call parity) of KF_MGS_MAXV rows of KF_NRED */
  double *mgs_part;
  uint64_t mcall;            /* resident MGS launches so far (sentinel numbering) */
  uint32_t *mgs_sync;        /* per parity: arrival counter and go/abort decision of the launch */
  uint64_t mgs_aborts;       /* resident launches that found their workgroups not co-resident */
  int mgs_nopoll;            /* 1: wait for the resident launch by synchronising, not by polling its flag */
  hipEvent_t mgs_ev;         /* marks the end of a resident launch whose flag is slow to show */
  int mgs_res;               /* 0: never, 1: wherever it fits, 3: always abort (tests the fallback) */
  int8_t mgs_ok[8];          /* per register-size variant: -1 unknown, 0 no, 1 co-resident */
  uint64_t ystate[2];       /* call whose forward sentinel y holds (FP64, FP32) */
  uint32_t *ticket;         /* monotonically increasing work tickets */
  uint32_t ticket_base[4];
  uint32_t epoch;
  /* BSR values */
  double *a64;
  float *a32;
  uint16_t *abf;
  double *jop64;            /* the assembled Jacobian kept for the BSR operator (FP64) ... */
  float *jop32;             /* ... or FP32 with a single-precision preconditioner */
  double *lold;             /* previous sweep's L blocks (exact ILU) */
  double *ud;               /* non-inverted U(i,i) (exact ILU) */
  /* state */
  double *q, *cdt;
  double *hscal;            /* pinned host mirror of d.scal */
  /* Kernel() work vectors */
  double *gx, *gy, *gz, *r, *w, *t, *xk, *vecs;
  int nvecs;
  float *f32a, *f32b;       /* single-precision forward / backward outputs */
  double *jac64;            /* Jacobi sweeps' ping-pong buffers (2 x 4N, FP64 ...) */
  float *jac32;             /* ... or FP32) */
  void *jsl_val;            /* the factors packed per 16-row slice for the Jacobi sweeps (jacobi.hip) */
  uint32_t *jsl_col;        /* ... their columns */
  size_t jsl_esize;         /* bytes per packed value */
  uint32_t jsl_k;           /* slots per row and side */
  int jsl_fresh;            /* 1: packed from the current factors */
  double *eterm, *sr;       /* per-edge time-step terms, per-vertex time-step sums */
  /* forces overlapped with the next pseudo-time step (kf_forces_async) */
  hipStream_t fstream;      /* default priority, like the main stream */
  hipEvent_t ev_upd, ev_forces;
  double *ftbuf;            /* per-face force terms */
  int forces_pending;
  void *noalloc;            /* this device's address of the fixed work array of kf_sptrsv_jacobi_noalloc */
  void *coarse;             /* the two-level preconditioner's coarse level (coarse.hip), built on first use */
  void *learn;              /* learned-preconditioner state of the driver (solver.c) */
};

static inline kf_ctx *kf_ctx_of(const GEOMETRY *g)
{
  if (!g || !g->matrix) {
    fprintf(stderr, "kf: GEOMETRY has no device context (Setup was not called)\n");
    abort();
  }
  return (kf_ctx *)g->matrix;
}

/* Workgroups are dispatched round robin over the 8 XCDs.  This remap gives
 * every XCD a contiguous run of logical blocks -- a contiguous slab of
 * vertices -- so that neighbour gathers mostly hit that XCD's own L2. */
__device__ __forceinline__ uint32_t kf_xcd_block()
{
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t per = (nb + 7u) / 8u;
  const uint32_t x = b % 8u, i = b / 8u;
  const uint32_t full = (nb % 8u) ? nb % 8u : 8u;
  return x < full ? x * per + i : full * per + (x - full) * (per - 1u) + i;
}

static inline uint32_t kf_blocks_for(size_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

static inline void kf_check_aligned(const void *p, const char *what)
{
  if (((uintptr_t)p) & 31u) {
    fprintf(stderr, "kf: %s must be 32-byte aligned (one vertex = 4 doubles)\n", what);
    abort();
  }
}

/* event pair around one launch; no-ops unless profiling is on (ctx.hip) */
int kf_prof_begin(kf_ctx *c, int id);
void kf_prof_end(kf_ctx *c, int slot);

/* launches (defined in the .hip files) */
/* sr != NULL: also the time-step sums of the state q (see k_residual) */
void kf_grad_setup(kf_ctx *c);
/* fd: iJacVec's combination formed by the residual's last kernel, whose
 * state q is then fd->w (k_residual<false, true>) */
struct kf_fdb;
void kf_launch_flux(kf_ctx *c, const double *q, double *gx, double *gy, double *gz, double *r, double *sr,
                    const kf_fdb *fd = NULL);
void kf_launch_timestep(kf_ctx *c, double cfl);
void kf_launch_timestep_div(kf_ctx *c, const double *sr, double cfl);
void kf_launch_forces(kf_ctx *c, double *forces_host);
void kf_forces_async(kf_ctx *c);
void kf_forces_collect(kf_ctx *c, double *forces_host);
void kf_launch_fill(kf_ctx *c);
void kf_launch_ilu(kf_ctx *c);
/* the finite-difference combination applied by the forward sweep as it
 * loads its right-hand side b = R(w): ((R(w) - r) + cdt (w - q0)) |x| / h,
 * |x| = *xnorm (src/kernel.c:33-44) */
struct kf_fdb {
  const double *r, *w, *q0, *cdt, *xnorm;
  double h;
};
void kf_launch_sptrsv(kf_ctx *c, const double *b, double *x, const kf_fdb *fd);
/* y = A x with the kept copy of the assembled Jacobian (spmv.hip) */
void kf_launch_spmv(kf_ctx *c, const double *x, double *y);
void kf_keep_jacobian(kf_ctx *c);
bool kf_trsv_capped(const kf_ctx *c);
/* x ~ U^-1 L^-1 b by opt.jacobi_sweeps Jacobi sweeps each way (jacobi.hip) */
void kf_launch_jacobi(kf_ctx *c, const double *b, double *x);
/* the multilevel preconditioner's fine solve: the sweeps on t (FP64 with FP64
 * factors, else FP32), then x += P z_c fused into the last sweep */
void kf_launch_jacobi_ml(kf_ctx *c, const void *t, double *x, const uint32_t *agg, const double *zc);
/* at context creation, with the context's device current: the device
 * address of kf_sptrsv_jacobi_noalloc's fixed work array (jacobi.hip) */
void kf_noalloc_setup(kf_ctx *c);
void kf_check_options(const kf_options *o);
/* a context holding a matrix only (pattern, plans, values, factors), on its
 * parent's stream and error word: the coarse level's (ctx.hip) */
kf_ctx *kf_ctx_create_matrix(const kf_prep *p, const kf_options *o, const kf_ctx *parent);
void kf_ctx_destroy_matrix(kf_ctx *c);
/* the two-level preconditioner (coarse.hip): the Galerkin coarse matrix and
 * its ILU(0) from the kept Jacobian, and x = M b */
void kf_coarse_factor(kf_ctx *c);
void kf_coarse_apply(kf_ctx *c, const double *b, double *x);
void kf_coarse_destroy(kf_ctx *c);

/* BLAS-1 on the context stream; reductions are deterministic (fixed
 * partition, fixed tree) and leave their result in d.scal[slot] */
void kf_dot_dev(kf_ctx *c, size_t n, const double *x, const double *y, int slot);
void kf_nrm2_dev(kf_ctx *c, size_t n, const double *x, int slot);
void kf_scal_fetch(kf_ctx *c, int slot, int count, double *out);
void kf_axpy(kf_ctx *c, size_t n, double a, const double *x, double *y);
void kf_newaxpy(kf_ctx *c, size_t n, double a, const double *x, const double *y, double *w);
void kf_scale_recip(kf_ctx *c, size_t n, double *x, int slot);
void kf_axpy_slot(kf_ctx *c, size_t n, int slot, const double *x, double *y);
int kf_mgs(kf_ctx *c, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot, int unit,
           double *out, int xslot, double fd_h, double *wpert);
void kf_mgs_setup(kf_ctx *c);
void kf_fd_perturb(kf_ctx *c, const double *x, double h, int slot_xnorm, double *w);
void kf_fd_combine(kf_ctx *c, const double *r, const double *w, double h, int slot_xnorm, double *y);
void kf_fill_zero(kf_ctx *c, size_t n, double *x);

#endif
