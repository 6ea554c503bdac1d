/*
 * kfun3d.h -- C-ABI of the MI355X (gfx950) pseudo-transient Newton-Krylov
 * path of KFUN3D.
 *
 * Drop-in boundary.  The entry points below carry the names, argument
 * meanings and call order of the reference:
 *   inc/core_kernel.h:6-22  ComputeA, ComputeTimeStep, ComputeFlux,
 *                           ComputeForces, ComputeNumericalILU,
 *                           ComputeSparseTriangularSolve
 *   src/kernel.c:24,47      iJacVec, Kernel
 *   inc/utils.h:18-28       Compute2ndNorm, ComputeAXPY, ComputeNewAXPY,
 *                           Normalize
 *   src/mesh.c:22           Setup
 * and the GEOMETRY container keeps the field layout of inc/geometry.h:4-106.
 *
 * Memory convention (the one thing a caller of the reference must know):
 *   - every `double *` vector argument of the entry points (q, r, x, y, w, b,
 *     the three GRADIENT arrays) is a DEVICE pointer of length 4N
 *     (AoS [p,u,v,w] per vertex, as in the reference);
 *   - g->q->q and g->n->cdt are DEVICE arrays, allocated by Setup;
 *   - every other GEOMETRY array (edges, normals, LSQ terms, boundary lists,
 *     coordinates) is HOST memory exactly as the reference reads it; Setup
 *     keeps a device mirror of each inside g->matrix;
 *   - g->matrix is the opaque device context (the reference stores its
 *     fun3d::Matrix there, src/mesh.c:121);
 *   - Compute2ndNorm and Normalize return host doubles (they synchronize);
 *     ComputeForces writes a host double[3].
 *
 * Errors: the reference returns void and asserts.  Here every HIP failure and
 * every invalid argument aborts with file/line on stderr, and kf_last_error()
 * reports non-finite pivots / norms detected on the device.
 *
 * Threading: one host thread per GEOMETRY; all work of one GEOMETRY is issued
 * on one HIP stream (kf_stream()).  Distinct GEOMETRYs are independent.
 */
#ifndef KFUN3D_H
#define KFUN3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- GEOMETRY: same layout as inc/geometry.h:4-106 ---------------------- */

struct xyzn { double *x0, *x1, *x2, *x3; };   /* unit normal + dual-face area */
struct xyz  { double *x0, *x1, *x2; };
typedef struct xyz GRADIENT;                   /* d/dx, d/dy, d/dz of (p,u,v,w) */
struct edge { unsigned int *n0, *n1; };
struct face { unsigned int *n0, *n1, *n2; };
struct weights { struct xyz *w0, *w1; };       /* LSQ terms seen from n0 / n1 */
struct etbl { unsigned long int sz; struct edge *eptr; struct xyzn *xyzn; struct weights *w; };
struct ntbl { unsigned long int sz; struct xyz *xyz; unsigned int *part; double *cdt; };
struct stbl { unsigned long int sz; unsigned int *i; };
struct bntbl { unsigned long int sz; unsigned int *nptr; struct xyz *xyz; };
struct bftbl { unsigned long int sz; struct face *fptr; };
struct btbl { struct bntbl *s; struct bntbl *f; struct bftbl *fc; };
struct ctbl { unsigned int b; unsigned int b2; unsigned long int sz; };
struct qtbl { double *q; };

typedef struct geometry {
  struct etbl *e;  /* edges (sorted by n0, n0 < n1), normals, LSQ terms */
  struct ntbl *n;  /* vertex coordinates, owner part (all 0), cdt (device) */
  struct btbl *b;  /* solid nodes, free nodes, solid faces (colour order) */
  struct ctbl *c;  /* b = 4, b2 = 16, sz = 4N */
  struct stbl *s;  /* edge ranges per owner: {0, E} (one owner: the GPU) */
  struct stbl *t;  /* solid-face colour offsets */
  struct qtbl *q;  /* state q[4N] (device) */
  void *matrix;    /* device context */
} GEOMETRY;

/* ---- reference entry points (inc/core_kernel.h:6-22) --------------------- */

/* src/fill.c:881.  A := analytic first-order Jacobian + diag(cdt) + BC blocks,
 * built from g->q->q and g->n->cdt into the context's BSR 4x4 storage. */
void ComputeA(GEOMETRY *g);
/* src/timestep.c:156.  g->n->cdt := local pseudo-time-step term / cfl. */
void ComputeTimeStep(const double cfl, GEOMETRY *g);
/* src/flux.c:782.  r := R(q) (LSQ gradient, Roe edge flux, slip wall,
 * free-stream nodes); grad := LSQ gradient of q.  q, grad->x*, r: device. */
void ComputeFlux(const GEOMETRY *g, const double *q, GRADIENT *grad, double *r);
/* src/forces.c:101.  forces[0..2] (host) := lift, drag, moment. */
void ComputeForces(const GEOMETRY *g, double *forces);
/* src/ilu.c:271.  In-place block ILU(0) of A (L\U with D^-1 on the diagonal).
 * ComputeA must precede every call, as in the reference. */
void ComputeNumericalILU(GEOMETRY *g);
/* src/sptrsv.c:86.  x := U^-1 L^-1 b with the factors of the last ILU.
 * b, x: device, must not alias. */
void ComputeSparseTriangularSolve(const GEOMETRY *g, const double *b, double *x);

/* ---- src/kernel.c --------------------------------------------------------- */

/* src/kernel.c:24.  y := ((R(q + h x/|x|) - r) + cdt (w - q)) |x| / h with
 * h = 1.490116119384766e-08 * norm; w receives the perturbed state and grad
 * its gradient.  All vectors device. */
void iJacVec(const GEOMETRY *g, const double *x, const double *r, const double norm,
             GRADIENT *grad, double *w, double *y);
/* src/kernel.c:47.  The pseudo-transient Newton-Krylov loop with the options
 * of kf_default_options() (reference profile) unless kf_set_options() ran. */
void Kernel(GEOMETRY *g);

/* ---- inc/utils.h:18-28 (device vectors, host scalars) ---------------------- */

double Compute2ndNorm(const size_t sz, const double *v);
void ComputeAXPY(const size_t sz, const double a, const double *x, double *y);
void ComputeNewAXPY(const size_t sz, const double a, const double *x, const double *y, double *w);
double Normalize(const size_t sz, double *x);

/* ---- src/mesh.c:22 --------------------------------------------------------- */

/* Reads a big-endian .msh file (layout of src/mesh.c, src/edges.c,
 * src/nodes.c, src/boundaries.c), runs the host preprocessing of the
 * reference (edge sort src/index.c, LSQ weights src/weights.c, solid-face
 * colouring src/sbface_coloring.c, CSR + level sets matrix/init.cpp), sets the
 * free-stream state and uploads everything into a new device context. */
void Setup(const char *fpath, GEOMETRY *g);

/* ---- extensions (no reference counterpart) ---------------------------------- */

typedef enum {
  KF_PREC_FP64 = 0,     /* Jacobian, ILU factors and triangular solves in FP64 */
  KF_PREC_FP32 = 1,     /* ... in FP32 (residual, Arnoldi, update stay FP64)   */
  KF_PREC_BF16 = 2      /* factors stored in BF16, FP32 arithmetic             */
} kf_precision;

typedef enum {
  KF_ORTHO_MGS = 0,     /* modified Gram-Schmidt, as src/kernel.c:129-139      */
  KF_ORTHO_CGS2 = 1     /* batched classical Gram-Schmidt, two passes          */
} kf_ortho;

typedef enum {
  KF_KRYLOV_GMRES = 0,  /* left-preconditioned restarted GMRES, src/kernel.c:102-200 */
  KF_KRYLOV_FGMRES = 1  /* right-preconditioned flexible GMRES: z_j = M(v_j),
                           x += Z y; converges on the unpreconditioned residual */
} kf_krylov;

typedef struct kf_options {
  double alpha_deg;     /* angle of attack (reference 3, inc/phy.h:8-9)       */
  double cfl_init;      /* CFL_INIT (reference 50, src/kernel.c:19)           */
  double cfl_max;       /* CLF_MAX  (reference 1e5)                           */
  int    ts_max;        /* TS_MAX   (reference 15)                            */
  double fnorm_ratio;   /* stop when fnorm_init/fnorm >= this (reference 1e7,
                           evaluated in double, not through uint32)          */
  int    gmres_restart; /* GMRES_RESTART (reference 30; 1 to 31 accepted)    */
  double gmres_rtol;    /* GMRES_RELATIVE_TOLERANCE (0.01)                    */
  int    gmres_max_it;  /* GMRES_MAX_IT (10000)                               */
  int    first_order;   /* 1: gradients forced to zero (first-order flux)     */
  kf_precision prec;    /* preconditioner precision                          */
  kf_ortho ortho;       /* orthogonalization                                  */
  int    verbose;       /* print the reference's per-step lines              */
  int    ilu_exact;     /* FP64: 1 = the reference's one-thread sweep bit for
                           bit (matrix/ilu.cpp:49-138, including the rounding
                           of its L/Ut copy side effects); 0 = textbook ILU(0) */
  int    reduce_seq;    /* 1 = norms and dot products summed in index order
                           like the reference's one-thread loops (one
                           wavefront; for parity runs); 0 = fixed-order tree */
  kf_krylov krylov;     /* Krylov method (the reference: left GMRES)        */
  int    bsr_operator;  /* 0: the Krylov operator is the matrix-free finite-
                           difference product of the second-order residual
                           (src/kernel.c:24-45, the reference); 1: the
                           assembled first-order Jacobian A = J1 + diag(cdt)
                           by BSR SpMV (kf_bsr_spmv; ComputeA then keeps a
                           copy of A, which the ILU does not overwrite).  The
                           residual stays second order, so both converge to
                           the same steady state; 1 takes more, cheaper
                           Newton steps (defect correction). */
  int    jacobi_sweeps; /* 0: exact level-scheduled triangular solves
                           (matrix/sptrsv.cpp:34-73, the reference's
                           default); k in 1..64: k Jacobi sweeps each way
                           over the same factors instead
                           (fun3d::Matrix::sptrsv_Jacobi, matrix/sptrsv.cpp:
                           75-132, its jacobi_sweeps): an approximate,
                           fixed linear preconditioner whose sweeps are
                           independent HBM-bound passes */
  int    coarse;        /* 0: the preconditioner is block ILU(0) (the
                           reference's); 1: two-level -- a Galerkin coarse
                           correction over greedy aggregates of neighbouring
                           vertices (2 x 2 x 2 on a lattice), factored by
                           block ILU(0) in the preconditioner's precision
                           (Galerkin sums in FP32 for single-precision
                           factors), applied before the fine
                           ILU(0) solve: z = P Ac^-1 P^T b, then
                           x = z + (LU)^-1 (b - A z); L in 2..KF_COARSE_MAX:
                           the coarse system is preconditioned the same way
                           in turn, L levels below the fine one.  No
                           counterpart in the reference; it cuts the Krylov
                           count on large meshes.  jacobi_sweeps applies to
                           the fine level's ILU(0) solve. */
  int    coarse_sweeps; /* coarse > 0: 0 solves every coarse level's ILU(0)
                           exactly (level-scheduled); k in 1..64 by k Jacobi
                           sweeps each way (as jacobi_sweeps) */
  int    lag_factors;   /* Kernel(): 0 forms the Jacobian and factors it at
                           every pseudo-time step (src/kernel.c:91-94, the
                           reference); k >= 1 keeps the last factorisation
                           while the step's CFL is at least the CFL it was
                           formed at and below k times it (k = 1: while the
                           CFL is unchanged, the ramp at cfl_max); a lower
                           CFL refactors.  The Krylov operator stays the
                           exact matrix-free product; only the
                           preconditioner lags.  Ignored with bsr_operator
                           (whose operator is the fill itself) and with a
                           learned correction attached.  No counterpart in
                           the reference. */
} kf_options;

#define KF_COARSE_MAX 8   /* most levels below the fine one (kf_options.coarse) */

typedef struct kf_step_record {
  int    step;          /* pseudo-time step index                             */
  double cfl;
  double fnorm;         /* |R(q)| before the update                           */
  int    krylov;        /* GMRES iterations of this step                      */
  double forces[3];     /* lift, drag, moment after the update                */
  int    learned;       /* 1: the update came from FGMRES with the learned
                           correction; 0: ILU(0) alone                        */
  int    krylov_ilu;    /* iterations of the ILU(0)-alone solve of this step's
                           system, -1 when it did not run                    */
  int    krylov_net;    /* ... of the solve with the learned correction      */
} kf_step_record;

kf_options kf_default_options(void);              /* reference profile */
void kf_set_options(GEOMETRY *g, const kf_options *o);
kf_options kf_get_options(const GEOMETRY *g);
/* Kernel() with an explicit record buffer; returns the number of steps run. */
int  kf_kernel(GEOMETRY *g, kf_step_record *rec, int max_rec);

/* Device-resident copies / uploads of a state vector (4N doubles). */
void kf_set_state(GEOMETRY *g, const double *q_host);
void kf_get_state(const GEOMETRY *g, double *q_host);
void kf_memcpy_d2h(void *dst, const void *src, size_t bytes);
void kf_memcpy_h2d(void *dst, const void *src, size_t bytes);
void *kf_device_alloc(size_t bytes);
void kf_device_free(void *p);
void kf_synchronize(const GEOMETRY *g);
void *kf_stream(const GEOMETRY *g);               /* the context's hipStream_t */

/* Factored BSR blocks (after ComputeNumericalILU) or the Jacobian (after
 * ComputeA) in FP64, nnzb x 16 row-major blocks in CSR order of
 * kf_pattern(); and the pattern itself (host copies). */
size_t kf_nnzb(const GEOMETRY *g);
/* 1 when GMRES's inner step lets the forward triangular sweep form iJacVec's
 * finite-difference combination as it loads its right-hand side (chosen
 * where the sweeps are throughput-bound; KF_TRSV_FOLD=0/1 forces it), 0 when
 * the combination runs as its own kernel first (for tools and tests) */
int kf_trsv_fold(const GEOMETRY *g);
/* 1 when the residual's gradient forms each incidence's LSQ terms from the
 * end's seven moments (src/weights.c:299-366) and the edge vector, 0 when
 * it reads the GEOMETRY's per-edge terms (g->e->w).  Setup takes the first
 * form only once the moments reproduce every term bit for bit;
 * KF_GRAD_MOMENTS=0 forces the second (for tools and tests).  The gradient
 * is the same in both. */
int kf_grad_moments(const GEOMETRY *g);
/* y = A x (device vectors of 4N doubles, not aliased) with the assembled
 * Jacobian of the last ComputeA -- FP64, or FP32 with a single-precision
 * preconditioner -- which ComputeA keeps when kf_options.bsr_operator is
 * set (aborts otherwise).  Per row, the blocks in ascending column, each
 * component summed a[c][0] x0 + ... + a[c][3] x3 in that order, in FP64.
 * No counterpart in the reference (it never multiplies by A). */
void kf_bsr_spmv(const GEOMETRY *g, const double *x, double *y);
void kf_pattern(const GEOMETRY *g, unsigned int *rowptr, unsigned int *col, unsigned int *diag);
void kf_get_blocks(const GEOMETRY *g, double *blocks_host);

/* Synthetic Kuhn-lattice mesh: nx*ny*nz vertices on [0,3]x[0,1]x[0,1],
 * interior jitter +-0.15 h (splitmix64, seed), bottom slip wall with the
 * bump z = 0.1 sin^2(pi (x-1)) on 1<=x<=2, other faces free stream. */
int  kf_write_lattice_msh(const char *fpath, int nx, int ny, int nz, uint64_t seed);

/* Setup with explicit options (Setup uses kf_default_options()). */
void kf_setup_ex(const char *fpath, GEOMETRY *g, const kf_options *o);
/* Host part of Setup only (reader + preprocessing, g->matrix = NULL). */
void kf_setup_host(const char *fpath, GEOMETRY *g);
/* Compute2ndNorm / ComputeAXPY / ComputeNewAXPY / Normalize take no
 * GEOMETRY (inc/utils.h); they run on the context bound last (Setup binds). */
void kf_bind(GEOMETRY *g);

/* Host-side schedule of the triangular solves (for tests and tuning): ranges
 * of R rows (1, 2, 4, 8 or 16), clustered into tasks of up to 16 / R ranges
 * that read only from themselves and from tasks of lower level, in ticket
 * order (cluster = 0: tasks of independent ranges of one level instead).  For the forward (upper = 0) or backward sweep, returns the number
 * of tasks and, when cap covers them, writes the ranges of each task (16 / R
 * per task, ~0u = none), each task's local depth, each row's local level
 * inside its task (N entries), if rowtab is given, per task and row slot q
 * (row q % R of range q / R) the row and its two block bounds (16 x 3 per
 * task), and, if eslot is given, per block position of a row's sweep part
 * the slot of the column's row when it lies in the same task, else 0xff
 * (nnzb entries).  Works on a host-only GEOMETRY. */
size_t kf_trsv_plan_host(const GEOMETRY *g, int upper, unsigned R, int cluster, unsigned *task, unsigned char *tdep,
                         unsigned char *lvl, unsigned *rowtab, unsigned char *eslot, size_t cap);

/* Host-side schedule of the ILU(0) factorisation (for tests): returns the
 * number of update pairs and, when cap covers them, writes per ticket (rows
 * in level order) the record {row, first block, blocks, lower blocks, first
 * pair, pairs, 0, 0} (8 N entries), per BSR position of a lower block (i, m)
 * the diagonal position of row m and the offset of its first pair (nnzb
 * entries each), and per pair the block (m, k) of row m and the position of
 * column k in row i.  Works on a host-only GEOMETRY. */
size_t kf_ilu_plan_host(const GEOMETRY *g, unsigned *rec, unsigned *ldiag, unsigned *moff, unsigned *mjm,
                        unsigned short *mlo, size_t cap);

/* The coarse level of the two-level preconditioner (kf_options.coarse; no
 * counterpart in the reference), for tests.  kf_coarse_plan_host works on a
 * host-only GEOMETRY: returns the number of aggregates nc and, for non-NULL
 * arrays, writes *cnnzb, each vertex's aggregate (N entries) and the coarse
 * BSR pattern (nc + 1 row pointers, cnnzb columns ascending per row).
 * kf_coarse_matrix needs a GEOMETRY whose ComputeNumericalILU ran with
 * coarse = 1 (returns 0 otherwise): the same outputs as built on the device,
 * plus, if blocks is given, the Galerkin blocks P^T A P of the last ComputeA
 * (16 cnnzb doubles, row-major 4x4 per block; FP32 sums with a
 * single-precision preconditioner). */
size_t kf_coarse_plan_host(const GEOMETRY *g, unsigned *agg, size_t *cnnzb, unsigned *crowptr, unsigned *ccol);
size_t kf_coarse_matrix(const GEOMETRY *g, size_t *cnnzb, unsigned *agg, unsigned *crowptr, unsigned *ccol,
                        double *blocks);

/* ---- triangular solve that allocates nothing (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same as
 * ComputeSparseTriangularSolve with Jacobi sweeps (kf_options.jacobi_sweeps
 * >= 1): the same kernels and the same bits.  Its three work vectors (4N
 * values each, in the preconditioner's precision) live in one fixed device
 * array of KF_NOALLOC_BYTES per GPU, held until the process ends, instead
 * of context memory allocated at the first call.  So N <= 174,762 vertices
 * with FP64 factors and N <= 349,525 with FP32 or BF16 factors.  The call
 * makes no hipMalloc, hipFree or other allocating call; the caller passes b
 * and x (device, 4N doubles, 32-byte aligned, not aliased) and has run
 * ComputeA and ComputeNumericalILU in the context's precision.
 * Returns 0, KF_NOALLOC_ESIZE (the mesh is too large for the array; nothing
 * launched) or KF_NOALLOC_EMODE (jacobi_sweeps is 0, no factors yet, or b
 * aliases x).  All calls on one GPU share the array, which keeps the last
 * call's contents: calls that overlap on one GPU (contexts on different
 * streams, or several host threads) race, so the caller orders them. */
#define KF_NOALLOC_BYTES (16u << 20)
#define KF_NOALLOC_ESIZE 1
#define KF_NOALLOC_EMODE 2
int kf_sptrsv_jacobi_noalloc(const GEOMETRY *g, const double *b, double *x);

/* Releases everything Setup allocated, host and device. */
void kf_geometry_free(GEOMETRY *g);

/* Non-zero when the device flagged a non-finite norm or a singular pivot. */
int  kf_last_error(const GEOMETRY *g);

/* Per-kernel timing with HIP events on the context stream.  kf_profile(g, 1)
 * clears and starts it; kf_profile_get returns the total milliseconds and the
 * launch count of one kernel class ("grad", "edge_flux", "residual",
 * "timestep", "fill", "ilu", "trsv_lower", "trsv_upper", "spmv",
 * "jacobi_lower", "jacobi_upper", "coarse_factor", "coarse"); -1 if
 * unknown.  kf_profile_only(g, name) clears and starts it for that one class
 * (fewer events in the timed work); -1 if unknown.  No counterpart in the
 * reference. */
void kf_profile(GEOMETRY *g, int on);
int  kf_profile_only(GEOMETRY *g, const char *name);
int  kf_profile_get(GEOMETRY *g, const char *name, double *total_ms, long *count);

/* ---- learned preconditioner correction (no counterpart in the reference) ----
 * An edge message-passing network (3 rounds, hidden width 32, BF16 weights on
 * the matrix cores, FP32 accumulation) maps the ILU factor's diagonal blocks
 * and the mesh geometry to a sparse linear correction operator G: per
 * vertex a 4 x 4 block S_v, per edge direction a block T_vu, and
 *   delta_v = S_v z_v + (1 / deg v) sum_u T_vu z_u,   z = ILU(0)^-1 v.
 * Flexible GMRES preconditions with M v = z + G z, linear in v.  The
 * parameters are kf_mpnn_nparam() FP32 values; the last kf_mpnn_nhead() of
 * them are the operator heads (all zero: G = 0, M = ILU(0)).  A network is
 * bound to one GEOMETRY (its stream and mesh); vectors are device pointers
 * of length 4N.  Diagonal blocks (and, for training, assembled Jacobians)
 * are held in `slot`s 0 .. nslot-1, each a snapshot of one step. */
typedef struct kf_mpnn kf_mpnn;
size_t   kf_mpnn_nparam(void);
size_t   kf_mpnn_nhead(void);
/* initial parameters (host array of kf_mpnn_nparam()): scaled normals for
 * the network, zero operator heads (G = 0: M = ILU(0)) */
void     kf_mpnn_init_params(float *theta_host, uint64_t seed);
kf_mpnn *kf_mpnn_create(GEOMETRY *g, const float *theta_host, int nslot);
void     kf_mpnn_destroy(kf_mpnn *net);
void     kf_mpnn_set_params(kf_mpnn *net, const float *theta_host);  /* also clears the Adam moments */
void     kf_mpnn_get_params(const kf_mpnn *net, float *theta_host);
/* diagonal blocks of the current ILU factor (after ComputeNumericalILU) */
void     kf_mpnn_take_diag(kf_mpnn *net, int slot);
/* diagonal blocks from the host, FP64 [N][16] */
void     kf_mpnn_set_diag(kf_mpnn *net, int slot, const double *D_host);
/* the assembled Jacobian A (after ComputeA, before ComputeNumericalILU
 * factors it in place) into `slot`, for the A-norm loss */
void     kf_mpnn_take_matrix(kf_mpnn *net, int slot);
/* G of `slot` (the network on its diagonal blocks), once per factorisation;
 * kf_mpnn_forward prepares on demand */
void     kf_mpnn_prepare(kf_mpnn *net, int slot);
/* out = G z (add = 0) or out += G z (add = 1) with the operator of `slot` */
void     kf_mpnn_forward(kf_mpnn *net, int slot, const double *z, double *out, int add);
/* loss of a pair (v, base = ILU^-1 v, target) with the slot's assembled
 * Jacobian A: 2 (default) |A (base + G base) - v|^2 / |v|^2, the residual
 * of the preconditioned system; 1 |A e|^2 / |v|^2 and 0 |e|^2 / |target|^2
 * with e = base + G base - target.  The tight solution (target) sets the
 * size of each pair's Krylov work and enters losses 0 and 1 only. */
void     kf_mpnn_set_loss(kf_mpnn *net, int metric);
/* the loss of (v, base, target) and its gradient (FP32 host copy when
 * grad_host != NULL); synchronises */
double   kf_mpnn_loss_grad(kf_mpnn *net, int slot, const double *v, const double *base, const double *target,
                           float *grad_host);
/* Adam step with the last gradient; gclip > 0 clips its global 2-norm */
void     kf_mpnn_adam(kf_mpnn *net, float lr, float beta1, float beta2, float eps, float gclip);
/* loss + gradient + Adam (0.9, 0.999, 1e-8) queued without a host round
 * trip; the loss is appended to the loss log */
void     kf_mpnn_train_step(kf_mpnn *net, int slot, const double *v, const double *base, const double *target,
                            float lr, float gclip);
int      kf_mpnn_losses(kf_mpnn *net, float *host, int cap);
void     kf_mpnn_reset_losses(kf_mpnn *net);
/* activations of the last network pass (tests): 0 features (N x 16), 1 first
 * pre-activation, 2+5l h_l, 3+5l A h, 4+5l B h, 5+5l messages, 6+5l update
 * pre-activation of round l, 17 the final h, 18 [S | P], 19 [Q | 0]
 * (N x 32 each) */
void     kf_mpnn_activation(kf_mpnn *net, int which, float *host);

/* Learning inside Kernel: with a network attached, kf_kernel runs flexible
 * GMRES and, during the first collect_steps pseudo-time steps, solves the
 * Newton residual and rhs_per_step Gaussian right-hand sides tightly with
 * ILU(0)-FGMRES; each (v / |v|, ILU^-1 v / |v|, x / |v|) is a training pair
 * with that step's diagonal blocks and assembled Jacobian.  After the last
 * collecting step it runs train_iters Adam steps over the pairs (held-out
 * fraction excluded), and from the next step on M v = z + a G z, z =
 * ILU^-1 v, a = min(1, guard |z| / |G z|).  Judged steps solve the Newton
 * system with ILU(0) alone and then with the correction, stopped once it
 * has used as many iterations, and keep the update of the one that needed
 * fewer; between judged steps a solve with the correction that needs more
 * than twice the last judged ILU(0) count is redone with ILU(0) alone.
 * Once ILU(0) alone wins it is used for the rest of the run. */
typedef struct kf_learn_options {
  int      collect_steps;  /* pseudo-time steps that collect pairs (10)          */
  int      rhs_per_step;   /* Gaussian right-hand sides per step (64)            */
  double   pair_rtol;      /* tight-solve relative tolerance (1e-10)             */
  int      pair_maxit;     /* iteration cap of one tight solve (300)             */
  int      train_iters;    /* Adam steps (2000)                                  */
  double   lr;             /* learning rate (1e-3)                               */
  double   gclip;          /* clip of the gradient's 2-norm, 0 = none (1)        */
  double   guard;          /* largest |delta| / |ILU^-1 v| applied (1)           */
  int      judge_every;    /* judge every this many learned steps (5); the first
                              learned step is always judged                    */
  double   holdout;        /* fraction of pairs held out of training (0.1)        */
  uint64_t seed;           /* right-hand sides and pair order (3)                */
  int      ablate;         /* 1: every learned step solves with both and keeps
                              ILU(0)'s update (the run is the plain run, with
                              both iteration counts recorded); 0               */
} kf_learn_options;

typedef struct kf_learn_stats {
  int    npairs, nheld;    /* pairs collected, of which held out                 */
  long   pair_iters;       /* FGMRES iterations of the tight solves              */
  double pair_ms, train_ms;/* wall time of collection and of training            */
  double loss_before;      /* held-out mean loss before training (~ ILU alone)   */
  double loss_after;       /* ... after training                                 */
  double loss_ilu;         /* held-out mean loss with delta = 0                  */
  int    steps_learned;    /* steps whose update used the network                */
  int    steps_judged, judged_won;
  int    fallback;         /* 1: ILU(0) alone took over                          */
} kf_learn_stats;

kf_learn_options kf_default_learn_options(void);
/* attach (net != NULL) or detach (net == NULL) a network; the network must
 * have at least 1 + collect_steps diagonal slots and outlive the attachment */
void kf_attach_mpnn(GEOMETRY *g, kf_mpnn *net, const kf_learn_options *lo);
void kf_learn_get_stats(const GEOMETRY *g, kf_learn_stats *st);


#ifdef __cplusplus
}
#endif

#endif /* KFUN3D_H */
