/*
 * solver.c -- the pseudo-transient Newton-Krylov driver of src/kernel.c on
 * the device context.  Host C: it issues the HIP entry points on the
 * context's stream and keeps only the scalar recurrences on the host (SER
 * CFL, Givens rotations of the (restart+1) x restart Hessenberg, back
 * substitution), exactly as the reference orders them.
 *
 *   iJacVec  src/kernel.c:24-45  matrix-free (J + diag(cdt)) x by forward
 *            difference: w = (h/|x|) x + q, y = ((R(w) - r) + cdt (w - q)) |x|/h
 *   Kernel   src/kernel.c:47-232 pseudo-time loop, left-preconditioned
 *            restarted GMRES with modified Gram-Schmidt (or batched CGS2);
 *            optionally right-preconditioned flexible GMRES (z_j = M(v_j),
 *            x += Z y), the form a varying preconditioner needs
 *
 * The only host/device round trips per Arnoldi step are the restart+2
 * Hessenberg scalars (one copy) needed by the Givens recurrence.
 *
 * With a learned correction attached (kf_attach_mpnn), flexible GMRES runs
 * with M(v) = ILU^-1 v + a delta(v); the same Krylov routine solves the
 * tight training systems of the collecting steps.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "kf_host.h"

#define FD_EPS 1.490116119384766e-08 /* src/kernel.c:12 */

/* iJacVec with |x| already in scal[KFS_SLOT_XNORM] (have >= 1) and the
 * perturbed state already in w (have == 2), left there by the
 * orthogonalisation that produced x */
static void ijacvec(GEOMETRY *g, const double *x, const double *r, const double norm, GRADIENT *grad, double *w,
                    double *y, int have)
{
  const size_t sz = g->c->sz;
  if (have < 1) kfs_nrm2(g, sz, x, KFS_SLOT_XNORM);
  const double h = FD_EPS * norm;
  if (have < 2) kfs_fd_perturb(g, x, h, KFS_SLOT_XNORM, w);
  if (y == r || y == w) {   /* aliased output: the two passes, as before */
    ComputeFlux(g, w, grad, y);
    kfs_fd_combine(g, r, w, h, KFS_SLOT_XNORM, y);
    return;
  }
  kfs_flux_fd(g, w, grad, y, r, h, KFS_SLOT_XNORM);
}

/* GMRES's inner step M^-1 (J x): iJacVec, then the preconditioner.  Where
 * the triangular sweeps are throughput-bound the forward sweep forms the
 * combination as it loads R(w) (kfs_sptrsv_fd); otherwise the residual's
 * last kernel forms it (kfs_flux_fd); both with the operations of
 * kfs_fd_combine.  y is scratch afterwards */
static void ijacvec_solve(GEOMETRY *g, const double *x, const double *r, const double norm, GRADIENT *grad,
                          double *w, double *y, int have, double *out)
{
  const size_t sz = g->c->sz;
  if (have < 1) kfs_nrm2(g, sz, x, KFS_SLOT_XNORM);
  const double h = FD_EPS * norm;
  if (have < 2) kfs_fd_perturb(g, x, h, KFS_SLOT_XNORM, w);
  if (kfs_fd_residual(g)) {
    kfs_flux_fd(g, w, grad, y, r, h, KFS_SLOT_XNORM);
    ComputeSparseTriangularSolve(g, y, out);
    return;
  }
  ComputeFlux(g, w, grad, y);
  kfs_sptrsv_fd(g, y, out, r, w, h, KFS_SLOT_XNORM);
}

void iJacVec(const GEOMETRY *gc, const double *x, const double *r, const double norm, GRADIENT *grad, double *w,
             double *y)
{
  ijacvec((GEOMETRY *)gc, x, r, norm, grad, w, y, 0);
}

static double fetch1(GEOMETRY *g, int slot)
{
  double v;
  kfs_fetch(g, slot, 1, &v);
  return v;
}

static double now_ms(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

/* ---- learned correction state ------------------------------------------------ */

typedef struct {
  kf_mpnn *net;
  kf_learn_options lo;
  kf_learn_stats st;
  int cap;               /* pairs the store holds */
  double *pv, *pb, *px;  /* per pair: v / |v|, ILU^-1 of it, tight solution */
  int *pslot;            /* diagonal slot of each pair */
  int enabled;           /* network in M(v) for the Newton solves */
  int since_judge;
  int last_ilu;          /* iterations of the last judged ILU(0) solve */
} learn_state;

kf_learn_options kf_default_learn_options(void)
{
  kf_learn_options lo;
  memset(&lo, 0, sizeof(lo));
  lo.collect_steps = 10;
  lo.rhs_per_step = 64;
  lo.pair_rtol = 1e-10;
  lo.pair_maxit = 300;
  lo.train_iters = 2000;
  lo.lr = 1e-3;
  lo.gclip = 1.0;
  lo.guard = 1.0;
  lo.judge_every = 5;
  lo.holdout = 0.1;
  lo.seed = 3;
  return lo;
}

void kfs_learn_detach(GEOMETRY *g)
{
  void **slot = kfs_learn_state(g);
  learn_state *L = (learn_state *)*slot;
  if (!L) return;
  kf_synchronize(g);
  if (L->pv) kf_device_free(L->pv);
  if (L->pb) kf_device_free(L->pb);
  if (L->px) kf_device_free(L->px);
  free(L->pslot);
  free(L);
  *slot = NULL;
}

void kf_attach_mpnn(GEOMETRY *g, kf_mpnn *net, const kf_learn_options *lo)
{
  kfs_learn_detach(g);
  if (!net) return;
  learn_state *L = (learn_state *)calloc(1, sizeof(learn_state));
  if (!L) abort();
  L->net = net;
  L->lo = lo ? *lo : kf_default_learn_options();
  if (L->lo.collect_steps < 0 || L->lo.rhs_per_step < 0 || L->lo.train_iters < 0 || L->lo.holdout < 0.0 ||
      L->lo.holdout >= 1.0) {
    fprintf(stderr, "kf: invalid learning options\n");
    abort();
  }
  L->cap = L->lo.collect_steps * (1 + L->lo.rhs_per_step);
  L->enabled = L->lo.collect_steps == 0; /* a pre-trained network is used at once */
  *kfs_learn_state(g) = L;
}

void kf_learn_get_stats(const GEOMETRY *g, kf_learn_stats *st)
{
  learn_state *L = (learn_state *)*kfs_learn_state((GEOMETRY *)g);
  if (L) *st = L->st;
  else memset(st, 0, sizeof(*st));
}

/* ---- the Krylov solve ---------------------------------------------------------- */

typedef struct {
  GEOMETRY *g;
  kf_options o;
  size_t sz;
  int restart, nb, flexible;
  GRADIENT grad;
  double *r, *w, *t, *vecs, *ghh;
  double qnorm;
  learn_state *L;
  int failed;          /* a Hessenberg value came back non-finite: the solve stopped */
} kspace;

#define VEC(k) (K->vecs + (size_t)(k) * K->sz)
#define ZVEC(k) (K->vecs + (size_t)(K->nb + (k)) * K->sz) /* FGMRES: z_k = M(v_k) */

/* z = ILU^-1 v on entry; z += a G z, a = min(1, guard |z| / |G z|); w is
 * scratch */
static void add_correction(kspace *K, double *z)
{
  GEOMETRY *g = K->g;
  kf_mpnn_forward(K->L->net, 0, z, K->w, 0);
  kfs_nrm2(g, K->sz, z, KFS_SLOT_GA);
  kfs_nrm2(g, K->sz, K->w, KFS_SLOT_GB);
  kfs_guard(g, KFS_SLOT_GA, KFS_SLOT_GB, K->L->lo.guard, KFS_SLOT_GC);
  kfs_axpy_slot(g, K->sz, KFS_SLOT_GC, K->w, z);
}

/* Solves (J + diag(cdt)) x = b at the current state from x = 0 (x is
 * zeroed here), restarting every `restart` steps, until the residual
 * estimate drops below rtol times the first cycle's |M b| (left GMRES) or
 * |b| (FGMRES) or more than maxit iterations ran; returns the iterations.
 * src/kernel.c:102-200 with b = r.  cap > 0 also stops the solve after cap
 * iterations (mid-cycle); *conv (if given) tells whether the tolerance was
 * met. */
static int ksolve(kspace *K, const double *b, double *x, double rtol, int maxit, int use_net, int cap, int *conv)
{
  GEOMETRY *g = K->g;
  const kf_options *o = &K->o;
  const size_t sz = K->sz;
  const int restart = K->restart, nb = K->nb, flexible = K->flexible;
  double *ghh = K->ghh, *w = K->w, *t = K->t;
  double cc[64], ss[64], rs[64], hcol[64];
  kfs_zero(g, sz, x);
  if (!flexible) ComputeSparseTriangularSolve(g, b, VEC(0));
  else kfs_newaxpy(g, sz, 0.0, b, b, VEC(0)); /* v = b (x = 0) */
  double tol = 0.0;
  int total = 0, flag = 0;
  const int bsr = o->bsr_operator;   /* the assembled first-order Jacobian as the operator */
  while (total <= maxit) {
    if (total > 0) {
      if (bsr) kf_bsr_spmv(g, x, t);
      else iJacVec(g, x, K->r, K->qnorm, &K->grad, w, t);
      if (!flexible) {
        kfs_newaxpy(g, sz, -1.0, t, b, w);
        ComputeSparseTriangularSolve(g, w, VEC(0));
      } else {
        kfs_newaxpy(g, sz, -1.0, t, b, VEC(0)); /* v = b - A x */
      }
    }
    kfs_nrm2(g, sz, VEC(0), KFS_SLOT_BETA);
    kfs_scale_recip(g, sz, VEC(0), KFS_SLOT_BETA);
    rs[0] = fetch1(g, KFS_SLOT_BETA);
    tol = (total == 0) ? rtol * rs[0] : tol;
    int it = 0, have_xnorm = 0;   /* |VEC(it)| left by the previous orthogonalisation */
    while (it < restart) {
      if (!flexible) {
        if (bsr) {
          kf_bsr_spmv(g, VEC(it), t);
          ComputeSparseTriangularSolve(g, t, VEC(it + 1));
        } else {
          ijacvec_solve(g, VEC(it), K->r, K->qnorm, &K->grad, w, t, have_xnorm, VEC(it + 1));
        }
      } else {
        ComputeSparseTriangularSolve(g, VEC(it), ZVEC(it)); /* z = M(v) */
        if (use_net) add_correction(K, ZVEC(it));
        if (bsr) kf_bsr_spmv(g, ZVEC(it), VEC(it + 1));
        else iJacVec(g, ZVEC(it), K->r, K->qnorm, &K->grad, w, VEC(it + 1));
      }
      double *hh = ghh + (size_t)it * (nb + 1);
      if (o->ortho == KF_ORTHO_CGS2) {
        /* two passes of batched classical Gram-Schmidt */
        kfs_cgs_dots(g, sz, VEC(0), it + 1, VEC(it + 1), KFS_SLOT_H);
        kfs_cgs_update(g, sz, VEC(0), it + 1, KFS_SLOT_H, VEC(it + 1));
        kfs_fetch(g, KFS_SLOT_H, it + 1, hcol);
        kfs_cgs_dots(g, sz, VEC(0), it + 1, VEC(it + 1), KFS_SLOT_H);
        kfs_cgs_update(g, sz, VEC(0), it + 1, KFS_SLOT_H, VEC(it + 1));
        kfs_nrm2(g, sz, VEC(it + 1), KFS_SLOT_H + it + 1);
        kfs_scale_recip(g, sz, VEC(it + 1), KFS_SLOT_H + it + 1);
        double h2[64];
        kfs_fetch(g, KFS_SLOT_H, it + 2, h2);
        for (int v = 0; v <= it; v++) hh[v] = hcol[v] + h2[v];
        hh[it + 1] = h2[it + 1];
      } else {
        /* modified Gram-Schmidt, src/kernel.c:129-141 */
        /* then v = w / |w|, and |v| and the perturbed state for the next
         * iJacVec */
        const int fd_next = !flexible && !bsr;   /* the next operator is iJacVec on this vector */
        have_xnorm = kfs_mgs_unit(g, sz, VEC(0), sz, it + 1, VEC(it + 1), KFS_SLOT_H, hh,
                                  fd_next ? KFS_SLOT_XNORM : -1, FD_EPS * K->qnorm, fd_next ? w : NULL);
      }
      /* Givens, src/kernel.c:143-173 */
      for (int i = 0; i < it; i++) {
        const double tt = *hh;
        *hh = cc[i] * *hh + ss[i] * *(hh + 1);
        hh++;
        *hh = cc[i] * *hh - ss[i] * tt;
      }
      const double tt = sqrt(*hh * *hh + *(hh + 1) * *(hh + 1));
      const double ct = *hh / tt, st = *(hh + 1) / tt;
      cc[it] = ct;
      ss[it] = st;
      rs[it + 1] = -(st * rs[it]);
      rs[it] = ct * rs[it];
      *hh = ct * *hh + st * *(hh + 1);
      const double res = fabs(rs[it + 1]);
      it++;
      if (!isfinite(res)) {
        /* a non-finite orthogonalisation (a device wait that expired, or a
         * singular preconditioner): no update is formed from it */
        K->failed = 1;
        break;
      }
      if (res <= tol) {
        flag = 1;
        break;
      }
      if (cap > 0 && total + it >= cap) break;
    }
    total += it;
    if (K->failed) break;
    /* back substitution, src/kernel.c:178-188 */
    rs[it - 1] /= ghh[(size_t)(it - 1) * (nb + 1) + (it - 1)];
    for (int i = 1; i <= it - 1; i++) {
      const int k = (it - 1) - i;
      double tt = rs[k];
      for (int j = k + 1; j <= it - 1; j++) tt -= ghh[(size_t)j * (nb + 1) + k] * rs[j];
      rs[k] = tt / ghh[(size_t)k * (nb + 1) + k];
    }
    for (int v = 0; v < it; v++) kfs_axpy(g, sz, rs[v], flexible ? ZVEC(v) : VEC(v), x);
    if (flag || (cap > 0 && total >= cap)) break;
  }
  if (conv) *conv = flag;
  return total;
}

/* ---- training pairs and training ---------------------------------------------------- */

/* pairs of pseudo-time step `its` (factor and state of this step): the
 * Newton residual r and rhs_per_step Gaussian vectors, each normalised,
 * preconditioned by ILU(0) and solved tightly with ILU(0)-FGMRES */
static void collect_pairs(kspace *K, int its)
{
  learn_state *L = K->L;
  GEOMETRY *g = K->g;
  const size_t sz = K->sz;
  const double t0 = now_ms();
  if (!L->pv) {
    L->pv = (double *)kf_device_alloc((size_t)L->cap * sz * sizeof(double));
    L->pb = (double *)kf_device_alloc((size_t)L->cap * sz * sizeof(double));
    L->px = (double *)kf_device_alloc((size_t)L->cap * sz * sizeof(double));
    L->pslot = (int *)calloc((size_t)L->cap, sizeof(int));
    if (!L->pslot) abort();
  }
  const int slot = 1 + its;
  kf_mpnn_take_diag(L->net, slot);
  for (int k = 0; k <= L->lo.rhs_per_step; k++) {
    const int p = L->st.npairs;
    if (p >= L->cap) abort();
    double *v = L->pv + (size_t)p * sz, *base = L->pb + (size_t)p * sz, *x = L->px + (size_t)p * sz;
    if (k == 0) kfs_newaxpy(g, sz, 0.0, K->r, K->r, v);
    else kfs_gaussian(g, sz, L->lo.seed, (uint64_t)its * 100003u + (uint64_t)k, v);
    kfs_nrm2(g, sz, v, KFS_SLOT_API);
    kfs_scale_recip(g, sz, v, KFS_SLOT_API);
    ComputeSparseTriangularSolve(g, v, base);
    L->st.pair_iters += ksolve(K, v, x, L->lo.pair_rtol, L->lo.pair_maxit, 0, 0, NULL);
    L->pslot[p] = slot;
    L->st.npairs++;
  }
  kf_synchronize(g);
  L->st.pair_ms += now_ms() - t0;
  if (K->o.verbose) {
    printf("LEARN: step %d: %d pairs so far, %ld tight-solve iterations so far, %.1f s collecting\n", its,
           L->st.npairs, L->st.pair_iters, L->st.pair_ms / 1e3);
    fflush(stdout);
  }
}

static uint64_t mix64(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* mean loss over the held-out pairs (the last nheld collected) */
static double heldout_loss(learn_state *L, size_t sz)
{
  if (L->st.nheld == 0) return 0.0;
  double s = 0.0;
  for (int p = L->st.npairs - L->st.nheld; p < L->st.npairs; p++)
    s += kf_mpnn_loss_grad(L->net, L->pslot[p], L->pv + (size_t)p * sz, L->pb + (size_t)p * sz,
                           L->px + (size_t)p * sz, NULL);
  return s / L->st.nheld;
}

/* train_iters Adam steps over the training pairs in shuffled epochs */
static void train(kspace *K)
{
  learn_state *L = K->L;
  const size_t sz = K->sz;
  const double t0 = now_ms();
  const int n = L->st.npairs;
  int nheld = (int)floor(L->lo.holdout * n);
  if (nheld >= n) nheld = n - 1;
  if (nheld < 0) nheld = 0;
  L->st.nheld = nheld;
  const int ntrain = n - nheld;
  /* the loss of ILU(0) alone: the decoder switched off */
  const size_t np = kf_mpnn_nparam();
  float *theta = (float *)malloc(np * sizeof(float)), *th0 = (float *)malloc(np * sizeof(float));
  if (!theta || !th0) abort();
  kf_mpnn_get_params(L->net, theta);
  memcpy(th0, theta, np * sizeof(float));
  for (size_t i = np - kf_mpnn_nhead(); i < np; i++) th0[i] = 0.0f; /* the operator heads: G = 0 */
  kf_mpnn_set_params(L->net, th0);
  L->st.loss_ilu = heldout_loss(L, sz);
  kf_mpnn_set_params(L->net, theta);
  L->st.loss_before = heldout_loss(L, sz);
  free(th0);
  free(theta);
  int *perm = (int *)malloc((size_t)(ntrain > 0 ? ntrain : 1) * sizeof(int));
  if (!perm) abort();
  uint64_t rs = L->lo.seed ^ 0x5851F42D4C957F2Dull;
  int pos = ntrain;
  for (int it = 0; it < L->lo.train_iters && ntrain > 0; it++) {
    if (pos == ntrain) { /* a new epoch: Fisher-Yates */
      for (int i = 0; i < ntrain; i++) perm[i] = i;
      for (int i = ntrain - 1; i > 0; i--) {
        const int j = (int)(mix64(&rs) % (uint64_t)(i + 1));
        const int tmp = perm[i];
        perm[i] = perm[j];
        perm[j] = tmp;
      }
      pos = 0;
    }
    const int p = perm[pos++];
    kf_mpnn_train_step(L->net, L->pslot[p], L->pv + (size_t)p * sz, L->pb + (size_t)p * sz, L->px + (size_t)p * sz,
                       (float)L->lo.lr, (float)L->lo.gclip);
  }
  free(perm);
  L->st.loss_after = heldout_loss(L, sz);
  L->st.train_ms += now_ms() - t0;
  if (K->o.verbose) {
    printf("LEARN: %d Adam steps on %d pairs (%d held out) in %.1f s; held-out loss: ILU(0) %g, start %g, "
           "trained %g\n", L->lo.train_iters, ntrain, nheld, L->st.train_ms / 1e3, L->st.loss_ilu,
           L->st.loss_before, L->st.loss_after);
    fflush(stdout);
  }
}

/* ---- the driver ------------------------------------------------------------------ */

/* a pseudo-time step whose forces are still being computed */
typedef struct {
  int valid, idx;
  kf_step_record r;
} step_pending;

/* the step's forces (the bits of ComputeForces on the updated state), then
 * the reference's per-step lines (src/kernel.c:198-212) and the record */
static void finish_step(GEOMETRY *g, const kf_options *o, step_pending *p, kf_step_record *rec, int max_rec)
{
  kfs_forces_collect(g, p->r.forces);
  if (o->verbose) {
    printf(">> GMRES: TOTAL ITERATIONS %d\n", p->r.krylov);
    printf("STEP: %d CFL: %g NORM: %g\n", p->r.step, p->r.cfl, p->r.fnorm);
    printf("LIFT: %g DRAG: %g MOMN: %g\n", p->r.forces[0], p->r.forces[1], p->r.forces[2]);
    fflush(stdout);
  }
  if (rec && p->idx < max_rec) rec[p->idx] = p->r;
  p->valid = 0;
}

int kf_kernel(GEOMETRY *g, kf_step_record *rec, int max_rec)
{
  const kf_options o = kf_get_options(g);
  learn_state *L = (learn_state *)*kfs_learn_state(g);
  kspace KS, *K = &KS;
  memset(K, 0, sizeof(KS));
  K->g = g;
  K->o = o;
  K->sz = g->c->sz;
  K->restart = o.gmres_restart;
  K->nb = o.gmres_restart + 1;
  K->flexible = o.krylov == KF_KRYLOV_FGMRES || L != NULL;
  K->L = L;
  const size_t sz = K->sz;
  double *gxyz[3], *x;
  kfs_work(g, K->flexible ? K->nb + K->restart : K->nb, gxyz, &K->r, &K->w, &K->t, &x, &K->vecs);
  K->grad.x0 = gxyz[0];
  K->grad.x1 = gxyz[1];
  K->grad.x2 = gxyz[2];
  K->ghh = (double *)calloc((size_t)K->nb * K->nb + K->nb, sizeof(double));
  double *x2 = NULL; /* judged steps: the second solution */
  if (L) {
    if (L->lo.judge_every < 1) L->lo.judge_every = 1;
    x2 = (double *)kf_device_alloc(sz * sizeof(double));
  }
  step_pending pend;
  memset(&pend, 0, sizeof(pend));
  double fnorm_init = 0.0, cfl = o.cfl_init, cfl_factored = -1.0;
  int its = 0, nrec = 0;
  while (its < o.ts_max) {
    /* the residual also leaves the time-step sums of this state */
    kfs_flux_spectral(g, g->q->q, &K->grad, K->r);
    kfs_nrm2(g, sz, K->r, KFS_SLOT_FNORM);
    const double fnorm = fetch1(g, KFS_SLOT_FNORM);
    const int derr = kf_last_error(g);
    if (derr) {
      fprintf(stderr, "kf: device error word %d at pseudo-step %d (%s%s); the solve stops\n", derr, its,
              (derr & 1) ? "non-finite pivot " : "", (derr & 2) ? "a device wait expired" : "");
      break;
    }
    if (!isfinite(fnorm)) {
      fprintf(stderr, "kf: non-finite residual norm at pseudo-step %d\n", its);
      break;
    }
    if (its == 0) fnorm_init = fnorm;
    else {
      double cfl0 = 1.1 * o.cfl_init * fnorm_init;
      cfl0 /= fnorm;
      cfl = (cfl0 < o.cfl_max) ? cfl0 : o.cfl_max;
    }
    kfs_timestep_spectral(g, cfl); /* = ComputeTimeStep(cfl, g) on the same state */
    /* lag_factors = k: the previous factorisation is kept while the CFL
     * stays in [its CFL, k times it) -- k = 1: while unchanged (the ramp at
     * its cap); the operator stays exact.  Not with the BSR operator, whose
     * operator is the fill itself. */
    const int lag = o.lag_factors > 0 && its > 0 && !L && !o.bsr_operator && cfl >= cfl_factored &&
                    (cfl == cfl_factored || cfl < o.lag_factors * cfl_factored);
    if (!lag) ComputeA(g);
    if (L && its < L->lo.collect_steps) kf_mpnn_take_matrix(L->net, 1 + its); /* the loss metric */
    kfs_nrm2(g, sz, g->q->q, KFS_SLOT_QNORM);
    K->qnorm = sqrt(1.0 + fetch1(g, KFS_SLOT_QNORM));
    if (!lag) {
      ComputeNumericalILU(g);
      cfl_factored = cfl;
    }
    int total, learned = 0, k_ilu = -1, k_net = -1;
    if (L && L->enabled) {
      kf_mpnn_take_diag(L->net, 0);
      kf_mpnn_prepare(L->net, 0); /* G of this factorisation */
      if (L->lo.ablate || L->since_judge % L->lo.judge_every == 0) {
        /* judged: ILU(0) alone, then with the correction on the same system,
         * stopped once it has used as many iterations (it has lost then) */
        k_ilu = ksolve(K, K->r, x2, o.gmres_rtol, o.gmres_max_it, 0, 0, NULL);
        int conv = 0;
        k_net = ksolve(K, K->r, x, o.gmres_rtol, o.gmres_max_it, 1, L->lo.ablate ? 0 : k_ilu, &conv);
        L->st.steps_judged++;
        L->last_ilu = k_ilu;
        const int won = conv && k_net < k_ilu;
        if (won) L->st.judged_won++;
        if (won && !L->lo.ablate) {
          learned = 1;
          total = k_net;
        } else {
          kfs_newaxpy(g, sz, 0.0, x2, x2, x); /* x = x2: keep ILU(0)'s update */
          total = k_ilu;
          if (!L->lo.ablate) {
            L->enabled = 0;
            L->st.fallback = 1;
          }
        }
      } else {
        /* between judged steps: the correction alone, redone with ILU(0)
         * alone (which then takes over) if it needs more than twice the
         * iterations of the last judged ILU(0) solve */
        int conv = 0;
        k_net = ksolve(K, K->r, x, o.gmres_rtol, o.gmres_max_it, 1, 2 * L->last_ilu, &conv);
        if (conv) {
          total = k_net;
          learned = 1;
        } else {
          total = k_ilu = ksolve(K, K->r, x, o.gmres_rtol, o.gmres_max_it, 0, 0, NULL);
          L->enabled = 0;
          L->st.fallback = 1;
        }
      }
      L->since_judge++;
      L->st.steps_learned += learned;
    } else {
      total = ksolve(K, K->r, x, o.gmres_rtol, o.gmres_max_it, 0, 0, NULL);
      if (L) k_ilu = total;
    }
    if (L && its < L->lo.collect_steps) {
      collect_pairs(K, its);
      if (its == L->lo.collect_steps - 1) {
        train(K);
        L->enabled = 1;
        L->since_judge = 0;
      }
    }
    if (K->failed) {
      fprintf(stderr, "kf: non-finite Krylov process at pseudo-step %d (device error word %d); the solve stops "
              "before the update\n", its, kf_last_error(g));
      break;
    }
    /* the previous step's forces (src/kernel.c:202) are collected before the
     * state changes again; they ran beside this step's work */
    if (pend.valid) finish_step(g, &o, &pend, rec, max_rec);
    kfs_axpy(g, sz, -1.0, x, g->q->q);
    kfs_forces_async(g);
    pend.valid = 1;
    pend.idx = nrec;
    pend.r.step = its;
    pend.r.cfl = cfl;
    pend.r.fnorm = fnorm;
    pend.r.krylov = total;
    pend.r.learned = learned;
    pend.r.krylov_ilu = k_ilu;
    pend.r.krylov_net = k_net;
    nrec++;
    /* src/kernel.c:214 rounds the ratio; compared in double (no uint32 wrap) */
    if (round(fnorm_init / fnorm) >= o.fnorm_ratio) break;
    its++;
  }
  if (pend.valid) finish_step(g, &o, &pend, rec, max_rec);
  if (x2) kf_device_free(x2);
  free(K->ghh);
  return nrec;
}

#undef VEC
#undef ZVEC

void Kernel(GEOMETRY *g) { kf_kernel(g, NULL, 0); }
