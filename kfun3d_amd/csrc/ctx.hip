/*
 * ctx.hip -- device context (g->matrix) and the C-ABI entry points that
 * replace the reference's inc/core_kernel.h, inc/utils.h and src/mesh.c:22.
 */
#include <math.h>
#include <string.h>

#include "kf_dev.h"

template <typename T>
static T *dalloc(size_t n)
{
  void *p = NULL;
  KF_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
  return (T *)p;
}

template <typename T>
static T *dupload(const T *h, size_t n)
{
  T *p = dalloc<T>(n);
  if (n) KF_CHECK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

/* launch-shape knobs of the level-scheduled kernels (tuning only) */
static void ctx_knobs(kf_ctx *c)
{
  const char *ev;
  /* triangular-solve waves: sized from the schedule unless set */
  c->trsv_waves = (ev = getenv("KF_TRSV_WAVES")) ? (uint32_t)atoi(ev) : 0;   /* 0: sized per launch */
  c->ilu_waves = (ev = getenv("KF_ILU_WAVES")) ? (uint32_t)atoi(ev) : 0;   /* 0: by precision */
  c->d.spin_sleep = (ev = getenv("KF_SPIN_SLEEP")) ? (uint32_t)atoi(ev) : 2;
  c->trsv_group = (ev = getenv("KF_TRSV_GROUP")) ? (uint32_t)atoi(ev) : 4;
  c->ilu_group = (ev = getenv("KF_ILU_GROUP")) ? (uint32_t)atoi(ev) : 4;
  if (c->trsv_group < 1 || c->trsv_group > 4) c->trsv_group = 4;
  if (c->ilu_group < 1 || c->ilu_group > 4) c->ilu_group = 4;
  int dev = 0, lds = 0;
  KF_CHECK(hipGetDevice(&dev));
  KF_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  c->lds_max = lds > 0 ? (size_t)lds : 64 * 1024;
}

/* the block matrix of a context: pattern, ILU(0) and triangular-solve plans,
 * tickets and the FP64 value array (a mesh's Jacobian, or the coarse
 * Galerkin matrix of the two-level preconditioner) */
static void ctx_matrix_setup(kf_ctx *c, const kf_prep *p)
{
  kf_dv &d = c->d;
  const size_t N = p->N;
  const char *ev;
  /* BSR pattern and level schedule */
  d.rowptr = dupload(p->rowptr, N + 1);
  d.col = dupload(p->col, p->nnzb);
  d.diag = dupload(p->diag, N);
  c->trsv_pf = 7;
  for (size_t i = 0; i < N; i++) {
    if (p->rowptr[i + 1] - p->rowptr[i] > c->maxrow) c->maxrow = p->rowptr[i + 1] - p->rowptr[i];
    if (p->diag[i] - p->rowptr[i] > 7 || p->rowptr[i + 1] - p->diag[i] - 1 > 7) c->trsv_pf = 8;
  }
  if ((ev = getenv("KF_TRSV_HELD")) && atoi(ev) == 8) c->trsv_pf = 8;   /* 8 held entries on any mesh */
  /* GMRES's combination formed by the forward sweep: -1 where the sweeps are
   * throughput-bound, 0 never, 1 always */
  c->trsv_fold = (ev = getenv("KF_TRSV_FOLD")) ? (atoi(ev) != 0) : -1;
  c->fd_res = (ev = getenv("KF_FD_RESIDUAL")) ? (atoi(ev) != 0) : 1;
  kf_phase("uploads");
  {
    kf_ilu_plan ip;
    kf_plan_ilu(p, &ip);
    kf_phase("ILU plan");
    c->ilu_rec = dupload(ip.rec, 8 * N);
    c->ilu_ldiag = dupload(ip.ldiag, p->nnzb);
    c->ilu_moff = dupload(ip.moff, p->nnzb);
    c->ilu_mjm = dupload(ip.mjm, ip.nmatch);
    c->ilu_mlo = dupload(ip.mlo, ip.nmatch);
    c->ilu_maxrow = ip.maxrow;
    c->ilu_maxlow = ip.maxlow;
    c->ilu_maxmatch = ip.maxmatch;
    kf_free_ilu_plan(&ip);
  }
  c->nlev_l = p->nlev_l;
  c->nlev_u = p->nlev_u;
  c->lrows = dupload(p->lrows, N);   /* row order of the ILU kernel without a plan */
  c->rdone = dalloc<uint32_t>(N);
  KF_CHECK(hipMemset(c->rdone, 0, N * sizeof(uint32_t)));
  {
    /* triangular solves: ranges of R rows (tuning knob, a power of two
     * up to 16) */
    uint32_t R = (ev = getenv("KF_TRSV_ROWS")) ? (uint32_t)atoi(ev) : 4;
    if (R != 1 && R != 2 && R != 4 && R != 8 && R != 16) R = 4;
    c->trsv_R = R;
    c->trsv_S = KF_GROUP / R;
    /* ranges clustered across levels into tasks (1) or tasks of one level (0) */
    const int cluster = (ev = getenv("KF_TRSV_CLUSTER")) ? atoi(ev) != 0 : 1;
    kf_trsv_plan t, tu;
    kf_phase("ILU plan upload");
    kf_plan_trsv_both(p, R, cluster, &t, &tu);
    kf_phase("solve plans");
    c->trow_l = dupload(t.trow, (size_t)t.ntask * KF_GROUP * 3);
    c->tlvl_l = dupload(t.tlvl, (size_t)t.ntask * KF_GROUP);
    c->tdep_l = dupload(t.tdep, t.ntask);
    c->ntask_l = t.ntask;
    {
      /* about 3.5 levels of tasks in flight (measured best at C2: 1024
       * waves there, 0.34 / 0.37 ms against 0.35 / 0.39 ms with 2048):
       * enough waves to start every task of a level as soon as it is ready,
       * few enough that waves waiting on later levels do not crowd the
       * memory-side polls.  The launch caps it by precision. */
      const uint32_t w = t.nlevel ? (uint32_t)(7ull * t.ntask / (2ull * t.nlevel)) : 1;
      c->trsv_auto = w < 512 ? 512 : w;
    }
    /* a block position is read by one sweep only (left of the diagonal by
     * the forward, right of it by the backward sweep) and holds 0xff in
     * the other's table: the two tables merge by bitwise and */
    uint8_t *es = t.eslot;
    t.eslot = NULL;
    kf_free_trsv_plan(&t);
    t = tu;
    c->trow_u = dupload(t.trow, (size_t)t.ntask * KF_GROUP * 3);
    c->tlvl_u = dupload(t.tlvl, (size_t)t.ntask * KF_GROUP);
    c->tdep_u = dupload(t.tdep, t.ntask);
    c->ntask_u = t.ntask;
    for (size_t j = 0; j < p->nnzb; j++) es[j] &= t.eslot[j];
    c->eslot = dupload(es, p->nnzb);
    free(es);
    kf_free_trsv_plan(&t);
    c->ystate[0] = c->ystate[1] = ~(uint64_t)0;
  }
  c->ticket = dalloc<uint32_t>(4);
  KF_CHECK(hipMemset(c->ticket, 0, 4 * sizeof(uint32_t)));
  c->a64 = dalloc<double>(16 * p->nnzb);
  KF_CHECK(hipMemset(c->a64, 0, 16 * p->nnzb * sizeof(double)));

}

/* rows of each level cut into groups of at most KF_GROUP */
extern "C" void kf_ctx_create(GEOMETRY *g, const kf_prep *p, const kf_options *o)
{
  kf_ctx *c = new kf_ctx();
  memset((void *)c, 0, sizeof(*c));
  kf_check_options(o);
  KF_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->opt = *o;
  kf_dv &d = c->d;
  const char *ev;
  ctx_knobs(c);
  const size_t N = p->N, E = p->E;
  if (N >= (1u << 31) || E >= (1u << 30) || p->nnzb >= (1u << 31)) {
    fprintf(stderr, "kf: mesh too large for 32-bit indices\n");
    abort();
  }
  d.N = (uint32_t)N;
  d.E = (uint32_t)E;
  d.S = (uint32_t)g->b->fc->sz;
  d.NS = (uint32_t)g->b->s->sz;
  d.NF = (uint32_t)g->b->f->sz;
  kf_free_stream(o->alpha_deg, &d.U, &d.V);
  c->nnzb = p->nnzb;

  {
    /* per-edge records: (n0, n1) and (unit normal, area) interleaved */
    uint32_t *hn = (uint32_t *)kf_xmalloc((2 * E + 2) * sizeof(uint32_t));
    double *hg = (double *)kf_xmalloc((4 * E + 4) * sizeof(double));
    kf_edge_records(g, hn, hg);
    d.en = dupload(hn, 2 * E);
    d.egeo = dupload(hg, 4 * E);
    free(hn);
    free(hg);
  }
  d.x = dupload(g->n->xyz->x0, N);
  d.y = dupload(g->n->xyz->x1, N);
  d.z = dupload(g->n->xyz->x2, N);
  {
    double *h = (double *)kf_xmalloc(4 * N * sizeof(double));
    for (size_t i = 0; i < N; i++) {
      h[4 * i] = g->n->xyz->x0[i];
      h[4 * i + 1] = g->n->xyz->x1[i];
      h[4 * i + 2] = g->n->xyz->x2[i];
      h[4 * i + 3] = 0.0;
    }
    d.xyzw = dupload(h, 4 * N);
    free(h);
  }

  /* sliced ELL of the incidence lists (slice height 64), built on the host */
  kf_sell sl;
  kf_build_sell(g, p, &sl);
  const uint32_t nslice = sl.nslice;
  d.nslice = nslice;
  const size_t ne = sl.ne;
  uint32_t *sptr = sl.sptr, *deg = sl.deg, *nbr = sl.nbr, *ent = sl.ent, *blk = sl.blk;
  double *wx = sl.wx, *wy = sl.wy, *wz = sl.wz;
  kf_phase("sliced ELL (host)");
  d.sell_ptr = dupload(sptr, nslice + 1);
  d.deg = dupload(deg, N);
  d.sell_nbr = dupload(nbr, ne);
  d.sell_ent = dupload(ent, ne);
  d.sell_blk = dupload(blk, ne);
  d.sell_wx = dupload(wx, ne);
  d.sell_wy = dupload(wy, ne);
  d.sell_wz = dupload(wz, ne);
  free(sptr); free(deg); free(nbr); free(ent); free(blk); free(wx); free(wy); free(wz);
  kf_grad_setup(c);
  kf_phase("LSQ moments (device)");

  /* boundary */
  d.sf0 = dupload(g->b->fc->fptr->n0, d.S);
  d.sf1 = dupload(g->b->fc->fptr->n1, d.S);
  d.sf2 = dupload(g->b->fc->fptr->n2, d.S);
  d.ncolor = (uint32_t)g->t->sz;
  d.color = dupload(g->t->i, (size_t)d.ncolor + 1);
  d.sfptr = dupload(p->sfptr, N + 1);
  d.sfent = dupload(p->sfent, 3 * (size_t)d.S);
  d.snx = dupload(g->b->s->xyz->x0, d.NS);
  d.sny = dupload(g->b->s->xyz->x1, d.NS);
  d.snz = dupload(g->b->s->xyz->x2, d.NS);
  d.snptr = dupload(p->snptr, N + 1);
  d.snent = dupload(p->snent, d.NS);
  d.fnx = dupload(g->b->f->xyz->x0, d.NF);
  d.fny = dupload(g->b->f->xyz->x1, d.NF);
  d.fnz = dupload(g->b->f->xyz->x2, d.NF);
  d.fnptr = dupload(p->fnptr, N + 1);
  d.fnent = dupload(p->fnent, d.NF);

  ctx_matrix_setup(c, p);

  /* state and work */
  c->q = dalloc<double>(4 * N);
  c->cdt = dalloc<double>(N);
  KF_CHECK(hipMemset(c->cdt, 0, N * sizeof(double)));
  d.eres = dalloc<double>(4 * E + 8);
  d.red = dalloc<double>(2 * KF_NRED);   /* two sets of partial sums (fused MGS) */
  c->mgs_part = dalloc<double>(2 * (size_t)KF_MGS_MAXV * KF_NRED);
  c->mgs_sync = dalloc<uint32_t>(4);
  c->mgs_res = (ev = getenv("KF_MGS_RESIDENT")) ? atoi(ev) : 1;
  if (c->mgs_res == 2) c->mgs_res = 1;
  if (c->mgs_res < 0 || c->mgs_res > 3) c->mgs_res = 1;
  for (int i = 0; i < 8; i++) c->mgs_ok[i] = -1;
  kf_mgs_setup(c);   /* the first resident launch's table */
  d.scal = dalloc<double>(KF_NSCAL);
  KF_CHECK(hipMemset(d.scal, 0, KF_NSCAL * sizeof(double)));
  d.err = dalloc<uint32_t>(1);
  KF_CHECK(hipMemset(d.err, 0, sizeof(uint32_t)));
  /* coherent: the resident orthogonalisation's stores reach the host while
   * the launch still runs (its ready flag is polled) */
  KF_CHECK(hipHostMalloc((void **)&c->hscal, KF_NSCAL * sizeof(double), hipHostMallocCoherent));
  /* KF_MGS_NOPOLL=1: wait for the resident launch by synchronising the
   * stream instead of polling its flag (the values are the same) */
  c->mgs_nopoll = (ev = getenv("KF_MGS_NOPOLL")) ? atoi(ev) != 0 : 0;

  /* free-stream state, src/guess.c:8-20 */
  double *qh = (double *)kf_xmalloc(4 * N * sizeof(double));
  for (size_t i = 0; i < N; i++) {
    qh[4 * i + 0] = 1.0; qh[4 * i + 1] = d.U; qh[4 * i + 2] = d.V; qh[4 * i + 3] = 0.0;
  }
  KF_CHECK(hipMemcpy(c->q, qh, 4 * N * sizeof(double), hipMemcpyHostToDevice));
  free(qh);

  kf_noalloc_setup(c);   /* the device address of the no-allocation solve's fixed array */

  g->q->q = c->q;
  g->n->cdt = c->cdt;
  g->matrix = c;
}

static void kf_unbind_ctx(kf_ctx *c);

static void dfree(void *p)
{
  if (p) KF_CHECK(hipFree(p));
}

/* A context holding only a block matrix -- the coarse level of the
 * two-level preconditioner: pattern, plans, tickets, values and factors.
 * It shares the parent's stream (so its launches are ordered with the
 * parent's) and the parent's device error word. */
kf_ctx *kf_ctx_create_matrix(const kf_prep *p, const kf_options *o, const kf_ctx *parent)
{
  kf_ctx *c = new kf_ctx();
  memset((void *)c, 0, sizeof(*c));
  c->opt = *o;
  c->stream = parent->stream;
  ctx_knobs(c);
  c->d.N = (uint32_t)p->N;
  c->d.err = parent->d.err;
  c->nnzb = p->nnzb;
  ctx_matrix_setup(c, p);
  c->trsv_fold = 0;   /* its right-hand sides are formed before each solve */
  return c;
}

void kf_ctx_destroy_matrix(kf_ctx *c)
{
  if (!c) return;
  KF_CHECK(hipStreamSynchronize(c->stream));
  void *ptrs[] = {c->d.rowptr, c->d.col, c->d.diag, c->lrows, c->rdone, c->trow_l, c->trow_u, c->tlvl_l,
                  c->tlvl_u, c->tdep_l, c->tdep_u, c->eslot, c->ty64, c->ilu_rec, c->ilu_ldiag, c->ilu_moff,
                  c->ilu_mjm, c->ilu_mlo, c->ilu_apub, c->ilu_udpub, c->ticket, c->a64, c->a32, c->abf, c->lold,
                  c->ud, c->f32a, c->f32b, c->jac64, c->jac32, c->jsl_val, c->jsl_col};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) dfree(ptrs[i]);
  delete c;
}

extern "C" void kf_ctx_destroy(GEOMETRY *g)
{
  if (!g->matrix) return;
  kf_ctx *c = (kf_ctx *)g->matrix;
  kfs_learn_detach(g);
  kf_coarse_destroy(c);   /* on the context's stream, with its error word */
  KF_CHECK(hipStreamSynchronize(c->stream));
  if (c->fstream) {
    KF_CHECK(hipStreamSynchronize(c->fstream));
    KF_CHECK(hipEventDestroy(c->ev_upd));
    KF_CHECK(hipEventDestroy(c->ev_forces));
    KF_CHECK(hipStreamDestroy(c->fstream));
  }
  kf_dv &d = c->d;
  void *ptrs[] = {d.en, d.egeo, d.x, d.y, d.z, d.xyzw, d.sell_ptr, d.deg, d.sell_nbr, d.sell_ent,
                  d.sell_blk, d.sell_wx, d.sell_wy, d.sell_wz, d.lsq, d.sell_nbs, d.sf0, d.sf1, d.sf2, d.color, d.sfptr, d.sfent, d.snx, d.sny,
                  d.snz, d.snptr, d.snent, d.fnx, d.fny, d.fnz, d.fnptr, d.fnent, d.rowptr, d.col, d.diag, d.eres,
                  d.red, d.scal, d.err, c->lrows, c->rdone,
                  c->trow_l, c->trow_u, c->tlvl_l, c->tlvl_u, c->tdep_l, c->tdep_u, c->eslot, c->ty64, c->ilu_rec, c->ilu_ldiag, c->ilu_moff, c->ilu_mjm, c->ilu_mlo, c->ilu_apub, c->ilu_udpub, c->ticket, c->a64, c->a32,
                  c->abf, c->jop64, c->jop32, c->ftbuf, c->lold, c->ud, c->q, c->cdt, c->gx, c->gy, c->gz, c->r, c->w, c->t, c->xk, c->vecs, c->f32a, c->f32b, c->jac64, c->jac32, c->jsl_val, c->jsl_col, c->eterm, c->sr, c->mgs_part, c->mgs_sync};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) dfree(ptrs[i]);
  KF_CHECK(hipHostFree(c->hscal));
  if (c->mgs_ev) KF_CHECK(hipEventDestroy(c->mgs_ev));
  KF_CHECK(hipStreamDestroy(c->stream));
  kf_unbind_ctx(c);
  delete c;
  g->matrix = NULL;
  if (g->q) g->q->q = NULL;
  if (g->n) g->n->cdt = NULL;
}

/* ---- options ---------------------------------------------------------------- */

extern "C" kf_options kf_default_options(void)
{
  kf_options o;
  memset(&o, 0, sizeof(o));
  o.alpha_deg = 3.0;        /* inc/phy.h:8-9 */
  o.cfl_init = 50.0;        /* src/kernel.c:19 */
  o.cfl_max = 100000.0;     /* src/kernel.c:20 */
  o.ts_max = 15;            /* src/kernel.c:21 */
  o.fnorm_ratio = 1e7;      /* src/kernel.c:22 */
  o.gmres_restart = 30;     /* src/kernel.c:15 */
  o.gmres_rtol = 0.01;      /* src/kernel.c:14 */
  o.gmres_max_it = 10000;   /* src/kernel.c:17 */
  o.first_order = 0;
  o.prec = KF_PREC_FP64;
  o.ortho = KF_ORTHO_MGS;
  o.verbose = 0;
  o.ilu_exact = 1;
  o.reduce_seq = 0;
  o.krylov = KF_KRYLOV_GMRES;
  o.bsr_operator = 0;
  o.jacobi_sweeps = 0;
  o.coarse = 0;
  o.coarse_sweeps = 0;
  o.lag_factors = 0;
  return o;
}

static void set_free_stream_state(kf_ctx *c);

void kf_check_options(const kf_options *o)
{
  if (o->gmres_restart < 1 || o->gmres_restart > 31) {
    fprintf(stderr, "kf: gmres_restart must be in [1, 31]\n");
    abort();
  }
  if (o->jacobi_sweeps < 0 || o->jacobi_sweeps > KF_JACOBI_MAX) {
    fprintf(stderr, "kf: jacobi_sweeps must be in [0, %d]\n", KF_JACOBI_MAX);
    abort();
  }
  if (o->lag_factors < 0 || o->lag_factors > 1000) {
    fprintf(stderr, "kf: lag_factors must be in [0, 1000]\n");
    abort();
  }
  if (o->coarse_sweeps < 0 || o->coarse_sweeps > KF_JACOBI_MAX) {
    fprintf(stderr, "kf: coarse_sweeps must be in [0, %d]\n", KF_JACOBI_MAX);
    abort();
  }
  if (o->coarse < 0 || o->coarse > KF_COARSE_MAX) {
    fprintf(stderr, "kf: coarse must be in [0, %d]\n", KF_COARSE_MAX);
    abort();
  }
}

extern "C" void kf_set_options(GEOMETRY *g, const kf_options *o)
{
  kf_ctx *c = kf_ctx_of(g);
  kf_check_options(o);
  const int alpha_changed = o->alpha_deg != c->opt.alpha_deg;
  c->opt = *o;
  if (alpha_changed) {
    kf_free_stream(o->alpha_deg, &c->d.U, &c->d.V);
    set_free_stream_state(c);
  }
}

extern "C" kf_options kf_get_options(const GEOMETRY *g) { return kf_ctx_of(g)->opt; }

static void set_free_stream_state(kf_ctx *c)
{
  const size_t N = c->d.N;
  double *qh = (double *)kf_xmalloc(4 * N * sizeof(double));
  for (size_t i = 0; i < N; i++) {
    qh[4 * i + 0] = 1.0; qh[4 * i + 1] = c->d.U; qh[4 * i + 2] = c->d.V; qh[4 * i + 3] = 0.0;
  }
  KF_CHECK(hipMemcpyAsync(c->q, qh, 4 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  KF_CHECK(hipStreamSynchronize(c->stream));
  free(qh);
}

/* ---- setup ------------------------------------------------------------------ */

extern "C" void kf_setup_host(const char *fpath, GEOMETRY *g)
{
  kf_phase(NULL);
  memset(g, 0, sizeof(*g));
  kf_read_msh(fpath, g);
  kf_host_setup(g);
}

extern "C" void kf_bind(GEOMETRY *g);

extern "C" void kf_setup_ex(const char *fpath, GEOMETRY *g, const kf_options *o)
{
  kf_setup_host(fpath, g);
  kf_prep p;
  kf_build_prep(g, &p);
  kf_ctx_create(g, &p, o);
  kf_free_prep(&p);
  kf_bind(g);
  kf_phase("device context");
}

extern "C" void Setup(const char *fpath, GEOMETRY *g)
{
  kf_options o = kf_default_options();
  kf_setup_ex(fpath, g, &o);
}

extern "C" void kf_geometry_free(GEOMETRY *g)
{
  kf_ctx_destroy(g);
  kf_free_host_geometry(g);
}

extern "C" int kf_last_error(const GEOMETRY *g)
{
  kf_ctx *c = kf_ctx_of(g);
  uint32_t e = 0;
  KF_CHECK(hipMemcpyAsync(&e, c->d.err, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  KF_CHECK(hipStreamSynchronize(c->stream));
  return (int)e;
}

/* ---- memory helpers ---------------------------------------------------------- */

extern "C" void kf_set_state(GEOMETRY *g, const double *q_host)
{
  kf_ctx *c = kf_ctx_of(g);
  KF_CHECK(hipMemcpyAsync(c->q, q_host, 4 * (size_t)c->d.N * sizeof(double), hipMemcpyHostToDevice, c->stream));
  KF_CHECK(hipStreamSynchronize(c->stream));
}

extern "C" void kf_get_state(const GEOMETRY *g, double *q_host)
{
  kf_ctx *c = kf_ctx_of(g);
  KF_CHECK(hipMemcpyAsync(q_host, c->q, 4 * (size_t)c->d.N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KF_CHECK(hipStreamSynchronize(c->stream));
}

extern "C" void kf_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
  KF_CHECK(hipDeviceSynchronize());
  KF_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

extern "C" void kf_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
  KF_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  KF_CHECK(hipDeviceSynchronize());
}

extern "C" void *kf_device_alloc(size_t bytes)
{
  void *p = NULL;
  KF_CHECK(hipMalloc(&p, bytes ? bytes : 8));
  KF_CHECK(hipMemset(p, 0, bytes ? bytes : 8));
  return p;
}

extern "C" void kf_device_free(void *p) { dfree(p); }
extern "C" void kf_synchronize(const GEOMETRY *g) { KF_CHECK(hipStreamSynchronize(kf_ctx_of(g)->stream)); }
extern "C" void *kf_stream(const GEOMETRY *g) { return (void *)kf_ctx_of(g)->stream; }
extern "C" size_t kf_nnzb(const GEOMETRY *g) { return kf_ctx_of(g)->nnzb; }

/* whether GMRES's inner step has the forward sweep form the finite-difference
 * combination at the current precision (where the sweeps are
 * throughput-bound, or KF_TRSV_FOLD=1) */
extern "C" int kf_trsv_fold(const GEOMETRY *g)
{
  kf_ctx *c = kf_ctx_of(g);
  if (c->opt.jacobi_sweeps > 0) return 0;   /* the combination runs first (every forward sweep reads b) */
  return !(c->trsv_fold == 0 || (c->trsv_fold < 0 && !kf_trsv_capped(c)));
}

/* whether the gradient forms its LSQ terms from per-vertex moments */
extern "C" int kf_grad_moments(const GEOMETRY *g) { return kf_ctx_of(g)->grad_mom; }

extern "C" void kf_pattern(const GEOMETRY *g, unsigned int *rowptr, unsigned int *col, unsigned int *diag)
{
  kf_ctx *c = kf_ctx_of(g);
  KF_CHECK(hipStreamSynchronize(c->stream));
  if (rowptr) KF_CHECK(hipMemcpy(rowptr, c->d.rowptr, ((size_t)c->d.N + 1) * 4, hipMemcpyDeviceToHost));
  if (col) KF_CHECK(hipMemcpy(col, c->d.col, c->nnzb * 4, hipMemcpyDeviceToHost));
  if (diag) KF_CHECK(hipMemcpy(diag, c->d.diag, (size_t)c->d.N * 4, hipMemcpyDeviceToHost));
}

extern "C" void kf_get_blocks(const GEOMETRY *g, double *blocks_host)
{
  kf_ctx *c = kf_ctx_of(g);
  KF_CHECK(hipStreamSynchronize(c->stream));
  if (c->opt.prec == KF_PREC_FP64) {
    KF_CHECK(hipMemcpy(blocks_host, c->a64, c->nnzb * 16 * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    float *tmp = (float *)kf_xmalloc(c->nnzb * 16 * sizeof(float));
    KF_CHECK(hipMemcpy(tmp, c->a32, c->nnzb * 16 * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < c->nnzb * 16; i++) blocks_host[i] = tmp[i];
    free(tmp);
  }
}

/* ---- reference entry points ---------------------------------------------------- */

extern "C" void ComputeFlux(const GEOMETRY *g, const double *q, GRADIENT *grad, double *r)
{
  kf_ctx *c = kf_ctx_of(g);
  kf_launch_flux(c, q, grad->x0, grad->x1, grad->x2, r, NULL);
}

/* iJacVec's residual and combination in one pass (src/kernel.c:27-44):
 * y = ((R(w) - r) + cdt (w - q)) |x| / h, the residual's last kernel
 * forming it before its store; |x| in scal[slot] */
extern "C" void kfs_flux_fd(GEOMETRY *g, const double *w, GRADIENT *grad, double *y, const double *r, double h,
                            int slot)
{
  kf_ctx *c = kf_ctx_of(g);
  if (y == w || y == r) {
    fprintf(stderr, "kf: kfs_flux_fd: y must not alias w or r\n");
    abort();
  }
  kf_check_aligned(r, "r");
  if (!c->fd_res) {   /* KF_FD_RESIDUAL=0: the two passes */
    kf_launch_flux(c, w, grad->x0, grad->x1, grad->x2, y, NULL);
    kf_fd_combine(c, r, w, h, slot, y);
    return;
  }
  const kf_fdb fd = {r, w, c->q, c->cdt, c->d.scal + slot, h};
  kf_launch_flux(c, w, grad->x0, grad->x1, grad->x2, y, NULL, &fd);
}

extern "C" void ComputeTimeStep(const double cfl, GEOMETRY *g) { kf_launch_timestep(kf_ctx_of(g), cfl); }

/* Driver-internal pair: the residual of the current state together with its
 * time-step sums, then the time step from those sums (the state must not
 * change in between; Kernel() calls them back to back). */
extern "C" void kfs_flux_spectral(GEOMETRY *g, const double *q, GRADIENT *grad, double *r)
{
  kf_ctx *c = kf_ctx_of(g);
  if (!c->sr) KF_CHECK(hipMalloc((void **)&c->sr, (size_t)c->d.N * sizeof(double)));
  kf_launch_flux(c, q, grad->x0, grad->x1, grad->x2, r, c->sr);
}

extern "C" void kfs_timestep_spectral(GEOMETRY *g, double cfl)
{
  kf_ctx *c = kf_ctx_of(g);
  if (!c->sr) {
    fprintf(stderr, "kf: time step from sums requested before a residual with sums\n");
    abort();
  }
  kf_launch_timestep_div(c, c->sr, cfl);
}

extern "C" void ComputeForces(const GEOMETRY *g, double *forces) { kf_launch_forces(kf_ctx_of(g), forces); }

/* Kernel's forces: started after an update on a second stream, collected
 * (bits of ComputeForces) before the state changes again */
extern "C" void kfs_forces_async(GEOMETRY *g) { kf_forces_async(kf_ctx_of(g)); }
extern "C" void kfs_forces_collect(GEOMETRY *g, double *forces) { kf_forces_collect(kf_ctx_of(g), forces); }

extern "C" void ComputeA(GEOMETRY *g)
{
  kf_ctx *c = kf_ctx_of(g);
  kf_launch_fill(c);
  if (c->opt.bsr_operator) kf_keep_jacobian(c);   /* the ILU factors the fill's array in place */
}

/* y = A x with the Jacobian of the last ComputeA (kept when bsr_operator is
 * set); no counterpart in the reference, whose Krylov operator is the
 * matrix-free product of src/kernel.c:24-45 */
extern "C" void kf_bsr_spmv(const GEOMETRY *g, const double *x, double *y)
{
  kf_ctx *c = kf_ctx_of(g);
  if (!c->jop64 && !c->jop32) {
    fprintf(stderr, "kf: kf_bsr_spmv needs a Jacobian kept by ComputeA (set kf_options.bsr_operator = 1 first)\n");
    abort();
  }
  if (x == y) {
    fprintf(stderr, "kf: kf_bsr_spmv: x and y must not alias\n");
    abort();
  }
  kf_launch_spmv(c, x, y);
}

extern "C" void ComputeNumericalILU(GEOMETRY *g)
{
  kf_ctx *c = kf_ctx_of(g);
  if (c->opt.coarse) kf_coarse_factor(c);   /* from the fill's values, before they are factored */
  kf_launch_ilu(c);
}

extern "C" void ComputeSparseTriangularSolve(const GEOMETRY *g, const double *b, double *x)
{
  if (b == x) {
    fprintf(stderr, "kf: ComputeSparseTriangularSolve: b and x must not alias\n");
    abort();
  }
  kf_ctx *c = kf_ctx_of(g);
  if (c->opt.coarse) kf_coarse_apply(c, b, x);
  else if (c->opt.jacobi_sweeps > 0) kf_launch_jacobi(c, b, x);   /* matrix/sptrsv.cpp:75-132 */
  else kf_launch_sptrsv(c, b, x, NULL);
}

/* GMRES's inner step: x = U^-1 L^-1 y with y = ((R(w) - r) + cdt (w - q))
 * |x| / h, b = R(w) -- iJacVec's combination (src/kernel.c:33-44), then
 * ComputeSparseTriangularSolve.  Where the sweeps are throughput-bound the
 * forward sweep forms y as it loads b (one pass less); b may be overwritten
 * with y.  |x| in scal[slot_xnorm] */
static bool fd_separate(const kf_ctx *c)
{
  return c->opt.coarse || c->opt.jacobi_sweeps > 0 || c->trsv_fold == 0 || (c->trsv_fold < 0 && !kf_trsv_capped(c));
}

/* 1 when GMRES's inner step has the residual form the combination
 * (kfs_flux_fd, then ComputeSparseTriangularSolve): where kfs_sptrsv_fd
 * would form it on its own before the solve (multilevel, Jacobi sweeps, or
 * latency-bound triangular sweeps), unless KF_FD_RESIDUAL=0 */
extern "C" int kfs_fd_residual(GEOMETRY *g)
{
  const kf_ctx *c = kf_ctx_of(g);
  return c->fd_res && fd_separate(c);
}

extern "C" void kfs_sptrsv_fd(GEOMETRY *g, const double *b, double *x, const double *r, const double *w, double h,
                              int slot_xnorm)
{
  kf_ctx *c = kf_ctx_of(g);
  if (b == x) {
    fprintf(stderr, "kf: kfs_sptrsv_fd: b and x must not alias\n");
    abort();
  }
  if (x == r || x == w) {
    /* the forward sweep presets x while other waves still read r and w */
    fprintf(stderr, "kf: kfs_sptrsv_fd: x must not alias r or w\n");
    abort();
  }
  kf_check_aligned(r, "r"); kf_check_aligned(w, "w");
  if (c->opt.coarse) {   /* restriction and remainder read b: form it once */
    kf_fd_combine(c, r, w, h, slot_xnorm, (double *)b);
    kf_coarse_apply(c, b, x);
    return;
  }
  if (c->opt.jacobi_sweeps > 0) {   /* every forward sweep reads b: form it once */
    kf_fd_combine(c, r, w, h, slot_xnorm, (double *)b);
    kf_launch_jacobi(c, b, x);
    return;
  }
  if (c->trsv_fold == 0 || (c->trsv_fold < 0 && !kf_trsv_capped(c))) {
    /* latency-bound sweeps: the four extra loads at each task's start
     * would lengthen the chain (0.748 against 0.754 s per solve on the
     * 0.33M-vertex lattice), so the combination runs on its own first */
    kf_fd_combine(c, r, w, h, slot_xnorm, (double *)b);
    kf_launch_sptrsv(c, b, x, NULL);
    return;
  }
  const kf_fdb fd = {r, w, c->q, c->cdt, c->d.scal + slot_xnorm, h};
  kf_launch_sptrsv(c, b, x, &fd);
}

/* inc/utils.h:18-28.  The reference's BLAS-1 take no GEOMETRY; these run on
 * the context of the most recent Setup (or kf_bind). */
static kf_ctx *g_current = NULL;

extern "C" void kf_bind(GEOMETRY *g) { g_current = kf_ctx_of(g); }

static void kf_unbind_ctx(kf_ctx *c)
{
  if (g_current == c) g_current = NULL;
}

static kf_ctx *current(void)
{
  if (!g_current) {
    fprintf(stderr, "kf: no device context bound (call Setup or kf_bind first)\n");
    abort();
  }
  return g_current;
}

extern "C" double Compute2ndNorm(const size_t sz, const double *v)
{
  kf_ctx *c = current();
  double out;
  kf_nrm2_dev(c, sz, v, 0);
  kf_scal_fetch(c, 0, 1, &out);
  return out;
}

extern "C" void ComputeAXPY(const size_t sz, const double a, const double *x, double *y) { kf_axpy(current(), sz, a, x, y); }

extern "C" void ComputeNewAXPY(const size_t sz, const double a, const double *x, const double *y, double *w)
{
  kf_newaxpy(current(), sz, a, x, y, w);
}

extern "C" double Normalize(const size_t sz, double *x)
{
  kf_ctx *c = current();
  double out;
  kf_nrm2_dev(c, sz, x, 0);
  kf_scale_recip(c, sz, x, 0);
  kf_scal_fetch(c, 0, 1, &out);
  return out;
}

/* ---- per-kernel timing with HIP events on the context stream ----------------- */

static const char *kf_prof_names[KP_N] = {"grad", "edge_flux", "residual", "timestep", "fill", "ilu",
                                          "trsv_lower", "trsv_upper", "spmv", "jacobi_lower", "jacobi_upper",
                                          "coarse_factor", "coarse"};

static void prof_drain(kf_ctx *c)
{
  kf_prof &p = c->prof;
  if (!p.used) return;
  KF_CHECK(hipStreamSynchronize(c->stream));
  for (int s = 0; s < p.used; s++) {
    float ms = 0.f;
    KF_CHECK(hipEventElapsedTime(&ms, p.ev[2 * s], p.ev[2 * s + 1]));
    p.total_ms[p.id[s]] += ms;
    p.count[p.id[s]]++;
  }
  p.used = 0;
}

int kf_prof_begin(kf_ctx *c, int id)
{
  kf_prof &p = c->prof;
  if (!p.on || (p.on > 1 && id != p.on - 2)) return -1;   /* on = 2 + id: that class only */
  if (p.used == KF_PROF_POOL) prof_drain(c);
  const int s = p.used++;
  p.id[s] = id;
  KF_CHECK(hipEventRecord(p.ev[2 * s], c->stream));
  return s;
}

void kf_prof_end(kf_ctx *c, int slot)
{
  if (slot < 0) return;
  KF_CHECK(hipEventRecord(c->prof.ev[2 * slot + 1], c->stream));
}

/* enable (1) / disable (0) kernel timing; enabling clears the totals */
extern "C" void kf_profile(GEOMETRY *g, int on)
{
  kf_ctx *c = kf_ctx_of(g);
  kf_prof &p = c->prof;
  if (on && !p.ev[0])
    for (int i = 0; i < 2 * KF_PROF_POOL; i++) KF_CHECK(hipEventCreate(&p.ev[i]));
  prof_drain(c);
  if (on)
    for (int k = 0; k < KP_N; k++) { p.total_ms[k] = 0.0; p.count[k] = 0; }
  p.on = on;
}

/* events for one kernel class only (the others run unrecorded): -1 for an
 * unknown name */
extern "C" int kf_profile_only(GEOMETRY *g, const char *name)
{
  for (int k = 0; k < KP_N; k++)
    if (!strcmp(name, kf_prof_names[k])) {
      kf_profile(g, 1);
      kf_ctx_of(g)->prof.on = 2 + k;
      return 0;
    }
  return -1;
}

/* total milliseconds and launch count of one kernel class since kf_profile(g, 1) */
extern "C" int kf_profile_get(GEOMETRY *g, const char *name, double *total_ms, long *count)
{
  kf_ctx *c = kf_ctx_of(g);
  prof_drain(c);
  for (int k = 0; k < KP_N; k++)
    if (!strcmp(name, kf_prof_names[k])) {
      *total_ms = c->prof.total_ms[k];
      *count = c->prof.count[k];
      return 0;
    }
  return -1;
}
