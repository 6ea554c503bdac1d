/*
 * coarse.hip -- the two-level preconditioner (kf_options.coarse) on gfx950.
 *
 * The reference preconditions GMRES with block ILU(0) alone
 * (matrix/ilu.cpp:49-138, matrix/sptrsv.cpp:34-73).  Its Krylov count grows
 * with the mesh: the error modes that ILU(0) leaves are global.  This adds
 * a coarse level over aggregates of neighbouring vertices (host/aggregate.c):
 *   Ac = P^T A P                (Galerkin; P piecewise constant per component)
 *   z  = P Ac^-1 P^T b          (Ac^-1: block ILU(0) of Ac, the same kernels
 *                                as the fine level, in a matrix-only context,
 *                                solved exactly or by coarse_sweeps Jacobi
 *                                sweeps; with coarse = L > 1, the same scheme
 *                                again on Ac, L levels below the fine one)
 *   x  = z + (LU)^-1 (b - A z)  (the fine ILU(0) solve on the remainder,
 *                                exact or by jacobi_sweeps sweeps)
 * a fixed linear preconditioner, so plain left-preconditioned GMRES keeps
 * working.  A is the fill's Jacobian, read before the fine factorisation
 * overwrites it, and enters only as A P: the fine rows with their blocks
 * summed by the aggregate of their column (about 6 blocks a row on a
 * lattice, against 15 in A).  A P serves both the Galerkin product
 * (P^T (A P)) and the remainder b - (A P) z_c, so the full Jacobian is
 * neither kept nor re-read per application.  A P and the Galerkin sums are
 * FP64 with an FP64 preconditioner and FP32 otherwise; the coarse factors
 * take the fine level's precision (FP64, FP32, or BF16-stored with FP32
 * arithmetic).
 *
 * Kernels (a quad of lanes per block row, lane c forming component c, as
 * spmv.hip; workgroups remapped XCD by XCD):
 *   k_gather_sum   a block = the sum of listed blocks in ascending position:
 *                  A P from the fine blocks, then P^T A P from A P
 *   k_restrict     coarse vector = sum of the aggregate's members, ascending
 *   k_coarse_resid t = b - (A P) z_c (FP32 where the fine solve is FP32
 *                  sweeps); k_coarse_resid_sl the same over a sliced BF16
 *                  copy of A P with a BF16 preconditioner
 *   k_prolong_add  x += P z_c (with Jacobi sweeps: inside their last sweep)
 * Every sum runs in a fixed order, so the preconditioner is deterministic.
 */
#include <cstring>

#include "kf_dev.h"

#define CG_WG 256   /* 64 block rows per workgroup */

struct kf_coarse {
  kf_ctx *cc;          /* the coarse matrix: pattern, plans, values, factors */
  uint32_t nc;         /* aggregates */
  size_t cnnzb, nap;
  uint32_t *agg;       /* [N] aggregate of each fine vertex */
  uint32_t *mptr, *mlist;   /* members of each aggregate */
  uint32_t *aprp, *apcol;   /* A P pattern: per fine row, its aggregates */
  uint32_t *apgptr, *apglist;   /* fine block positions of each A P block */
  uint32_t *gptr, *glist;   /* A P positions of each coarse block */
  double *ap64;        /* A P values (FP64 preconditioner) ... */
  float *ap32;         /* ... or FP32 */
  void *apsl;          /* A P packed per 16-row slice and slot in BF16, for the
                          remainder product with a BF16 preconditioner */
  uint32_t *apsl_col;  /* ... the slots' aggregates */
  double *rc, *zc;     /* coarse right-hand side and solution, 4 nc */
  double *t;           /* the fine remainder b - A z, 4 N */
  float *t32;          /* ... in FP32, where the fine solve is FP32 sweeps */
};

template <typename T>
static T *cupload(const T *h, size_t n)
{
  T *p = NULL;
  KF_CHECK(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)));
  if (n) KF_CHECK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

template <typename T>
__global__ __launch_bounds__(CG_WG) void k_gather_sum(uint32_t nb, const uint32_t *__restrict__ gptr,
                                                    const uint32_t *__restrict__ glist, const T *__restrict__ a,
                                                    T *__restrict__ ac)
{
  const uint32_t b = kf_xcd_block() * (CG_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (b >= nb) return;
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
  for (uint32_t k = gptr[b]; k < gptr[b + 1]; k++) {
    const T *p = a + 16 * (size_t)glist[k] + 4 * c;
    s0 += p[0];
    s1 += p[1];
    s2 += p[2];
    s3 += p[3];
  }
  T *o = ac + 16 * (size_t)b + 4 * c;
  o[0] = s0;
  o[1] = s1;
  o[2] = s2;
  o[3] = s3;
}

__global__ __launch_bounds__(CG_WG) void k_restrict(uint32_t nc, const uint32_t *__restrict__ mptr,
                                                    const uint32_t *__restrict__ mlist,
                                                    const double *__restrict__ v, double *__restrict__ rc)
{
  const uint32_t I = kf_xcd_block() * (CG_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (I >= nc) return;
  double s = 0.0;
  for (uint32_t m = mptr[I]; m < mptr[I + 1]; m++) s += v[4 * (size_t)mlist[m] + c];
  rc[4 * (size_t)I + c] = s;
}

template <typename T>
__device__ __forceinline__ void cg_row4(const T *a, size_t j, int c, double &a0, double &a1, double &a2, double &a3)
{
  const T *p = a + 16 * j + 4 * c;
  a0 = (double)p[0];
  a1 = (double)p[1];
  a2 = (double)p[2];
  a3 = (double)p[3];
}

/* t = b - (A P) z_c; per row the A P blocks in ascending aggregate, each
 * component summed a[c][0] z0 + ... + a[c][3] z3 in FP64; stored as TT
 * (float where the fine sweeps round it to FP32 on reading anyway) */
template <typename T, typename TT>
__global__ __launch_bounds__(CG_WG) void k_coarse_resid(uint32_t N, const uint32_t *__restrict__ aprp,
                                                        const uint32_t *__restrict__ apcol, const T *__restrict__ ap,
                                                        const double *__restrict__ zc,
                                                        const double *__restrict__ b, TT *__restrict__ t)
{
  const uint32_t r = kf_xcd_block() * (CG_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  double acc = 0.0;
  for (uint32_t k = aprp[r]; k < aprp[r + 1]; k++) {
    double a0, a1, a2, a3;
    cg_row4<T>(ap, k, c, a0, a1, a2, a3);
    const double *z = zc + 4 * (size_t)apcol[k];
    acc = acc + a0 * z[0];
    acc = acc + a1 * z[1];
    acc = acc + a2 * z[2];
    acc = acc + a3 * z[3];
  }
  const size_t o = 4 * (size_t)r + c;
  t[o] = (TT)(b[o] - acc);
}

/* ---- A P per 16-row slice, for the remainder product ------------------------
 *
 * As the Jacobi sweeps' factors (jacobi.hip): slot k of a slice's 16 rows is
 * 16 consecutive blocks, so one load instruction of a wavefront reads one
 * contiguous run; AP_SLOTS = 8 slots a row (a lattice row has 5 or 8 A P
 * blocks), a slot past the row's end holds zeros and aggregate 0 and is
 * selected away, entries past the last slot are read from the CSR copy.
 * Used with a BF16 preconditioner only, the values rounded to nearest even
 * from the FP32 sums: padded to 8 slots, an FP32 or FP64 copy would read
 * more bytes than the CSR form (5.6 blocks a row), and measured slower. */
#define AP_SLOTS 8

__device__ __forceinline__ void cg_st(uint16_t *o, double v)
{
  uint32_t u = __float_as_uint((float)v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  *o = (uint16_t)(u >> 16);
}

template <typename TS, typename TA>
__global__ __launch_bounds__(CG_WG) void k_ap_pack(uint32_t N, const uint32_t *__restrict__ aprp,
                                                   const uint32_t *__restrict__ apcol, const TA *__restrict__ ap,
                                                   TS *__restrict__ sv, uint32_t *__restrict__ sc)
{
  const uint32_t r = kf_xcd_block() * (CG_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  const uint32_t j0 = aprp[r], n = aprp[r + 1] - j0, q = r & 15u;
  for (int k = 0; k < AP_SLOTS; k++) {
    const size_t slot = (size_t)(r >> 4) * AP_SLOTS + k;
    TS *o = sv + (slot * 16 + q) * 16 + 4 * c;
    const TA *src = ap + 16 * (size_t)(j0 + k) + 4 * c;
    for (int i = 0; i < 4; i++) cg_st(o + i, (uint32_t)k < n ? (double)src[i] : 0.0);
    if (c == 0) sc[slot * 16 + q] = (uint32_t)k < n ? apcol[j0 + k] : 0u;
  }
}

__device__ __forceinline__ void cg_ld4(const double *p, double a[4])
{
  const double2 lo = reinterpret_cast<const double2 *>(p)[0], hi = reinterpret_cast<const double2 *>(p)[1];
  a[0] = lo.x; a[1] = lo.y; a[2] = hi.x; a[3] = hi.y;
}
__device__ __forceinline__ void cg_ld4(const uint16_t *p, double a[4])
{
  const uint2 v = *reinterpret_cast<const uint2 *>(p);
  a[0] = __uint_as_float(v.x << 16);
  a[1] = __uint_as_float(v.x & 0xffff0000u);
  a[2] = __uint_as_float(v.y << 16);
  a[3] = __uint_as_float(v.y & 0xffff0000u);
}

/* t = b - (A P) z_c over the sliced copy: per row the A P blocks in
 * ascending aggregate (its slots, then its CSR tail), each component summed
 * a[c][0] z0 + ... + a[c][3] z3 in FP64, the slots loaded in one batch;
 * stored as TT (float where the fine sweeps round it to FP32 on reading
 * anyway) */
template <typename TS, typename TA, typename TT>
__global__ __launch_bounds__(CG_WG) void k_coarse_resid_sl(uint32_t N, const uint32_t *__restrict__ aprp,
                                                           const uint32_t *__restrict__ apcol,
                                                           const TA *__restrict__ ap, const TS *__restrict__ sv,
                                                           const uint32_t *__restrict__ sc,
                                                           const double *__restrict__ zc,
                                                           const double *__restrict__ b, TT *__restrict__ t)
{
  const uint32_t r = kf_xcd_block() * (CG_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  const uint32_t j0 = aprp[r], n = aprp[r + 1] - j0, q = r & 15u;
  const size_t base = (size_t)(r >> 4) * AP_SLOTS;
  uint32_t kk[AP_SLOTS];
#pragma unroll
  for (int k = 0; k < AP_SLOTS; k++) kk[k] = sc[(base + k) * 16 + q];
  double a[AP_SLOTS][4];
#pragma unroll
  for (int k = 0; k < AP_SLOTS; k++) cg_ld4(sv + ((base + k) * 16 + q) * 16 + 4 * c, a[k]);
  double z[AP_SLOTS][4];
#pragma unroll
  for (int k = 0; k < AP_SLOTS; k++) cg_ld4(zc + 4 * (size_t)kk[k], z[k]);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < AP_SLOTS; k++) {
    double s = acc + a[k][0] * z[k][0];
    s = s + a[k][1] * z[k][1];
    s = s + a[k][2] * z[k][2];
    s = s + a[k][3] * z[k][3];
    acc = (uint32_t)k < n ? s : acc;
  }
  for (uint32_t k = AP_SLOTS; k < n; k++) {
    double a0, a1, a2, a3;
    cg_row4<TA>(ap, j0 + k, c, a0, a1, a2, a3);
    const double *zz = zc + 4 * (size_t)apcol[j0 + k];
    acc = acc + a0 * zz[0];
    acc = acc + a1 * zz[1];
    acc = acc + a2 * zz[2];
    acc = acc + a3 * zz[3];
  }
  const size_t o = 4 * (size_t)r + c;
  t[o] = (TT)(b[o] - acc);
}

__global__ __launch_bounds__(CG_WG) void k_prolong_add(size_t n4, const uint32_t *__restrict__ agg,
                                                       const double *__restrict__ zc, double *__restrict__ x)
{
  const size_t i = (size_t)blockIdx.x * CG_WG + threadIdx.x;
  if (i < n4) x[i] = x[i] + zc[4 * (size_t)agg[i >> 2] + (i & 3)];
}

/* the coarse level, built on first use from the context's own pattern */
static kf_coarse *coarse_build(kf_ctx *c)
{
  const size_t N = c->d.N, nnzb = c->nnzb;
  kf_prep fp;
  memset(&fp, 0, sizeof(fp));
  fp.N = N;
  fp.nnzb = nnzb;
  fp.rowptr = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  fp.col = (uint32_t *)kf_xmalloc((nnzb + 1) * sizeof(uint32_t));
  fp.diag = (uint32_t *)kf_xmalloc((N + 1) * sizeof(uint32_t));
  KF_CHECK(hipMemcpy(fp.rowptr, c->d.rowptr, (N + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  KF_CHECK(hipMemcpy(fp.col, c->d.col, nnzb * sizeof(uint32_t), hipMemcpyDeviceToHost));
  KF_CHECK(hipMemcpy(fp.diag, c->d.diag, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
  kf_coarse_host h;
  kf_build_coarse(&fp, &h);
  free(fp.rowptr);
  free(fp.col);
  free(fp.diag);

  kf_options co = c->opt;
  co.prec = c->opt.prec;
  co.ilu_exact = 0;
  co.jacobi_sweeps = c->opt.coarse_sweeps;   /* the coarse levels' own solves */
  co.coarse = c->opt.coarse - 1;   /* further levels below this one */
  kf_coarse *k = new kf_coarse();
  k->cc = kf_ctx_create_matrix(&h.prep, &co, c);
  k->nc = (uint32_t)h.prep.N;
  k->cnnzb = h.prep.nnzb;
  k->nap = h.nap;
  k->agg = cupload(h.agg, N);
  k->mptr = cupload(h.mptr, (size_t)k->nc + 1);
  k->mlist = cupload(h.mlist, N);
  k->aprp = cupload(h.aprp, N + 1);
  k->apcol = cupload(h.apcol, h.nap);
  k->apgptr = cupload(h.apgptr, h.nap + 1);
  k->apglist = cupload(h.apglist, nnzb);
  k->gptr = cupload(h.gptr, k->cnnzb + 1);
  k->glist = cupload(h.glist, h.nap);
  KF_CHECK(hipMalloc((void **)&k->rc, 4 * (size_t)k->nc * sizeof(double)));
  KF_CHECK(hipMalloc((void **)&k->zc, 4 * (size_t)k->nc * sizeof(double)));
  KF_CHECK(hipMalloc((void **)&k->t, 4 * N * sizeof(double)));
  kf_free_coarse(&h);
  return k;
}

/* the sliced BF16 copy of A P, with a BF16 preconditioner */
static void ap_pack(kf_ctx *c, kf_coarse *k)
{
  if (c->opt.prec != KF_PREC_BF16) return;
  const uint32_t N = c->d.N, nslice = (N + 15) / 16;
  if (!k->apsl) KF_CHECK(hipMalloc(&k->apsl, (size_t)nslice * AP_SLOTS * 256 * sizeof(uint16_t)));
  if (!k->apsl_col) KF_CHECK(hipMalloc((void **)&k->apsl_col, (size_t)nslice * AP_SLOTS * 16 * sizeof(uint32_t)));
  hipLaunchKernelGGL((k_ap_pack<uint16_t, float>), dim3(kf_blocks_for(N, CG_WG / 4)), dim3(CG_WG), 0, c->stream, N,
                     k->aprp, k->apcol, (const float *)k->ap32, (uint16_t *)k->apsl, k->apsl_col);
  KF_LAUNCH_CHECK();
}

/* P^T (A P) from A P into a coarse value array */
template <typename T>
static void galerkin_from_ap(kf_ctx *c, const kf_coarse *k, const T *ap, T *out)
{
  hipLaunchKernelGGL(k_gather_sum<T>, dim3(kf_blocks_for(k->cnnzb, CG_WG / 4)), dim3(CG_WG), 0, c->stream,
                     (uint32_t)k->cnnzb, k->gptr, k->glist, ap, out);
  KF_LAUNCH_CHECK();
}

/* A P from the fill's values (before the fine factorisation overwrites
 * them), then P^T A P into the coarse context and its ILU(0) */
void kf_coarse_factor(kf_ctx *c)
{
  if (!c->coarse) c->coarse = coarse_build(c);
  kf_coarse *k = (kf_coarse *)c->coarse;
  kf_ctx *cc = k->cc;
  const bool f64 = c->opt.prec == KF_PREC_FP64;
  cc->opt.prec = c->opt.prec;   /* the coarse factors follow the fine level's precision */
  cc->opt.jacobi_sweeps = cc->opt.coarse_sweeps = c->opt.coarse_sweeps;
  if (cc->opt.coarse != c->opt.coarse - 1) {   /* the number of levels changed */
    if (c->opt.coarse == 1) kf_coarse_destroy(cc);
    cc->opt.coarse = c->opt.coarse - 1;
  }
  if (f64 ? k->ap32 != NULL : k->ap64 != NULL) {
    /* the precision changed: drop the other copy of A P (its last reader is
     * done once the stream is), so that a stale one is never read */
    KF_CHECK(hipStreamSynchronize(c->stream));
    KF_CHECK(hipFree(f64 ? (void *)k->ap32 : (void *)k->ap64));
    if (f64) k->ap32 = NULL;
    else k->ap64 = NULL;
  }
  const int s = kf_prof_begin(c, KP_COARSE_FACTOR);
  const uint32_t nb = kf_blocks_for(k->nap, CG_WG / 4);
  if (f64) {
    if (!k->ap64) KF_CHECK(hipMalloc((void **)&k->ap64, 16 * k->nap * sizeof(double)));
    hipLaunchKernelGGL(k_gather_sum<double>, dim3(nb), dim3(CG_WG), 0, c->stream, (uint32_t)k->nap, k->apgptr,
                       k->apglist, (const double *)c->a64, k->ap64);
    KF_LAUNCH_CHECK();
    galerkin_from_ap<double>(c, k, k->ap64, cc->a64);
  } else {
    if (!c->a32) {
      fprintf(stderr, "kf: two-level preconditioner: no single-precision Jacobian (ComputeA must run first)\n");
      abort();
    }
    if (!k->ap32) KF_CHECK(hipMalloc((void **)&k->ap32, 16 * k->nap * sizeof(float)));
    if (!cc->a32) KF_CHECK(hipMalloc((void **)&cc->a32, 16 * k->cnnzb * sizeof(float)));
    hipLaunchKernelGGL(k_gather_sum<float>, dim3(nb), dim3(CG_WG), 0, c->stream, (uint32_t)k->nap, k->apgptr,
                       k->apglist, (const float *)c->a32, k->ap32);
    KF_LAUNCH_CHECK();
    galerkin_from_ap<float>(c, k, k->ap32, cc->a32);
  }
  ap_pack(c, k);
  if (cc->opt.coarse) kf_coarse_factor(cc);   /* the next level, from the Galerkin values */
  kf_launch_ilu(cc);
  kf_prof_end(c, s);
}

/* x = M b with the two-level preconditioner (b, x device, 4N, not aliased) */
void kf_coarse_apply(kf_ctx *c, const double *b, double *x)
{
  kf_coarse *k = (kf_coarse *)c->coarse;
  if (!k) {
    fprintf(stderr, "kf: two-level preconditioner: ComputeNumericalILU has not run with coarse = 1\n");
    abort();
  }
  const bool f64 = c->opt.prec == KF_PREC_FP64, bf = c->opt.prec == KF_PREC_BF16;
  if ((f64 ? !k->ap64 : !k->ap32) || (bf && !k->apsl)) {
    fprintf(stderr, "kf: two-level preconditioner: the preconditioner's precision changed since "
                    "ComputeNumericalILU\n");
    abort();
  }
  kf_check_aligned(b, "b");
  kf_check_aligned(x, "x");
  const kf_dv &d = c->d;
  const int s = kf_prof_begin(c, KP_COARSE);
  hipLaunchKernelGGL(k_restrict, dim3(kf_blocks_for(k->nc, CG_WG / 4)), dim3(CG_WG), 0, c->stream, k->nc, k->mptr,
                     k->mlist, b, k->rc);
  KF_LAUNCH_CHECK();
  if (k->cc->opt.coarse) kf_coarse_apply(k->cc, k->rc, k->zc);   /* the coarse system by the same scheme */
  else if (k->cc->opt.jacobi_sweeps > 0) kf_launch_jacobi(k->cc, k->rc, k->zc);
  else kf_launch_sptrsv(k->cc, k->rc, k->zc, NULL);
  const uint32_t nb = kf_blocks_for(d.N, CG_WG / 4);
  const size_t n4 = 4 * (size_t)d.N;
  const bool sweeps = c->opt.jacobi_sweeps > 0;
  const bool t32 = sweeps && !f64;
  if (t32 && !k->t32) KF_CHECK(hipMalloc((void **)&k->t32, n4 * sizeof(float)));
  if (f64) {
    hipLaunchKernelGGL((k_coarse_resid<double, double>), dim3(nb), dim3(CG_WG), 0, c->stream, d.N, k->aprp, k->apcol,
                       (const double *)k->ap64, k->zc, b, k->t);
  } else if (!bf) {
    if (t32)
      hipLaunchKernelGGL((k_coarse_resid<float, float>), dim3(nb), dim3(CG_WG), 0, c->stream, d.N, k->aprp, k->apcol,
                         (const float *)k->ap32, k->zc, b, k->t32);
    else
      hipLaunchKernelGGL((k_coarse_resid<float, double>), dim3(nb), dim3(CG_WG), 0, c->stream, d.N, k->aprp, k->apcol,
                         (const float *)k->ap32, k->zc, b, k->t);
  } else {
    if (t32)
      hipLaunchKernelGGL((k_coarse_resid_sl<uint16_t, float, float>), dim3(nb), dim3(CG_WG), 0, c->stream, d.N,
                         k->aprp, k->apcol, (const float *)k->ap32, (const uint16_t *)k->apsl, k->apsl_col, k->zc, b,
                         k->t32);
    else
      hipLaunchKernelGGL((k_coarse_resid_sl<uint16_t, float, double>), dim3(nb), dim3(CG_WG), 0, c->stream, d.N,
                         k->aprp, k->apcol, (const float *)k->ap32, (const uint16_t *)k->apsl, k->apsl_col, k->zc, b,
                         k->t);
  }
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  if (sweeps) {
    /* the fine solve by sweeps, x += P z_c in its last sweep */
    kf_launch_jacobi_ml(c, f64 ? (const void *)k->t : (const void *)k->t32, x, k->agg, k->zc);
    return;
  }
  kf_launch_sptrsv(c, k->t, x, NULL);
  hipLaunchKernelGGL(k_prolong_add, dim3(kf_blocks_for(n4, CG_WG)), dim3(CG_WG), 0, c->stream, n4, k->agg, k->zc, x);
  KF_LAUNCH_CHECK();
}

void kf_coarse_destroy(kf_ctx *c)
{
  kf_coarse *k = (kf_coarse *)c->coarse;
  if (!k) return;
  kf_coarse_destroy(k->cc);
  kf_ctx_destroy_matrix(k->cc);
  void *ptrs[] = {k->agg, k->mptr, k->mlist, k->aprp, k->apcol, k->apgptr, k->apglist, k->gptr, k->glist,
                  k->ap64, k->ap32, k->apsl, k->apsl_col, k->rc, k->zc, k->t, k->t32};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
    if (ptrs[i]) KF_CHECK(hipFree(ptrs[i]));
  delete k;
  c->coarse = NULL;
}

/* the coarse level's size, its aggregate map and its Galerkin blocks (for
 * tests): returns the number of aggregates; with non-NULL arrays also copies
 * agg [N], the coarse pattern rowptr [nc + 1], col [cnnzb] and the blocks
 * [16 cnnzb] as the last factorisation formed them (before their ILU) */
extern "C" size_t kf_coarse_matrix(const GEOMETRY *g, size_t *cnnzb, unsigned *agg, unsigned *rowptr, unsigned *col,
                                   double *blocks)
{
  kf_ctx *c = kf_ctx_of(g);
  kf_coarse *k = (kf_coarse *)c->coarse;
  if (!k) return 0;
  KF_CHECK(hipStreamSynchronize(c->stream));
  if (cnnzb) *cnnzb = k->cnnzb;
  if (agg) KF_CHECK(hipMemcpy(agg, k->agg, (size_t)c->d.N * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (rowptr) KF_CHECK(hipMemcpy(rowptr, k->cc->d.rowptr, ((size_t)k->nc + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (col) KF_CHECK(hipMemcpy(col, k->cc->d.col, k->cnnzb * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (blocks) {
    /* the Galerkin product again from A P (the coarse values were factored
     * in place), into the host array */
    const size_t n = 16 * k->cnnzb;
    const bool f64 = c->opt.prec == KF_PREC_FP64;
    if (f64 ? !k->ap64 : !k->ap32) {
      fprintf(stderr, "kf: kf_coarse_matrix: the preconditioner's precision changed since ComputeNumericalILU\n");
      abort();
    }
    void *tmp = NULL;
    KF_CHECK(hipMalloc(&tmp, n * (f64 ? sizeof(double) : sizeof(float))));
    if (f64) galerkin_from_ap<double>(c, k, k->ap64, (double *)tmp);
    else galerkin_from_ap<float>(c, k, k->ap32, (float *)tmp);
    KF_CHECK(hipStreamSynchronize(c->stream));
    if (f64) {
      KF_CHECK(hipMemcpy(blocks, tmp, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      float *h = (float *)kf_xmalloc(n * sizeof(float));
      KF_CHECK(hipMemcpy(h, tmp, n * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; i++) blocks[i] = (double)h[i];
      free(h);
    }
    KF_CHECK(hipFree(tmp));
  }
  return k->nc;
}
