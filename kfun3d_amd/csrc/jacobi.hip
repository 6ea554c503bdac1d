/*
 * jacobi.hip -- x ~ U^-1 L^-1 b by Jacobi sweeps over the in-place ILU(0)
 * factors: fun3d::Matrix::sptrsv_Jacobi (matrix/sptrsv.cpp:75-132), the
 * reference's alternative to its level-scheduled solve (selected at
 * matrix/wrapper.cpp:86-91; its sweep count is Matrix::jacobi_sweeps,
 * matrix/matrix.h:35).
 *
 * With k sweeps (k >= 1; the reference's do-while runs at least one):
 *   forward   y0 = b;     y_s[r] = b[r] - sum_{j < r} L(r,j) y_{s-1}[j]
 *   backward  x0 = D^-1 y_k;  x_s[r] = D^-1 (y_k[r] - sum_{j > r} U(r,j) x_{s-1}[j])
 * (D^-1 is the inverted diagonal block the factorisation leaves in place).
 * Each sweep reads every row's old values only, so the rows are independent:
 * no level schedule, no waits, every sweep one HBM-bound pass over its half
 * of the factors.  The result approaches the exact triangular solves as k
 * grows and equals them (up to rounding) once k reaches the depth of the
 * dependency DAG.
 *
 * Layout and order follow k_bsr_spmv (spmv.hip): a quad of lanes per row,
 * lane c forming component c; per row the reference's order exactly:
 *   t = dgemv(block, x_col)   t_c = ((0 + a_c0 x0) + a_c1 x1) + a_c2 x2) + a_c3 x3
 *   s = t + s                 over the row's entries in ascending column, s from 0
 *   forward   y = (-s) + b
 *   backward  u = (-s) + y,   x = dgemv(D^-1, u)   (the quad exchanges u)
 * so the FP64 mode is bit-identical to a one-thread run of the reference's
 * loops.  Single-precision modes: FP32 or BF16 factors, FP32 vectors and
 * arithmetic, b widened from FP64 as read, the last backward sweep writing
 * FP64.  The last forward sweep also writes x0 = D^-1 y_k (fused; the
 * reference's separate dgemv pass).  Workgroups are remapped XCD by XCD.
 * Bytes per sweep: (half the factor blocks) (16 s + 4) + vector traffic.
 */
#include <mutex>

#include "kf_dev.h"

#define JAC_WG 256   /* 64 rows per workgroup */

namespace {

__device__ __forceinline__ double jmat(const double *a, size_t i) { return a[i]; }
__device__ __forceinline__ float jmat(const float *a, size_t i) { return a[i]; }
__device__ __forceinline__ float jmat(const uint16_t *a, size_t i) { return __uint_as_float((uint32_t)a[i] << 16); }

template <typename TM, typename TV>
__device__ __forceinline__ void jrow(const TM *__restrict__ a, uint32_t j, int c, TV m[4])
{
  const size_t o = 16 * (size_t)j + 4 * c;
  m[0] = (TV)jmat(a, o);
  m[1] = (TV)jmat(a, o + 1);
  m[2] = (TV)jmat(a, o + 2);
  m[3] = (TV)jmat(a, o + 3);
}

template <typename TV, typename TX>
__device__ __forceinline__ void jvec(const TX *__restrict__ x, uint32_t col, TV v[4])
{
  const TX *p = x + 4 * (size_t)col;
  v[0] = (TV)p[0];
  v[1] = (TV)p[1];
  v[2] = (TV)p[2];
  v[3] = (TV)p[3];
}

/* dgemv row: summed from zero in column order (matrix/init.cpp:96-103) */
template <typename TV>
__device__ __forceinline__ TV jdot(const TV m[4], const TV x[4])
{
  TV s = TV(0);
  s += m[0] * x[0];
  s += m[1] * x[1];
  s += m[2] * x[2];
  s += m[3] * x[3];
  return s;
}

template <int I>
__device__ __forceinline__ float jquad(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), I * 0x55, 0xf, 0xf, false));
}

template <int I>
__device__ __forceinline__ double jquad(double v)
{
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), I * 0x55, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), I * 0x55, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

/* s = sum over entries [j0, j1) of the row, in ascending column:
 * s = dgemv(A_j, x_col) + s (matrix/sptrsv.cpp:91-96 / 116-121).  The first
 * JAC_HELD entries are loaded in one batch -- column indices, then every
 * block and vector at once, slots past the row's end reading the diagonal
 * block jd (always in bounds) -- and summed in order with the empty slots
 * selected away, so a row's loads overlap instead of chaining; longer rows
 * finish in a loop. */
#define JAC_HELD 8
template <typename TM, typename TV, typename TX>
__device__ __forceinline__ TV jsum(const uint32_t *__restrict__ col, const TM *__restrict__ a, const TX *__restrict__ x,
                                   uint32_t j0, uint32_t j1, uint32_t jd, int c)
{
  const uint32_t n = j1 - j0;
  uint32_t jj[JAC_HELD], kk[JAC_HELD];
#pragma unroll
  for (int i = 0; i < JAC_HELD; i++) jj[i] = (uint32_t)i < n ? j0 + i : jd;
#pragma unroll
  for (int i = 0; i < JAC_HELD; i++) kk[i] = col[jj[i]];
  TV m[JAC_HELD][4], xv[JAC_HELD][4];
#pragma unroll
  for (int i = 0; i < JAC_HELD; i++) jrow<TM, TV>(a, jj[i], c, m[i]);
#pragma unroll
  for (int i = 0; i < JAC_HELD; i++) jvec<TV, TX>(x, kk[i], xv[i]);
  TV s = TV(0);
#pragma unroll
  for (int i = 0; i < JAC_HELD; i++) {
    const TV t = jdot(m[i], xv[i]);
    s = (uint32_t)i < n ? t + s : s;
  }
  for (uint32_t j = j0 + JAC_HELD; j < j1; j++) {
    TV m0[4], x0[4];
    jrow<TM, TV>(a, j, c, m0);
    jvec<TV, TX>(x, col[j], x0);
    const TV t0 = jdot(m0, x0);
    s = t0 + s;
  }
  return s;
}

/* quad exchange, then dgemv(D^-1, u) row c */
template <typename TM, typename TV>
__device__ __forceinline__ TV jdinv(const TM *__restrict__ a, uint32_t jd, int c, TV u)
{
  TV dg[4], uv[4];
  jrow<TM, TV>(a, jd, c, dg);
  uv[0] = jquad<0>(u);
  uv[1] = jquad<1>(u);
  uv[2] = jquad<2>(u);
  uv[3] = jquad<3>(u);
  return jdot(dg, uv);
}

}  // namespace

/* One forward sweep: y[r] = b[r] - sum_{j < r} L(r,j) xin[j]; with DIAG also
 * x0[r] = D^-1 y[r].  xin is b itself (TX = double) on the first sweep. */
template <typename TM, typename TV, typename TX, bool DIAG>
__global__ __launch_bounds__(JAC_WG) void k_jac_lower(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ diag, const TM *__restrict__ a,
                                                      const double *__restrict__ b, const TX *__restrict__ xin,
                                                      TV *__restrict__ y, TV *__restrict__ x0)
{
  const uint32_t r = kf_xcd_block() * (JAC_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;   /* whole quads leave together (N rows of 4 lanes) */
  const uint32_t jd = diag[r];
  const TV s = jsum<TM, TV, TX>(col, a, xin, rowptr[r], jd, jd, c);
  const TV v = (-s) + (TV)b[4 * (size_t)r + c];
  y[4 * (size_t)r + c] = v;
  if (DIAG) x0[4 * (size_t)r + c] = jdinv<TM, TV>(a, jd, c, v);
}

/* One backward sweep: x[r] = D^-1 (y[r] - sum_{j > r} U(r,j) xin[j]) */
template <typename TM, typename TV, typename TO>
__global__ __launch_bounds__(JAC_WG) void k_jac_upper(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ diag, const TM *__restrict__ a,
                                                      const TV *__restrict__ y, const TV *__restrict__ xin,
                                                      TO *__restrict__ x)
{
  const uint32_t r = kf_xcd_block() * (JAC_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  const uint32_t jd = diag[r];
  const TV s = jsum<TM, TV, TV>(col, a, xin, jd + 1, rowptr[r + 1], jd, c);
  const TV u = (-s) + y[4 * (size_t)r + c];
  x[4 * (size_t)r + c] = (TO)jdinv<TM, TV>(a, jd, c, u);
}

/* ---- the sliced copy of the factors ------------------------------------------
 *
 * The CSR factor array stores a row's blocks side by side, so one load
 * instruction of a wavefront (16 rows, a quad each) touches 16 blocks a row
 * apart.  For the sweeps the factors are also packed per slice of 16 rows
 * (one wavefront) and per entry slot k < K: slot k of the slice's 16 rows
 * is 16 consecutive blocks (512 bytes in BF16), so that instruction reads
 * one contiguous run.  Layout, in the factors' type (jsl_val):
 *   L:  [slice][k][row of slice][16]    the k-th block left of the diagonal
 *   U:  [slice][k][row of slice][16]    the k-th block right of it
 *   D:  [slice][row of slice][16]       the inverted diagonal block
 * and the slots' columns (jsl_col): L then U, [slice][k][row of slice].  A
 * slot past a row's end holds zeros and the row's own column; the sweeps
 * select it away (the counts come from rowptr and diag), so the sums are
 * the CSR kernels' to the bit.  Entries past slot K (rows longer than K,
 * none on a lattice) are read from the CSR array.  K is the most held
 * entries of the triangular solves (7, or 8 when a row has more than 7 on a
 * side).  Packed after each factorisation, on the first sweep that needs
 * it. */
template <typename TM, typename TV>
__device__ __forceinline__ void jld4(const TM *p, TV m[4])
{
  m[0] = (TV)jmat(p, 0);
  m[1] = (TV)jmat(p, 1);
  m[2] = (TV)jmat(p, 2);
  m[3] = (TV)jmat(p, 3);
}
template <>
__device__ __forceinline__ void jld4<uint16_t, float>(const uint16_t *p, float m[4])
{
  const uint2 v = *reinterpret_cast<const uint2 *>(p);   /* 8-byte aligned: 4 values of a block row */
  m[0] = __uint_as_float(v.x << 16);
  m[1] = __uint_as_float(v.x & 0xffff0000u);
  m[2] = __uint_as_float(v.y << 16);
  m[3] = __uint_as_float(v.y & 0xffff0000u);
}
template <>
__device__ __forceinline__ void jld4<float, float>(const float *p, float m[4])
{
  const float4 v = *reinterpret_cast<const float4 *>(p);
  m[0] = v.x;
  m[1] = v.y;
  m[2] = v.z;
  m[3] = v.w;
}

/* s over a row's part: slots [0, min(n, K)) from the slice, the rest of the
 * n entries from the CSR array from position jcsr */
template <int K, typename TM, typename TV, typename TX>
__device__ __forceinline__ TV jsum_sl(const TM *__restrict__ sv, const uint32_t *__restrict__ sc, uint32_t base,
                                      uint32_t q, int c, uint32_t n, const uint32_t *__restrict__ col,
                                      const TM *__restrict__ a, uint32_t jcsr, const TX *__restrict__ x)
{
  uint32_t kk[K];
#pragma unroll
  for (int k = 0; k < K; k++) kk[k] = sc[(size_t)(base + k) * 16 + q];
  TV m[K][4], xv[K][4];
#pragma unroll
  for (int k = 0; k < K; k++) jld4<TM, TV>(sv + ((size_t)(base + k) * 16 + q) * 16 + 4 * c, m[k]);
#pragma unroll
  for (int k = 0; k < K; k++) jvec<TV, TX>(x, kk[k], xv[k]);
  TV s = TV(0);
#pragma unroll
  for (int k = 0; k < K; k++) {
    const TV t = jdot(m[k], xv[k]);
    s = (uint32_t)k < n ? t + s : s;
  }
  for (uint32_t k = K; k < n; k++) {
    TV m0[4], x0[4];
    jrow<TM, TV>(a, jcsr + k, c, m0);
    jvec<TV, TX>(x, col[jcsr + k], x0);
    const TV t0 = jdot(m0, x0);
    s = t0 + s;
  }
  return s;
}

template <typename TM, typename TV>
__device__ __forceinline__ TV jdinv_sl(const TM *__restrict__ dv, size_t o, TV u)
{
  TV dg[4], uv[4];
  jld4<TM, TV>(dv + o, dg);
  uv[0] = jquad<0>(u);
  uv[1] = jquad<1>(u);
  uv[2] = jquad<2>(u);
  uv[3] = jquad<3>(u);
  return jdot(dg, uv);
}

/* the sliced layout's parts: nslice slices of K slots */
struct jsl_view {
  uint32_t nslice;
  const void *val;        /* L, U, D */
  const uint32_t *col;    /* L, U */
};

/* row r of one forward sweep (lane c of its quad) */
template <int K, typename TM, typename TV, typename TB, typename TX, bool DIAG>
__device__ __forceinline__ void jac_lower_row(uint32_t r, int c, const uint32_t *__restrict__ rowptr,
                                              const uint32_t *__restrict__ col, const uint32_t *__restrict__ diag,
                                              const TM *__restrict__ a, const jsl_view &v, const TB *__restrict__ b,
                                              const TX *__restrict__ xin, TV *__restrict__ y, TV *__restrict__ x0)
{
  const uint32_t sl = r >> 4, q = r & 15u, j0 = rowptr[r], jd = diag[r];
  const TM *sv = static_cast<const TM *>(v.val);
  const TV s = jsum_sl<K, TM, TV, TX>(sv, v.col, sl * K, q, c, jd - j0, col, a, j0, xin);
  const TV u = (-s) + (TV)b[4 * (size_t)r + c];
  y[4 * (size_t)r + c] = u;
  if (DIAG) {
    const TM *dv = sv + (size_t)2 * v.nslice * K * 256;
    x0[4 * (size_t)r + c] = jdinv_sl<TM, TV>(dv, ((size_t)sl * 16 + q) * 16 + 4 * c, u);
  }
}

template <int K, typename TM, typename TV, typename TB, typename TX, bool DIAG>
__global__ __launch_bounds__(JAC_WG) void k_jac_lower_sl(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col,
                                                         const uint32_t *__restrict__ diag, const TM *__restrict__ a,
                                                         jsl_view v, const TB *__restrict__ b,
                                                         const TX *__restrict__ xin, TV *__restrict__ y,
                                                         TV *__restrict__ x0)
{
  const uint32_t r = kf_xcd_block() * (JAC_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  jac_lower_row<K, TM, TV, TB, TX, DIAG>(r, c, rowptr, col, diag, a, v, b, xin, y, x0);
}

/* row r of one backward sweep (lane c of its quad) */
template <int K, typename TM, typename TV, typename TO>
__device__ __forceinline__ void jac_upper_row(uint32_t r, int c, const uint32_t *__restrict__ rowptr,
                                              const uint32_t *__restrict__ col, const uint32_t *__restrict__ diag,
                                              const TM *__restrict__ a, const jsl_view &v, const TV *__restrict__ y,
                                              const TV *__restrict__ xin, TO *__restrict__ x,
                                              const uint32_t *__restrict__ pagg, const double *__restrict__ pzc)
{
  const uint32_t sl = r >> 4, q = r & 15u, jd = diag[r];
  const TM *sv = static_cast<const TM *>(v.val);
  const size_t nk = (size_t)v.nslice * K;
  const TV s = jsum_sl<K, TM, TV, TV>(sv + nk * 256, v.col + nk * 16, sl * K, q, c, rowptr[r + 1] - jd - 1, col, a,
                                      jd + 1, xin);
  const TV u = (-s) + y[4 * (size_t)r + c];
  const TM *dv = sv + 2 * nk * 256;
  TO xo = (TO)jdinv_sl<TM, TV>(dv, ((size_t)sl * 16 + q) * 16 + 4 * c, u);
  /* the multilevel preconditioner's prolongation x += P z_c, fused into the
   * last sweep (TO = double; the same add as k_prolong_add) */
  if (pagg) xo = (TO)((double)xo + pzc[4 * (size_t)pagg[r] + c]);
  x[4 * (size_t)r + c] = xo;
}

template <int K, typename TM, typename TV, typename TO>
__global__ __launch_bounds__(JAC_WG) void k_jac_upper_sl(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ col,
                                                         const uint32_t *__restrict__ diag, const TM *__restrict__ a,
                                                         jsl_view v, const TV *__restrict__ y,
                                                         const TV *__restrict__ xin, TO *__restrict__ x,
                                                         const uint32_t *__restrict__ pagg,
                                                         const double *__restrict__ pzc)
{
  const uint32_t r = kf_xcd_block() * (JAC_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  jac_upper_row<K, TM, TV, TO>(r, c, rowptr, col, diag, a, v, y, xin, x, pagg, pzc);
}

/* A whole solve of jacobi_solve_sl in one workgroup, for systems of at most
 * JAC_SMALL_ROWS rows (the small levels of the multilevel preconditioner,
 * where each sweep's launch costs more than its work): the same sweeps,
 * rows, buffers and operations, a barrier between two sweeps instead of a
 * launch. */
#define JAC_SMALL_WG 512   /* 2 waves per SIMD: registers for a whole row, no spills */
#define JAC_SMALL_ROWS (JAC_SMALL_WG / 4)
template <int K, typename TM, typename TV, typename TB>
__global__ __launch_bounds__(JAC_SMALL_WG) void k_jac_small(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                            const uint32_t *__restrict__ col,
                                                            const uint32_t *__restrict__ diag,
                                                            const TM *__restrict__ a, jsl_view v,
                                                            const TB *__restrict__ b, double *__restrict__ x,
                                                            TV *__restrict__ y, TV *__restrict__ p0,
                                                            TV *__restrict__ p1, int k,
                                                            const uint32_t *__restrict__ pagg,
                                                            const double *__restrict__ pzc)
{
  const int c = threadIdx.x & 3;
  const uint32_t r = threadIdx.x >> 2;   /* one row per quad: N <= JAC_SMALL_ROWS */
  const bool on = r < N;
  TV *const xread = k >= 2 ? (((k - 1) & 1) ? p0 : p1) : NULL;
  TV *cur = xread == p0 ? p1 : p0;
  const TV *prev = NULL;
  for (int sw = 1; sw <= k; sw++) {
    TV *out = sw == k ? y : ((sw & 1) ? p0 : p1);
    if (on) {
      if (sw == 1 && sw == k) jac_lower_row<K, TM, TV, TB, TB, true>(r, c, rowptr, col, diag, a, v, b, b, out, cur);
      else if (sw == 1) jac_lower_row<K, TM, TV, TB, TB, false>(r, c, rowptr, col, diag, a, v, b, b, out, (TV *)NULL);
      else if (sw == k) jac_lower_row<K, TM, TV, TB, TV, true>(r, c, rowptr, col, diag, a, v, b, prev, out, cur);
      else jac_lower_row<K, TM, TV, TB, TV, false>(r, c, rowptr, col, diag, a, v, b, prev, out, (TV *)NULL);
    }
    prev = out;
    __syncthreads();
  }
  TV *other = cur == p0 ? p1 : p0;
  for (int sw = 1; sw <= k; sw++) {
    if (sw == k) {
      if (on) jac_upper_row<K, TM, TV, double>(r, c, rowptr, col, diag, a, v, y, cur, x, pagg, pzc);
    } else {
      if (on) jac_upper_row<K, TM, TV, TV>(r, c, rowptr, col, diag, a, v, y, cur, other, (const uint32_t *)NULL,
                                           (const double *)NULL);
      TV *t = cur;
      cur = other;
      other = t;
      __syncthreads();
    }
  }
}

/* the sliced copy from the CSR factors (a quad per row, lane c copying
 * block row c) */
template <int K, typename TM>
__global__ __launch_bounds__(JAC_WG) void k_jac_pack(uint32_t N, uint32_t nslice, const uint32_t *__restrict__ rowptr,
                                                     const uint32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ diag, const TM *__restrict__ a,
                                                     TM *__restrict__ val, uint32_t *__restrict__ scol)
{
  const uint32_t r = kf_xcd_block() * (JAC_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  const uint32_t sl = r >> 4, q = r & 15u, j0 = rowptr[r], jd = diag[r], j1 = rowptr[r + 1];
  const size_t nk = (size_t)nslice * K;
  for (int part = 0; part < 2; part++) {
    const uint32_t first = part ? jd + 1 : j0, n = part ? j1 - jd - 1 : jd - j0;
    TM *pv = val + (size_t)part * nk * 256;
    uint32_t *pc = scol + (size_t)part * nk * 16;
    for (int k = 0; k < K; k++) {
      const size_t slot = (size_t)sl * K + k;
      TM *o = pv + (slot * 16 + q) * 16 + 4 * c;
      if ((uint32_t)k < n) {
        const TM *src = a + 16 * (size_t)(first + k) + 4 * c;
        o[0] = src[0];
        o[1] = src[1];
        o[2] = src[2];
        o[3] = src[3];
      } else {
        o[0] = o[1] = o[2] = o[3] = TM(0);
      }
      if (c == 0) pc[slot * 16 + q] = (uint32_t)k < n ? col[first + k] : r;
    }
  }
  TM *o = val + 2 * nk * 256 + ((size_t)sl * 16 + q) * 16 + 4 * c;
  const TM *src = a + 16 * (size_t)jd + 4 * c;
  o[0] = src[0];
  o[1] = src[1];
  o[2] = src[2];
  o[3] = src[3];
}

/* The k sweeps each way.  Work vectors (4N of TV each): y (the forward
 * result, kept for the backward sweeps) and two ping-pong buffers p0, p1.
 * x (FP64) receives the last backward sweep.  This is the CSR form (the
 * no-allocation entry); jacobi_solve_sl below is the sliced one, with the
 * same buffers, order and bits. */
template <typename TM, typename TV>
static void jacobi_solve(kf_ctx *c, const TM *a, const double *b, double *x, TV *y, TV *p0, TV *p1, int k)
{
  const kf_dv &d = c->d;
  const uint32_t nb = kf_blocks_for(d.N, JAC_WG / 4);
  const dim3 grid(nb), block(JAC_WG);
  /* forward: sweeps 1 .. k-1 write p0, p1, p0, ...; sweep k reads the last
   * of them (b when k == 1), writes y, and writes the backward start D^-1 y
   * into the buffer it does not read */
  TV *const xread = k >= 2 ? (((k - 1) & 1) ? p0 : p1) : NULL;
  TV *cur = xread == p0 ? p1 : p0;
  int s = kf_prof_begin(c, KP_JAC_L);
  const TV *prev = NULL;
  for (int sw = 1; sw <= k; sw++) {
    const bool last = sw == k;
    TV *out = last ? y : ((sw & 1) ? p0 : p1);
    if (sw == 1 && last)
      hipLaunchKernelGGL((k_jac_lower<TM, TV, double, true>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag,
                         a, b, b, out, cur);
    else if (sw == 1)
      hipLaunchKernelGGL((k_jac_lower<TM, TV, double, false>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag,
                         a, b, b, out, (TV *)NULL);
    else if (last)
      hipLaunchKernelGGL((k_jac_lower<TM, TV, TV, true>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag, a,
                         b, prev, out, cur);
    else
      hipLaunchKernelGGL((k_jac_lower<TM, TV, TV, false>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag, a,
                         b, prev, out, (TV *)NULL);
    KF_LAUNCH_CHECK();
    prev = out;
  }
  kf_prof_end(c, s);
  /* backward: from cur, ping-pong with the other buffer, the last sweep
   * into x */
  TV *other = cur == p0 ? p1 : p0;
  s = kf_prof_begin(c, KP_JAC_U);
  for (int sw = 1; sw <= k; sw++) {
    if (sw == k) {
      hipLaunchKernelGGL((k_jac_upper<TM, TV, double>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag, a, y,
                         cur, x);
    } else {
      hipLaunchKernelGGL((k_jac_upper<TM, TV, TV>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag, a, y, cur,
                         other);
      TV *t = cur;
      cur = other;
      other = t;
    }
    KF_LAUNCH_CHECK();
  }
  kf_prof_end(c, s);
}

/* one forward sweep over the sliced factors: the first reads the
 * right-hand side b (type TB) as its old vector, the last also writes
 * D^-1 y into cur */
template <int K, typename TM, typename TV, typename TB>
static void lower_sl(kf_ctx *c, const TM *a, const jsl_view &v, const TB *b, const TV *prev, TV *out, TV *cur,
                     bool first, bool last)
{
  const kf_dv &d = c->d;
  const dim3 grid(kf_blocks_for(d.N, JAC_WG / 4)), block(JAC_WG);
  if (first && last)
    hipLaunchKernelGGL((k_jac_lower_sl<K, TM, TV, TB, TB, true>), grid, block, 0, c->stream, d.N, d.rowptr, d.col,
                       d.diag, a, v, b, b, out, cur);
  else if (first)
    hipLaunchKernelGGL((k_jac_lower_sl<K, TM, TV, TB, TB, false>), grid, block, 0, c->stream, d.N, d.rowptr, d.col,
                       d.diag, a, v, b, b, out, (TV *)NULL);
  else if (last)
    hipLaunchKernelGGL((k_jac_lower_sl<K, TM, TV, TB, TV, true>), grid, block, 0, c->stream, d.N, d.rowptr, d.col,
                       d.diag, a, v, b, prev, out, cur);
  else
    hipLaunchKernelGGL((k_jac_lower_sl<K, TM, TV, TB, TV, false>), grid, block, 0, c->stream, d.N, d.rowptr, d.col,
                       d.diag, a, v, b, prev, out, (TV *)NULL);
  KF_LAUNCH_CHECK();
}

/* one backward sweep over the sliced factors, into a TV buffer or (the
 * last) into x, optionally adding P z_c */
template <int K, typename TM, typename TV, typename TO>
static void upper_sl(kf_ctx *c, const TM *a, const jsl_view &v, const TV *y, const TV *cur, TO *out,
                     const uint32_t *pagg, const double *pzc)
{
  const kf_dv &d = c->d;
  const dim3 grid(kf_blocks_for(d.N, JAC_WG / 4)), block(JAC_WG);
  hipLaunchKernelGGL((k_jac_upper_sl<K, TM, TV, TO>), grid, block, 0, c->stream, d.N, d.rowptr, d.col, d.diag, a, v,
                     y, cur, out, pagg, pzc);
  KF_LAUNCH_CHECK();
}

/* jacobi_solve over the sliced factors, the right-hand side in TB (double,
 * or float where the sweeps compute in FP32 anyway: the same bits), and
 * with pagg non-NULL the prolongation x += P z_c fused into the last sweep */
template <int K, typename TM, typename TV, typename TB>
static void jacobi_solve_sl(kf_ctx *c, const TM *a, const TB *b, double *x, TV *y, TV *p0, TV *p1, int k,
                            const jsl_view &v, const uint32_t *pagg, const double *pzc)
{
  static const bool small_on = !(getenv("KF_JAC_SMALL") && atoi(getenv("KF_JAC_SMALL")) == 0);
  if (small_on && c->d.N <= JAC_SMALL_ROWS) {   /* the whole solve in one workgroup (KF_JAC_SMALL=0: never) */
    const kf_dv &d = c->d;
    const int s = kf_prof_begin(c, KP_JAC_L);
    hipLaunchKernelGGL((k_jac_small<K, TM, TV, TB>), dim3(1), dim3(JAC_SMALL_WG), 0, c->stream, d.N, d.rowptr, d.col,
                       d.diag, a, v, b, x, y, p0, p1, k, pagg, pzc);
    KF_LAUNCH_CHECK();
    kf_prof_end(c, s);
    return;
  }
  TV *const xread = k >= 2 ? (((k - 1) & 1) ? p0 : p1) : NULL;
  TV *cur = xread == p0 ? p1 : p0;
  int s = kf_prof_begin(c, KP_JAC_L);
  const TV *prev = NULL;
  for (int sw = 1; sw <= k; sw++) {
    TV *out = sw == k ? y : ((sw & 1) ? p0 : p1);
    lower_sl<K, TM, TV, TB>(c, a, v, b, prev, out, cur, sw == 1, sw == k);
    prev = out;
  }
  kf_prof_end(c, s);
  TV *other = cur == p0 ? p1 : p0;
  s = kf_prof_begin(c, KP_JAC_U);
  for (int sw = 1; sw <= k; sw++) {
    if (sw == k) {
      upper_sl<K, TM, TV, double>(c, a, v, y, cur, x, pagg, pzc);
    } else {
      upper_sl<K, TM, TV, TV>(c, a, v, y, cur, other, NULL, NULL);
      TV *t = cur;
      cur = other;
      other = t;
    }
  }
  kf_prof_end(c, s);
}

/* the sliced copy of the current factors, packed now unless the last
 * factorisation's is already there (kf_launch_ilu marks it stale) */
template <typename TM>
static void jsl_pack_as(kf_ctx *c, const TM *a, uint32_t nslice)
{
  const kf_dv &d = c->d;
  const dim3 grid(kf_blocks_for(d.N, JAC_WG / 4)), block(JAC_WG);
  if (c->jsl_k == 7)
    hipLaunchKernelGGL((k_jac_pack<7, TM>), grid, block, 0, c->stream, d.N, nslice, d.rowptr, d.col, d.diag, a,
                       (TM *)c->jsl_val, c->jsl_col);
  else
    hipLaunchKernelGGL((k_jac_pack<8, TM>), grid, block, 0, c->stream, d.N, nslice, d.rowptr, d.col, d.diag, a,
                       (TM *)c->jsl_val, c->jsl_col);
  KF_LAUNCH_CHECK();
}

static jsl_view jsl_pack(kf_ctx *c)
{
  const uint32_t nslice = (c->d.N + 15) / 16;
  const size_t es = c->opt.prec == KF_PREC_FP64 ? 8 : c->opt.prec == KF_PREC_FP32 ? 4 : 2;
  c->jsl_k = c->trsv_pf;   /* 7 or 8 */
  const size_t nk = (size_t)nslice * c->jsl_k;
  if (c->jsl_val && c->jsl_esize != es) {   /* the precision changed */
    KF_CHECK(hipStreamSynchronize(c->stream));
    KF_CHECK(hipFree(c->jsl_val));
    c->jsl_val = NULL;
  }
  if (!c->jsl_val) {
    KF_CHECK(hipMalloc(&c->jsl_val, (2 * nk + nslice) * 256 * es));
    c->jsl_esize = es;
    c->jsl_fresh = 0;
  }
  if (!c->jsl_col) KF_CHECK(hipMalloc((void **)&c->jsl_col, 2 * nk * 16 * sizeof(uint32_t)));
  if (!c->jsl_fresh) {
    if (es == 8) jsl_pack_as<double>(c, c->a64, nslice);
    else if (es == 4) jsl_pack_as<float>(c, c->a32, nslice);
    else jsl_pack_as<uint16_t>(c, c->abf, nslice);
    c->jsl_fresh = 1;
  }
  jsl_view v;
  v.nslice = nslice;
  v.val = c->jsl_val;
  v.col = c->jsl_col;
  return v;
}

/* the sweeps of kf_launch_jacobi / kf_launch_jacobi_ml in the context's
 * precision; b is double, or float (TB) with single-precision sweeps */
template <typename TB>
static void jacobi_launch(kf_ctx *c, const TB *b, double *x, const uint32_t *pagg, const double *pzc)
{
  kf_check_aligned(b, "b");
  kf_check_aligned(x, "x");
  if ((const void *)b == (const void *)x) {
    fprintf(stderr, "kf: Jacobi triangular sweeps: b and x must not alias\n");
    abort();
  }
  const int k = c->opt.jacobi_sweeps;
  if (k < 1 || k > KF_JACOBI_MAX) {
    fprintf(stderr, "kf: jacobi_sweeps %d outside 1..%d\n", k, KF_JACOBI_MAX);
    abort();
  }
  const size_t n = 4 * (size_t)c->d.N;
  const void *factors = c->opt.prec == KF_PREC_FP64 ? (const void *)c->a64
                        : c->opt.prec == KF_PREC_FP32 ? (const void *)c->a32 : (const void *)c->abf;
  if (!factors) {
    fprintf(stderr, "kf: Jacobi triangular sweeps: no ILU(0) factors in the context's precision yet\n");
    abort();
  }
  const jsl_view v = jsl_pack(c);
  const bool k7 = c->jsl_k == 7;
  if (c->opt.prec == KF_PREC_FP64) {
    if (!c->ty64) KF_CHECK(hipMalloc((void **)&c->ty64, n * sizeof(double)));
    if (!c->jac64) KF_CHECK(hipMalloc((void **)&c->jac64, 2 * n * sizeof(double)));
    if (k7) jacobi_solve_sl<7, double, double, TB>(c, c->a64, b, x, c->ty64, c->jac64, c->jac64 + n, k, v, pagg, pzc);
    else jacobi_solve_sl<8, double, double, TB>(c, c->a64, b, x, c->ty64, c->jac64, c->jac64 + n, k, v, pagg, pzc);
    c->ystate[0] = ~(uint64_t)0;   /* y no longer holds the level-scheduled sweep's sentinels */
    return;
  }
  if (!c->f32a) KF_CHECK(hipMalloc((void **)&c->f32a, n * sizeof(float)));
  if (!c->jac32) KF_CHECK(hipMalloc((void **)&c->jac32, 2 * n * sizeof(float)));
  float *y = c->f32a, *p0 = c->jac32, *p1 = c->jac32 + n;
  if (c->opt.prec == KF_PREC_FP32) {
    if (k7) jacobi_solve_sl<7, float, float, TB>(c, c->a32, b, x, y, p0, p1, k, v, pagg, pzc);
    else jacobi_solve_sl<8, float, float, TB>(c, c->a32, b, x, y, p0, p1, k, v, pagg, pzc);
  } else {
    if (k7) jacobi_solve_sl<7, uint16_t, float, TB>(c, c->abf, b, x, y, p0, p1, k, v, pagg, pzc);
    else jacobi_solve_sl<8, uint16_t, float, TB>(c, c->abf, b, x, y, p0, p1, k, v, pagg, pzc);
  }
  c->ystate[1] = ~(uint64_t)0;
}

/* x ~ U^-1 L^-1 b with c->opt.jacobi_sweeps sweeps each way, in the
 * preconditioner's precision; b, x: device, 4N doubles, not aliased */
void kf_launch_jacobi(kf_ctx *c, const double *b, double *x)
{
  jacobi_launch<double>(c, b, x, NULL, NULL);
}

/* the multilevel preconditioner's fine solve: as kf_launch_jacobi on t
 * (FP64 with FP64 factors, else FP32: the sweeps' own arithmetic type), then
 * x += P z_c (agg, zc) in the last sweep */
void kf_launch_jacobi_ml(kf_ctx *c, const void *t, double *x, const uint32_t *agg, const double *zc)
{
  if (c->opt.prec == KF_PREC_FP64) jacobi_launch<double>(c, (const double *)t, x, agg, zc);
  else jacobi_launch<float>(c, (const float *)t, x, agg, zc);
}

/* ---- the same solve with no allocation: kf_sptrsv_jacobi_noalloc ----------
 *
 * The three work vectors (y and the two ping-pong buffers) live in one fixed
 * device array, `scratch`, instead of context memory allocated at the first
 * call, so the entry point makes no hipMalloc, hipFree or other allocating
 * call and can run where the caller cannot allocate (HIP graph capture,
 * serving loops).  Each GPU the library runs on holds its own copy of the
 * array (KF_NOALLOC_BYTES) until the process ends.  Its device address is
 * looked up once per device, at context creation with the context's device
 * current (kf_noalloc_setup); host code uses that address, never the array's
 * name.  Every call on one GPU shares the array: calls that overlap on one
 * GPU (contexts on different streams, or different host threads) race, and
 * the caller orders them.  Nothing in the array needs to start at zero:
 * every sweep writes all rows before the next sweep reads them. */
static __device__ __align__(256) unsigned char scratch[KF_NOALLOC_BYTES];

#define KF_NOALLOC_DEVICES 64
static std::mutex noalloc_mu;
static void *noalloc_ptr[KF_NOALLOC_DEVICES];   /* per device: the address of its scratch */

void kf_noalloc_setup(kf_ctx *c)
{
  int dev = 0;
  KF_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= KF_NOALLOC_DEVICES) {
    c->noalloc = NULL;   /* kf_sptrsv_jacobi_noalloc then reports KF_NOALLOC_EMODE */
    return;
  }
  std::lock_guard<std::mutex> lock(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void *p = NULL;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      /* only the no-allocation solve depends on it: it then reports
       * KF_NOALLOC_EMODE; the error is cleared so later launch checks do
       * not see it */
      (void)hipGetLastError();
      fprintf(stderr, "kf: hipGetSymbolAddress of the fixed work array failed on device %d (%s); "
                      "kf_sptrsv_jacobi_noalloc is unavailable there\n", dev, hipGetErrorString(e));
      c->noalloc = NULL;
      return;
    }
    noalloc_ptr[dev] = p;
  }
  c->noalloc = noalloc_ptr[dev];
}

extern "C" int kf_sptrsv_jacobi_noalloc(const GEOMETRY *g, const double *b, double *x)
{
  kf_ctx *c = kf_ctx_of(g);
  if (b == x) {
    fprintf(stderr, "kf: kf_sptrsv_jacobi_noalloc: b and x must not alias\n");
    return KF_NOALLOC_EMODE;
  }
  kf_check_aligned(b, "b");
  kf_check_aligned(x, "x");
  const int k = c->opt.jacobi_sweeps;
  if (k < 1 || k > KF_JACOBI_MAX) {
    fprintf(stderr, "kf: kf_sptrsv_jacobi_noalloc: the context's jacobi_sweeps is %d (needs 1..%d)\n", k,
            KF_JACOBI_MAX);
    return KF_NOALLOC_EMODE;
  }
  const size_t n = 4 * (size_t)c->d.N;
  const size_t vb = c->opt.prec == KF_PREC_FP64 ? sizeof(double) : sizeof(float);
  if (3 * n * vb > (size_t)KF_NOALLOC_BYTES) {
    fprintf(stderr,
            "kf: kf_sptrsv_jacobi_noalloc: %u vertices need %zu bytes of work space in this precision, more than "
            "the fixed array's %u (at most %zu vertices)\n",
            c->d.N, 3 * n * vb, (unsigned)KF_NOALLOC_BYTES, (size_t)KF_NOALLOC_BYTES / (12 * vb));
    return KF_NOALLOC_ESIZE;
  }
  const void *factors = c->opt.prec == KF_PREC_FP64 ? (const void *)c->a64
                        : c->opt.prec == KF_PREC_FP32 ? (const void *)c->a32 : (const void *)c->abf;
  if (!factors || !c->noalloc) {
    fprintf(stderr, "kf: kf_sptrsv_jacobi_noalloc: no ILU(0) factors in the context's precision yet (call "
                    "ComputeA and ComputeNumericalILU first)\n");
    return KF_NOALLOC_EMODE;
  }
  if (c->opt.prec == KF_PREC_FP64) {
    double *y = static_cast<double *>(c->noalloc);
    jacobi_solve<double, double>(c, c->a64, b, x, y, y + n, y + 2 * n, k);
  } else {
    float *y = static_cast<float *>(c->noalloc);
    if (c->opt.prec == KF_PREC_FP32)
      jacobi_solve<float, float>(c, c->a32, b, x, y, y + n, y + 2 * n, k);
    else
      jacobi_solve<uint16_t, float>(c, c->abf, b, x, y, y + n, y + 2 * n, k);
  }
  return 0;
}
