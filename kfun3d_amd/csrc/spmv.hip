/*
 * spmv.hip -- y = A x with the assembled block Jacobian (BSR 4x4, the
 * pattern of matrix/init.cpp:7-74, the values of ComputeA / src/fill.c) on
 * gfx950.
 *
 * The reference never multiplies by its assembled matrix: its Krylov
 * operator is the matrix-free finite-difference product (src/kernel.c:24-45)
 * and A only feeds the ILU preconditioner.  This product serves the
 * first-order-operator mode of GMRES (kf_options.bsr_operator) and tools.
 *
 * Layout and order.  A quad of lanes owns one block row; lane c forms
 * component c:  y[4r + c] = sum over the row's blocks j in ascending column
 * of  a_j[c][0] x0 + a_j[c][1] x1 + a_j[c][2] x2 + a_j[c][3] x3,  each
 * product added to the running sum in that order, in FP64 (FP32-stored
 * values are widened first).  The four lanes of a quad read one 128-byte
 * block (FP64) row by row, so a wave's loads are whole blocks; the column's
 * four x values are the same address for the quad.  Workgroups are remapped
 * XCD by XCD (kf_xcd_block) so that each XCD works on a contiguous slab of
 * rows and its x gathers stay in its L2.  Two blocks are in flight per lane
 * step.  Bytes per launch: nnzb (16 s + 4) + 64 N with s the stored value
 * size (HBM-bound).
 */
#include "kf_dev.h"

#define SPMV_WG 256   /* 64 rows per workgroup */

template <typename T>
__device__ __forceinline__ void load_row4(const T *a, size_t j, int c, double &a0, double &a1, double &a2, double &a3);

template <>
__device__ __forceinline__ void load_row4<double>(const double *a, size_t j, int c, double &a0, double &a1, double &a2,
                                                  double &a3)
{
  const double2 *p = reinterpret_cast<const double2 *>(a + 16 * j + 4 * c);
  const double2 u = p[0], v = p[1];
  a0 = u.x; a1 = u.y; a2 = v.x; a3 = v.y;
}

template <>
__device__ __forceinline__ void load_row4<float>(const float *a, size_t j, int c, double &a0, double &a1, double &a2,
                                                 double &a3)
{
  const float4 u = *reinterpret_cast<const float4 *>(a + 16 * j + 4 * c);
  a0 = (double)u.x; a1 = (double)u.y; a2 = (double)u.z; a3 = (double)u.w;
}

template <typename T>
__global__ __launch_bounds__(SPMV_WG) void k_bsr_spmv(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ col, const T *__restrict__ a,
                                                      const double *__restrict__ x, double *__restrict__ y)
{
  const uint32_t r = kf_xcd_block() * (SPMV_WG / 4) + (threadIdx.x >> 2);
  const int c = threadIdx.x & 3;
  if (r >= N) return;
  const uint32_t j0 = rowptr[r], j1 = rowptr[r + 1];
  double acc = 0.0;
  uint32_t j = j0;
  for (; j + 1 < j1; j += 2) {
    const uint32_t k0 = col[j], k1 = col[j + 1];
    double p0, p1, p2, p3, q0, q1, q2, q3;
    load_row4<T>(a, j, c, p0, p1, p2, p3);
    load_row4<T>(a, j + 1, c, q0, q1, q2, q3);
    const double4 u = *reinterpret_cast<const double4 *>(x + 4 * (size_t)k0);
    const double4 v = *reinterpret_cast<const double4 *>(x + 4 * (size_t)k1);
    acc = acc + p0 * u.x; acc = acc + p1 * u.y; acc = acc + p2 * u.z; acc = acc + p3 * u.w;
    acc = acc + q0 * v.x; acc = acc + q1 * v.y; acc = acc + q2 * v.z; acc = acc + q3 * v.w;
  }
  if (j < j1) {
    const uint32_t k0 = col[j];
    double p0, p1, p2, p3;
    load_row4<T>(a, j, c, p0, p1, p2, p3);
    const double4 u = *reinterpret_cast<const double4 *>(x + 4 * (size_t)k0);
    acc = acc + p0 * u.x; acc = acc + p1 * u.y; acc = acc + p2 * u.z; acc = acc + p3 * u.w;
  }
  y[4 * (size_t)r + c] = acc;
}

/* y = A x with the Jacobian ComputeA assembled last (the copy the context
 * keeps for the operator: FP64, or FP32 with a single-precision
 * preconditioner); x, y device vectors of 4N doubles, 32-byte aligned, not
 * aliased */
void kf_launch_spmv(kf_ctx *c, const double *x, double *y)
{
  kf_check_aligned(x, "x");
  kf_check_aligned(y, "y");
  const uint32_t nb = kf_blocks_for(c->d.N, SPMV_WG / 4);
  const int s = kf_prof_begin(c, KP_SPMV);
  if (c->jop64)
    hipLaunchKernelGGL(k_bsr_spmv<double>, dim3(nb), dim3(SPMV_WG), 0, c->stream, c->d.N, c->d.rowptr, c->d.col,
                       c->jop64, x, y);
  else
    hipLaunchKernelGGL(k_bsr_spmv<float>, dim3(nb), dim3(SPMV_WG), 0, c->stream, c->d.N, c->d.rowptr, c->d.col,
                       c->jop32, x, y);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
}

/* after a fill: the operator's copy of the assembled Jacobian (the ILU then
 * factors the fill's array in place) */
void kf_keep_jacobian(kf_ctx *c)
{
  const size_t n = 16 * c->nnzb;
  const bool f64 = c->opt.prec == KF_PREC_FP64;
  if ((f64 && c->jop32) || (!f64 && c->jop64)) {
    /* the precision changed: drop the other copy (its last reader is done
     * once the stream is) */
    KF_CHECK(hipStreamSynchronize(c->stream));
    KF_CHECK(hipFree(f64 ? (void *)c->jop32 : (void *)c->jop64));
    c->jop32 = f64 ? NULL : c->jop32;
    c->jop64 = f64 ? c->jop64 : NULL;
  }
  if (f64) {
    if (!c->jop64) KF_CHECK(hipMalloc((void **)&c->jop64, n * sizeof(double)));
    KF_CHECK(hipMemcpyAsync(c->jop64, c->a64, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  } else {
    if (!c->jop32) KF_CHECK(hipMalloc((void **)&c->jop32, n * sizeof(float)));
    KF_CHECK(hipMemcpyAsync(c->jop32, c->a32, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  }
}
