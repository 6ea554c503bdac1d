/*
 * sptrsv.hip -- x = U^-1 L^-1 b with the in-place ILU(0) factors
 * (matrix/sptrsv.cpp:34-73, level scheduling of matrix/init.cpp:194-274).
 *
 * A quad of lanes owns one block row (lane c computes component c); a
 * wavefront owns up to 16 rows of one level, handed out by ticket in level
 * order (sync.h).  Per row, the block products are summed in ascending column
 * order exactly as the reference:
 *   forward   x_r = b_r;  x_r -= A(r,j) x_j  for j < r           (dgemv, then axpy -1)
 *   backward  x_r -= A(r,j) x_j  for j > r;  x_r = inv(D_r) x_r
 * so the FP64 result is bit-identical.  In the single-precision modes the
 * factors are FP32 or BF16 (converted to FP32 on load), the vectors FP32,
 * and the result is widened to FP64 on the last store.
 */
#include "sync.h"

__device__ __forceinline__ double mat_ld(const double *a, size_t i) { return a[i]; }
__device__ __forceinline__ float mat_ld(const float *a, size_t i) { return a[i]; }
__device__ __forceinline__ float mat_ld(const uint16_t *a, size_t i) { return __uint_as_float((uint32_t)a[i] << 16); }

template <typename TM, typename TV>
__global__ __launch_bounds__(64) void k_trsv_lower(kf_dv d, const uint32_t *__restrict__ rows,
                                                   const uint32_t *__restrict__ grp, uint32_t ngrp, uint32_t *ticket,
                                                   uint32_t base, uint32_t *rdone, uint32_t epoch,
                                                   const TM *__restrict__ a, const double *__restrict__ b, TV *xv)
{
  const uint32_t lane = threadIdx.x, qd = lane >> 2, c = lane & 3;
  bool aborted = false;
  for (;;) {
    const uint32_t t = next_ticket(ticket, base);
    if (t >= ngrp) break;
    const uint32_t g0 = grp[t], g1 = grp[t + 1];
    const bool active = g0 + qd < g1;
    const uint32_t r = active ? rows[g0 + qd] : 0;
    const uint32_t j0 = active ? d.rowptr[r] : 0, jd = active ? d.diag[r] : 0;
    /* the quad's four lanes share the waits on the row's lower neighbours */
    for (uint32_t j = j0 + c; j < jd; j += 4) wait_row(rdone, d.col[j], epoch, d.err, aborted, d.spin_sleep);
    aborted = __any(aborted);
    consume_fence();
    if (active) {
      TV xr = (TV)b[4 * (size_t)r + c];
      for (uint32_t j = j0; j < jd; j++) {
        const TV *xc = xv + 4 * (size_t)d.col[j];
        const size_t o = 16 * (size_t)j + 4 * c;
        TV s = TV(0);
        s += (TV)mat_ld(a, o) * xc[0];
        s += (TV)mat_ld(a, o + 1) * xc[1];
        s += (TV)mat_ld(a, o + 2) * xc[2];
        s += (TV)mat_ld(a, o + 3) * xc[3];
        xr = xr - s;
      }
      xv[4 * (size_t)r + c] = xr;
    }
    publish_fence();
    if (active && c == 0) flag_store(rdone + r, epoch);
  }
}

template <typename TM, typename TV>
__global__ __launch_bounds__(64) void k_trsv_upper(kf_dv d, const uint32_t *__restrict__ rows,
                                                   const uint32_t *__restrict__ grp, uint32_t ngrp, uint32_t *ticket,
                                                   uint32_t base, uint32_t *rdone, uint32_t epoch,
                                                   const TM *__restrict__ a, TV *xv, double *__restrict__ xout)
{
  const uint32_t lane = threadIdx.x, qd = lane >> 2, c = lane & 3;
  bool aborted = false;
  for (;;) {
    const uint32_t t = next_ticket(ticket, base);
    if (t >= ngrp) break;
    const uint32_t g0 = grp[t], g1 = grp[t + 1];
    const bool active = g0 + qd < g1;
    const uint32_t r = active ? rows[g0 + qd] : 0;
    const uint32_t jd = active ? d.diag[r] : 0, j1 = active ? d.rowptr[r + 1] : 0;
    for (uint32_t j = jd + 1 + c; j < j1; j += 4) wait_row(rdone, d.col[j], epoch, d.err, aborted, d.spin_sleep);
    aborted = __any(aborted);
    consume_fence();
    TV xr = TV(0);
    if (active) {
      xr = xv[4 * (size_t)r + c];
      for (uint32_t j = jd + 1; j < j1; j++) {
        const TV *xc = xv + 4 * (size_t)d.col[j];
        const size_t o = 16 * (size_t)j + 4 * c;
        TV s = TV(0);
        s += (TV)mat_ld(a, o) * xc[0];
        s += (TV)mat_ld(a, o + 1) * xc[1];
        s += (TV)mat_ld(a, o + 2) * xc[2];
        s += (TV)mat_ld(a, o + 3) * xc[3];
        xr = xr - s;
      }
    }
    /* x_r = inv(D_r) x_r: the quad exchanges its four components */
    const TV y0 = __shfl(xr, 0, 4), y1 = __shfl(xr, 1, 4), y2 = __shfl(xr, 2, 4), y3 = __shfl(xr, 3, 4);
    if (active) {
      const size_t o = 16 * (size_t)jd + 4 * c;
      TV s = TV(0);
      s += (TV)mat_ld(a, o) * y0;
      s += (TV)mat_ld(a, o + 1) * y1;
      s += (TV)mat_ld(a, o + 2) * y2;
      s += (TV)mat_ld(a, o + 3) * y3;
      xv[4 * (size_t)r + c] = s;
      if ((void *)xout != (void *)xv) xout[4 * (size_t)r + c] = (double)s;
    }
    publish_fence();
    if (active && c == 0) flag_store(rdone + r, epoch);
  }
}

template <typename TM, typename TV>
static void launch_trsv(kf_ctx *c, const TM *a, const double *b, TV *xv, double *x)
{
  const uint32_t nwl = c->nlgrp < c->trsv_waves ? (c->nlgrp ? c->nlgrp : 1) : c->trsv_waves;
  const uint32_t nwu = c->nugrp < c->trsv_waves ? (c->nugrp ? c->nugrp : 1) : c->trsv_waves;
  const uint32_t el = ++c->epoch;
  int s = kf_prof_begin(c, KP_TRSV_L);
  hipLaunchKernelGGL((k_trsv_lower<TM, TV>), dim3(nwl), dim3(64), 0, c->stream, c->d, c->lrows, c->lgrp, c->nlgrp,
                     c->ticket + 0, c->ticket_base[0], c->rdone, el, a, b, xv);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  c->ticket_base[0] += c->nlgrp + nwl;
  const uint32_t eu = ++c->epoch;
  s = kf_prof_begin(c, KP_TRSV_U);
  hipLaunchKernelGGL((k_trsv_upper<TM, TV>), dim3(nwu), dim3(64), 0, c->stream, c->d, c->urows, c->ugrp, c->nugrp,
                     c->ticket + 1, c->ticket_base[1], c->rdone, eu, a, xv, x);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  c->ticket_base[1] += c->nugrp + nwu;
}

void kf_launch_sptrsv(kf_ctx *c, const double *b, double *x)
{
  kf_check_aligned(b, "b"); kf_check_aligned(x, "x");
  if (c->opt.prec == KF_PREC_FP64) {
    launch_trsv<double, double>(c, c->a64, b, x, x);
    return;
  }
  if (!c->f32a) KF_CHECK(hipMalloc((void **)&c->f32a, 4 * (size_t)c->d.N * sizeof(float)));
  if (c->opt.prec == KF_PREC_FP32) launch_trsv<float, float>(c, c->a32, b, c->f32a, x);
  else launch_trsv<uint16_t, float>(c, c->abf, b, c->f32a, x);
}
