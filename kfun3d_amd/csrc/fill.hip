/*
 * fill.hip -- analytic first-order Jacobian A = J1 + diag(cdt) + boundary
 * blocks (src/fill.c:115-917, matrix/matrix.cpp:137-207) into BSR 4x4.
 *
 * One thread per block row.  The reference scatters the two edge blocks
 * v0, v1 of every edge into (n0,n0) += v0, (n0,n1) += v1, (n1,n0) -= v0,
 * (n1,n1) -= v1 in edge order, then adds the slip-wall column and the
 * free-stream block.  Here row v walks its incidences in ascending edge id,
 * recomputes v0 and v1 of each edge (cheaper than an HBM round trip of
 * 256 bytes per edge), accumulates its diagonal block in the reference's
 * order and writes each off-diagonal block exactly once.  The FP64 blocks
 * are bit-identical to the reference's; T = float gives the FP32
 * preconditioner matrix (inputs rounded to float, arithmetic in float).
 */
#include "kf_dev.h"

#define FLOAT_095 ((double)0.95f)

template <typename T>
__device__ __forceinline__ void edge_blocks(T xn, T yn, T zn, T ln, T uL, T vL, T wL, T uR, T vR, T wR, T v0[16],
                                            T v1[16])
{
  const T B = T(15), h = T(0.5), one = T(1);
  T dot = xn, X1, Y1, Z1;
  if (fabs((double)dot) < FLOAT_095) { X1 = one - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn; }
  else { dot = yn; X1 = -dot * xn; Y1 = one - dot * yn; Z1 = -dot * zn; }
  T size = X1 * X1; size += Y1 * Y1; size += Z1 * Z1; size = sqrt(size);
  X1 /= size; Y1 /= size; Z1 /= size;
  T X2 = yn * Z1; X2 -= zn * Y1;
  T Y2 = zn * X1; Y2 -= xn * Z1;
  T Z2 = xn * Y1; Z2 -= yn * X1;
  T ubarL = xn * uL; ubarL += yn * vL; ubarL += zn * wL;
  T ubarR = xn * uR; ubarR += yn * vR; ubarR += zn * wR;
  const T u = h * (uL + uR), v = h * (vL + vR), w = h * (wL + wR);
  T ubar = xn * u; ubar += yn * v; ubar += zn * w;
  const T c2 = ubar * ubar + B;
  const T c = sqrt(c2);
  const T eig1 = ln * fabs(ubar), eig2 = ln * fabs(ubar), eig3 = ln * fabs(ubar + c), eig4 = ln * fabs(ubar - c);
  T phi1 = xn * B; phi1 += u * ubar;
  T phi2 = yn * B; phi2 += v * ubar;
  T phi3 = zn * B; phi3 += w * ubar;
  T phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
  T phi5 = Z2 * phi1; phi5 -= X2 * phi3;
  T phi6 = X2 * phi2; phi6 -= Y2 * phi1;
  T phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
  T phi8 = X1 * phi3; phi8 -= Z1 * phi1;
  T phi9 = Y1 * phi1; phi9 -= X1 * phi2;
  const T c2inv = one / c2;
  T y11 = u * phi4; y11 += v * phi5; y11 += w * phi6; y11 = -c2inv * y11 / B;
  T y21 = u * phi7; y21 += v * phi8; y21 += w * phi9; y21 = -c2inv * y21 / B;
  T y31 = c2inv * (c - ubar); y31 = h * y31 / B;
  T y41 = c2inv * (c + ubar); y41 = -h * y41 / B;
  const T y12 = c2inv * phi4, y22 = c2inv * phi7, y32 = c2inv * h * xn, y42 = c2inv * h * xn;
  const T y13 = c2inv * phi5, y23 = c2inv * phi8, y33 = c2inv * h * yn, y43 = c2inv * h * yn;
  const T y14 = c2inv * phi6, y24 = c2inv * phi9, y34 = c2inv * h * zn, y44 = c2inv * h * zn;
  const T t13 = c * B;
  T t23 = u * (ubar + c); t23 += xn * B;
  T t33 = v * (ubar + c); t33 += yn * B;
  T t43 = w * (ubar + c); t43 += zn * B;
  const T t14 = -c * B;
  T t24 = u * (ubar - c); t24 += xn * B;
  T t34 = v * (ubar - c); t34 += yn * B;
  T t44 = w * (ubar - c); t44 += zn * B;
  T a11 = eig3 * t13 * y31; a11 += eig4 * t14 * y41;
  T a12 = eig3 * t13 * y32; a12 += eig4 * t14 * y42;
  T a13 = eig3 * t13 * y33; a13 += eig4 * t14 * y43;
  T a14 = eig3 * t13 * y34; a14 += eig4 * t14 * y44;
  T a21 = eig1 * X1 * y11; a21 += eig2 * X2 * y21; a21 += eig3 * t23 * y31; a21 += eig4 * t24 * y41;
  T a22 = eig1 * X1 * y12; a22 += eig2 * X2 * y22; a22 += eig3 * t23 * y32; a22 += eig4 * t24 * y42;
  T a23 = eig1 * X1 * y13; a23 += eig2 * X2 * y23; a23 += eig3 * t23 * y33; a23 += eig4 * t24 * y43;
  T a24 = eig1 * X1 * y14; a24 += eig2 * X2 * y24; a24 += eig3 * t23 * y34; a24 += eig4 * t24 * y44;
  T a31 = eig1 * Y1 * y11; a31 += eig2 * Y2 * y21; a31 += eig3 * t33 * y31; a31 += eig4 * t34 * y41;
  T a32 = eig1 * Y1 * y12; a32 += eig2 * Y2 * y22; a32 += eig3 * t33 * y32; a32 += eig4 * t34 * y42;
  T a33 = eig1 * Y1 * y13; a33 += eig2 * Y2 * y23; a33 += eig3 * t33 * y33; a33 += eig4 * t34 * y43;
  T a34 = eig1 * Y1 * y14; a34 += eig2 * Y2 * y24; a34 += eig3 * t33 * y34; a34 += eig4 * t34 * y44;
  T a41 = eig1 * Z1 * y11; a41 += eig2 * Z2 * y21; a41 += eig3 * t43 * y31; a41 += eig4 * t44 * y41;
  T a42 = eig1 * Z1 * y12; a42 += eig2 * Z2 * y22; a42 += eig3 * t43 * y32; a42 += eig4 * t44 * y42;
  T a43 = eig1 * Z1 * y13; a43 += eig2 * Z2 * y23; a43 += eig3 * t43 * y33; a43 += eig4 * t44 * y43;
  T a44 = eig1 * Z1 * y14; a44 += eig2 * Z2 * y24; a44 += eig3 * t43 * y34; a44 += eig4 * t44 * y44;
  const T lb = ln * B, lx = ln * xn, ly = ln * yn, lz = ln * zn;
  v0[0] = h * a11; v0[4] = h * (lx + a21); v0[8] = h * (ly + a31); v0[12] = h * (lz + a41);
  v0[1] = h * ((lb * xn) + a12); v0[5] = h * ((ln * (ubarL + xn * uL)) + a22);
  v0[9] = h * ((lx * vL) + a32); v0[13] = h * ((lx * wL) + a42);
  v0[2] = h * ((lb * yn) + a13); v0[6] = h * ((ly * uL) + a23);
  v0[10] = h * ((ln * (ubarL + yn * vL)) + a33); v0[14] = h * ((ly * wL) + a43);
  v0[3] = h * ((lb * zn) + a14); v0[7] = h * ((lz * uL) + a24);
  v0[11] = h * ((lz * vL) + a34); v0[15] = h * ((ln * (ubarL + zn * wL)) + a44);
  v1[0] = h * -a11; v1[4] = h * (lx - a21); v1[8] = h * (ly - a31); v1[12] = h * (lz - a41);
  v1[1] = h * ((lb * xn) - a12); v1[5] = h * ((ln * (ubarR + xn * uR)) - a22);
  v1[9] = h * ((lx * vR) - a32); v1[13] = h * ((lx * wR) - a42);
  v1[2] = h * ((lb * yn) - a13); v1[6] = h * ((ly * uR) - a23);
  v1[10] = h * ((ln * (ubarR + yn * vR)) - a33); v1[14] = h * ((ly * wR) - a43);
  v1[3] = h * ((lb * zn) - a14); v1[7] = h * ((lz * uR) - a24);
  v1[11] = h * ((lz * vR) - a34); v1[15] = h * ((ln * (ubarR + zn * wR)) - a44);
}

/* free-stream characteristic block: src/fill.c:528-873 */
template <typename T>
__device__ __forceinline__ void freestream_block(T xn, T yn, T zn, T U, T V, T pi, T ui, T vi, T wi, T v[16])
{
  const T B = T(15), h = T(0.5), one = T(1), zero = T(0), W = T(0), P = T(1);
  const T ln = sqrt(xn * xn + yn * yn + zn * zn);
  xn /= ln; yn /= ln; zn /= ln;
  T dot = xn, X1, Y1, Z1;
  if (fabs((double)dot) < FLOAT_095) { X1 = one - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn; }
  else { dot = yn; X1 = -dot * xn; Y1 = one - dot * yn; Z1 = -dot * zn; }
  const T size = sqrt(X1 * X1 + Y1 * Y1 + Z1 * Z1);
  X1 /= size; Y1 /= size; Z1 /= size;
  const T X2 = yn * Z1 - zn * Y1, Y2 = zn * X1 - xn * Z1, Z2 = xn * Y1 - yn * X1;
  T ubar0 = xn * U; ubar0 += yn * V; ubar0 += zn * W;
  const T c20 = ubar0 * ubar0 + B;
  const T c0 = sqrt(c20);
  T phi1 = xn * B; phi1 += U * ubar0;
  T phi2 = yn * B; phi2 += V * ubar0;
  T phi3 = zn * B; phi3 += W * ubar0;
  T phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
  T phi5 = Z2 * phi1; phi5 -= X2 * phi3;
  T phi6 = X2 * phi2; phi6 -= Y2 * phi1;
  T phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
  T phi8 = X1 * phi3; phi8 -= Z1 * phi1;
  T phi9 = Y1 * phi1; phi9 -= X1 * phi2;
  const T t13 = c0 * B;
  T t23 = U * (ubar0 + c0); t23 += xn * B;
  T t33 = V * (ubar0 + c0); t33 += yn * B;
  T t43 = W * (ubar0 + c0); t43 += zn * B;
  const T t14 = -c0 * B;
  T t24 = U * (ubar0 - c0); t24 += xn * B;
  T t34 = V * (ubar0 - c0); t34 += yn * B;
  T t44 = W * (ubar0 - c0); t44 += zn * B;
  T ti11 = U * phi4; ti11 += V * phi5; ti11 += W * phi6; ti11 = -ti11 / B / c20;
  T ti21 = U * phi7; ti21 += V * phi8; ti21 += W * phi9; ti21 = -ti21 / B / c20;
  const T ti31 = (c0 - ubar0) / (T(2) * B * c20);
  const T ti41 = -(c0 + ubar0) / (T(2) * B * c20);
  const T ti12 = phi4 / c20, ti22 = phi7 / c20, ti32 = h * xn / c20, ti42 = h * xn / c20;
  const T ti13 = phi5 / c20, ti23 = phi8 / c20, ti33 = h * yn / c20, ti43 = h * yn / c20;
  const T ti14 = phi6 / c20, ti24 = phi9 / c20, ti34 = h * zn / c20, ti44 = h * zn / c20;
  T un = xn * ui; un += yn * vi; un += zn * wi;
  T pr, prp, ur, uru, vr, vrv, wr, wrw;
  if (un > zero) { pr = pi; prp = one; ur = ui; uru = one; vr = vi; vrv = one; wr = wi; wrw = one; }
  else { pr = P; prp = zero; ur = U; uru = zero; vr = V; vrv = zero; wr = W; wrw = zero; }
  T rhs1 = ti11 * pr; rhs1 += ti12 * ur; rhs1 += ti13 * vr; rhs1 += ti14 * wr;
  const T rhs1p = ti11 * prp, rhs1u = ti12 * uru, rhs1v = ti13 * vrv, rhs1w = ti14 * wrw;
  T rhs2 = ti21 * pr; rhs2 += ti22 * ur; rhs2 += ti23 * vr; rhs2 += ti24 * wr;
  const T rhs2p = ti21 * prp, rhs2u = ti22 * uru, rhs2v = ti23 * vrv, rhs2w = ti24 * wrw;
  T rhs3 = ti31 * pi; rhs3 += ti32 * ui; rhs3 += ti33 * vi; rhs3 += ti34 * wi;
  T rhs4 = ti41 * P; rhs4 += ti42 * U; rhs4 += ti43 * V; rhs4 += ti44 * W;
  const T pbp = t13 * ti31, pbu = t13 * ti32, pbv = t13 * ti33, pbw = t13 * ti34;
  T ub = X1 * rhs1; ub += X2 * rhs2; ub += t23 * rhs3; ub += t24 * rhs4;
  T ubp = X1 * rhs1p; ubp += X2 * rhs2p; ubp += t23 * ti31;
  T ubu = X1 * rhs1u; ubu += X2 * rhs2u; ubu += t23 * ti32;
  T ubv = X1 * rhs1v; ubv += X2 * rhs2v; ubv += t23 * ti33;
  T ubw = X1 * rhs1w; ubw += X2 * rhs2w; ubw += t23 * ti34;
  T vb = Y1 * rhs1; vb += Y2 * rhs2; vb += t33 * rhs3; vb += t34 * rhs4;
  T vbp = Y1 * rhs1p; vbp += Y2 * rhs2p; vbp += t33 * ti31;
  T vbu = Y1 * rhs1u; vbu += Y2 * rhs2u; vbu += t33 * ti32;
  T vbv = Y1 * rhs1v; vbv += Y2 * rhs2v; vbv += t33 * ti33;
  T vbw = Y1 * rhs1w; vbw += Y2 * rhs2w; vbw += t33 * ti34;
  T wb = Z1 * rhs1; wb += Z2 * rhs2; wb += t43 * rhs3; wb += t44 * rhs4;
  T wbp = Z1 * rhs1p; wbp += Z2 * rhs2p; wbp += t43 * ti31;
  T wbu = Z1 * rhs1u; wbu += Z2 * rhs2u; wbu += t43 * ti32;
  T wbv = Z1 * rhs1v; wbv += Z2 * rhs2v; wbv += t43 * ti33;
  T wbw = Z1 * rhs1w; wbw += Z2 * rhs2w; wbw += t43 * ti34;
  T unb = xn * ub; unb += yn * vb; unb += zn * wb;
  T unbp = xn * ubp; unbp += yn * vbp; unbp += zn * wbp;
  T unbu = xn * ubu; unbu += yn * vbu; unbu += zn * wbu;
  T unbv = xn * ubv; unbv += yn * vbv; unbv += zn * wbv;
  T unbw = xn * ubw; unbw += yn * vbw; unbw += zn * wbw;
  v[0] = ln * B * unbp;
  v[4] = ln * (ub * unbp + unb * ubp + xn * pbp);
  v[8] = ln * (vb * unbp + unb * vbp + yn * pbp);
  v[12] = ln * (wb * unbp + unb * wbp + zn * pbp);
  v[1] = ln * B * unbu;
  v[5] = ln * (ub * unbu + unb * ubu + xn * pbu);
  v[9] = ln * (vb * unbu + unb * vbu + yn * pbu);
  v[13] = ln * (wb * unbu + unb * wbu + zn * pbu);
  v[2] = ln * B * unbv;
  v[6] = ln * (ub * unbv + unb * ubv + xn * pbv);
  v[10] = ln * (vb * unbv + unb * vbv + yn * pbv);
  v[14] = ln * (wb * unbv + unb * wbv + zn * pbv);
  v[3] = ln * B * unbw;
  v[7] = ln * (ub * unbw + unb * ubw + xn * pbw);
  v[11] = ln * (vb * unbw + unb * vbw + yn * pbw);
  v[15] = ln * (wb * unbw + unb * wbw + zn * pbw);
}

template <typename T>
__device__ __forceinline__ void store_block(T *__restrict__ a, uint32_t blk, const T v[16]);

template <>
__device__ __forceinline__ void store_block<double>(double *__restrict__ a, uint32_t blk, const double v[16])
{
  double2 *o = reinterpret_cast<double2 *>(a + 16 * (size_t)blk);
#pragma unroll
  for (int k = 0; k < 8; k++) o[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

template <>
__device__ __forceinline__ void store_block<float>(float *__restrict__ a, uint32_t blk, const float v[16])
{
  float4 *o = reinterpret_cast<float4 *>(a + 16 * (size_t)blk);
#pragma unroll
  for (int k = 0; k < 4; k++) o[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}

template <typename T>
__global__ __launch_bounds__(128) void k_fill(kf_dv d, const double *__restrict__ q, const double *__restrict__ cdt,
                                              T *__restrict__ a)
{
  const uint32_t r = kf_xcd_block() * 128u + threadIdx.x;
  if (r >= d.N) return;
  const T zero = T(0), mone = T(-1);
  T dg[16];
#pragma unroll
  for (int k = 0; k < 16; k++) dg[k] = zero;
  const T cd = (T)cdt[r];
  dg[0] += cd; dg[5] += cd; dg[10] += cd; dg[15] += cd;
  const double *qr = q + 4 * (size_t)r;
  const T ur = (T)qr[1], vr = (T)qr[2], wr = (T)qr[3];
  const uint32_t deg = d.deg[r];
  size_t idx = (size_t)d.sell_ptr[r >> 6] * 64u + (r & 63u);
  for (uint32_t k = 0; k < deg; k++, idx += 64) {
    const uint32_t ent = d.sell_ent[idx], e = ent >> 1, nb = d.sell_nbr[idx];
    const double *qn = q + 4 * (size_t)nb;
    const T un = (T)qn[1], vn = (T)qn[2], wn = (T)qn[3];
    T v0[16], v1[16], off[16];
    const kf_egeo eg = kf_edge_geo(d.egeo, e);
    if (ent & 1u) {   /* r is n1: (n1,n0) -= v0, (n1,n1) -= v1 */
      edge_blocks<T>((T)eg.xn, (T)eg.yn, (T)eg.zn, (T)eg.ln, un, vn, wn, ur, vr, wr, v0, v1);
#pragma unroll
      for (int j = 0; j < 16; j++) { off[j] = zero + mone * v0[j]; dg[j] += mone * v1[j]; }
    } else {          /* r is n0: (n0,n0) += v0, (n0,n1) += v1 */
      edge_blocks<T>((T)eg.xn, (T)eg.yn, (T)eg.zn, (T)eg.ln, ur, vr, wr, un, vn, wn, v0, v1);
#pragma unroll
      for (int j = 0; j < 16; j++) { off[j] = zero + v1[j]; dg[j] += v0[j]; }
    }
    store_block<T>(a, d.sell_blk[idx], off);
  }
  for (uint32_t s = d.snptr[r]; s < d.snptr[r + 1]; s++) {   /* matrix/matrix.cpp:282-296 */
    const uint32_t i = d.snent[s];
    dg[4] += (T)d.snx[i]; dg[8] += (T)d.sny[i]; dg[12] += (T)d.snz[i];
  }
  for (uint32_t s = d.fnptr[r]; s < d.fnptr[r + 1]; s++) {
    const uint32_t i = d.fnent[s];
    T v[16];
    freestream_block<T>((T)d.fnx[i], (T)d.fny[i], (T)d.fnz[i], (T)d.U, (T)d.V, (T)qr[0], ur, vr, wr, v);
#pragma unroll
    for (int j = 0; j < 16; j++) dg[j] += v[j];
  }
  store_block<T>(a, d.diag[r], dg);
}

static void ensure_f32(kf_ctx *c)
{
  if (!c->a32) KF_CHECK(hipMalloc((void **)&c->a32, 16 * c->nnzb * sizeof(float)));
}

void kf_launch_fill(kf_ctx *c)
{
  const uint32_t nb = kf_blocks_for(c->d.N, 128);
  const int s = kf_prof_begin(c, KP_FILL);
  if (c->opt.prec == KF_PREC_FP64) {
    hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(128), 0, c->stream, c->d, c->q, c->cdt, c->a64);
  } else {
    ensure_f32(c);
    hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(128), 0, c->stream, c->d, c->q, c->cdt, c->a32);
  }
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
}
