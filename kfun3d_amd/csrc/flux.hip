/*
 * flux.hip -- residual R(q) (src/flux.c:15-825), local time step
 * (src/timestep.c:21-188) and wall forces (src/forces.c:8-120) on gfx950.
 *
 * The reference accumulates every vertex sum in edge order with scatter
 * updates from one thread.  Here every sum is a gather: the vertex walks its
 * incidence list (ascending edge id) and adds the same terms in the same
 * order, so the FP64 result is bit-identical and the kernels are atomic free
 * and run-to-run deterministic.  The per-edge Roe flux is evaluated once per
 * edge (k_edge_flux) and read back by both end points.
 *
 * Compiled with -ffp-contract=off: no FMA contraction, IEEE division and
 * square root, so each expression rounds as in the reference.
 */
#include "kf_dev.h"

#define B_ 15.0
#define FLOAT_095 ((double)0.95f)

struct d4 { double p, u, v, w; };

__device__ __forceinline__ d4 ld4(const double *__restrict__ a, uint32_t v)
{
  const double2 *p = reinterpret_cast<const double2 *>(a + 4 * (size_t)v);
  const double2 lo = p[0], hi = p[1];
  return {lo.x, lo.y, hi.x, hi.y};
}

__device__ __forceinline__ void st4(double *__restrict__ a, uint32_t v, double p, double u, double w_, double z)
{
  double2 *o = reinterpret_cast<double2 *>(a + 4 * (size_t)v);
  o[0] = make_double2(p, u);
  o[1] = make_double2(w_, z);
}

/* basis (V1, V2) orthogonal to the unit normal: src/flux.c:179-217 */
__device__ __forceinline__ void tangent_basis(double xn, double yn, double zn, double &X1, double &Y1, double &Z1,
                                              double &X2, double &Y2, double &Z2)
{
  double dot = xn;
  if (fabs(dot) < FLOAT_095) {
    X1 = 1.0 - dot * xn; Y1 = -dot * yn; Z1 = -dot * zn;
  } else {
    dot = yn;
    X1 = -dot * xn; Y1 = 1.0 - dot * yn; Z1 = -dot * zn;
  }
  double size = X1 * X1;
  size += Y1 * Y1;
  size += Z1 * Z1;
  size = sqrt(size);
  X1 /= size; Y1 /= size; Z1 /= size;
  X2 = yn * Z1; X2 -= zn * Y1;
  Y2 = zn * X1; Y2 -= xn * Z1;
  Z2 = xn * Y1; Z2 -= yn * X1;
}

/* ---- LSQ weights: src/weights.c:9-398 -------------------------------------- */

/* The terms of one incidence of vertex v (src/weights.c:9-100) from v's seven
 * moments m = (w00, w01, w02, w11, w12, w22, w33) and the edge vector
 * (dx, dy, dz) = x(n1) - x(n0): side 0 when v is the edge's n0, side 1 when
 * it is n1.  The same expressions, in the same order, as the host's
 * per-edge terms, so the same operands give the same bits. */
__device__ __forceinline__ void lsq_terms(const double m[7], double dx, double dy, double dz, bool side1,
                                          double &tx, double &ty, double &tz)
{
  double c0, c1;
  if (!side1) {
    c0 = -dx * m[1] + dy;
    c1 = -dx * m[2] + dz;
  } else {
    c0 = dx * m[1] - dy;
    c1 = dx * m[2] - dz;
  }
  c1 = -m[4] * c0 + c1;
  tx = m[3] * m[1] * c0;
  tx = (side1 ? -dx : dx) * m[0] - tx;
  tx += m[6] * c1;
  ty = m[4] * m[5] * c1;
  ty = m[3] * c0 - ty;
  tz = m[5] * c1;
}

/* (dx, dy, dz) = x(n1) - x(n0) of the edge between v (at cv) and nb (at cn) */
__device__ __forceinline__ void edge_vector(const d4 &cv, const d4 &cn, bool side1, double &dx, double &dy,
                                            double &dz)
{
  if (!side1) { dx = cn.p - cv.p; dy = cn.u - cv.u; dz = cn.v - cv.v; }
  else { dx = cv.p - cn.p; dy = cv.u - cn.u; dz = cv.v - cn.v; }
}

/* Per-vertex moments (src/weights.c:103-366): three passes over the
 * incidences in ascending edge id, each reading only the vertex's own sums
 * of the earlier passes, then the magnitudes.  Also writes sell_nbs. */
__global__ __launch_bounds__(256) void k_lsq_moments(kf_dv d, double *__restrict__ m, uint32_t *__restrict__ nbs)
{
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= d.N) return;
  const d4 cv = ld4(d.xyzw, v);
  const uint32_t deg = d.deg[v];
  const size_t base = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0, w4 = 0.0, w5 = 0.0;
  /* w00, w01, w02: src/weights.c:237-297 (r = x(nb) - x(v) from either end) */
  for (uint32_t k = 0; k < deg; k++) {
    const size_t idx = base + 64 * (size_t)k;
    const uint32_t nb = d.sell_nbr[idx], side = d.sell_ent[idx] & 1u;
    nbs[idx] = nb | side << 31;
    const d4 cn = ld4(d.xyzw, nb);
    const double rx = cn.p - cv.p, ry = cn.u - cv.u, rz = cn.v - cv.v;
    w0 += rx * rx; w1 += rx * ry; w2 += rx * rz;
  }
  w0 = sqrt(w0);
  w1 /= w0;
  w2 /= w0;
  /* w11, w12: src/weights.c:173-232 */
  for (uint32_t k = 0; k < deg; k++) {
    const size_t idx = base + 64 * (size_t)k;
    const bool side1 = d.sell_ent[idx] & 1u;
    const d4 cn = ld4(d.xyzw, d.sell_nbr[idx]);
    double dx, dy, dz;
    edge_vector(cv, cn, side1, dx, dy, dz);
    const double dd = w1 / w0;
    if (!side1) {
      const double d_ = dy - dx * dd;
      w3 += d_ * d_;
      w4 += d_ * dz;
    } else {
      const double d_ = -dy + dx * dd;
      w3 += d_ * d_;
      w4 -= d_ * dz;
    }
  }
  w3 = sqrt(w3);
  w4 /= w3;
  /* w22: src/weights.c:103-170 */
  for (uint32_t k = 0; k < deg; k++) {
    const size_t idx = base + 64 * (size_t)k;
    const bool side1 = d.sell_ent[idx] & 1u;
    const d4 cn = ld4(d.xyzw, d.sell_nbr[idx]);
    double dx, dy, dz;
    edge_vector(cv, cn, side1, dx, dy, dz);
    const double d0 = w1 / w0, d1 = w2 / w0, d2 = w4 / w3;
    double d0_, d1_;
    if (!side1) { d0_ = dy - dx * d0; d1_ = dz - dx * d1; }
    else { d0_ = -dy + dx * d0; d1_ = -dz + dx * d1; }
    const double d2_ = d1_ - d2 * d0_;
    w5 += d2_ * d2_;
  }
  /* magnitudes: src/weights.c:313-366 */
  w5 = sqrt(w5);
  const double sw00 = w0 * w0, sw11 = w3 * w3, sw22 = w5 * w5;
  double m0 = w1 * w4;
  m0 -= w2 * w3;
  const double m1 = w0 * w3 * sw22;
  const size_t N = d.N;
  m[v] = 1.0 / sw00;
  m[N + v] = w1 / w0;
  m[2 * N + v] = w2 / w0;
  m[3 * N + v] = 1.0 / sw11;
  m[4 * N + v] = w4 / w3;
  m[5 * N + v] = 1.0 / sw22;
  m[6 * N + v] = m0 / m1;
}

__device__ __forceinline__ void lsq_load(const double *__restrict__ m, uint32_t N, uint32_t v, double out[7])
{
#pragma unroll
  for (int k = 0; k < 7; k++) out[k] = m[(size_t)k * N + v];
}

/* every incidence's terms from the moments against the uploaded ones, bit
 * for bit; any difference sets *bad */
__global__ __launch_bounds__(256) void k_lsq_check(kf_dv d, const double *__restrict__ m, uint32_t *__restrict__ bad)
{
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v >= d.N) return;
  double mv[7];
  lsq_load(m, d.N, v, mv);
  const d4 cv = ld4(d.xyzw, v);
  const uint32_t deg = d.deg[v];
  const size_t base = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  bool ok = true;
  for (uint32_t k = 0; k < deg; k++) {
    const size_t idx = base + 64 * (size_t)k;
    const bool side1 = d.sell_ent[idx] & 1u;
    const d4 cn = ld4(d.xyzw, d.sell_nbr[idx]);
    double dx, dy, dz, tx, ty, tz;
    edge_vector(cv, cn, side1, dx, dy, dz);
    lsq_terms(mv, dx, dy, dz, side1, tx, ty, tz);
    ok = ok && __double_as_longlong(tx) == __double_as_longlong(d.sell_wx[idx]) &&
         __double_as_longlong(ty) == __double_as_longlong(d.sell_wy[idx]) &&
         __double_as_longlong(tz) == __double_as_longlong(d.sell_wz[idx]);
  }
  if (!ok) atomicOr(bad, 1u);
}

/* k_grad's LSQ terms from per-vertex moments (56 bytes a vertex, and the
 * neighbour's coordinates beside its state) rather than 24 bytes per
 * incidence, once every incidence's terms are shown bit-identical to the
 * GEOMETRY's (KF_GRAD_MOMENTS=0 keeps the per-incidence terms) */
void kf_grad_setup(kf_ctx *c)
{
  kf_dv &d = c->d;
  const char *ev = getenv("KF_GRAD_MOMENTS");
  c->grad_mom = 0;
  if ((ev && atoi(ev) == 0) || d.N == 0 || d.N > 0x7fffffffu) return;
  uint32_t nslice_rows;
  KF_CHECK(hipMemcpy(&nslice_rows, d.sell_ptr + d.nslice, sizeof(uint32_t), hipMemcpyDeviceToHost));
  KF_CHECK(hipMalloc((void **)&d.lsq, 7 * (size_t)d.N * sizeof(double)));
  KF_CHECK(hipMalloc((void **)&d.sell_nbs, (size_t)nslice_rows * 64 * sizeof(uint32_t) + sizeof(uint32_t)));
  uint32_t *bad;
  KF_CHECK(hipMalloc((void **)&bad, sizeof(uint32_t)));
  KF_CHECK(hipMemsetAsync(bad, 0, sizeof(uint32_t), c->stream));
  const dim3 gv(kf_blocks_for(d.N, 256)), b(256);
  hipLaunchKernelGGL(k_lsq_moments, gv, b, 0, c->stream, d, d.lsq, d.sell_nbs);
  KF_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_lsq_check, gv, b, 0, c->stream, d, (const double *)d.lsq, bad);
  KF_LAUNCH_CHECK();
  uint32_t hbad = 1;
  KF_CHECK(hipMemcpyAsync(&hbad, bad, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  KF_CHECK(hipStreamSynchronize(c->stream));
  KF_CHECK(hipFree(bad));
  if (hbad) {
    KF_CHECK(hipFree(d.lsq));
    KF_CHECK(hipFree(d.sell_nbs));
    d.lsq = NULL;
    d.sell_nbs = NULL;
    return;
  }
  KF_CHECK(hipFree(d.sell_wx));
  KF_CHECK(hipFree(d.sell_wy));
  KF_CHECK(hipFree(d.sell_wz));
  d.sell_wx = d.sell_wy = d.sell_wz = NULL;
  c->grad_mom = 1;
}

/* ---- LSQ gradient: src/flux.c:60-145 as a per-vertex gather -------------- */

/* a += t (q(nb) - q(v)), component by component */
__device__ __forceinline__ void grad_add(double a0[4], double a1[4], double a2[4], const double qs[4], const d4 &qn,
                                         double tx, double ty, double tz)
{
  const double qb[4] = {qn.p, qn.u, qn.v, qn.w};
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const double dq = qb[c] - qs[c];
    a0[c] += tx * dq;
    a1[c] += ty * dq;
    a2[c] += tz * dq;
  }
}

/* MOM: the terms from the moments (lsq_terms, the bits of the per-incidence
 * ones, which kf_grad_setup checked), two incidences a round with all four
 * gathers issued before the sums, which keep the edge order (288x96x96:
 * 0.299 ms with the per-incidence terms, 0.264 ms one incidence a round,
 * 0.243 ms two; four a round at 118 VGPRs, or two capped at 96 VGPRs for a
 * fifth wave, 0.31-0.32 ms) */
template <bool MOM>
__global__ __launch_bounds__(256) void k_grad(kf_dv d, const double *__restrict__ q, double *__restrict__ gx,
                                              double *__restrict__ gy, double *__restrict__ gz, int first_order)
{
  const uint32_t v = kf_xcd_block() * 256u + threadIdx.x;
  if (v >= d.N) return;
  double a0[4] = {0.0, 0.0, 0.0, 0.0}, a1[4] = {0.0, 0.0, 0.0, 0.0}, a2[4] = {0.0, 0.0, 0.0, 0.0};
  if (!first_order) {
    const d4 qv = ld4(q, v);
    const double qs[4] = {qv.p, qv.u, qv.v, qv.w};
    const uint32_t deg = d.deg[v];
    size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
    uint32_t k = 0;
    if (MOM) {
      double mv[7];
      lsq_load(d.lsq, d.N, v, mv);
      const d4 cv = ld4(d.xyzw, v);
      double dx, dy, dz, tx, ty, tz;
      for (; k + 1 < deg; k += 2, idx += 128) {
        const uint32_t sa = d.sell_nbs[idx], sb = d.sell_nbs[idx + 64];
        const uint32_t na = sa & 0x7fffffffu, nb = sb & 0x7fffffffu;
        const d4 ca = ld4(d.xyzw, na), cb = ld4(d.xyzw, nb);
        const d4 qa = ld4(q, na), qb = ld4(q, nb);
        edge_vector(cv, ca, sa >> 31, dx, dy, dz);
        lsq_terms(mv, dx, dy, dz, sa >> 31, tx, ty, tz);
        grad_add(a0, a1, a2, qs, qa, tx, ty, tz);
        edge_vector(cv, cb, sb >> 31, dx, dy, dz);
        lsq_terms(mv, dx, dy, dz, sb >> 31, tx, ty, tz);
        grad_add(a0, a1, a2, qs, qb, tx, ty, tz);
      }
      if (k < deg) {
        const uint32_t sa = d.sell_nbs[idx], na = sa & 0x7fffffffu;
        const d4 ca = ld4(d.xyzw, na), qa = ld4(q, na);
        edge_vector(cv, ca, sa >> 31, dx, dy, dz);
        lsq_terms(mv, dx, dy, dz, sa >> 31, tx, ty, tz);
        grad_add(a0, a1, a2, qs, qa, tx, ty, tz);
      }
    } else {
      for (; k < deg; k++, idx += 64) {
        const uint32_t nb = d.sell_nbr[idx];
        grad_add(a0, a1, a2, qs, ld4(q, nb), d.sell_wx[idx], d.sell_wy[idx], d.sell_wz[idx]);
      }
    }
  }
  st4(gx, v, a0[0], a0[1], a0[2], a0[3]);
  st4(gy, v, a1[0], a1[1], a1[2], a1[3]);
  st4(gz, v, a2[0], a2[1], a2[2], a2[3]);
}

/* ---- Roe-type edge flux: src/flux.c:151-448, one thread per edge --------- */

/* edge term of the local time step (src/timestep.c:58-101), from the states
 * at the two end points: |u_avg . n ln| + c ln */
__device__ __forceinline__ double spectral_term(const d4 &qa, const d4 &qb, double xn, double yn, double zn,
                                                double ln)
{
  const double xnorm = xn * ln, ynorm = yn * ln, znorm = zn * ln;
  const double u = 0.5 * (qa.u + qb.u), vv = 0.5 * (qa.v + qb.v), w = 0.5 * (qa.w + qb.w);
  const double ubar = xn * u + yn * vv + zn * w;
  const double c = sqrt(ubar * ubar + B_);
  double term = u * xnorm;
  term += vv * ynorm;
  term += w * znorm;
  return fabs(term) + c * ln;
}

/* boundary-node terms of the local time step (src/timestep.c:103-150): solid
 * nodes, then free nodes, added to acc in that order */
__device__ __forceinline__ double spectral_boundary(const kf_dv &d, uint32_t v, const d4 &qv, double acc)
{
  for (int sec = 0; sec < 2; sec++) {
    const uint32_t *ptr = sec ? d.fnptr : d.snptr, *ent = sec ? d.fnent : d.snent;
    const double *nx = sec ? d.fnx : d.snx, *ny = sec ? d.fny : d.sny, *nz = sec ? d.fnz : d.snz;
    for (uint32_t s = ptr[v]; s < ptr[v + 1]; s++) {
      const uint32_t i = ent[s];
      const double xn = nx[i], yn = ny[i], zn = nz[i];
      const double ln = sqrt(xn * xn + yn * yn + zn * zn);
      const double ubar = qv.u * xn + qv.v * yn + qv.w * zn;
      const double ubar_ = ubar / ln;
      const double c = sqrt(ubar_ * ubar_ + B_);
      const double Vn = fabs(ubar) + c * ln;
      acc += Vn;
    }
  }
  return acc;
}

/* SPECTRAL: also the edge's time-step term into eterm (the driver's first
 * residual of a pseudo-time step, whose state the time step then uses).
 * The end points' coordinates come from the interleaved copy (one 32-byte
 * gather each instead of three; 0.507 -> 0.479 ms on the 2.65M-vertex
 * lattice).  Interleaving the gradient too (12 values per vertex) made the
 * gradient kernel's stores slower than the gathers it saved here. */
template <bool SPECTRAL>
__global__ __launch_bounds__(256) void k_edge_flux(kf_dv d, const double *__restrict__ q,
                                                   const double *__restrict__ gx, const double *__restrict__ gy,
                                                   const double *__restrict__ gz, double *__restrict__ eres,
                                                   double *__restrict__ eterm)
{
  const uint32_t i = kf_xcd_block() * 256u + threadIdx.x;
  if (i >= d.E) return;
  const uint2 nn = reinterpret_cast<const uint2 *>(d.en)[i];
  const uint32_t node0 = nn.x, node1 = nn.y;
  const kf_egeo eg = kf_edge_geo(d.egeo, i);
  const double xn = eg.xn, yn = eg.yn, zn = eg.zn, ln = eg.ln;
  const d4 c0 = ld4(d.xyzw, node0), c1 = ld4(d.xyzw, node1);
  const double x0 = c0.p, y0 = c0.u, z0 = c0.v;
  const double x1 = c1.p, y1 = c1.u, z1 = c1.v;
  const double xmean = 0.5 * (x0 + x1);
  const double ymean = 0.5 * (y0 + y1);
  const double zmean = 0.5 * (z0 + z1);
  double X1, Y1, Z1, X2, Y2, Z2;
  tangent_basis(xn, yn, zn, X1, Y1, Z1, X2, Y2, Z2);

  const d4 q0 = ld4(q, node0), q1 = ld4(q, node1);
  const d4 g0x = ld4(gx, node0), g0y = ld4(gy, node0), g0z = ld4(gz, node0);
  const d4 g1x = ld4(gx, node1), g1y = ld4(gy, node1), g1z = ld4(gz, node1);

  double rx = xmean - x0, ry = ymean - y0, rz = zmean - z0;
  double pL = q0.p + g0x.p * rx; pL += g0y.p * ry; pL += g0z.p * rz;
  double uL = q0.u + g0x.u * rx; uL += g0y.u * ry; uL += g0z.u * rz;
  double vL = q0.v + g0x.v * rx; vL += g0y.v * ry; vL += g0z.v * rz;
  double wL = q0.w + g0x.w * rx; wL += g0y.w * ry; wL += g0z.w * rz;
  double ubarL = xn * uL; ubarL += yn * vL; ubarL += zn * wL;

  rx = xmean - x1; ry = ymean - y1; rz = zmean - z1;
  double pR = q1.p + g1x.p * rx; pR += g1y.p * ry; pR += g1z.p * rz;
  double uR = q1.u + g1x.u * rx; uR += g1y.u * ry; uR += g1z.u * rz;
  double vR = q1.v + g1x.v * rx; vR += g1y.v * ry; vR += g1z.v * rz;
  double wR = q1.w + g1x.w * rx; wR += g1y.w * ry; wR += g1z.w * rz;
  double ubarR = xn * uR; ubarR += yn * vR; ubarR += zn * wR;

  const double u = 0.5 * (uL + uR), v = 0.5 * (vL + vR), w = 0.5 * (wL + wR);
  double ubar = xn * u; ubar += yn * v; ubar += zn * w;
  double phi1 = xn * B_; phi1 += u * ubar;
  double phi2 = yn * B_; phi2 += v * ubar;
  double phi3 = zn * B_; phi3 += w * ubar;
  double phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
  double phi5 = Z2 * phi1; phi5 -= X2 * phi3;
  double phi6 = X2 * phi2; phi6 -= Y2 * phi1;
  double phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
  double phi8 = X1 * phi3; phi8 -= Z1 * phi1;
  double phi9 = Y1 * phi1; phi9 -= X1 * phi2;
  const double c2 = ubar * ubar + B_;
  const double c = sqrt(c2);
  const double eig1 = fabs(ubar), eig2 = fabs(ubar), eig3 = fabs(ubar + c), eig4 = fabs(ubar - c);
  const double dp = pR - pL, du = uR - uL, dv = vR - vL, dw = wR - wL;
  double ti11 = u * phi4; ti11 += v * phi5; ti11 += w * phi6; ti11 = -ti11 / B_;
  double ti21 = u * phi7; ti21 += v * phi8; ti21 += w * phi9; ti21 = -ti21 / B_;
  double ti31 = 0.5 * (c - ubar); ti31 /= B_;
  double ti41 = -0.5 * (c + ubar); ti41 /= B_;
  double dv1 = ti11 * dp; dv1 += phi4 * du; dv1 += phi5 * dv; dv1 += phi6 * dw; dv1 /= c2;
  double dv2 = ti21 * dp; dv2 += phi7 * du; dv2 += phi8 * dv; dv2 += phi9 * dw; dv2 /= c2;
  double dv3 = 2.0 * ti31 * dp; dv3 += xn * du; dv3 += yn * dv; dv3 += zn * dw; dv3 *= 0.5 / c2;
  double dv4 = 2.0 * ti41 * dp; dv4 += xn * du; dv4 += yn * dv; dv4 += zn * dw; dv4 *= 0.5 / c2;
  const double r13 = c * B_;
  double r23 = u * (ubar + c); r23 += xn * B_;
  double r33 = v * (ubar + c); r33 += yn * B_;
  double r43 = w * (ubar + c); r43 += zn * B_;
  const double r14 = -c * B_;
  double r24 = u * (ubar - c); r24 += xn * B_;
  double r34 = v * (ubar - c); r34 += yn * B_;
  double r44 = w * (ubar - c); r44 += zn * B_;
  const double t1 = eig3 * r13 * dv3 + eig4 * r14 * dv4;
  double t2 = eig1 * X1 * dv1 + eig2 * X2 * dv2; t2 += eig3 * r23 * dv3 + eig4 * r24 * dv4;
  double t3 = eig1 * Y1 * dv1 + eig2 * Y2 * dv2; t3 += eig3 * r33 * dv3 + eig4 * r34 * dv4;
  double t4 = eig1 * Z1 * dv1 + eig2 * Z2 * dv2; t4 += eig3 * r43 * dv3 + eig4 * r44 * dv4;
  const double fluxp1 = ln * B_ * ubarL;
  const double fluxp2 = ln * (uL * ubarL + xn * pL);
  const double fluxp3 = ln * (vL * ubarL + yn * pL);
  const double fluxp4 = ln * (wL * ubarL + zn * pL);
  const double fluxm1 = ln * B_ * ubarR;
  const double fluxm2 = ln * (uR * ubarR + xn * pR);
  const double fluxm3 = ln * (vR * ubarR + yn * pR);
  const double fluxm4 = ln * (wR * ubarR + zn * pR);
  const double res1 = 0.5 * (fluxp1 + fluxm1 - ln * t1);
  const double res2 = 0.5 * (fluxp2 + fluxm2 - ln * t2);
  const double res3 = 0.5 * (fluxp3 + fluxm3 - ln * t3);
  const double res4 = 0.5 * (fluxp4 + fluxm4 - ln * t4);
  st4(eres, i, res1, res2, res3, res4);
  if (SPECTRAL) eterm[i] = spectral_term(q0, q1, xn, yn, zn, ln);
}

/* ---- slip-wall face pressure terms: src/flux.c:477-537 ------------------- */

__device__ __forceinline__ void wall_term(const kf_dv &d, const double *__restrict__ q, uint32_t f, uint32_t corner,
                                          double &xn, double &yn, double &zn, double &pc)
{
  const uint32_t node0 = d.sf0[f], node1 = d.sf1[f], node2 = d.sf2[f];
  const double p1 = q[4 * (size_t)node0], p2 = q[4 * (size_t)node1], p3 = q[4 * (size_t)node2];
  const double ax = d.x[node1] - d.x[node0], ay = d.y[node1] - d.y[node0], az = d.z[node1] - d.z[node0];
  const double bx = d.x[node2] - d.x[node0], by = d.y[node2] - d.y[node0], bz = d.z[node2] - d.z[node0];
  const double mag0 = 0.5 / 3, mag1 = -(0.5 / 3);
  xn = ay * bz; xn -= az * by; xn *= mag1;
  yn = ax * bz; yn -= az * bx; yn *= mag0;
  zn = ax * by; zn -= ay * bx; zn *= mag1;
  if (corner == 0) { pc = 0.125 * (p2 + p3); pc += 0.75 * p1; }
  else if (corner == 1) { pc = 0.125 * (p3 + p1); pc += 0.75 * p2; }
  else { pc = 0.125 * (p1 + p2); pc += 0.75 * p3; }
}

/* ---- free-stream characteristic node flux: src/flux.c:543-778 ------------ */

__device__ __forceinline__ void freestream_term(const kf_dv &d, uint32_t i, const d4 &qi, double out[4])
{
  const double U = d.U, V = d.V, W = 0.0, P = 1.0;
  double xn = d.fnx[i], yn = d.fny[i], zn = d.fnz[i];
  double area = xn * xn; area += yn * yn; area += zn * zn; area = sqrt(area);
  xn /= area; yn /= area; zn /= area;
  double X1, Y1, Z1, X2, Y2, Z2;
  tangent_basis(xn, yn, zn, X1, Y1, Z1, X2, Y2, Z2);
  double ubar0 = xn * U; ubar0 += yn * V; ubar0 += zn * W;
  const double c20 = ubar0 * ubar0 + B_;
  const double c0 = sqrt(c20);
  double phi1 = xn * B_; phi1 += U * ubar0;
  double phi2 = yn * B_; phi2 += V * ubar0;
  double phi3 = zn * B_; phi3 += W * ubar0;
  double phi4 = Y2 * phi3; phi4 -= Z2 * phi2;
  double phi5 = Z2 * phi1; phi5 -= X2 * phi3;
  double phi6 = X2 * phi2; phi6 -= Y2 * phi1;
  double phi7 = Z1 * phi2; phi7 -= Y1 * phi3;
  double phi8 = X1 * phi3; phi8 -= Z1 * phi1;
  double phi9 = Y1 * phi1; phi9 -= X1 * phi2;
  const double t13 = c0 * B_;
  double t23 = U * (ubar0 + c0); t23 += xn * B_;
  double t33 = V * (ubar0 + c0); t33 += yn * B_;
  double t43 = W * (ubar0 + c0); t43 += zn * B_;
  const double t14 = -c0 * B_;
  double t24 = U * (ubar0 - c0); t24 += xn * B_;
  double t34 = V * (ubar0 - c0); t34 += yn * B_;
  double t44 = W * (ubar0 - c0); t44 += zn * B_;
  double ti11 = U * phi4; ti11 += V * phi5; ti11 += W * phi6; ti11 = -ti11 / B_;
  double ti21 = U * phi7; ti21 += V * phi8; ti21 += W * phi9; ti21 = -ti21 / B_;
  double ti31 = 0.5 * (c0 - ubar0); ti31 /= B_;
  double ti41 = -0.5 * (c0 + ubar0); ti41 /= B_;
  const double pi = qi.p, ui = qi.u, vi = qi.v, wi = qi.w;
  double un = xn * ui; un += yn * vi; un += zn * wi;
  double pr, ur, vr, wr;
  if (un > 0.0) { pr = pi; ur = ui; vr = vi; wr = wi; }
  else { pr = P; ur = U; vr = V; wr = W; }
  double rhs1 = ti11 * pr; rhs1 += phi4 * ur; rhs1 += phi5 * vr; rhs1 += phi6 * wr; rhs1 /= c20;
  double rhs2 = ti21 * pr; rhs2 += phi7 * ur; rhs2 += phi8 * vr; rhs2 += phi9 * wr; rhs2 /= c20;
  double rhs3 = 2.0 * ti31 * pi; rhs3 += xn * ui; rhs3 += yn * vi; rhs3 += zn * wi; rhs3 = 0.5 * rhs3 / c20;
  double rhs4 = 2.0 * ti41 * P; rhs4 += xn * U; rhs4 += yn * V; rhs4 += zn * W; rhs4 = 0.5 * rhs4 / c20;
  double pb = t13 * rhs3; pb += t14 * rhs4;
  double ub = X1 * rhs1; ub += X2 * rhs2; ub += t23 * rhs3; ub += t24 * rhs4;
  double vb = Y1 * rhs1; vb += Y2 * rhs2; vb += t33 * rhs3; vb += t34 * rhs4;
  double wb = Z1 * rhs1; wb += Z2 * rhs2; wb += t43 * rhs3; wb += t44 * rhs4;
  double ubar = xn * ub; ubar += yn * vb; ubar += zn * wb;
  out[0] = area * B_ * ubar;
  out[1] = area * (ub * ubar + xn * pb);
  out[2] = area * (vb * ubar + yn * pb);
  out[3] = area * (wb * ubar + zn * pb);
}

/* ---- per-vertex residual: edges (edge order), walls (colour order), free
 * stream, exactly the reference's order of updates to r[v] --------------- */

/* SPECTRAL: also the vertex's time-step sum (edge terms in edge order, then
 * the boundary-node terms: k_timestep's order) into sr, before the division
 * by the CFL */
/* FD: q is iJacVec's perturbed state w, and r receives its combination
 * ((R(w) - fd.r) + cdt (w - q0)) |x| / h (src/kernel.c:33-44) instead of
 * R(w): the operations of k_fd_combine on the same operands, so the same
 * bits, without R(w)'s round trip through memory */
template <bool SPECTRAL, bool FD>
__global__ __launch_bounds__(256) void k_residual(kf_dv d, const double *__restrict__ q,
                                                  const double *__restrict__ eres, double *__restrict__ r,
                                                  const double *__restrict__ eterm, double *__restrict__ sr, kf_fdb fd)
{
  const uint32_t v = kf_xcd_block() * 256u + threadIdx.x;
  if (v >= d.N) return;
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0, acc = 0.0;
  const uint32_t deg = d.deg[v];
  const size_t idx0 = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  /* incidences in chunks of RC: a chunk's edge results are all loaded
   * before its sums, which run in incidence (edge) order */
  constexpr uint32_t RC = 8;
  for (uint32_t k0 = 0; k0 < deg; k0 += RC) {
    uint32_t ent[RC];
    d4 e[RC];
    double et[RC];
#pragma unroll
    for (uint32_t i = 0; i < RC; i++)
      if (k0 + i < deg) ent[i] = d.sell_ent[idx0 + 64u * (k0 + i)];
#pragma unroll
    for (uint32_t i = 0; i < RC; i++)
      if (k0 + i < deg) {
        e[i] = ld4(eres, ent[i] >> 1);
        if (SPECTRAL) et[i] = eterm[ent[i] >> 1];
      }
#pragma unroll
    for (uint32_t i = 0; i < RC; i++)
      if (k0 + i < deg) {
        if (ent[i] & 1u) { r0 = r0 - e[i].p; r1 = r1 - e[i].u; r2 = r2 - e[i].v; r3 = r3 - e[i].w; }
        else { r0 = r0 + e[i].p; r1 = r1 + e[i].u; r2 = r2 + e[i].v; r3 = r3 + e[i].w; }
        if (SPECTRAL) acc = acc + et[i];
      }
  }
  if (SPECTRAL) sr[v] = spectral_boundary(d, v, ld4(q, v), acc);
  for (uint32_t s = d.sfptr[v]; s < d.sfptr[v + 1]; s++) {
    const uint32_t ent = d.sfent[s];
    double xn, yn, zn, pc;
    wall_term(d, q, ent >> 2, ent & 3u, xn, yn, zn, pc);
    r1 += xn * pc; r2 += yn * pc; r3 += zn * pc;
  }
  const uint32_t f0 = d.fnptr[v], f1 = d.fnptr[v + 1];
  if (f0 < f1) {
    const d4 qi = ld4(q, v);
    for (uint32_t s = f0; s < f1; s++) {
      double o[4];
      freestream_term(d, d.fnent[s], qi, o);
      r0 += o[0]; r1 += o[1]; r2 += o[2]; r3 += o[3];
    }
  }
  if (FD) {
    const double a = *fd.xnorm / fd.h;
    const double cd = fd.cdt[v];
    const d4 rb = ld4(fd.r, v), wv = ld4(q, v), q0 = ld4(fd.q0, v);
    r0 = (r0 - rb.p + cd * (wv.p - q0.p)) * a;
    r1 = (r1 - rb.u + cd * (wv.u - q0.u)) * a;
    r2 = (r2 - rb.v + cd * (wv.v - q0.v)) * a;
    r3 = (r3 - rb.w + cd * (wv.w - q0.w)) * a;
  }
  st4(r, v, r0, r1, r2, r3);
}

void kf_launch_flux(kf_ctx *c, const double *q, double *gx, double *gy, double *gz, double *r, double *sr,
                    const kf_fdb *fd)
{
  kf_check_aligned(q, "q"); kf_check_aligned(r, "r");
  kf_check_aligned(gx, "grad->x0"); kf_check_aligned(gy, "grad->x1"); kf_check_aligned(gz, "grad->x2");
  const kf_dv &d = c->d;
  const dim3 gv(kf_blocks_for(d.N, 256)), ge(kf_blocks_for(d.E, 256)), b(256);
  int s = kf_prof_begin(c, KP_GRAD);
  if (c->grad_mom) hipLaunchKernelGGL(k_grad<true>, gv, b, 0, c->stream, d, q, gx, gy, gz, c->opt.first_order);
  else hipLaunchKernelGGL(k_grad<false>, gv, b, 0, c->stream, d, q, gx, gy, gz, c->opt.first_order);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  s = kf_prof_begin(c, KP_EDGE);
  if (sr && !c->eterm) KF_CHECK(hipMalloc((void **)&c->eterm, (size_t)d.E * sizeof(double)));
  if (sr) hipLaunchKernelGGL(k_edge_flux<true>, ge, b, 0, c->stream, d, q, gx, gy, gz, d.eres, c->eterm);
  else hipLaunchKernelGGL(k_edge_flux<false>, ge, b, 0, c->stream, d, q, gx, gy, gz, d.eres, (double *)NULL);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  s = kf_prof_begin(c, KP_RES);
  const kf_fdb none = {};
  if (sr && fd) {
    fprintf(stderr, "kf: the residual with time-step sums takes no combination\n");
    abort();
  }
  if (sr)
    hipLaunchKernelGGL((k_residual<true, false>), dim3(kf_blocks_for(d.N, 256)), dim3(256), 0, c->stream, d, q,
                       d.eres, r, c->eterm, sr, none);
  else if (fd)
    hipLaunchKernelGGL((k_residual<false, true>), dim3(kf_blocks_for(d.N, 256)), dim3(256), 0, c->stream, d, q,
                       d.eres, r, (const double *)NULL, (double *)NULL, *fd);
  else
    hipLaunchKernelGGL((k_residual<false, false>), dim3(kf_blocks_for(d.N, 256)), dim3(256), 0, c->stream, d, q,
                       d.eres, r, (const double *)NULL, (double *)NULL, none);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
}

/* ---- local time step: src/timestep.c:58-152 as a per-vertex gather ------- */

__global__ __launch_bounds__(256) void k_timestep(kf_dv d, const double *__restrict__ q, double cfl,
                                                  double *__restrict__ cdt)
{
  const uint32_t v = kf_xcd_block() * 256u + threadIdx.x;
  if (v >= d.N) return;
  double acc = 0.0;
  const uint32_t deg = d.deg[v];
  size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  const d4 qv = ld4(q, v);
  for (uint32_t k = 0; k < deg; k++, idx += 64) {
    const uint32_t ent = d.sell_ent[idx], e = ent >> 1;
    const d4 qn = ld4(q, d.sell_nbr[idx]);
    const d4 qa = (ent & 1u) ? qn : qv, qb = (ent & 1u) ? qv : qn;   /* (n0, n1) */
    const kf_egeo eg = kf_edge_geo(d.egeo, e);
    acc = acc + spectral_term(qa, qb, eg.xn, eg.yn, eg.zn, eg.ln);
  }
  cdt[v] = spectral_boundary(d, v, qv, acc) / cfl;
}

/* cdt = sr / cfl from the sums the driver's residual left (same rounding as
 * k_timestep: the sum, then one division) */
__global__ __launch_bounds__(256) void k_timestep_div(uint32_t N, const double *__restrict__ sr, double cfl,
                                                      double *__restrict__ cdt)
{
  const uint32_t v = blockIdx.x * 256u + threadIdx.x;
  if (v < N) cdt[v] = sr[v] / cfl;
}

void kf_launch_timestep(kf_ctx *c, double cfl)
{
  const int s = kf_prof_begin(c, KP_TSTEP);
  hipLaunchKernelGGL(k_timestep, dim3(kf_blocks_for(c->d.N, 256)), dim3(256), 0, c->stream, c->d, c->q, cfl, c->cdt);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
}

void kf_launch_timestep_div(kf_ctx *c, const double *sr, double cfl)
{
  const int s = kf_prof_begin(c, KP_TSTEP);
  hipLaunchKernelGGL(k_timestep_div, dim3(kf_blocks_for(c->d.N, 256)), dim3(256), 0, c->stream, c->d.N, sr, cfl,
                     c->cdt);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
}

/* ---- wall forces: src/forces.c:27-98 --------------------------------------
 * per-face terms in parallel, then one lane sums them in colour order so the
 * three sums round exactly as the reference's one-thread loop */

__global__ __launch_bounds__(256) void k_force_terms(kf_dv d, const double *__restrict__ q, double *__restrict__ ft)
{
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= d.S) return;
  const uint32_t n0 = d.sf0[j], n1 = d.sf1[j], n2 = d.sf2[j];
  const double x0 = d.x[n0], y0 = d.y[n0], z0 = d.z[n0];
  const double x1 = d.x[n1], y1 = d.y[n1], z1 = d.z[n1];
  const double x2 = d.x[n2], y2 = d.y[n2], z2 = d.z[n2];
  const double ax = x1 - x0, ay = y1 - y0, az = z1 - z0;
  const double bx = x2 - x0, by = y2 - y0, bz = z2 - z0;
  double xnorm = ay * bz; xnorm -= az * by; xnorm = -0.5 * xnorm;
  double ynorm = ax * bz; ynorm -= az * bx; ynorm = 0.5 * ynorm;
  const double p0 = q[4 * (size_t)n0], p1 = q[4 * (size_t)n1], p2 = q[4 * (size_t)n2];
  const double press = (p0 + p1 + p2) / 3.0;
  const double cp = 2.0 * (press - 1.0);
  const double dcx = cp * xnorm, dcy = cp * ynorm;
  const double xmid = x0 + x1 + x2, ymid = y0 + y1 + y2;
  double *o = ft + 6 * (size_t)j;
  o[0] = dcx * d.V; o[1] = dcy * d.U;
  o[2] = dcx * d.U; o[3] = dcy * d.V;
  o[4] = (xmid - 0.25) * dcy; o[5] = ymid * dcx;
}

/* src/forces.c:33 reduces each colour into a zero-initialised private sum
 * that is then added to the running total.  The colours' sums are
 * independent: k_force_part forms colour blockIdx.x's, sequential in face
 * order as the reference's (the workgroup stages 256 faces' terms in LDS
 * with coalesced loads and one lane per component adds them up), and
 * k_force_total adds them to the total in colour order. */
__global__ __launch_bounds__(256) void k_force_part(kf_dv d, const double *__restrict__ ft, double *__restrict__ parts)
{
  __shared__ double s[6 * 256];
  const uint32_t cl = blockIdx.x, j0 = d.color[cl], j1 = d.color[cl + 1];
  /* lane c < 3 carries component c (lift, drag, moment) through the faces
   * in order: lift (l - t0) + t1, drag (dr + t2) + t3, moment (m + t4) - t5;
   * a sign of -1 is exact, so (x + (-1) t) == x - t bit for bit */
  const uint32_t c = threadIdx.x;
  const double s0 = c == 0 ? -1.0 : 1.0, s1 = c == 2 ? -1.0 : 1.0;
  double part = 0.0;
  for (uint32_t b = j0; b < j1; b += 256) {
    const uint32_t n = j1 - b < 256u ? j1 - b : 256u;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 6 * n; k += 256) s[k] = ft[6 * (size_t)b + k];
    __syncthreads();
    if (c < 3) {
      const double *o = s + 2 * c;
      uint32_t i = 0;
      for (; i + 8 <= n; i += 8) {
        double x[16];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          x[2 * q] = o[6 * (i + q)];
          x[2 * q + 1] = o[6 * (i + q) + 1];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) part = (part + s0 * x[2 * q]) + s1 * x[2 * q + 1];
      }
      for (; i < n; i++) part = (part + s0 * o[6 * i]) + s1 * o[6 * i + 1];
    }
  }
  if (c < 3) parts[3 * (size_t)cl + c] = part;
}

__global__ __launch_bounds__(64) void k_force_total(uint32_t nc, const double *__restrict__ parts,
                                                    double *__restrict__ out)
{
  const uint32_t c = threadIdx.x;
  if (c >= 3) return;
  double acc = 0.0;
  for (uint32_t cl = 0; cl < nc; cl++) acc = acc + parts[3 * (size_t)cl + c];
  out[c] = acc;
}

/* the face terms (ft[6 S]) summed per colour into parts[3 ncolor], then
 * into out[3] */
static void force_sums(const kf_dv &d, hipStream_t st, const double *ft, double *parts, double *out)
{
  if (d.ncolor) {
    hipLaunchKernelGGL(k_force_part, dim3(d.ncolor), dim3(256), 0, st, d, ft, parts);
    KF_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_force_total, dim3(1), dim3(64), 0, st, d.ncolor, (const double *)parts, out);
  KF_LAUNCH_CHECK();
}

void kf_launch_forces(kf_ctx *c, double *forces_host)
{
  const kf_dv &d = c->d;
  /* the per-edge flux buffer (4E doubles) is free between residuals */
  double *ft = d.eres;
  if (6 * (size_t)d.S + 3 * (size_t)d.ncolor > 4 * (size_t)d.E) {
    fprintf(stderr, "kf: force scratch too small\n");
    abort();
  }
  if (d.S) {
    hipLaunchKernelGGL(k_force_terms, dim3(kf_blocks_for(d.S, 256)), dim3(256), 0, c->stream, d, c->q, ft);
    KF_LAUNCH_CHECK();
  }
  force_sums(d, c->stream, ft, ft + 6 * (size_t)d.S, d.scal + (KF_NSCAL - 3));
  kf_scal_fetch(c, KF_NSCAL - 3, 3, forces_host);
}

/* The forces of the state as it stands once the work queued so far is done
 * (Kernel: right after a pseudo-time step's update), computed on a second
 * stream so that they overlap the next step's work: the same two kernels as
 * kf_launch_forces, with their own term buffer and scalar slots, the result
 * copied to the pinned host mirror.  kf_forces_collect waits for them and
 * orders the main stream after them; it must run before anything changes
 * the state again. */
void kf_forces_async(kf_ctx *c)
{
  const kf_dv &d = c->d;
  if (!c->fstream) {
    KF_CHECK(hipStreamCreateWithFlags(&c->fstream, hipStreamNonBlocking));
    KF_CHECK(hipEventCreateWithFlags(&c->ev_upd, hipEventDisableTiming));
    KF_CHECK(hipEventCreateWithFlags(&c->ev_forces, hipEventDisableTiming));
  }
  if (!c->ftbuf) KF_CHECK(hipMalloc((void **)&c->ftbuf, (6 * (size_t)d.S + 3 * (size_t)d.ncolor + 1) * sizeof(double)));
  KF_CHECK(hipEventRecord(c->ev_upd, c->stream));
  KF_CHECK(hipStreamWaitEvent(c->fstream, c->ev_upd, 0));
  if (d.S) {
    hipLaunchKernelGGL(k_force_terms, dim3(kf_blocks_for(d.S, 256)), dim3(256), 0, c->fstream, d, c->q, c->ftbuf);
    KF_LAUNCH_CHECK();
  }
  force_sums(d, c->fstream, c->ftbuf, c->ftbuf + 6 * (size_t)d.S, d.scal + KF_SLOT_FASYNC);
  KF_CHECK(hipMemcpyAsync(c->hscal + KF_SLOT_FASYNC, d.scal + KF_SLOT_FASYNC, 3 * sizeof(double),
                          hipMemcpyDeviceToHost, c->fstream));
  KF_CHECK(hipEventRecord(c->ev_forces, c->fstream));
  c->forces_pending = 1;
}

void kf_forces_collect(kf_ctx *c, double *forces_host)
{
  if (!c->forces_pending) {
    fprintf(stderr, "kf: no forces pending\n");
    abort();
  }
  KF_CHECK(hipStreamWaitEvent(c->stream, c->ev_forces, 0));
  KF_CHECK(hipEventSynchronize(c->ev_forces));
  for (int k = 0; k < 3; k++) forces_host[k] = c->hscal[KF_SLOT_FASYNC + k];
  c->forces_pending = 0;
}
