/*
 * mpnn.hip -- the learned preconditioner's edge message-passing network on
 * gfx950: forward, backward (parameter gradients) and Adam.
 *
 * The network turns the ILU(0) factor's diagonal blocks and the mesh
 * geometry into a sparse linear correction operator G; the preconditioner
 * is M v = z + G z with z = ILU(0)^-1 v, linear in v.  Context (per vertex
 * v; u over v's neighbours in ascending edge order, e_vu the edge's unit
 * normal times its area, + when v is the edge's first vertex):
 *   x_v   = D_v / sd                                  16 features
 *   h0_v  = relu(W0 x_v + b0)                          32
 *   round l = 0..2:
 *     m_v = mean_u relu(A_l h_v + B_l h_u + C_l e_vu + c_l)
 *     h_v = h_v + relu(U_l h_v + V_l m_v + u_l)
 * Operator heads (16 values = one 4 x 4 block, row-major):
 *   [S_v; P_v] = WSA h_v + bSA,   Q_v = Wb h_v,   T_vu = P_v + Q_u + We e_vu
 * Correction:
 *   delta_v = S_v z_v + (1 / deg v) sum_u T_vu z_u
 * D_v is the diagonal block as stored (the inverse of U(v,v)) and sd the
 * RMS of all D blocks.  The context runs once per factorisation
 * (kf_mpnn_prepare); applying G is one gather over the incidence lists.
 *
 * Dense layers -- forward (W0, A_l, B_l, [U_l V_l], WSA, Wb), the data
 * gradients of backward and every weight gradient -- run on the matrix
 * cores: v_mfma_f32_32x32x8_bf16 with both operands rounded to bfloat16 and
 * FP32 accumulation.  Messages, the operator application and its backward
 * and the loss are FP32 vector code; every sum has a fixed order (per-vertex
 * gathers over the sliced-ELL incidence in ascending edge order, fixed
 * vertex partitions, fixed-tree and index-order reductions), so the network,
 * its gradient and its training are deterministic.
 *
 * Parameters are one FP32 vector in the order W0 b0 {A B C c U V u}x3
 * WSA bSA Wb We; it is the master copy Adam updates, and the bf16 copies
 * the matrix cores read are repacked after every update.
 */
#include <math.h>
#include <string.h>

#include "kf_dev.h"

#define MP_H 32
#define MP_NIN 16 /* inputs: one diagonal block (a multiple of the MFMA k-step) */
#define MP_NR 3

/* offsets into the parameter vector */
#define MP_W0 0
#define MP_B0 (MP_W0 + MP_H * MP_NIN)
#define MP_R0 (MP_B0 + MP_H)
#define MP_RSZ (4 * MP_H * MP_H + MP_H * 3 + 2 * MP_H)
#define MP_A(l) (MP_R0 + (l) * MP_RSZ)
#define MP_B(l) (MP_A(l) + MP_H * MP_H)
#define MP_C(l) (MP_B(l) + MP_H * MP_H)
#define MP_CB(l) (MP_C(l) + MP_H * 3)
#define MP_U(l) (MP_CB(l) + MP_H)
#define MP_V(l) (MP_U(l) + MP_H * MP_H)
#define MP_UB(l) (MP_V(l) + MP_H * MP_H)
#define MP_WSA (MP_R0 + MP_NR * MP_RSZ) /* 32 x 32: S (rows 0-15), P (16-31) */
#define MP_BSA (MP_WSA + 32 * MP_H)
#define MP_WB (MP_BSA + 32)             /* 16 x 32: Q */
#define MP_WE (MP_WB + 16 * MP_H)       /* 16 x 3 */
#define MP_NPARAM (MP_WE + 16 * 3)
#define MP_HEADS (MP_NPARAM - MP_WSA)

static_assert(MP_NPARAM == 14928, "parameter layout of the network");

/* bf16 copies for the matrix cores, row-major [out][k].  Forward: W0
 * (32 x 16), per round A, B (32 x 32) and [U V] (32 x 64), then WSA and Wb
 * padded to 32 rows.  Backward (transposed): per round U^T, V^T (32 x 32)
 * and [A^T B^T] (32 x 64), then [WSA^T Wb^T] (32 x 64, Wb^T padded). */
#define MPB_W0 0
#define MPB_A(l) (MP_H * MP_NIN + (l) * 4 * MP_H * MP_H)
#define MPB_B(l) (MPB_A(l) + MP_H * MP_H)
#define MPB_UV(l) (MPB_B(l) + MP_H * MP_H)
#define MPB_WSA (MP_H * MP_NIN + MP_NR * 4 * MP_H * MP_H)
#define MPB_WB (MPB_WSA + MP_H * MP_H)
#define MPB_FWD (MPB_WB + MP_H * MP_H)
#define MPB_UT(l) (MPB_FWD + (l) * 4 * MP_H * MP_H)
#define MPB_VT(l) (MPB_UT(l) + MP_H * MP_H)
#define MPB_ABT(l) (MPB_VT(l) + MP_H * MP_H)
#define MPB_HT (MPB_FWD + MP_NR * 4 * MP_H * MP_H)
#define MPB_SIZE (MPB_HT + 2 * MP_H * MP_H)

#define MP_PARTS 1024 /* vertex partitions of the gradient reductions */
#define MP_RB 256

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float x)
{
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u); /* round to nearest, ties to even */
  return (unsigned short)(u >> 16);
}

/* ---- dense layer on the matrix cores ----------------------------------------
 * Y[n][o] = epi(sum_k X[n][k] W[o][k] + bias[o]) for o < 32 and k < K1 + K2:
 * the first K1 columns come from X1 (row stride K1), the rest from X2 (row
 * stride K2).  MASK: X1[n][k] counts only where M1[n][k] > 0 (the ReLU
 * derivative in backward).  One wavefront computes 32 rows.
 * Operand layout of v_mfma_f32_32x32x8_bf16 (lane l, element j < 4):
 *   A: X[row0 + l % 32][8 s + 4 (l / 32) + j]
 *   B: W[l % 32][8 s + 4 (l / 32) + j]
 *   D: register r holds Y[row0 + 8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
 * EPI 0: z   1: relu(z)   2: R + relu(z)   3: R + z   4: z / max(deg, 1)
 * Z (optional) receives z. */
template <int K1, int K2, int EPI, bool MASK>
__global__ __launch_bounds__(64) void k_dense32(uint32_t N, const float *__restrict__ X1, const float *__restrict__ X2,
                                                const float *__restrict__ M1, const unsigned short *__restrict__ W,
                                                const float *__restrict__ bias, const float *__restrict__ R,
                                                const uint32_t *__restrict__ deg, float *__restrict__ Z,
                                                float *__restrict__ Y)
{
  constexpr int K = K1 + K2;
  static_assert(K1 % 8 == 0 && K2 % 8 == 0, "k-steps of 8");
  const uint32_t lane = threadIdx.x, li = lane & 31u, lk = (lane >> 5) * 4u;
  const uint32_t row0 = blockIdx.x * 32u, n = row0 + li;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
  for (int s = 0; s < K / 8; s++) {
    const int k = 8 * s + (int)lk;
    bf16x4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      float x = 0.0f;
      if (n < N) {
        if (k < K1) {
          x = X1[(size_t)n * K1 + k + j];
          if (MASK && !(M1[(size_t)n * K1 + k + j] > 0.0f)) x = 0.0f;
        } else {
          x = X2[(size_t)n * K2 + (k - K1) + j];
        }
      }
      av[j] = (short)f2bf(x);
      bv[j] = (short)W[li * K + k + j];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  const float b = bias ? bias[li] : 0.0f;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const uint32_t row = row0 + 8u * (r / 4) + lk + (r % 4);
    if (row >= N) continue;
    const size_t o = (size_t)row * MP_H + li;
    const float z = acc[r] + b;
    if (Z) Z[o] = z;
    float y;
    if (EPI == 0) y = z;
    else if (EPI == 1) y = fmaxf(z, 0.0f);
    else if (EPI == 2) y = R[o] + fmaxf(z, 0.0f);
    else if (EPI == 3) y = R[o] + z;
    else y = z * (1.0f / (float)max(deg[row], 1u));
    Y[o] = y;
  }
}

/* ---- weight gradients on the matrix cores --------------------------------------
 * For vertex partition p = blockIdx.x (vertices [p chunk, (p+1) chunk)):
 *   part[p][offw + o KX + k] = sum_n dY[n][o] X[n][k]      o < OY, k < KX
 *   part[p][offb + o]        = sum_n dY[n][o]              (offb >= 0)
 * dY has row stride DYW, X row stride XW; MASK: dY[n][o] counts only where
 * Mz[n][o] > 0 (Mz with dY's layout).  The k dimension of the MFMA runs
 * over 8 vertices at a time:
 *   A: dY[n0 + 4 (l / 32) + j][l % 32]      (rows: o)
 *   B: X[n0 + 4 (l / 32) + j][l % 32]       (columns: k)
 * The bias sums are taken of the unrounded FP32 values. */
template <int DYW, int XW, int OY, int KX, bool MASK>
__global__ __launch_bounds__(64) void k_wgrad(uint32_t N, uint32_t chunk, const float *__restrict__ dY,
                                              const float *__restrict__ Mz, const float *__restrict__ X,
                                              float *__restrict__ part, uint32_t offw, int offb)
{
  static_assert(OY <= 32 && OY <= DYW && KX <= 32 && KX <= XW, "one 32 x 32 tile");
  const uint32_t lane = threadIdx.x, li = lane & 31u, lk = (lane >> 5) * 4u;
  const uint32_t p = blockIdx.x, n0 = p * chunk, n1 = min(N, n0 + chunk);
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.0f;
  float bsum = 0.0f;
  for (uint32_t nb = n0; nb < n1; nb += 8) {
    bf16x4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t n = nb + lk + j;
      float y = 0.0f, x = 0.0f;
      if (n < n1) {
        if (li < (uint32_t)OY) {
          y = dY[(size_t)n * DYW + li];
          if (MASK && !(Mz[(size_t)n * DYW + li] > 0.0f)) y = 0.0f;
        }
        if (li < (uint32_t)KX) x = X[(size_t)n * XW + li];
      }
      bsum += y;
      av[j] = (short)f2bf(y);
      bv[j] = (short)f2bf(x);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  bsum += __shfl_xor(bsum, 32);
  float *pp = part + (size_t)p * MP_NPARAM;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const uint32_t o = 8u * (r / 4) + lk + (r % 4);
    if (o < (uint32_t)OY && li < (uint32_t)KX) pp[offw + o * KX + li] = acc[r];
  }
  if (offb >= 0 && lane < (uint32_t)OY) pp[offb + lane] = bsum;
}

/* grad[i] = sum over partitions of part[p][i], in partition order (FP64) */
__global__ __launch_bounds__(MP_RB) void k_mp_grad_reduce(const float *__restrict__ part, uint32_t nparts,
                                                          float *__restrict__ grad)
{
  const uint32_t i = blockIdx.x * MP_RB + threadIdx.x;
  if (i >= MP_NPARAM) return;
  double acc = 0.0;
  for (uint32_t p = 0; p < nparts; p++) acc += (double)part[(size_t)p * MP_NPARAM + i];
  grad[i] = (float)acc;
}

/* ---- reductions and features ---------------------------------------------------- */

__device__ __forceinline__ double mp_block_sum(double v, double *sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = MP_RB / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

/* KF_NRED partial sums of squares (fixed grid and stride, fixed tree) */
template <typename T>
__global__ __launch_bounds__(MP_RB) void k_mp_sumsq(size_t n, const T *__restrict__ x, double *__restrict__ part)
{
  __shared__ double sh[MP_RB];
  double acc = 0.0;
  const size_t stride = (size_t)gridDim.x * MP_RB;
  for (size_t i = (size_t)blockIdx.x * MP_RB + threadIdx.x; i < n; i += stride) {
    const double v = (double)x[i];
    acc += v * v;
  }
  const double s = mp_block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

/* One lane sums the KF_NRED partials in index order into *out = 1 / RMS
 * over `count` values (1 when zero or not finite). */
__global__ void k_mp_finish(const double *__restrict__ part, double count, float *__restrict__ out)
{
  if (threadIdx.x != 0) return;
  double a = 0.0;
  for (int i = 0; i < KF_NRED; i++) a += part[i];
  double s = sqrt(a / count);
  if (!(s > 0.0) || !isfinite(s)) s = 1.0;
  *out = (float)(1.0 / s);
}

/* the ILU factor's diagonal blocks, as FP32 [N][16] */
template <typename TD>
__global__ __launch_bounds__(MP_RB) void k_mp_take_diag(uint32_t N, const uint32_t *__restrict__ diag,
                                                        const TD *__restrict__ a, float *__restrict__ D)
{
  const size_t k = (size_t)blockIdx.x * MP_RB + threadIdx.x;
  if (k >= 16 * (size_t)N) return;
  D[k] = (float)a[16 * (size_t)diag[k >> 4] + (k & 15)];
}

/* x_n = D_n / sd */
__global__ __launch_bounds__(MP_RB) void k_mp_features(size_t n, const float *__restrict__ D,
                                                       const float *__restrict__ isd_p, float *__restrict__ X)
{
  const size_t k = (size_t)blockIdx.x * MP_RB + threadIdx.x;
  if (k < n) X[k] = D[k] * *isd_p;
}

/* signed edge vector of every sliced-ELL slot: (nx, ny, nz) area, negated
 * where the vertex is the edge's second end */
__global__ __launch_bounds__(MP_RB) void k_mp_edge_vectors(kf_dv d, float4 *__restrict__ ev)
{
  const uint32_t v = blockIdx.x * MP_RB + threadIdx.x;
  if (v >= d.N) return;
  size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  for (uint32_t k = 0; k < d.deg[v]; k++, idx += 64) {
    const uint32_t ent = d.sell_ent[idx], e = ent >> 1;
    const kf_egeo eg = kf_edge_geo(d.egeo, e);
    const double ar = (ent & 1u) ? -eg.ln : eg.ln;
    ev[idx] = make_float4((float)(eg.xn * ar), (float)(eg.yn * ar), (float)(eg.zn * ar), 0.0f);
  }
}

/* ---- messages ------------------------------------------------------------------ */

/* m_v = mean over v's incidences of relu(P_v + Q_u + C e_vu + c); two
 * vertices per wavefront, lane c of a half computes channel c */
__global__ __launch_bounds__(64) void k_mp_message(kf_dv d, const float4 *__restrict__ ev,
                                                   const float *__restrict__ P, const float *__restrict__ Q,
                                                   const float *__restrict__ theta, uint32_t offC, uint32_t offc,
                                                   float *__restrict__ M)
{
  const uint32_t lane = threadIdx.x, c = lane & 31u, v = blockIdx.x * 2u + (lane >> 5);
  if (v >= d.N) return;
  const float c0 = theta[offC + 3 * c], c1 = theta[offC + 3 * c + 1], c2 = theta[offC + 3 * c + 2];
  const float cb = theta[offc + c], pv = P[(size_t)v * MP_H + c];
  const uint32_t deg = d.deg[v];
  size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  float acc = 0.0f;
  for (uint32_t k = 0; k < deg; k++, idx += 64) {
    const uint32_t u = d.sell_nbr[idx];
    const float4 e = ev[idx];
    const float z = pv + Q[(size_t)u * MP_H + c] + (c0 * e.x + c1 * e.y + c2 * e.z) + cb;
    acc += fmaxf(z, 0.0f);
  }
  M[(size_t)v * MP_H + c] = acc * (1.0f / (float)max(deg, 1u));
}

/* Backward of the messages for partition p = blockIdx.x: 8 vertices at a
 * time (32 lanes each, lane = channel c).  With dm already divided by the
 * degree:
 *   dP_v = sum_{(v,u)} dm_v [z_vu > 0]
 *   dQ_v = sum_{(v,u)} dm_u [z_uv > 0]     (the edge seen from u: e_uv = -e_vu)
 * and the partition's sums of dm_v [z_vu > 0] e_vu and of dm_v [z_vu > 0]
 * (the gradients of C and c), reduced over the 8 vertex groups in a fixed
 * order. */
__global__ __launch_bounds__(MP_RB) void k_mp_message_bwd(kf_dv d, uint32_t chunk, const float4 *__restrict__ ev,
                                                          const float *__restrict__ P, const float *__restrict__ Q,
                                                          const float *__restrict__ dm,
                                                          const float *__restrict__ theta, uint32_t offC,
                                                          uint32_t offc, float *__restrict__ dP,
                                                          float *__restrict__ dQ, float *__restrict__ part)
{
  __shared__ float sh[8][MP_H][4];
  const uint32_t c = threadIdx.x & 31u, grp = threadIdx.x >> 5;
  const uint32_t p = blockIdx.x, n0 = p * chunk, n1 = min(d.N, n0 + chunk);
  const float c0 = theta[offC + 3 * c], c1 = theta[offC + 3 * c + 1], c2 = theta[offC + 3 * c + 2];
  const float cb = theta[offc + c];
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, gb = 0.0f;
  for (uint32_t v = n0 + grp; v < n1; v += 8) {
    const float pv = P[(size_t)v * MP_H + c], qv = Q[(size_t)v * MP_H + c], dmv = dm[(size_t)v * MP_H + c];
    const uint32_t deg = d.deg[v];
    size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
    float ap = 0.0f, aq = 0.0f;
    for (uint32_t k = 0; k < deg; k++, idx += 64) {
      const uint32_t u = d.sell_nbr[idx];
      const float4 e = ev[idx];
      const float zvu = pv + Q[(size_t)u * MP_H + c] + (c0 * e.x + c1 * e.y + c2 * e.z) + cb;
      const float zuv = P[(size_t)u * MP_H + c] + qv + (c0 * -e.x + c1 * -e.y + c2 * -e.z) + cb;
      const float gv = zvu > 0.0f ? dmv : 0.0f;
      ap += gv;
      g0 += gv * e.x;
      g1 += gv * e.y;
      g2 += gv * e.z;
      gb += gv;
      aq += zuv > 0.0f ? dm[(size_t)u * MP_H + c] : 0.0f;
    }
    dP[(size_t)v * MP_H + c] = ap;
    dQ[(size_t)v * MP_H + c] = aq;
  }
  sh[grp][c][0] = g0;
  sh[grp][c][1] = g1;
  sh[grp][c][2] = g2;
  sh[grp][c][3] = gb;
  __syncthreads();
  if (threadIdx.x < MP_H * 4) {
    const uint32_t cc = threadIdx.x >> 2, j = threadIdx.x & 3u;
    float s = 0.0f;
    for (int q = 0; q < 8; q++) s += sh[q][cc][j];
    float *pp = part + (size_t)p * MP_NPARAM;
    if (j < 3) pp[offC + 3 * cc + j] = s;
    else pp[offc + cc] = s;
  }
}

/* ---- the correction operator ------------------------------------------------------ */

__device__ __forceinline__ float4 ld4(const double *z, uint32_t v)
{
  return make_float4((float)z[4 * (size_t)v], (float)z[4 * (size_t)v + 1], (float)z[4 * (size_t)v + 2],
                     (float)z[4 * (size_t)v + 3]);
}

/* delta_v = S_v z_v + (1/deg v) [(P_v + b) zs_v + We : Ez_v + sum_u Q_u z_u]
 * with zs_v = sum_u z_u and Ez_v[j][k] = sum_u z_u[j] e_vu[k] (the edge
 * blocks T_vu = P_v + b + Q_u + We e_vu factored), one thread per vertex.
 * OP1 = [S | P + b] (N x 32), OPQ = [Q | 0] (N x 32).  Output: FP64 out
 * (stored, or added when add) or, when outf is given, FP32 outf. */
__global__ __launch_bounds__(MP_RB) void k_mp_apply(kf_dv d, const float4 *__restrict__ ev,
                                                    const float *__restrict__ OP1, const float *__restrict__ OPQ,
                                                    const float *__restrict__ theta, const double *__restrict__ z,
                                                    double *__restrict__ out, int add, float *__restrict__ outf)
{
  __shared__ float we[48];
  if (threadIdx.x < 48) we[threadIdx.x] = theta[MP_WE + threadIdx.x];
  __syncthreads();
  const uint32_t v = blockIdx.x * MP_RB + threadIdx.x;
  if (v >= d.N) return;
  const uint32_t deg = d.deg[v];
  size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  float qz[4] = {0.0f, 0.0f, 0.0f, 0.0f}, zs[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ez[12];
#pragma unroll
  for (int i = 0; i < 12; i++) ez[i] = 0.0f;
  for (uint32_t k = 0; k < deg; k++, idx += 64) {
    const uint32_t u = d.sell_nbr[idx];
    const float4 e = ev[idx], zu4 = ld4(z, u);
    const float zu[4] = {zu4.x, zu4.y, zu4.z, zu4.w};
    const float4 *q = (const float4 *)(OPQ + (size_t)u * MP_H);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float4 qi = q[i];
      qz[i] += qi.x * zu[0] + qi.y * zu[1] + qi.z * zu[2] + qi.w * zu[3];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      zs[j] += zu[j];
      ez[3 * j] += zu[j] * e.x;
      ez[3 * j + 1] += zu[j] * e.y;
      ez[3 * j + 2] += zu[j] * e.z;
    }
  }
  const float4 zv4 = ld4(z, v);
  const float zv[4] = {zv4.x, zv4.y, zv4.z, zv4.w};
  const float *op = OP1 + (size_t)v * MP_H;
  const float idg = 1.0f / (float)max(deg, 1u);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float sv = 0.0f, t = qz[i];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      sv += op[4 * i + j] * zv[j];
      t += op[16 + 4 * i + j] * zs[j];
      t += we[3 * (4 * i + j)] * ez[3 * j] + we[3 * (4 * i + j) + 1] * ez[3 * j + 1] +
           we[3 * (4 * i + j) + 2] * ez[3 * j + 2];
    }
    const float dl = sv + t * idg;
    if (outf) outf[4 * (size_t)v + i] = dl;
    else out[4 * (size_t)v + i] = add ? out[4 * (size_t)v + i] + (double)dl : (double)dl;
  }
}

/* Backward of the correction given gv = dL/d delta (N x 4, FP32) and z:
 *   dOP1_v = [g_v z_v^T | (g_v / deg v) zs_v^T]       (N x 32)
 *   dOPQ_v = [gbar_v z_v^T | 0],  gbar_v = sum_u g_u / deg u   (N x 32)
 *   gd_v   = g_v / deg v (N x 4),  Ez_v (N x 12, [j][k] as 3 j + k)
 * (the gradient of We is then sum_v gd_v Ez_v^T, a weight gradient). */
__global__ __launch_bounds__(MP_RB) void k_mp_apply_bwd(kf_dv d, const float4 *__restrict__ ev,
                                                        const double *__restrict__ z, const float *__restrict__ gv,
                                                        float *__restrict__ dOP1, float *__restrict__ dOPQ,
                                                        float *__restrict__ gd, float *__restrict__ Ez)
{
  const uint32_t v = blockIdx.x * MP_RB + threadIdx.x;
  if (v >= d.N) return;
  const uint32_t deg = d.deg[v];
  size_t idx = (size_t)d.sell_ptr[v >> 6] * 64u + (v & 63u);
  float zs[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gb[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ez[12];
#pragma unroll
  for (int i = 0; i < 12; i++) ez[i] = 0.0f;
  for (uint32_t k = 0; k < deg; k++, idx += 64) {
    const uint32_t u = d.sell_nbr[idx];
    const float4 e = ev[idx], zu4 = ld4(z, u), gu = *(const float4 *)(gv + 4 * (size_t)u);
    const float iu = 1.0f / (float)max(d.deg[u], 1u);
    const float zu[4] = {zu4.x, zu4.y, zu4.z, zu4.w};
    gb[0] += gu.x * iu;
    gb[1] += gu.y * iu;
    gb[2] += gu.z * iu;
    gb[3] += gu.w * iu;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      zs[j] += zu[j];
      ez[3 * j] += zu[j] * e.x;
      ez[3 * j + 1] += zu[j] * e.y;
      ez[3 * j + 2] += zu[j] * e.z;
    }
  }
  const float4 zv4 = ld4(z, v), g4 = *(const float4 *)(gv + 4 * (size_t)v);
  const float zv[4] = {zv4.x, zv4.y, zv4.z, zv4.w}, g[4] = {g4.x, g4.y, g4.z, g4.w};
  const float idg = 1.0f / (float)max(deg, 1u);
  float *o1 = dOP1 + (size_t)v * MP_H, *oq = dOPQ + (size_t)v * MP_H;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float gi = g[i] * idg;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      o1[4 * i + j] = g[i] * zv[j];
      o1[16 + 4 * i + j] = gi * zs[j];
      oq[4 * i + j] = gb[i] * zv[j];
      oq[16 + 4 * i + j] = 0.0f;
    }
    gd[4 * (size_t)v + i] = gi;
  }
#pragma unroll
  for (int i = 0; i < 12; i++) Ez[12 * (size_t)v + i] = ez[i];
}

/* ---- loss ---------------------------------------------------------------------------- */

/* e = base + delta - target (target may be NULL: e = base + delta), FP32
 * computed in FP64 */
__global__ __launch_bounds__(MP_RB) void k_mp_err(size_t n, const float *__restrict__ dl,
                                                  const double *__restrict__ base, const double *__restrict__ target,
                                                  float *__restrict__ e)
{
  const size_t k = (size_t)blockIdx.x * MP_RB + threadIdx.x;
  if (k < n) e[k] = (float)(base[k] + (double)dl[k] - (target ? target[k] : 0.0));
}

/* y -= v (FP32 y, FP64 v) */
__global__ __launch_bounds__(MP_RB) void k_mp_sub(size_t n, const double *__restrict__ v, float *__restrict__ y)
{
  const size_t k = (size_t)blockIdx.x * MP_RB + threadIdx.x;
  if (k < n) y[k] = (float)((double)y[k] - v[k]);
}

/* y_i = sum over row i's blocks A(i, j) x_j (one thread per block row) */
__global__ __launch_bounds__(MP_RB) void k_mp_bsr_mv(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                     const uint32_t *__restrict__ col, const float *__restrict__ A,
                                                     const float *__restrict__ x, float *__restrict__ y)
{
  const uint32_t i = blockIdx.x * MP_RB + threadIdx.x;
  if (i >= N) return;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  for (uint32_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
    const float4 xv = *(const float4 *)(x + 4 * (size_t)col[k]);
    const float *b = A + 16 * (size_t)k;
    a0 += b[0] * xv.x + b[1] * xv.y + b[2] * xv.z + b[3] * xv.w;
    a1 += b[4] * xv.x + b[5] * xv.y + b[6] * xv.z + b[7] * xv.w;
    a2 += b[8] * xv.x + b[9] * xv.y + b[10] * xv.z + b[11] * xv.w;
    a3 += b[12] * xv.x + b[13] * xv.y + b[14] * xv.z + b[15] * xv.w;
  }
  *(float4 *)(y + 4 * (size_t)i) = make_float4(a0, a1, a2, a3);
}

/* g_j = sum over row j's positions k (column i = col[k]) of A(i, j)^T y_i,
 * A(i, j) at tpos[k] (the pattern is structurally symmetric) */
__global__ __launch_bounds__(MP_RB) void k_mp_bsr_mvt(uint32_t N, const uint32_t *__restrict__ rowptr,
                                                      const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ tpos, const float *__restrict__ A,
                                                      const float *__restrict__ y, float *__restrict__ g)
{
  const uint32_t j = blockIdx.x * MP_RB + threadIdx.x;
  if (j >= N) return;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
  for (uint32_t k = rowptr[j]; k < rowptr[j + 1]; k++) {
    const float4 yv = *(const float4 *)(y + 4 * (size_t)col[k]);
    const float *b = A + 16 * (size_t)tpos[k];
    a0 += b[0] * yv.x + b[4] * yv.y + b[8] * yv.z + b[12] * yv.w;
    a1 += b[1] * yv.x + b[5] * yv.y + b[9] * yv.z + b[13] * yv.w;
    a2 += b[2] * yv.x + b[6] * yv.y + b[10] * yv.z + b[14] * yv.w;
    a3 += b[3] * yv.x + b[7] * yv.y + b[11] * yv.z + b[15] * yv.w;
  }
  *(float4 *)(g + 4 * (size_t)j) = make_float4(a0, a1, a2, a3);
}

__global__ __launch_bounds__(MP_RB) void k_mp_sq4(uint32_t N, const float *__restrict__ x, double *__restrict__ lpart)
{
  __shared__ double sh[MP_RB];
  const uint32_t n = blockIdx.x * MP_RB + threadIdx.x;
  double acc = 0.0;
  if (n < N) {
#pragma unroll
    for (int o = 0; o < 4; o++) {
      const double v = (double)x[4 * (size_t)n + o];
      acc += v * v;
    }
  }
  const double s = mp_block_sum(acc, sh);
  if (threadIdx.x == 0) lpart[blockIdx.x] = s;
}

/* loss = (block partials summed in order) / (reference partials summed in
 * order) into out[0]; the gradient scale 2 / that reference sum into sc[3] */
__global__ void k_mp_loss_finish(const double *__restrict__ lpart, uint32_t nb, const double *__restrict__ tpart,
                                 float *__restrict__ out, float *__restrict__ sc)
{
  if (threadIdx.x != 0) return;
  double a = 0.0, t = 0.0;
  for (uint32_t i = 0; i < nb; i++) a += lpart[i];
  for (int i = 0; i < KF_NRED; i++) t += tpart[i];
  out[0] = t > 0.0 ? (float)(a / t) : 0.0f;
  sc[3] = t > 0.0 ? (float)(2.0 / t) : 0.0f;
}

/* g = f x, f = sc[3] */
__global__ __launch_bounds__(MP_RB) void k_mp_scale(size_t n, const float *__restrict__ x,
                                                    const float *__restrict__ sc, float *__restrict__ g)
{
  const size_t k = (size_t)blockIdx.x * MP_RB + threadIdx.x;
  if (k < n) g[k] = x[k] * sc[3];
}

/* A (FP64 or FP32 BSR values) as FP32 */
template <typename TA>
__global__ __launch_bounds__(MP_RB) void k_mp_to_f32(size_t n, const TA *__restrict__ a, float *__restrict__ out)
{
  const size_t k = (size_t)blockIdx.x * MP_RB + threadIdx.x;
  if (k < n) out[k] = (float)a[k];
}

/* ---- Adam ---------------------------------------------------------------------- */

/* One block: global gradient norm (fixed tree), clip to gclip (<= 0: none),
 * then m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 * theta -= lr (m c1) / (sqrt(v c2) + eps), c1 = 1/(1-b1^t), c2 = 1/(1-b2^t). */
__global__ __launch_bounds__(1024) void k_mp_adam(float *__restrict__ theta, float *__restrict__ m,
                                                  float *__restrict__ v, const float *__restrict__ grad, float lr,
                                                  float b1, float b2, float eps, float c1, float c2, float gclip,
                                                  float *__restrict__ gnorm_out)
{
  __shared__ double sh[1024];
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int i = t; i < MP_NPARAM; i += 1024) acc += (double)grad[i] * (double)grad[i];
  sh[t] = acc;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  const double gn = sqrt(sh[0]);
  const float scale = (gclip > 0.0f && gn > (double)gclip) ? (float)((double)gclip / gn) : 1.0f;
  if (t == 0 && gnorm_out) *gnorm_out = (float)gn;
  for (int i = t; i < MP_NPARAM; i += 1024) {
    const float g = grad[i] * scale;
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * (g * g);
    m[i] = mi;
    v[i] = vi;
    theta[i] = theta[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

/* bf16 copies of the dense weights (forward and transposed) */
__global__ void k_mp_pack(const float *__restrict__ theta, unsigned short *__restrict__ wb)
{
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= MPB_SIZE) return;
  float w = 0.0f;
  const uint32_t HH = MP_H * MP_H;
  if (t < MPB_A(0)) {
    w = theta[MP_W0 + t]; /* 32 x 16 as stored */
  } else if (t < MPB_WSA) {
    const uint32_t r = t - MPB_A(0), l = r / (4 * HH), q = r % (4 * HH);
    if (q < HH) w = theta[MP_A(l) + q];
    else if (q < 2 * HH) w = theta[MP_B(l) + q - HH];
    else {
      const uint32_t qq = q - 2 * HH, o = qq / (2 * MP_H), k = qq % (2 * MP_H);
      w = k < MP_H ? theta[MP_U(l) + o * MP_H + k] : theta[MP_V(l) + o * MP_H + (k - MP_H)];
    }
  } else if (t < MPB_WB) {
    w = theta[MP_WSA + (t - MPB_WSA)];
  } else if (t < MPB_FWD) {
    const uint32_t q = t - MPB_WB, o = q / MP_H, k = q % MP_H;
    w = o < 16 ? theta[MP_WB + o * MP_H + k] : 0.0f;
  } else if (t < MPB_HT) {
    const uint32_t r = t - MPB_FWD, l = r / (4 * HH), q = r % (4 * HH);
    if (q < 2 * HH) {
      /* U^T then V^T: row o, column k = U[k][o] (V[k][o]) */
      const uint32_t which = q / HH, qq = q % HH, o = qq / MP_H, k = qq % MP_H;
      w = theta[(which ? MP_V(l) : MP_U(l)) + k * MP_H + o];
    } else {
      const uint32_t qq = q - 2 * HH, o = qq / (2 * MP_H), k = qq % (2 * MP_H);
      w = k < MP_H ? theta[MP_A(l) + k * MP_H + o] : theta[MP_B(l) + (k - MP_H) * MP_H + o];
    }
  } else {
    /* [WSA^T Wb^T]: row o (hidden channel), column k < 32: WSA[k][o], then Wb[k - 32][o] */
    const uint32_t q = t - MPB_HT, o = q / (2 * MP_H), k = q % (2 * MP_H);
    if (k < MP_H) w = theta[MP_WSA + k * MP_H + o];
    else if (k - MP_H < 16) w = theta[MP_WB + (k - MP_H) * MP_H + o];
  }
  wb[t] = f2bf(w);
}

/* ---- the network object ------------------------------------------------------------ */

struct kf_mpnn {
  kf_ctx *c;
  uint32_t N, nparts, chunk, nslot, nlblk;
  float *theta, *am, *av, *grad, *part, *sc, *gnorm, *losses;
  uint32_t nloss_cap, nloss;
  unsigned short *wb;
  float4 *ev;
  float *X, *Z0, *H[MP_NR + 1], *P[MP_NR], *Q[MP_NR], *M[MP_NR], *ZU[MP_NR];
  float *OP1, *OPQ;          /* operator coefficients of the prepared slot */
  float *dh, *dhn, *dm, *dP, *dQ;
  float *dOP1, *dOPQ, *gd, *Ez;
  float *dstore, *isd;       /* per slot: FP32 [N][16] diagonal blocks, 1 / their RMS */
  double *red, *tred, *lpart;
  float *dl, *ebuf, *ybuf, *gbuf; /* loss: correction, error, A error, gradient (N x 4) */
  float **mslot;             /* per slot: the assembled Jacobian (FP32 BSR) or NULL */
  uint32_t *tpos;            /* BSR position of the transposed block */
  int t;                     /* Adam steps so far */
  int metric;                /* loss: 0 |e|^2/|target|^2, 1 |A e|^2/|v|^2,
                                2 |A (base + G base) - v|^2/|v|^2 */
  int prepared;              /* slot whose operator OP1/OPQ hold, -1: none */
};

template <typename T>
static T *mp_alloc(size_t n)
{
  void *p = NULL;
  KF_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
  return (T *)p;
}

static void mp_check_slot(const kf_mpnn *net, int slot)
{
  if (slot < 0 || (uint32_t)slot >= net->nslot) {
    fprintf(stderr, "kf: network slot %d out of range [0, %u)\n", slot, net->nslot);
    abort();
  }
}

static void mp_repack(kf_mpnn *net)
{
  hipLaunchKernelGGL(k_mp_pack, dim3(kf_blocks_for(MPB_SIZE, 256)), dim3(256), 0, net->c->stream, net->theta,
                     net->wb);
  KF_LAUNCH_CHECK();
  net->prepared = -1;
}

/* the context network on slot `slot`'s diagonal blocks and the operator
 * coefficients OP1, OPQ; every activation stays in the object for backward */
static void mp_context(kf_mpnn *net, int slot)
{
  kf_ctx *c = net->c;
  hipStream_t st = c->stream;
  const uint32_t N = net->N, nw = kf_blocks_for(N, 32);
  const float *nf = NULL;
  const uint32_t *nu = NULL;
  float *nz = NULL;
  hipLaunchKernelGGL(k_mp_features, dim3(kf_blocks_for(16 * (size_t)N, MP_RB)), dim3(MP_RB), 0, st,
                     16 * (size_t)N, net->dstore + (size_t)slot * 16 * N, net->isd + slot, net->X);
  hipLaunchKernelGGL((k_dense32<MP_NIN, 0, 1, false>), dim3(nw), dim3(64), 0, st, N, net->X, nf, nf,
                     net->wb + MPB_W0, net->theta + MP_B0, nf, nu, net->Z0, net->H[0]);
  for (int l = 0; l < MP_NR; l++) {
    hipLaunchKernelGGL((k_dense32<MP_H, 0, 0, false>), dim3(nw), dim3(64), 0, st, N, net->H[l], nf, nf,
                       net->wb + MPB_A(l), nf, nf, nu, nz, net->P[l]);
    hipLaunchKernelGGL((k_dense32<MP_H, 0, 0, false>), dim3(nw), dim3(64), 0, st, N, net->H[l], nf, nf,
                       net->wb + MPB_B(l), nf, nf, nu, nz, net->Q[l]);
    hipLaunchKernelGGL(k_mp_message, dim3(kf_blocks_for(N, 2)), dim3(64), 0, st, c->d, net->ev, net->P[l], net->Q[l],
                       net->theta, (uint32_t)MP_C(l), (uint32_t)MP_CB(l), net->M[l]);
    hipLaunchKernelGGL((k_dense32<MP_H, MP_H, 2, false>), dim3(nw), dim3(64), 0, st, N, net->H[l], net->M[l], nf,
                       net->wb + MPB_UV(l), net->theta + MP_UB(l), net->H[l], nu, net->ZU[l], net->H[l + 1]);
  }
  hipLaunchKernelGGL((k_dense32<MP_H, 0, 0, false>), dim3(nw), dim3(64), 0, st, N, net->H[MP_NR], nf, nf,
                     net->wb + MPB_WSA, net->theta + MP_BSA, nf, nu, nz, net->OP1);
  hipLaunchKernelGGL((k_dense32<MP_H, 0, 0, false>), dim3(nw), dim3(64), 0, st, N, net->H[MP_NR], nf, nf,
                     net->wb + MPB_WB, nf, nf, nu, nz, net->OPQ);
  KF_LAUNCH_CHECK();
  net->prepared = slot;
}

/* loss of (v, base, target) and the parameter gradient into net->grad; the
 * loss lands in *loss_dev */
static void mp_loss_grad(kf_mpnn *net, int slot, const double *v, const double *base, const double *target,
                         float *loss_dev)
{
  kf_ctx *c = net->c;
  hipStream_t st = c->stream;
  const uint32_t N = net->N, nw = kf_blocks_for(N, 32), P = net->nparts, ch = net->chunk, nb = net->nlblk;
  const size_t n4 = 4 * (size_t)N;
  const uint32_t nb4 = kf_blocks_for(n4, MP_RB);
  const float *nf = NULL;
  const uint32_t *nu = NULL;
  float *nz = NULL;
  const int metric = net->metric;
  const float *A = net->mslot[slot];
  if (metric >= 1 && !A) {
    fprintf(stderr, "kf: the A-norm loss needs the Jacobian of slot %d (kf_mpnn_take_matrix)\n", slot);
    abort();
  }
  mp_context(net, slot);
  hipLaunchKernelGGL(k_mp_apply, dim3(nb), dim3(MP_RB), 0, st, c->d, net->ev, net->OP1, net->OPQ, net->theta, base,
                     (double *)NULL, 0, net->dl);
  hipLaunchKernelGGL(k_mp_err, dim3(nb4), dim3(MP_RB), 0, st, n4, net->dl, base, metric == 2 ? NULL : target,
                     net->ebuf);
  hipLaunchKernelGGL(k_mp_sumsq<double>, dim3(KF_NRED), dim3(MP_RB), 0, st, n4, metric ? v : target, net->tred);
  if (metric >= 1)
    hipLaunchKernelGGL(k_mp_bsr_mv, dim3(nb), dim3(MP_RB), 0, st, N, c->d.rowptr, c->d.col, A, net->ebuf, net->ybuf);
  if (metric == 2) hipLaunchKernelGGL(k_mp_sub, dim3(nb4), dim3(MP_RB), 0, st, n4, v, net->ybuf);
  hipLaunchKernelGGL(k_mp_sq4, dim3(nb), dim3(MP_RB), 0, st, N, metric ? net->ybuf : net->ebuf, net->lpart);
  hipLaunchKernelGGL(k_mp_loss_finish, dim3(1), dim3(64), 0, st, net->lpart, nb, net->tred, loss_dev, net->sc);
  if (metric >= 1)
    hipLaunchKernelGGL(k_mp_bsr_mvt, dim3(nb), dim3(MP_RB), 0, st, N, c->d.rowptr, c->d.col, net->tpos, A, net->ybuf,
                       net->ebuf);
  hipLaunchKernelGGL(k_mp_scale, dim3(nb4), dim3(MP_RB), 0, st, n4, net->ebuf, net->sc, net->gbuf);
  /* the operator heads */
  hipLaunchKernelGGL(k_mp_apply_bwd, dim3(nb), dim3(MP_RB), 0, st, c->d, net->ev, base, net->gbuf, net->dOP1,
                     net->dOPQ, net->gd, net->Ez);
  hipLaunchKernelGGL((k_wgrad<MP_H, MP_H, 32, MP_H, false>), dim3(P), dim3(64), 0, st, N, ch, net->dOP1, nf,
                     net->H[MP_NR], net->part, (uint32_t)MP_WSA, (int)MP_BSA);
  hipLaunchKernelGGL((k_wgrad<MP_H, MP_H, 16, MP_H, false>), dim3(P), dim3(64), 0, st, N, ch, net->dOPQ, nf,
                     net->H[MP_NR], net->part, (uint32_t)MP_WB, -1);
  hipLaunchKernelGGL((k_wgrad<4, 12, 4, 12, false>), dim3(P), dim3(64), 0, st, N, ch, net->gd, nf, net->Ez,
                     net->part, (uint32_t)MP_WE, -1);
  hipLaunchKernelGGL((k_dense32<MP_H, MP_H, 0, false>), dim3(nw), dim3(64), 0, st, N, net->dOP1, net->dOPQ, nf,
                     net->wb + MPB_HT, nf, nf, nu, nz, net->dh);
  for (int l = MP_NR - 1; l >= 0; l--) {
    /* dzu = dh [zu > 0]: gradients of U, u and V */
    hipLaunchKernelGGL((k_wgrad<MP_H, MP_H, MP_H, MP_H, true>), dim3(P), dim3(64), 0, st, N, ch, net->dh, net->ZU[l],
                       net->H[l], net->part, (uint32_t)MP_U(l), (int)MP_UB(l));
    hipLaunchKernelGGL((k_wgrad<MP_H, MP_H, MP_H, MP_H, true>), dim3(P), dim3(64), 0, st, N, ch, net->dh, net->ZU[l],
                       net->M[l], net->part, (uint32_t)MP_V(l), -1);
    /* dhn = dh + dzu U;  dm = (dzu V) / deg */
    hipLaunchKernelGGL((k_dense32<MP_H, 0, 3, true>), dim3(nw), dim3(64), 0, st, N, net->dh, nf, net->ZU[l],
                       net->wb + MPB_UT(l), nf, net->dh, nu, nz, net->dhn);
    hipLaunchKernelGGL((k_dense32<MP_H, 0, 4, true>), dim3(nw), dim3(64), 0, st, N, net->dh, nf, net->ZU[l],
                       net->wb + MPB_VT(l), nf, nf, c->d.deg, nz, net->dm);
    hipLaunchKernelGGL(k_mp_message_bwd, dim3(P), dim3(MP_RB), 0, st, c->d, ch, net->ev, net->P[l], net->Q[l],
                       net->dm, net->theta, (uint32_t)MP_C(l), (uint32_t)MP_CB(l), net->dP, net->dQ, net->part);
    hipLaunchKernelGGL((k_wgrad<MP_H, MP_H, MP_H, MP_H, false>), dim3(P), dim3(64), 0, st, N, ch, net->dP, nf,
                       net->H[l], net->part, (uint32_t)MP_A(l), -1);
    hipLaunchKernelGGL((k_wgrad<MP_H, MP_H, MP_H, MP_H, false>), dim3(P), dim3(64), 0, st, N, ch, net->dQ, nf,
                       net->H[l], net->part, (uint32_t)MP_B(l), -1);
    /* dh = dhn + dP A + dQ B */
    hipLaunchKernelGGL((k_dense32<MP_H, MP_H, 3, false>), dim3(nw), dim3(64), 0, st, N, net->dP, net->dQ, nf,
                       net->wb + MPB_ABT(l), nf, net->dhn, nu, nz, net->dh);
  }
  hipLaunchKernelGGL((k_wgrad<MP_H, MP_NIN, MP_H, MP_NIN, true>), dim3(P), dim3(64), 0, st, N, ch, net->dh, net->Z0,
                     net->X, net->part, (uint32_t)MP_W0, (int)MP_B0);
  hipLaunchKernelGGL(k_mp_grad_reduce, dim3(kf_blocks_for(MP_NPARAM, MP_RB)), dim3(MP_RB), 0, st, net->part, P,
                     net->grad);
  KF_LAUNCH_CHECK();
}

/* Initial parameters: normals scaled by sqrt(2 / fan-in) for W0, A, B, C,
 * U and V (U and V halved), zero biases, zero operator heads (G = 0, so
 * training starts from ILU(0) itself).  Normals from splitmix64 and the
 * Box-Muller cosine branch. */
static uint64_t mp_mix(uint64_t *s)
{
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double mp_normal(uint64_t *s)
{
  const double u1 = ((double)(mp_mix(s) >> 11) + 1.0) * 0x1.0p-53, u2 = (double)(mp_mix(s) >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

extern "C" void kf_mpnn_init_params(float *theta, uint64_t seed)
{
  uint64_t st = seed * 0xD1B54A32D192ED03ull + 1;
  memset(theta, 0, sizeof(float) * MP_NPARAM);
  auto fill = [&](int off, int n, int fan_in, double scale) {
    const double sd = sqrt(2.0 / fan_in) * scale;
    for (int i = 0; i < n; i++) theta[off + i] = (float)(mp_normal(&st) * sd);
  };
  fill(MP_W0, MP_H * MP_NIN, MP_NIN, 1.0);
  for (int l = 0; l < MP_NR; l++) {
    fill(MP_A(l), MP_H * MP_H, MP_H, 1.0);
    fill(MP_B(l), MP_H * MP_H, MP_H, 1.0);
    fill(MP_C(l), MP_H * 3, 3, 1.0);
    fill(MP_U(l), MP_H * MP_H, MP_H, 0.5);
    fill(MP_V(l), MP_H * MP_H, MP_H, 0.5);
  }
}

extern "C" size_t kf_mpnn_nparam(void) { return MP_NPARAM; }
extern "C" size_t kf_mpnn_nhead(void) { return MP_HEADS; }

extern "C" void kf_mpnn_set_params(kf_mpnn *net, const float *theta_host);

extern "C" kf_mpnn *kf_mpnn_create(GEOMETRY *g, const float *theta_host, int nslot)
{
  kf_ctx *c = kf_ctx_of(g);
  if (!theta_host || nslot < 1) {
    fprintf(stderr, "kf: kf_mpnn_create needs initial parameters and at least one slot\n");
    abort();
  }
  kf_mpnn *net = new kf_mpnn();
  memset((void *)net, 0, sizeof(*net));
  net->c = c;
  net->metric = 2;
  net->prepared = -1;
  const uint32_t N = c->d.N;
  net->N = N;
  net->nslot = (uint32_t)nslot;
  uint32_t ch = (N + MP_PARTS - 1) / MP_PARTS;
  ch = ((ch + 7u) / 8u) * 8u;
  if (ch < 8) ch = 8;
  net->chunk = ch;
  net->nparts = (N + ch - 1) / ch;
  net->nlblk = kf_blocks_for(N, MP_RB);
  net->theta = mp_alloc<float>(MP_NPARAM);
  net->am = mp_alloc<float>(MP_NPARAM);
  net->av = mp_alloc<float>(MP_NPARAM);
  net->grad = mp_alloc<float>(MP_NPARAM);
  net->part = mp_alloc<float>((size_t)net->nparts * MP_NPARAM);
  net->sc = mp_alloc<float>(8);
  net->gnorm = mp_alloc<float>(1);
  net->nloss_cap = 1u << 16;
  net->losses = mp_alloc<float>(net->nloss_cap);
  net->wb = mp_alloc<unsigned short>(MPB_SIZE);
  /* sliced-ELL slots: the last slice pointer counts them (in units of 64) */
  uint32_t nsl_ent = 0;
  KF_CHECK(hipMemcpy(&nsl_ent, c->d.sell_ptr + c->d.nslice, sizeof(uint32_t), hipMemcpyDeviceToHost));
  net->ev = mp_alloc<float4>((size_t)nsl_ent * 64u);
  const size_t NH = (size_t)N * MP_H;
  net->X = mp_alloc<float>((size_t)N * MP_NIN);
  net->Z0 = mp_alloc<float>(NH);
  for (int l = 0; l <= MP_NR; l++) net->H[l] = mp_alloc<float>(NH);
  for (int l = 0; l < MP_NR; l++) {
    net->P[l] = mp_alloc<float>(NH);
    net->Q[l] = mp_alloc<float>(NH);
    net->M[l] = mp_alloc<float>(NH);
    net->ZU[l] = mp_alloc<float>(NH);
  }
  net->OP1 = mp_alloc<float>(NH);
  net->OPQ = mp_alloc<float>(NH);
  net->dh = mp_alloc<float>(NH);
  net->dhn = mp_alloc<float>(NH);
  net->dm = mp_alloc<float>(NH);
  net->dP = mp_alloc<float>(NH);
  net->dQ = mp_alloc<float>(NH);
  net->dOP1 = mp_alloc<float>(NH);
  net->dOPQ = mp_alloc<float>(NH);
  net->gd = mp_alloc<float>(4 * (size_t)N);
  net->Ez = mp_alloc<float>(12 * (size_t)N);
  net->dstore = mp_alloc<float>((size_t)nslot * 16 * N);
  net->isd = mp_alloc<float>(nslot);
  net->red = mp_alloc<double>(KF_NRED);
  net->tred = mp_alloc<double>(KF_NRED);
  net->lpart = mp_alloc<double>(net->nlblk);
  net->dl = mp_alloc<float>(4 * (size_t)N);
  net->ebuf = mp_alloc<float>(4 * (size_t)N);
  net->ybuf = mp_alloc<float>(4 * (size_t)N);
  net->gbuf = mp_alloc<float>(4 * (size_t)N);
  net->mslot = (float **)calloc((size_t)nslot, sizeof(float *));
  if (!net->mslot) abort();
  {
    /* transposed-block positions: (i, j) at k  ->  (j, i) */
    const size_t nnzb = c->nnzb;
    uint32_t *rp = (uint32_t *)malloc(sizeof(uint32_t) * (N + 1)), *cl = (uint32_t *)malloc(sizeof(uint32_t) * nnzb);
    uint32_t *tp = (uint32_t *)malloc(sizeof(uint32_t) * (nnzb ? nnzb : 1));
    if (!rp || !cl || !tp) abort();
    KF_CHECK(hipMemcpy(rp, c->d.rowptr, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost));
    KF_CHECK(hipMemcpy(cl, c->d.col, sizeof(uint32_t) * nnzb, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < N; i++)
      for (uint32_t k = rp[i]; k < rp[i + 1]; k++) {
        const uint32_t j = cl[k];
        uint32_t lo = rp[j], hi = rp[j + 1];
        while (lo < hi) { /* columns of a row are sorted */
          const uint32_t mid = (lo + hi) / 2;
          if (cl[mid] < i) lo = mid + 1;
          else hi = mid;
        }
        if (lo == rp[j + 1] || cl[lo] != i) {
          fprintf(stderr, "kf: block pattern is not structurally symmetric\n");
          abort();
        }
        tp[k] = lo;
      }
    net->tpos = mp_alloc<uint32_t>(nnzb ? nnzb : 1);
    KF_CHECK(hipMemcpy(net->tpos, tp, sizeof(uint32_t) * nnzb, hipMemcpyHostToDevice));
    free(rp);
    free(cl);
    free(tp);
  }
  KF_CHECK(hipMemsetAsync(net->dstore, 0, sizeof(float) * (size_t)nslot * 16 * N, c->stream));
  float *ones = (float *)malloc(sizeof(float) * nslot);
  for (int i = 0; i < nslot; i++) ones[i] = 1.0f;
  KF_CHECK(hipMemcpyAsync(net->isd, ones, sizeof(float) * nslot, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_mp_edge_vectors, dim3(kf_blocks_for(N, MP_RB)), dim3(MP_RB), 0, c->stream, c->d, net->ev);
  KF_LAUNCH_CHECK();
  KF_CHECK(hipStreamSynchronize(c->stream));
  free(ones);
  kf_mpnn_set_params(net, theta_host);
  return net;
}

extern "C" void kf_mpnn_set_params(kf_mpnn *net, const float *theta_host)
{
  hipStream_t st = net->c->stream;
  KF_CHECK(hipMemcpyAsync(net->theta, theta_host, sizeof(float) * MP_NPARAM, hipMemcpyHostToDevice, st));
  KF_CHECK(hipMemsetAsync(net->am, 0, sizeof(float) * MP_NPARAM, st));
  KF_CHECK(hipMemsetAsync(net->av, 0, sizeof(float) * MP_NPARAM, st));
  net->t = 0;
  mp_repack(net);
  KF_CHECK(hipStreamSynchronize(st));
}

extern "C" void kf_mpnn_get_params(const kf_mpnn *net, float *theta_host)
{
  KF_CHECK(hipStreamSynchronize(net->c->stream));
  KF_CHECK(hipMemcpy(theta_host, net->theta, sizeof(float) * MP_NPARAM, hipMemcpyDeviceToHost));
}

extern "C" void kf_mpnn_destroy(kf_mpnn *net)
{
  if (!net) return;
  KF_CHECK(hipStreamSynchronize(net->c->stream));
  void *bufs[] = {net->theta, net->am,   net->av,   net->grad,  net->part, net->sc,   net->gnorm, net->losses,
                  net->wb,    net->ev,   net->X,    net->Z0,    net->OP1,  net->OPQ,  net->dh,    net->dhn,
                  net->dm,    net->dP,   net->dQ,   net->dOP1,  net->dOPQ, net->gd,   net->Ez,    net->dstore,
                  net->isd,   net->red,  net->tred, net->lpart, net->dl,   net->ebuf, net->ybuf,  net->gbuf,
                  net->tpos};
  for (void *p : bufs)
    if (p) KF_CHECK(hipFree(p));
  for (uint32_t k = 0; k < net->nslot; k++)
    if (net->mslot[k]) KF_CHECK(hipFree(net->mslot[k]));
  free(net->mslot);
  for (int l = 0; l <= MP_NR; l++) KF_CHECK(hipFree(net->H[l]));
  for (int l = 0; l < MP_NR; l++) {
    KF_CHECK(hipFree(net->P[l]));
    KF_CHECK(hipFree(net->Q[l]));
    KF_CHECK(hipFree(net->M[l]));
    KF_CHECK(hipFree(net->ZU[l]));
  }
  delete net;
}

/* the current ILU factor's diagonal blocks (and 1 / their RMS) into `slot` */
extern "C" void kf_mpnn_take_diag(kf_mpnn *net, int slot)
{
  mp_check_slot(net, slot);
  kf_ctx *c = net->c;
  hipStream_t st = c->stream;
  const uint32_t N = net->N;
  float *D = net->dstore + (size_t)slot * 16 * N;
  const uint32_t nb = kf_blocks_for(16 * (size_t)N, MP_RB);
  if (c->opt.prec == KF_PREC_FP64)
    hipLaunchKernelGGL(k_mp_take_diag<double>, dim3(nb), dim3(MP_RB), 0, st, N, c->d.diag, (const double *)c->a64, D);
  else
    hipLaunchKernelGGL(k_mp_take_diag<float>, dim3(nb), dim3(MP_RB), 0, st, N, c->d.diag, (const float *)c->a32, D);
  hipLaunchKernelGGL(k_mp_sumsq<float>, dim3(KF_NRED), dim3(MP_RB), 0, st, 16 * (size_t)N, (const float *)D,
                     net->red);
  hipLaunchKernelGGL(k_mp_finish, dim3(1), dim3(64), 0, st, net->red, 16.0 * N, net->isd + slot);
  KF_LAUNCH_CHECK();
  if (net->prepared == slot) net->prepared = -1;
}

/* the assembled Jacobian (after ComputeA, before ComputeNumericalILU
 * factors it in place) into `slot`, as FP32 */
extern "C" void kf_mpnn_take_matrix(kf_mpnn *net, int slot)
{
  mp_check_slot(net, slot);
  kf_ctx *c = net->c;
  const size_t n = 16 * c->nnzb;
  if (!net->mslot[slot]) net->mslot[slot] = mp_alloc<float>(n);
  if (c->opt.prec == KF_PREC_FP64)
    hipLaunchKernelGGL(k_mp_to_f32<double>, dim3(kf_blocks_for(n, MP_RB)), dim3(MP_RB), 0, c->stream, n,
                       (const double *)c->a64, net->mslot[slot]);
  else
    KF_CHECK(hipMemcpyAsync(net->mslot[slot], c->a32, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
  KF_LAUNCH_CHECK();
}

/* 0: |e|^2 / |target|^2;  1: |A e|^2 / |v|^2 with the slot's A;
 * 2 (default): |A (base + G base) - v|^2 / |v|^2 (the target is not read) */
extern "C" void kf_mpnn_set_loss(kf_mpnn *net, int metric)
{
  if (metric < 0 || metric > 2) {
    fprintf(stderr, "kf: kf_mpnn_set_loss: no loss %d\n", metric);
    abort();
  }
  net->metric = metric;
}

/* diagonal blocks given by the caller (host FP64 [N][16]) into `slot` */
extern "C" void kf_mpnn_set_diag(kf_mpnn *net, int slot, const double *D_host)
{
  mp_check_slot(net, slot);
  kf_ctx *c = net->c;
  const size_t n = 16 * (size_t)net->N;
  float *tmp = (float *)malloc(sizeof(float) * n);
  for (size_t i = 0; i < n; i++) tmp[i] = (float)D_host[i];
  float *D = net->dstore + (size_t)slot * n;
  KF_CHECK(hipMemcpyAsync(D, tmp, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_mp_sumsq<float>, dim3(KF_NRED), dim3(MP_RB), 0, c->stream, n, (const float *)D, net->red);
  hipLaunchKernelGGL(k_mp_finish, dim3(1), dim3(64), 0, c->stream, net->red, (double)n, net->isd + slot);
  KF_LAUNCH_CHECK();
  KF_CHECK(hipStreamSynchronize(c->stream));
  free(tmp);
  if (net->prepared == slot) net->prepared = -1;
}

/* the correction operator of `slot` (the context network on its diagonal
 * blocks); kf_mpnn_forward prepares on demand */
extern "C" void kf_mpnn_prepare(kf_mpnn *net, int slot)
{
  mp_check_slot(net, slot);
  mp_context(net, slot);
}

/* out = G z (add = 0) or out += G z (add = 1), z = ILU^-1 v; device vectors
 * of 4N */
extern "C" void kf_mpnn_forward(kf_mpnn *net, int slot, const double *z, double *out, int add)
{
  mp_check_slot(net, slot);
  if (net->prepared != slot) mp_context(net, slot);
  hipLaunchKernelGGL(k_mp_apply, dim3(net->nlblk), dim3(MP_RB), 0, net->c->stream, net->c->d, net->ev, net->OP1,
                     net->OPQ, net->theta, z, out, add, (float *)NULL);
  KF_LAUNCH_CHECK();
}

/* loss of (v, base, target) and its gradient (host FP32 copy when grad_host
 * is given); device vectors of 4N.  Synchronises. */
extern "C" double kf_mpnn_loss_grad(kf_mpnn *net, int slot, const double *v, const double *base, const double *target,
                                    float *grad_host)
{
  mp_check_slot(net, slot);
  mp_loss_grad(net, slot, v, base, target, net->sc + 4);
  float loss;
  KF_CHECK(hipMemcpyAsync(&loss, net->sc + 4, sizeof(float), hipMemcpyDeviceToHost, net->c->stream));
  if (grad_host)
    KF_CHECK(hipMemcpyAsync(grad_host, net->grad, sizeof(float) * MP_NPARAM, hipMemcpyDeviceToHost, net->c->stream));
  KF_CHECK(hipStreamSynchronize(net->c->stream));
  return (double)loss;
}

static void mp_adam(kf_mpnn *net, float lr, float b1, float b2, float eps, float gclip)
{
  net->t++;
  const float c1 = (float)(1.0 / (1.0 - pow((double)b1, net->t)));
  const float c2 = (float)(1.0 / (1.0 - pow((double)b2, net->t)));
  hipLaunchKernelGGL(k_mp_adam, dim3(1), dim3(1024), 0, net->c->stream, net->theta, net->am, net->av, net->grad, lr,
                     b1, b2, eps, c1, c2, gclip, net->gnorm);
  KF_LAUNCH_CHECK();
  mp_repack(net);
}

/* one Adam step with the gradient of the last kf_mpnn_loss_grad */
extern "C" void kf_mpnn_adam(kf_mpnn *net, float lr, float beta1, float beta2, float eps, float gclip)
{
  mp_adam(net, lr, beta1, beta2, eps, gclip);
  KF_CHECK(hipStreamSynchronize(net->c->stream));
}

/* loss + gradient + Adam (0.9, 0.999, 1e-8) on one pair, queued on the
 * stream with no host round trip; the loss goes to the loss log */
extern "C" void kf_mpnn_train_step(kf_mpnn *net, int slot, const double *v, const double *base, const double *target,
                                   float lr, float gclip)
{
  mp_check_slot(net, slot);
  float *ld = net->losses + (net->nloss % net->nloss_cap);
  net->nloss++;
  mp_loss_grad(net, slot, v, base, target, ld);
  mp_adam(net, lr, 0.9f, 0.999f, 1e-8f, gclip);
}

/* the loss log: returns the steps logged since the last reset and copies the
 * first min(cap, 65536) of them */
extern "C" int kf_mpnn_losses(kf_mpnn *net, float *host, int cap)
{
  KF_CHECK(hipStreamSynchronize(net->c->stream));
  const uint32_t n = net->nloss < net->nloss_cap ? net->nloss : net->nloss_cap;
  const uint32_t k = (cap > 0 && (uint32_t)cap < n) ? (uint32_t)cap : n;
  if (k && host) KF_CHECK(hipMemcpy(host, net->losses, sizeof(float) * k, hipMemcpyDeviceToHost));
  return (int)net->nloss;
}

extern "C" void kf_mpnn_reset_losses(kf_mpnn *net) { net->nloss = 0; }

/* activations of the last context pass, for tests: 0 x (N x 16), 1 z0, then
 * per round l: 2+5l h_l, 3+5l P, 4+5l Q, 5+5l m, 6+5l zu; 17 the final h,
 * 18 OP1, 19 OPQ (N x 32 each) */
extern "C" void kf_mpnn_activation(kf_mpnn *net, int which, float *host)
{
  KF_CHECK(hipStreamSynchronize(net->c->stream));
  const size_t NH = (size_t)net->N * MP_H;
  const float *src = NULL;
  size_t n = NH;
  if (which == 0) {
    src = net->X;
    n = (size_t)net->N * MP_NIN;
  } else if (which == 1) src = net->Z0;
  else if (which == 17) src = net->H[MP_NR];
  else if (which == 18) src = net->OP1;
  else if (which == 19) src = net->OPQ;
  else if (which >= 2 && which < 17) {
    const int l = (which - 2) / 5, k = (which - 2) % 5;
    const float *tab[5] = {net->H[l], net->P[l], net->Q[l], net->M[l], net->ZU[l]};
    src = tab[k];
  } else {
    fprintf(stderr, "kf: kf_mpnn_activation: no activation %d\n", which);
    abort();
  }
  KF_CHECK(hipMemcpy(host, src, sizeof(float) * n, hipMemcpyDeviceToHost));
}
