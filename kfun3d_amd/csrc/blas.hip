/*
 * blas.hip -- BLAS-1 of src/utils.c:53-123 and the FD Jacobian-vector
 * combination of src/kernel.c:24-45 on gfx950.
 *
 * Reductions are deterministic: a fixed grid of KF_NRED blocks walks the
 * vector in element pairs (2p, 2p + 1) with a fixed stride of pairs, each
 * thread summing its pairs' products in order (so a lane moves 16 bytes per
 * load), each block reduces through LDS in a fixed tree, and one block
 * combines the KF_NRED partials in a fixed tree.  The
 * sum order differs from the reference's one-thread loop (it is a tree), the
 * result is the same on every run.  Scalars stay on the device (d.scal) so
 * that dependent vector updates need no host round trip.
 */
#include "kf_dev.h"
#include "sync.h"
#include "handoff.h"

#define RB 256

/* The fixed tree over the block's RB values: at each halving s, value t
 * becomes v[t] + v[t + s].  The levels s >= 64 pair threads of different
 * waves and go through LDS; the levels s <= 32 pair lanes of wave 0 and go
 * through lane shifts, with the same operands in the same order. */
static_assert(RB == 256, "block_sum's LDS levels are s = 128, 64");
__device__ __forceinline__ double block_sum(double v, double *sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  if (t < 128) sh[t] = sh[t] + sh[t + 128];
  __syncthreads();
  if (t < 64) {
    double x = sh[t] + sh[t + 64];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) x = x + __shfl_down(x, s, 64);
    if (t == 0) sh[0] = x;
  }
  __syncthreads();
  const double r = sh[0];
  __syncthreads();   /* every thread has read sh[0] before sh is reused */
  return r;
}

/* elements i, i + 1 of x: one 16-byte load when the vector is 16-byte
 * aligned (VEC), two 8-byte loads otherwise; the same values either way */
template <bool VEC>
__device__ __forceinline__ void ld_pair(const double *x, size_t i, double &a, double &b)
{
  if (VEC) {
    const double2 t = *reinterpret_cast<const double2 *>(x + i);
    a = t.x;
    b = t.y;
  } else {
    a = x[i];
    b = x[i + 1];
  }
}

static inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

template <bool VEC>
__global__ __launch_bounds__(RB) void k_dot_part(size_t n, const double *__restrict__ x, const double *__restrict__ y,
                                                 double *__restrict__ part)
{
  __shared__ double sh[RB];
  double acc = 0.0;
  const size_t stride = (size_t)gridDim.x * RB, np = n / 2;
  size_t p = (size_t)blockIdx.x * RB + threadIdx.x;
  for (; p < np; p += stride) {
    double x0, x1, y0, y1;
    ld_pair<VEC>(x, 2 * p, x0, x1);
    ld_pair<VEC>(y, 2 * p, y0, y1);
    acc += x0 * y0;
    acc += x1 * y1;
  }
  if (p == np && (n & 1)) acc += x[n - 1] * y[n - 1];   /* the half pair of an odd length */
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

static void launch_dot_part(kf_ctx *c, size_t n, const double *x, const double *y, double *part)
{
  if (aligned16(x) && aligned16(y))
    hipLaunchKernelGGL(k_dot_part<true>, dim3(KF_NRED), dim3(RB), 0, c->stream, n, x, y, part);
  else
    hipLaunchKernelGGL(k_dot_part<false>, dim3(KF_NRED), dim3(RB), 0, c->stream, n, x, y, part);
  KF_LAUNCH_CHECK();
}

__global__ __launch_bounds__(RB) void k_red_final(const double *__restrict__ part, double *__restrict__ out, int take_sqrt)
{
  __shared__ double sh[RB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < KF_NRED; i += RB) acc += part[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = take_sqrt ? sqrt(s) : s;
}

/* Index-order sum, for parity runs: the reference's one-thread loops add
 * x[0] y[0], x[1] y[1], ... in that order (src/utils.c:61-62,
 * src/kernel.c:133-134) into an OpenMP reduction copy that starts at zero and
 * is then added to 0.  One wavefront forms 64 products at a time and adds
 * them lane by lane; every lane carries the same running sum. */
__device__ __forceinline__ double lane_value(double v, int k)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__global__ __launch_bounds__(64) void k_dot_seq(size_t n, const double *__restrict__ x, const double *__restrict__ y,
                                                double *__restrict__ out, int take_sqrt)
{
  const int lane = threadIdx.x;
  double acc = 0.0;
  for (size_t base = 0; base < n; base += 64) {
    const size_t i = base + lane;
    const double p = i < n ? x[i] * y[i] : 0.0;
    const int cnt = (int)(n - base < 64 ? n - base : 64);
    for (int k = 0; k < cnt; k++) acc = acc + lane_value(p, k);
  }
  acc = 0.0 + acc;
  if (lane == 0) *out = take_sqrt ? sqrt(acc) : acc;
}

static void dot_any(kf_ctx *c, size_t n, const double *x, const double *y, int slot, int take_sqrt)
{
  if (c->opt.reduce_seq) {
    hipLaunchKernelGGL(k_dot_seq, dim3(1), dim3(64), 0, c->stream, n, x, y, c->d.scal + slot, take_sqrt);
    KF_LAUNCH_CHECK();
    return;
  }
  launch_dot_part(c, n, x, y, c->d.red);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(RB), 0, c->stream, c->d.red, c->d.scal + slot, take_sqrt);
  KF_LAUNCH_CHECK();
}

void kf_dot_dev(kf_ctx *c, size_t n, const double *x, const double *y, int slot) { dot_any(c, n, x, y, slot, 0); }

void kf_nrm2_dev(kf_ctx *c, size_t n, const double *x, int slot) { dot_any(c, n, x, x, slot, 1); }

/* One modified Gram-Schmidt step fused with its neighbours (src/kernel.c:
 * 129-141): h = <w, v> from the KF_NRED partials of the previous pass,
 * reduced in k_red_final's tree (every block forms the same value; block 0
 * stores it), then w = (-h) v + w, and the partials of <w, u> for the next
 * step (u = the next basis vector, or w itself for the closing norm).  The
 * grid and the element order of each thread are k_dot_part's, so the
 * partials are bit-identical to an unfused dot after an unfused update. */
template <bool VEC>
__global__ __launch_bounds__(RB) void k_mgs_step(size_t n, const double *__restrict__ part_in, double *hslot,
                                                 const double *v, double *w, const double *u,
                                                 double *__restrict__ part_out)
{
  __shared__ double sh[RB];
  const bool self = u == w;
  const size_t stride = (size_t)gridDim.x * RB, np = n / 2;
  size_t p = (size_t)blockIdx.x * RB + threadIdx.x;
  /* pairs in chunks of two: a chunk's operands are loaded ahead of its
   * updates (u is either w itself or a vector that overlaps neither v nor
   * w), and the first chunk's loads are in flight while the previous dot's
   * partials are summed; updates and accumulation keep the element order */
  double vv[4], ww[4], uu[4];
  bool chunk = p + stride < np;
  if (chunk) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const size_t i = 2 * (p + k * stride);
      ld_pair<VEC>(v, i, vv[2 * k], vv[2 * k + 1]);
      ld_pair<VEC>(w, i, ww[2 * k], ww[2 * k + 1]);
      if (!self) ld_pair<VEC>(u, i, uu[2 * k], uu[2 * k + 1]);
    }
  }
  double acc = 0.0;
  for (int j = threadIdx.x; j < KF_NRED; j += RB) acc += part_in[j];
  const double h = block_sum(acc, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) *hslot = h;
  const double a = -h;
  acc = 0.0;
  while (chunk) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const size_t i = 2 * (p + k * stride);
      const double w0 = a * vv[2 * k] + ww[2 * k];   /* contraction off: the product, then the sum */
      const double w1 = a * vv[2 * k + 1] + ww[2 * k + 1];
      if (VEC) {
        *reinterpret_cast<double2 *>(w + i) = make_double2(w0, w1);
      } else {
        w[i] = w0;
        w[i + 1] = w1;
      }
      acc += w0 * (self ? w0 : uu[2 * k]);
      acc += w1 * (self ? w1 : uu[2 * k + 1]);
    }
    p += 2 * stride;
    chunk = p + stride < np;
    if (chunk) {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        const size_t i = 2 * (p + k * stride);
        ld_pair<VEC>(v, i, vv[2 * k], vv[2 * k + 1]);
        ld_pair<VEC>(w, i, ww[2 * k], ww[2 * k + 1]);
        if (!self) ld_pair<VEC>(u, i, uu[2 * k], uu[2 * k + 1]);
      }
    }
  }
  for (; p < np; p += stride) {
    const size_t i = 2 * p;
    double v0, v1, x0, x1;
    ld_pair<VEC>(v, i, v0, v1);
    ld_pair<VEC>(w, i, x0, x1);
    const double w0 = a * v0 + x0, w1 = a * v1 + x1;
    w[i] = w0;
    w[i + 1] = w1;
    double u0 = w0, u1 = w1;
    if (!self) ld_pair<VEC>(u, i, u0, u1);
    acc += w0 * u0;
    acc += w1 * u1;
  }
  if (p == np && (n & 1)) {   /* the half pair of an odd length */
    const double wn = a * v[n - 1] + w[n - 1];
    w[n - 1] = wn;
    acc += wn * (self ? wn : u[n - 1]);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part_out[blockIdx.x] = s;
}

/* The resident launch's ready flag in the pinned (coherent) host mirror: a
 * system-scope release store, so every earlier store of the thread (the
 * Hessenberg values) is visible to the host before the flag is. */
__device__ __forceinline__ void publish_flag(double *p, double v)
{
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v),
                     __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* A partial sum published by another workgroup of the launch: polled until
 * it no longer holds the call's sentinel (bounded; a thread that gives up
 * records KF_ERR_SPIN_LIMIT and stops waiting for the rest of the launch). */
__device__ __forceinline__ double poll_part(const double *p, unsigned long long s, uint32_t *err, bool &aborted,
                                            uint32_t sleep)
{
  double v = ld_coherent(p);
  for (uint32_t k = 0; sentinel<double>::raw(v) == s && !aborted; k++) {
    if (k >= KF_SPIN_LIMIT) {
      atomicOr(err, KF_ERR_SPIN_LIMIT);
      aborted = true;
      break;
    }
    for (uint32_t z = 0; z < sleep; z++) __builtin_amdgcn_s_sleep(1);
    v = ld_coherent(p);
  }
  return v;
}

/* The whole modified Gram-Schmidt of w against nv basis vectors in one
 * launch whose workgroups are all resident at once (checked by the kernel
 * itself, see below), w held in registers throughout (src/kernel.c:129-141
 * and the closing norm, src/kernel.c:138-141).
 *
 * The grid is k_dot_part's KF_NRED x RB threads folded two to one: thread t
 * of workgroup b plays thread t of blocks 2b and 2b + 1 (the virtual sets
 * 0 and 1), owns exactly the element pairs those threads walk (pairs
 * g, g + S, g + 2S, ... with S = KF_NRED RB, and the half pair of an odd
 * length), and keeps them in registers, at most MP pairs per set.  Step k:
 *   the partials of <w, V_k> (of <w, w> after the last vector), summed per
 *   set in the pair order of k_dot_part and reduced in block_sum's tree,
 *   published to row k of the partial table (sentinel-preset per call);
 *   h_k from that row in k_red_final's tree, formed by every workgroup;
 *   w = (-h_k) V_k + w on the registers.
 * So every partial, every h_k and the norm are bit-identical to the
 * k_dot_part / k_mgs_step / k_red_final sequence.  The update reuses the
 * V_k values the dot loaded: from registers (HOLDV), else from LDS for the
 * first NL slots and from memory again for the rest; w is read once and
 * written once. */
template <int MP, bool HOLDV>
__global__ __launch_bounds__(RB, 2) void k_mgs_res(size_t n, int nv, const double *__restrict__ vbase, size_t vstride,
                                                   double *__restrict__ w, double *__restrict__ hslot, double *part,
                                                   unsigned long long sent, uint32_t *err, uint32_t sleep, int unit,
                                                   uint32_t *sync, int force_abort, double *part_next,
                                                   unsigned long long sent_next, uint32_t *sync_next,
                                                   double *xnorm_out, double *hhost, double fd_h,
                                                   const double *__restrict__ qstate, double *__restrict__ wpert)
{
  /* without HOLDV the first NL slots of V_k are kept in LDS between the dot
   * and the update (72 KB for NL = 9: two workgroups still fit a CU) */
  constexpr int NL = HOLDV ? 0 : (MP < 9 ? MP : 9);
  __shared__ double sh[RB];
  __shared__ double2 vl[2][NL > 0 ? NL : 1][RB];
  /* np = q S + r: every virtual thread owns pairs m < q, those with g < r
   * also pair q, and the thread g == r the half pair of an odd length.
   * Every lane walks MP >= q + (r > 0) pair slots without branching: a slot
   * the lane does not own loads pair 0 in its place and adds +0 for it,
   * which leaves the sum's bits unchanged (a sum that starts at +0 is never
   * -0), and is never stored. */
  /* Co-residency, decided once per launch: sync[0] counts arrivals, sync[1]
   * goes from 0 to 1 (go: the last workgroup arrived) or to 2 (abort: a
   * workgroup waited KF_MGS_ARRIVE polls first), by compare-and-swap so that
   * every workgroup follows the same decision.  On abort nothing is written
   * but the flag hslot[nv + 1] = 1 (the host then repeats the call with the
   * per-step kernels); on go block 0 leaves hslot[nv + 1] = 0. */
  /* The tables alternate between calls: the other one (last used by the
   * previous launch, which has finished) is made ready for the next call --
   * every workgroup presets its own two columns of it, block 0 its arrival
   * words -- so no separate preset launch is needed. */
  if (threadIdx.x < 2 * KF_MGS_MAXV)
    part_next[(size_t)(threadIdx.x >> 1) * KF_NRED + 2 * blockIdx.x + (threadIdx.x & 1)] =
        sentinel<double>::val(sent_next);
  if (blockIdx.x == 0 && threadIdx.x < 2) sync_next[threadIdx.x] = 0;
  __shared__ uint32_t decision;
  if (threadIdx.x == 0) {
    uint32_t st = 0;
    if (force_abort) {
      atomicCAS(sync + 1, 0u, 2u);
      st = 2;
    } else if (atomicAdd(sync, 1u) + 1 == gridDim.x) {
      const uint32_t old = atomicCAS(sync + 1, 0u, 1u);
      st = old == 0 ? 1u : old;
    } else {
      st = __hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (uint32_t k = 0; st == 0; k++) {
        if (k >= KF_MGS_ARRIVE) {
          const uint32_t old = atomicCAS(sync + 1, 0u, 2u);
          st = old == 0 ? 2u : old;
          break;
        }
        for (uint32_t z = 0; z < sleep; z++) __builtin_amdgcn_s_sleep(1);
        st = __hip_atomic_load(sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (st != 1) {
      hslot[nv + 1] = 1.0;
      publish_flag(hhost + nv + 1, 1.0);
    }
    decision = st;
  }
  __syncthreads();
  if (decision != 1) return;   /* uniform over the workgroup */
  const size_t S = (size_t)KF_NRED * RB, np = n / 2;
  const int q = (int)(np / S);
  const uint32_t r = (uint32_t)(np % S), PB = KF_NRED * RB * 16u;   /* bytes between a thread's pairs */
  uint32_t voff[2];
  bool ext[2], tail[2];
  double wr[2][MP][2], wt[2] = {0.0, 0.0};
#pragma unroll
  for (int v = 0; v < 2; v++) {
    const uint32_t g = (2 * blockIdx.x + v) * RB + threadIdx.x;
    ext[v] = g < r;
    tail[v] = (n & 1) && g == r;
    voff[v] = 16u * g;
  }
  /* whether set v owns slot m, and the slot's byte offset: a uniform base
   * plus one 32-bit per-lane offset */
#define KF_ON(m, v) ((m) < q || ((m) == q && ext[v]))
#define KF_PAIR(x, m, v)                                                                                   \
  (*reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(x) +                                  \
                                      (KF_ON(m, v) ? voff[v] + (uint32_t)(m) * PB : 0u)))
#pragma unroll
  for (int v = 0; v < 2; v++) {
#pragma unroll
    for (int m = 0; m < MP; m++) {
      const double2 t = KF_PAIR(w, m, v);
      wr[v][m][0] = t.x;
      wr[v][m][1] = t.y;
    }
    if (tail[v]) wt[v] = w[n - 1];
  }
  bool aborted = false;
  double nrm2 = 0.0;
  for (int k = 0; k <= nv; k++) {
    /* the offsets pass through an empty asm each step, so the pair
     * addresses are formed where they are used instead of being kept in
     * registers across the steps */
    asm volatile("" : "+v"(voff[0]), "+v"(voff[1]));
    const bool self = k == nv;
    const double *u = vbase + (size_t)k * vstride;
    double vr[2][HOLDV ? MP : 1][2];
    double vt[2] = {0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 2; v++) {
      double acc = 0.0;
      if (self) {
#pragma unroll
        for (int m = 0; m < MP; m++) {
          const double p0 = wr[v][m][0] * wr[v][m][0], p1 = wr[v][m][1] * wr[v][m][1];
          acc += KF_ON(m, v) ? p0 : 0.0;
          acc += KF_ON(m, v) ? p1 : 0.0;
        }
        if (tail[v]) acc += wt[v] * wt[v];
      } else {
        /* chunks of four pairs: a chunk's loads are in flight together */
#pragma unroll
        for (int m0 = 0; m0 < MP; m0 += 4) {
          double uu[4][2];
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (m0 + j < MP) {
              const double2 t = KF_PAIR(u, m0 + j, v);
              uu[j][0] = t.x;
              uu[j][1] = t.y;
            }
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int m = m0 + j;
            if (m < MP) {
              const double p0 = wr[v][m][0] * uu[j][0], p1 = wr[v][m][1] * uu[j][1];
              acc += KF_ON(m, v) ? p0 : 0.0;
              acc += KF_ON(m, v) ? p1 : 0.0;
              if (HOLDV) {
                vr[v][HOLDV ? m : 0][0] = uu[j][0];
                vr[v][HOLDV ? m : 0][1] = uu[j][1];
              } else if (m < NL) {
                vl[v][m < NL ? m : 0][threadIdx.x] = make_double2(uu[j][0], uu[j][1]);
              }
            }
          }
          if (!HOLDV) __builtin_amdgcn_sched_barrier(0);   /* one chunk's loads live at a time */
        }
        if (tail[v]) {
          vt[v] = u[n - 1];
          acc += wt[v] * vt[v];
        }
      }
      const double s = block_sum(acc, sh);
      if (threadIdx.x == 0) st_coherent(part + (size_t)k * KF_NRED + 2 * blockIdx.x + v, canon(s));
    }
    /* h_k (or |w|^2) from row k, in k_red_final's tree */
    double acc = 0.0;
    for (int j = threadIdx.x; j < KF_NRED; j += RB)
      acc += poll_part(part + (size_t)k * KF_NRED + j, sent, err, aborted, sleep);
    const double h = block_sum(acc, sh);
    /* whether any thread of this workgroup gave up a wait (its h_k hold a
     * placeholder then); uniform: self is the same for every thread */
    const int gaveup = self ? __syncthreads_or(aborted ? 1 : 0) : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      hslot[k] = hhost[k] = self ? sqrt(h) : h;
      if (self) {
        /* every value the host needs is written: the ready flag (0), or 2
         * when a wait expired and the values are not valid, by a
         * system-scope release store (the host polls it while the launch
         * finishes) */
        hslot[nv + 1] = gaveup ? 2.0 : 0.0;
        publish_flag(hhost + nv + 1, gaveup ? 2.0 : 0.0);
      }
    }
    if (self) {   /* the closing norm: the loop ends */
      nrm2 = h;
      continue;
    }
    const double a = -h;
    asm volatile("" : "+v"(voff[0]), "+v"(voff[1]));   /* addresses re-formed for the update */
#pragma unroll
    for (int v = 0; v < 2; v++) {
      if (HOLDV) {
#pragma unroll
        for (int m = 0; m < MP; m++) {
          wr[v][m][0] = a * vr[v][HOLDV ? m : 0][0] + wr[v][m][0];   /* contraction off */
          wr[v][m][1] = a * vr[v][HOLDV ? m : 0][1] + wr[v][m][1];
        }
      } else {
#pragma unroll
        for (int m0 = 0; m0 < MP; m0 += 4) {
          double uu[4][2];
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (m0 + j < MP) {
              /* the same values the dot read: from LDS, else again from memory */
              const double2 t = m0 + j < NL ? vl[v][m0 + j < NL ? m0 + j : 0][threadIdx.x] : KF_PAIR(u, m0 + j, v);
              uu[j][0] = t.x;
              uu[j][1] = t.y;
            }
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int m = m0 + j;
            if (m < MP) {
              wr[v][m][0] = a * uu[j][0] + wr[v][m][0];
              wr[v][m][1] = a * uu[j][1] + wr[v][m][1];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      if (tail[v]) wt[v] = a * vt[v] + wt[v];
    }
  }
  /* unit: w scaled to unit length as k_scale_recip does, x *= 1.0 / |w| */
  if (unit) {
    const double f = 1.0 / sqrt(nrm2);
#pragma unroll
    for (int v = 0; v < 2; v++) {
#pragma unroll
      for (int m = 0; m < MP; m++) {
        wr[v][m][0] *= f;
        wr[v][m][1] *= f;
      }
      wt[v] *= f;
    }
  }
  /* xnorm_out: |w| of the stored (scaled) w as kf_nrm2_dev forms it --
   * k_dot_part's partials, then k_red_final's tree and square root -- in one
   * more hand-off round on the table's row nv + 1 */
  if (xnorm_out) {
#pragma unroll
    for (int v = 0; v < 2; v++) {
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < MP; m++) {
        const double p0 = wr[v][m][0] * wr[v][m][0], p1 = wr[v][m][1] * wr[v][m][1];
        acc += KF_ON(m, v) ? p0 : 0.0;
        acc += KF_ON(m, v) ? p1 : 0.0;
      }
      if (tail[v]) acc += wt[v] * wt[v];
      const double s = block_sum(acc, sh);
      if (threadIdx.x == 0) st_coherent(part + (size_t)(nv + 1) * KF_NRED + 2 * blockIdx.x + v, canon(s));
    }
    double acc = 0.0;
    for (int j = threadIdx.x; j < KF_NRED; j += RB)
      acc += poll_part(part + (size_t)(nv + 1) * KF_NRED + j, sent, err, aborted, sleep);
    const double h = block_sum(acc, sh);
    const double xn = sqrt(h);
    if (blockIdx.x == 0 && threadIdx.x == 0) *xnorm_out = xn;
    /* wpert: the next iJacVec's perturbed state, as k_fd_perturb forms it:
     * a = h_fd / |x|, wpert = a x + q (the product, then the sum) */
    if (wpert) {
      const double a = fd_h / xn;
      asm volatile("" : "+v"(voff[0]), "+v"(voff[1]));
#pragma unroll
      for (int v = 0; v < 2; v++) {
#pragma unroll
        for (int m0 = 0; m0 < MP; m0 += 4) {
          double qq[4][2];
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (m0 + j < MP) {
              const double2 t = KF_PAIR(qstate, m0 + j, v);
              qq[j][0] = t.x;
              qq[j][1] = t.y;
            }
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int m = m0 + j;
            if (m < MP && KF_ON(m, v)) {
              const double ax0 = a * wr[v][m][0], ax1 = a * wr[v][m][1];
              *reinterpret_cast<double2 *>(reinterpret_cast<char *>(wpert) + voff[v] + (uint32_t)m * PB) =
                  make_double2(ax0 + qq[j][0], ax1 + qq[j][1]);
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        if (tail[v]) {
          const double ax = a * wt[v];
          wpert[n - 1] = ax + qstate[n - 1];
        }
      }
    }
  }
  asm volatile("" : "+v"(voff[0]), "+v"(voff[1]));
#pragma unroll
  for (int v = 0; v < 2; v++) {
#pragma unroll
    for (int m = 0; m < MP; m++)
      if (KF_ON(m, v))
        *reinterpret_cast<double2 *>(reinterpret_cast<char *>(w) + voff[v] + (uint32_t)m * PB) =
            make_double2(wr[v][m][0], wr[v][m][1]);
    if (tail[v]) w[n - 1] = wt[v];
  }
#undef KF_PAIR
#undef KF_ON
}

/* a partial table preset to a call's sentinel, and its arrival counter and
 * decision cleared (set-up of the first call's table) */
__global__ void k_mgs_preset(double *part, size_t rows, unsigned long long sent, uint32_t *sync)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows) part[i] = sentinel<double>::val(sent);
  if (i < 2) sync[i] = 0;
}

/* the resident variants: pairs per virtual thread at most, and whether the
 * basis vector's pairs stay in registers between the dot and the update */
#define KF_MGS_VARIANTS 8
static const int mgs_mp[KF_MGS_VARIANTS] = {1, 2, 3, 4, 8, 12, 16, 21};

static const void *mgs_kernel(int i)
{
  switch (i) {
  case 0: return (const void *)k_mgs_res<1, true>;
  case 1: return (const void *)k_mgs_res<2, true>;
  case 2: return (const void *)k_mgs_res<3, true>;
  case 3: return (const void *)k_mgs_res<4, true>;
  case 4: return (const void *)k_mgs_res<8, false>;
  case 5: return (const void *)k_mgs_res<12, false>;
  case 6: return (const void *)k_mgs_res<16, false>;
  default: return (const void *)k_mgs_res<21, false>;
  }
}

/* One launch of k_mgs_res when w and the basis are 16-byte aligned, the
 * pairs fit a variant and the occupancy of its KF_NRED / 2 workgroups allows
 * them all on this device at once; false (nothing launched) otherwise.  The
 * launch is an ordinary one -- co-residency is checked by the kernel itself,
 * which writes the abort flag to scal[slot + nv + 1] -- because a
 * cooperative launch cost about 30 us more per call on MI355X. */
static bool mgs_resident(kf_ctx *c, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot,
                         int unit, int xslot, double fd_h, double *wpert)
{
  if (!c->mgs_res || nv + 2 > KF_MGS_MAXV || n < 2) return false;   /* a lane without pairs reads pair 0 */
  if (slot + nv + 1 >= KF_NSCAL || xslot >= KF_NSCAL) return false;
  if (!(aligned16(vbase) && aligned16(w) && (stride & 1) == 0)) return false;
  const size_t S = (size_t)KF_NRED * RB, np = n / 2, need = (np + S - 1) / S;   /* pairs per thread, at most */
  int vi = 0;
  while (vi < KF_MGS_VARIANTS && (size_t)mgs_mp[vi] < need) vi++;
  if (vi == KF_MGS_VARIANTS) return false;
  const void *fn = mgs_kernel(vi);
  if (c->mgs_ok[vi] < 0) {
    int nb = 0, dev = 0, ncu = 0;
    KF_CHECK(hipGetDevice(&dev));
    KF_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    KF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, RB, 0));
    c->mgs_ok[vi] = ((size_t)nb * (size_t)ncu >= KF_NRED / 2) ? 1 : 0;
  }
  if (!c->mgs_ok[vi]) return false;
  /* this call's table and arrival words (preset by the previous launch, or
   * at set-up for the first), and the other parity's, for the next call */
  const uint64_t k = c->mcall;
  const unsigned long long sent = sentinel<double>::of(k), sent_next = sentinel<double>::of(k + 1);
  double *part = c->mgs_part + (k & 1) * (size_t)KF_MGS_MAXV * KF_NRED;
  double *part_next = c->mgs_part + ((k + 1) & 1) * (size_t)KF_MGS_MAXV * KF_NRED;
  uint32_t *sync = c->mgs_sync + 2 * (k & 1), *sync_next = c->mgs_sync + 2 * ((k + 1) & 1);
  double *hslot = c->d.scal + slot;
  uint32_t *err = c->d.err;
  uint32_t sleep = c->d.spin_sleep;
  const int force_abort = c->mgs_res == 3;
  double *xnorm_out = xslot >= 0 ? c->d.scal + xslot : NULL;
  double *hhost = c->hscal + slot;   /* the pinned host mirror, polled by the host */
  c->hscal[slot + nv + 1] = -1.0;    /* not ready: the launch writes 0 (go, values written), 1 (abort:
                                        nothing written) or 2 (a wait expired: the values are not valid) */
  const double *qstate = c->q;
  if (!xnorm_out || n != 4 * (size_t)c->d.N || !aligned16(wpert)) wpert = NULL;   /* needs the norm round */
  void *args[] = {&n,   &nv,    &vbase, &stride, &w,           &hslot,     &part,      (void *)&sent,
                  &err, &sleep, &unit,  &sync,   (void *)&force_abort, &part_next, (void *)&sent_next, &sync_next,
                  &xnorm_out, &hhost, &fd_h, &qstate, &wpert};
  KF_CHECK(hipLaunchKernel(fn, dim3(KF_NRED / 2), dim3(RB), args, 0, c->stream));
  c->mcall++;
  return true;
}

/* the first resident launch's table and arrival words (call 0, parity 0) */
void kf_mgs_setup(kf_ctx *c)
{
  const size_t rows = (size_t)KF_MGS_MAXV * KF_NRED;
  hipLaunchKernelGGL(k_mgs_preset, dim3(kf_blocks_for(rows, RB)), dim3(RB), 0, c->stream, c->mgs_part, rows,
                     sentinel<double>::of(0), c->mgs_sync);
  KF_LAUNCH_CHECK();
}

/* the per-step modified Gram-Schmidt (k_dot_part, k_mgs_step per basis
 * vector, k_red_final), then the unit scaling */
static void mgs_steps(kf_ctx *c, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot, int unit)
{
  double *part[2] = {c->d.red, c->d.red + KF_NRED};
  launch_dot_part(c, n, w, vbase, part[0]);
  const bool vec = aligned16(vbase) && aligned16(w) && (stride & 1) == 0;
  for (int k = 0; k < nv; k++) {
    const double *u = k + 1 < nv ? vbase + (size_t)(k + 1) * stride : w;
    if (vec)
      hipLaunchKernelGGL(k_mgs_step<true>, dim3(KF_NRED), dim3(RB), 0, c->stream, n, part[k & 1],
                         c->d.scal + slot + k, vbase + (size_t)k * stride, w, u, part[(k + 1) & 1]);
    else
      hipLaunchKernelGGL(k_mgs_step<false>, dim3(KF_NRED), dim3(RB), 0, c->stream, n, part[k & 1],
                         c->d.scal + slot + k, vbase + (size_t)k * stride, w, u, part[(k + 1) & 1]);
    KF_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(RB), 0, c->stream, part[nv & 1], c->d.scal + slot + nv, 1);
  KF_LAUNCH_CHECK();
  if (unit) kf_scale_recip(c, n, w, slot + nv);
}

/* w orthogonalised against nv basis vectors vbase + k stride by modified
 * Gram-Schmidt: h_k in scal[slot + k], |w| in scal[slot + nv]; unit: w then
 * scaled to unit length (x *= 1.0 / |w|, kf_scale_recip); out (if not NULL):
 * the nv + 1 values fetched to the host.  A resident launch writes them to
 * the pinned host mirror as well, then a ready flag there, which the host
 * polls (the launch's tail overlaps the caller's work); if it aborted
 * (workgroups not co-resident), w is untouched and the call is repeated
 * with the per-step kernels, which the context then keeps; if one of its
 * waits expired, out receives NaN.  xslot >= 0 (with unit):
 * the resident launch also leaves |w| of the scaled w in scal[xslot], as
 * kf_nrm2_dev would; wpert (with xslot, w of the state's length): the launch
 * also writes the next iJacVec's perturbed state (fd_h / |w|) w + q as
 * kf_fd_perturb would.  Returns 2 if both were left, 1 if the norm only, 0
 * otherwise.  Index-order sums (reduce_seq) keep the unfused dot / update
 * sequence. */
int kf_mgs(kf_ctx *c, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot, int unit,
           double *out, int xslot, double fd_h, double *wpert)
{
  if (c->opt.reduce_seq) {
    for (int k = 0; k < nv; k++) {
      dot_any(c, n, w, vbase + (size_t)k * stride, slot + k, 0);
      kf_axpy_slot(c, n, slot + k, vbase + (size_t)k * stride, w);
    }
    dot_any(c, n, w, w, slot + nv, 1);
    if (unit) kf_scale_recip(c, n, w, slot + nv);
  } else if (mgs_resident(c, n, vbase, stride, nv, w, slot, unit, unit ? xslot : -1, fd_h, wpert)) {
    /* h_k, |w| and the flag are written to the pinned host mirror as soon
     * as they are known: poll the flag (bounded, then a stream
     * synchronisation) instead of waiting for the whole launch, whose tail
     * overlaps the caller's work; later launches follow it in the stream */
    const volatile double *flag = c->hscal + slot + nv + 1;
    if (c->mgs_nopoll) KF_CHECK(hipStreamSynchronize(c->stream));
    double fv = *flag;
    /* The wait ends when the flag shows, or when the launch has ended
     * without it: after a first stretch of polls an event is recorded (the
     * launch is the last work queued on the stream, so the event completes
     * when it ends) and queried now and then.  Work queued ahead of the
     * launch (a factorisation, training) only lengthens the wait; it never
     * expires it. */
    bool marked = false;
    for (uint64_t k = 1; fv == -1.0; k++) {
      if ((k & 0xffffu) == 0) {
        if (!marked) {
          if (!c->mgs_ev) KF_CHECK(hipEventCreateWithFlags(&c->mgs_ev, hipEventDisableTiming));
          KF_CHECK(hipEventRecord(c->mgs_ev, c->stream));
          marked = true;
        } else {
          const hipError_t q = hipEventQuery(c->mgs_ev);
          if (q != hipErrorNotReady) {
            KF_CHECK(q);
            fv = *flag;
            if (fv == -1.0) {
              /* the launch ended and its flag never showed in the host
               * mirror: synchronise, here and from now on (the values are
               * fetched by copy) */
              c->mgs_nopoll = 1;
              KF_CHECK(hipMemcpyAsync(c->hscal + slot, c->d.scal + slot, (nv + 2) * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream));
              KF_CHECK(hipStreamSynchronize(c->stream));
              fv = *flag;
            }
            break;
          }
        }
      }
      __builtin_ia32_pause();
      fv = *flag;
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    const double *tmp = c->hscal + slot;
    if (fv == 2.0) {
      /* a wait inside the launch expired (KF_ERR_SPIN_LIMIT is set): the
       * values are placeholders, which the caller sees as non-finite */
      KF_CHECK(hipStreamSynchronize(c->stream));
      fprintf(stderr, "kf: resident Gram-Schmidt: a wait for another workgroup's partial sum expired\n");
      if (out)
        for (int i = 0; i <= nv; i++) out[i] = NAN;
      return 0;
    }
    if (fv == 0.0) {
      if (out)
        for (int i = 0; i <= nv; i++) out[i] = tmp[i];
      if (!(unit && xslot >= 0)) return 0;
      return (wpert && n == 4 * (size_t)c->d.N && aligned16(wpert)) ? 2 : 1;
    }
    KF_CHECK(hipStreamSynchronize(c->stream));   /* the aborted launch has ended */
    c->mgs_res = 0;   /* not co-resident here: the per-step kernels from now on */
    c->mgs_aborts++;
    mgs_steps(c, n, vbase, stride, nv, w, slot, unit);
  } else {
    mgs_steps(c, n, vbase, stride, nv, w, slot, unit);
  }
  if (out) kf_scal_fetch(c, slot, nv + 1, out);
  return 0;
}

void kf_scal_fetch(kf_ctx *c, int slot, int count, double *out)
{
  KF_CHECK(hipMemcpyAsync(c->hscal + slot, c->d.scal + slot, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KF_CHECK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < count; i++) out[i] = c->hscal[slot + i];
}

/* y = a x + y with the reference's ax-then-add rounding (src/utils.c:79-83) */
__global__ __launch_bounds__(RB) void k_axpy(size_t n, double a, const double *__restrict__ x, double *__restrict__ y)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const double ax = a * x[i];
  y[i] = ax + y[i];
}

__global__ __launch_bounds__(RB) void k_newaxpy(size_t n, double a, const double *__restrict__ x,
                                                const double *__restrict__ y, double *__restrict__ w)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const double ax = a * x[i];
  w[i] = ax + y[i];
}

/* x *= 1/s, s = scal[slot] (src/utils.c:117-118: the reciprocal, then the product) */
__global__ __launch_bounds__(RB) void k_scale_recip(size_t n, double *__restrict__ x, const double *__restrict__ s)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const double f = 1.0 / *s;
  x[i] *= f;
}

/* y = (-h) x + y, h = scal[slot] (the MGS update, src/kernel.c:138) */
__global__ __launch_bounds__(RB) void k_axpy_slot(size_t n, const double *__restrict__ s, const double *__restrict__ x,
                                                  double *__restrict__ y)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const double a = -*s;
  const double ax = a * x[i];
  y[i] = ax + y[i];
}

/* w = (h / |x|) x + q (src/kernel.c:27-30) */
__global__ __launch_bounds__(RB) void k_fd_perturb(size_t n, double h, const double *__restrict__ xnorm,
                                                   const double *__restrict__ x, const double *__restrict__ q,
                                                   double *__restrict__ w)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const double a = h / *xnorm;
  const double ax = a * x[i];
  w[i] = ax + q[i];
}

/* y = ((y - r) + cdt (w - q)) |x| / h (src/kernel.c:33-44) */
__global__ __launch_bounds__(RB) void k_fd_combine(uint32_t N, double h, const double *__restrict__ xnorm,
                                                   const double *__restrict__ r, const double *__restrict__ w,
                                                   const double *__restrict__ q, const double *__restrict__ cdt,
                                                   double *__restrict__ y)
{
  const size_t idx = (size_t)blockIdx.x * RB + threadIdx.x;
  if (idx >= 4 * (size_t)N) return;
  const double a = *xnorm / h;
  const double cd = cdt[idx >> 2];
  y[idx] = (y[idx] - r[idx] + cd * (w[idx] - q[idx])) * a;
}

__global__ __launch_bounds__(RB) void k_fill_zero(size_t n, double *__restrict__ x)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

void kf_axpy(kf_ctx *c, size_t n, double a, const double *x, double *y)
{
  if (!n) return;
  hipLaunchKernelGGL(k_axpy, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, a, x, y);
  KF_LAUNCH_CHECK();
}

void kf_newaxpy(kf_ctx *c, size_t n, double a, const double *x, const double *y, double *w)
{
  if (!n) return;
  hipLaunchKernelGGL(k_newaxpy, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, a, x, y, w);
  KF_LAUNCH_CHECK();
}

void kf_scale_recip(kf_ctx *c, size_t n, double *x, int slot)
{
  if (!n) return;
  hipLaunchKernelGGL(k_scale_recip, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, x, c->d.scal + slot);
  KF_LAUNCH_CHECK();
}

void kf_axpy_slot(kf_ctx *c, size_t n, int slot, const double *x, double *y)
{
  if (!n) return;
  hipLaunchKernelGGL(k_axpy_slot, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, c->d.scal + slot, x, y);
  KF_LAUNCH_CHECK();
}

void kf_fd_perturb(kf_ctx *c, const double *x, double h, int slot_xnorm, double *w)
{
  const size_t n = 4 * (size_t)c->d.N;
  hipLaunchKernelGGL(k_fd_perturb, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, h, c->d.scal + slot_xnorm,
                     x, c->q, w);
  KF_LAUNCH_CHECK();
}

void kf_fd_combine(kf_ctx *c, const double *r, const double *w, double h, int slot_xnorm, double *y)
{
  const size_t n = 4 * (size_t)c->d.N;
  hipLaunchKernelGGL(k_fd_combine, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, c->d.N, h,
                     c->d.scal + slot_xnorm, r, w, c->q, c->cdt, y);
  KF_LAUNCH_CHECK();
}

void kf_fill_zero(kf_ctx *c, size_t n, double *x)
{
  if (!n) return;
  hipLaunchKernelGGL(k_fill_zero, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, x);
  KF_LAUNCH_CHECK();
}

/* ---- batched classical Gram-Schmidt: h_v = V_v . w for all v in one pass
 * over w, then w -= sum_v h_v V_v (one more pass); deterministic trees ---- */

#define CGS_MAXV 32
#define CGS_BLOCKS 512

__global__ __launch_bounds__(RB) void k_cgs_dots(size_t n, const double *__restrict__ vb, int nv,
                                                 const double *__restrict__ w, double *__restrict__ part)
{
  __shared__ double sh[RB];
  double acc[CGS_MAXV];
#pragma unroll
  for (int v = 0; v < CGS_MAXV; v++) acc[v] = 0.0;
  const size_t stride = (size_t)gridDim.x * RB;
  for (size_t i = (size_t)blockIdx.x * RB + threadIdx.x; i < n; i += stride) {
    const double wi = w[i];
#pragma unroll
    for (int v = 0; v < CGS_MAXV; v++)
      if (v < nv) acc[v] += vb[(size_t)v * n + i] * wi;
  }
  for (int v = 0; v < nv; v++) {
    const double s = block_sum(acc[v], sh);
    if (threadIdx.x == 0) part[(size_t)v * CGS_BLOCKS + blockIdx.x] = s;
    __syncthreads();
  }
}

__global__ __launch_bounds__(RB) void k_cgs_final(const double *__restrict__ part, int nv, double *__restrict__ out)
{
  __shared__ double sh[RB];
  for (int v = 0; v < nv; v++) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < CGS_BLOCKS; i += RB) acc += part[(size_t)v * CGS_BLOCKS + i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[v] = s;
    __syncthreads();
  }
}

__global__ __launch_bounds__(RB) void k_cgs_update(size_t n, const double *__restrict__ vb, int nv,
                                                   const double *__restrict__ h, double *__restrict__ w)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int v = 0; v < nv; v++) s += h[v] * vb[(size_t)v * n + i];
  w[i] = w[i] - s;
}

/* ---- C wrappers for the host driver --------------------------------------- */

extern "C" {

void kfs_nrm2(GEOMETRY *g, size_t n, const double *x, int slot) { kf_nrm2_dev(kf_ctx_of(g), n, x, slot); }
void kfs_dot(GEOMETRY *g, size_t n, const double *x, const double *y, int slot) { kf_dot_dev(kf_ctx_of(g), n, x, y, slot); }
void kfs_fetch(GEOMETRY *g, int slot, int count, double *out) { kf_scal_fetch(kf_ctx_of(g), slot, count, out); }

unsigned long long kfs_mgs_launches(GEOMETRY *g) { return kf_ctx_of(g)->mcall; }
void kfs_axpy(GEOMETRY *g, size_t n, double a, const double *x, double *y) { kf_axpy(kf_ctx_of(g), n, a, x, y); }
void kfs_newaxpy(GEOMETRY *g, size_t n, double a, const double *x, const double *y, double *w)
{
  kf_newaxpy(kf_ctx_of(g), n, a, x, y, w);
}
void kfs_scale_recip(GEOMETRY *g, size_t n, double *x, int slot) { kf_scale_recip(kf_ctx_of(g), n, x, slot); }
void kfs_axpy_slot(GEOMETRY *g, size_t n, int slot, const double *x, double *y) { kf_axpy_slot(kf_ctx_of(g), n, slot, x, y); }
void kfs_mgs(GEOMETRY *g, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot)
{
  kf_mgs(kf_ctx_of(g), n, vbase, stride, nv, w, slot, 0, NULL, -1, 0.0, NULL);
}

int kfs_mgs_unit(GEOMETRY *g, size_t n, const double *vbase, size_t stride, int nv, double *w, int slot,
                 double *out, int xslot, double fd_h, double *wpert)
{
  return kf_mgs(kf_ctx_of(g), n, vbase, stride, nv, w, slot, 1, out, xslot, fd_h, wpert);
}

unsigned long long kfs_mgs_aborts(GEOMETRY *g) { return kf_ctx_of(g)->mgs_aborts; }
void kfs_fd_perturb(GEOMETRY *g, const double *x, double h, int slot, double *w) { kf_fd_perturb(kf_ctx_of(g), x, h, slot, w); }
void kfs_fd_combine(GEOMETRY *g, const double *r, const double *w, double h, int slot, double *y)
{
  kf_fd_combine(kf_ctx_of(g), r, w, h, slot, y);
}
void kfs_zero(GEOMETRY *g, size_t n, double *x) { kf_fill_zero(kf_ctx_of(g), n, x); }

void kfs_cgs_dots(GEOMETRY *g, size_t n, const double *vbase, int nv, const double *w, int slot)
{
  kf_ctx *c = kf_ctx_of(g);
  if (nv > CGS_MAXV || KF_NRED < CGS_BLOCKS * 1) abort();
  double *part = c->d.eres;   /* free between residual evaluations: 4E >= 32 * 512 doubles */
  if (4 * (size_t)c->d.E < (size_t)CGS_MAXV * CGS_BLOCKS) {
    fprintf(stderr, "kf: mesh too small for the batched Gram-Schmidt scratch\n");
    abort();
  }
  hipLaunchKernelGGL(k_cgs_dots, dim3(CGS_BLOCKS), dim3(RB), 0, c->stream, n, vbase, nv, w, part);
  KF_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_cgs_final, dim3(1), dim3(RB), 0, c->stream, part, nv, c->d.scal + slot);
  KF_LAUNCH_CHECK();
}

void kfs_cgs_update(GEOMETRY *g, size_t n, const double *vbase, int nv, int slot, double *w)
{
  kf_ctx *c = kf_ctx_of(g);
  hipLaunchKernelGGL(k_cgs_update, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, vbase, nv,
                     c->d.scal + slot, w);
  KF_LAUNCH_CHECK();
}

/* standard normal deviates: x[i] from a counter-based generator keyed by
 * (seed, stream, i) -- splitmix64 of the counter for two uniforms, then the
 * Box-Muller cosine branch */
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(RB) void k_gaussian(size_t n, uint64_t key, double *__restrict__ x)
{
  const size_t i = (size_t)blockIdx.x * RB + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = splitmix64(key + 2 * (uint64_t)i), b = splitmix64(key + 2 * (uint64_t)i + 1);
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53; /* (0, 1] */
  const double u2 = (double)(b >> 11) * 0x1.0p-53;         /* [0, 1) */
  x[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* scal[out] = min(1, gamma scal[a] / scal[b]) (1 when scal[b] is 0) */
__global__ void k_guard(double *scal, int a, int b, double gamma, int out)
{
  if (threadIdx.x != 0) return;
  const double lim = gamma * scal[a], nb = scal[b];
  scal[out] = (nb > lim) ? lim / nb : 1.0;
}

void kfs_gaussian(GEOMETRY *g, size_t n, uint64_t seed, uint64_t stream, double *x)
{
  kf_ctx *c = kf_ctx_of(g);
  const uint64_t key = seed * 0xD1B54A32D192ED03ull ^ (stream * 0x8CB92BA72F3D8DD7ull);
  hipLaunchKernelGGL(k_gaussian, dim3(kf_blocks_for(n, RB)), dim3(RB), 0, c->stream, n, key, x);
  KF_LAUNCH_CHECK();
}

void kfs_guard(GEOMETRY *g, int slot_a, int slot_b, double gamma, int slot_out)
{
  kf_ctx *c = kf_ctx_of(g);
  hipLaunchKernelGGL(k_guard, dim3(1), dim3(64), 0, c->stream, c->d.scal, slot_a, slot_b, gamma, slot_out);
  KF_LAUNCH_CHECK();
}

void **kfs_learn_state(GEOMETRY *g) { return &kf_ctx_of(g)->learn; }

void kfs_work(GEOMETRY *g, int nbasis, double **grad3, double **r, double **w, double **t, double **x, double **vecs)
{
  kf_ctx *c = kf_ctx_of(g);
  const size_t sz = 4 * (size_t)c->d.N;
  double **bufs[7] = {&c->gx, &c->gy, &c->gz, &c->r, &c->w, &c->t, &c->xk};
  for (int k = 0; k < 7; k++)
    if (!*bufs[k]) {
      KF_CHECK(hipMalloc((void **)bufs[k], sz * sizeof(double)));
      KF_CHECK(hipMemset(*bufs[k], 0, sz * sizeof(double)));
    }
  if (c->nvecs < nbasis) {
    if (c->vecs) KF_CHECK(hipFree(c->vecs));
    KF_CHECK(hipMalloc((void **)&c->vecs, (size_t)nbasis * sz * sizeof(double)));
    KF_CHECK(hipMemset(c->vecs, 0, (size_t)nbasis * sz * sizeof(double)));
    c->nvecs = nbasis;
  }
  grad3[0] = c->gx; grad3[1] = c->gy; grad3[2] = c->gz;
  *r = c->r; *w = c->w; *t = c->t; *x = c->xk; *vecs = c->vecs;
}

}  // extern "C"
