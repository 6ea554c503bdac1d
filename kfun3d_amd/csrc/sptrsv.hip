/*
 * sptrsv.hip -- x = U^-1 L^-1 b with the in-place ILU(0) factors
 * (matrix/sptrsv.cpp:34-73; the reference schedules rows by level sets,
 * matrix/init.cpp:194-274).
 *
 * Work unit: a task of up to 16 rows, owned by one wavefront with a quad of
 * lanes per row (lane c computes component c).  The host (host/plan.c)
 * cuts the rows into contiguous ranges of R rows and clusters the ranges
 * into tasks that read only from themselves and from tasks of lower level;
 * tasks are handed out by ticket in level order, G consecutive tickets per
 * atomic to a workgroup of G wavefronts (one task each; the group takes its
 * next G tickets when all its tasks are done), so a wave only ever waits
 * for tasks whose tickets were issued before its own, to waves that are
 * running: no wait can deadlock.
 *
 * Hand-off without flags or fences.  Every result slot is preset to a
 * signalling NaN that names the sweep and the call (the sentinel); a
 * producer stores its results with agent-scope atomic stores, a consumer
 * loads the values it needs with agent-scope atomic loads and retries the
 * ones that still hold the sentinel.  Only single-location coherence is
 * relied on -- the value read is the value used -- so no release/acquire
 * (L2 write-back and invalidate on the 8 XCDs) is needed, and the data
 * arrives in the same round trip that signals it.  Arithmetic never yields a
 * signalling NaN and every stored result is canonicalised, so a result never
 * equals a sentinel.  Sentinels are set up for free: the forward waves preset
 * the backward output of their rows, the backward waves reset the forward
 * output of their rows for the next call.
 *
 *   external entries (columns in earlier tasks): polled as above, and their
 *     block products formed as soon as they arrive;
 *   internal entries (columns in the same task; eslot names the column's
 *     slot): resolved inside the wave through LDS, row by row in the local
 *     level order computed on the host (tlvl; the task's depth tdep).
 *
 * Columns, slots and factor blocks of the first PF off-diagonal entries of
 * every row are loaded into registers as soon as the wave takes its ticket;
 * PF is 7 when no row has more than 7 entries on either side of the
 * diagonal (the Kuhn lattices: 14 neighbours, 7 below and 7 above), else
 * KF_TRSV_PF, so that no level forms a product past the end of every row.
 *
 * Per row the block products are subtracted in ascending column order
 * exactly as the reference:
 *   forward   x_r = b_r;  x_r -= A(r,j) x_j  for j < r           (dgemv, then axpy -1)
 *   backward  x_r -= A(r,j) x_j  for j > r;  x_r = inv(D_r) x_r
 * internal and external entries interleaved as the columns come: a product
 * is the same whenever it is formed, and only the order of the subtractions
 * matters, so the FP64 result is bit-identical.  In the single-precision
 * modes the factors are FP32 or BF16 (converted to FP32 on load), the
 * vectors FP32, and the result is widened to FP64 on the last store.
 *
 * Every wait is bounded (KF_SPIN_LIMIT polls): a lane that gives up records
 * KF_ERR_SPIN_LIMIT and stops waiting for the rest of the launch.
 */
#include "sync.h"
#include "handoff.h"

static_assert(KF_GROUP == 16, "one quad of lanes per row, 16 rows per task");

#define KF_TRSV_GMAX 4 /* wavefronts per workgroup at most */
#define KF_TRSV_PF 8   /* off-diagonal entries per row held in registers, at most */

__device__ __forceinline__ double mat_ld(const double *a, size_t i) { return a[i]; }
__device__ __forceinline__ float mat_ld(const float *a, size_t i) { return a[i]; }
__device__ __forceinline__ float mat_ld(const uint16_t *a, size_t i) { return __uint_as_float((uint32_t)a[i] << 16); }

/* s = A(j) row c . x  (dgemv row; summed from zero in column order) */
template <typename TV>
__device__ __forceinline__ TV row_dot(const TV m[4], TV x0, TV x1, TV x2, TV x3)
{
  TV s = TV(0);
  s += m[0] * x0;
  s += m[1] * x1;
  s += m[2] * x2;
  s += m[3] * x3;
  return s;
}

/* value of lane I of this lane's quad (DPP quad_perm broadcast; the quad's
 * four lanes are active together) */
template <int I>
__device__ __forceinline__ float quad_get(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), I * 0x55, 0xf, 0xf, false));
}

template <int I>
__device__ __forceinline__ double quad_get(double v)
{
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), I * 0x55, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), I * 0x55, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <typename TM, typename TV>
__device__ __forceinline__ void ld_row(const TM *__restrict__ a, uint32_t j, uint32_t c, TV m[4])
{
  const size_t o = 16 * (size_t)j + 4 * c;
  m[0] = (TV)mat_ld(a, o);
  m[1] = (TV)mat_ld(a, o + 1);
  m[2] = (TV)mat_ld(a, o + 2);
  m[3] = (TV)mat_ld(a, o + 3);
}

/* the four components of row `col` of a vector being produced in this
 * launch, once none of them holds the sentinel; false if the lane gave up */
template <typename TV>
__device__ __forceinline__ bool wait_vec(const TV *v, uint32_t col, typename sentinel<TV>::bits s, TV x[4],
                                         uint32_t *err, uint32_t sleep)
{
  const TV *p = v + 4 * (size_t)col;
  for (uint32_t n = 0;; n++) {
    x[0] = ld_coherent(p); x[1] = ld_coherent(p + 1); x[2] = ld_coherent(p + 2); x[3] = ld_coherent(p + 3);
    if (sentinel<TV>::raw(x[0]) != s && sentinel<TV>::raw(x[1]) != s && sentinel<TV>::raw(x[2]) != s &&
        sentinel<TV>::raw(x[3]) != s)
      return true;
    if (n >= KF_SPIN_LIMIT) {
      atomicOr(err, KF_ERR_SPIN_LIMIT);
      return false;
    }
    for (uint32_t z = 0; z < sleep; z++) __builtin_amdgcn_s_sleep(1);
  }
}

/* the four values at p in the wave's LDS area (every operand offset is a
 * multiple of 4 values from a 16-byte aligned base): 128-bit reads */
__device__ __forceinline__ void lds_get4(const float *p, float x[4])
{
  const float4 v = *reinterpret_cast<const float4 *>(p);
  x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
}

__device__ __forceinline__ void lds_get4(const double *p, double x[4])
{
  const double2 *q = reinterpret_cast<const double2 *>(p);
  const double2 u = q[0], v = q[1];
  x[0] = u.x; x[1] = u.y; x[2] = v.x; x[3] = v.y;
}

/* Per-wave LDS: the task's row results by slot, the external operands of
 * the prefetched entries by row slot and entry, and a zero block that
 * entries past the end of a row read. */
#define KF_TRSV_EXT 64                           /* first external operand */
#define KF_TRSV_ZERO (64 + KF_GROUP * KF_TRSV_PF * 4)
#define KF_TRSV_LDS (KF_TRSV_ZERO + 4)

/* Component c of x at the prefetched columns in the mask pend, into xo,
 * retrying those that still hold the sentinel; all loads of one round are
 * in flight together.  Returns false if the lane gave up. */
template <int PF, typename TV>
__device__ __forceinline__ bool wait_own(const TV *v, const uint32_t cp[PF], uint32_t pend, uint32_t c,
                                         typename sentinel<TV>::bits s, TV xo[PF], uint32_t *err,
                                         uint32_t sleep)
{
  for (uint32_t n = 0; pend; n++) {
#pragma unroll
    for (int k = 0; k < PF; k++)
      if (pend & (1u << k)) xo[k] = ld_coherent(v + 4 * (size_t)cp[k] + c);
#pragma unroll
    for (int k = 0; k < PF; k++)
      if ((pend & (1u << k)) && sentinel<TV>::raw(xo[k]) != s) pend &= ~(1u << k);
    if (!pend) break;
    if (n >= KF_SPIN_LIMIT) {
      atomicOr(err, KF_ERR_SPIN_LIMIT);
      return false;
    }
    for (uint32_t z = 0; z < sleep; z++) __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

/* Set-up of a row's sweep part [j0, jend) for one task: blocks of the first
 * PF entries into mp (zero past the end), and per entry the LDS
 * offset of its operand -- the producing slot for an entry of the same task
 * (eslot), else the entry's place in the external area, which this lane
 * fills with component c of the polled value, or the zero block past the
 * end.  Returns false if the lane gave up waiting.  The caller orders the
 * external stores before any read with wave_lds_sync(). */
template <int PF, typename TM, typename TV>
__device__ __forceinline__ bool setup_row(const kf_dv &d, const uint8_t *__restrict__ eslot,
                                          const TM *__restrict__ a, const TV *v, typename sentinel<TV>::bits s,
                                          uint32_t j0, uint32_t jend, uint32_t qd, uint32_t c, TV *lds,
                                          uint32_t off[PF], TV mp[PF][4], bool aborted)
{
  static_assert(PF <= KF_TRSV_PF, "the external area holds KF_TRSV_PF operands per row");
  /* every entry's column, slot and block row are loaded unconditionally --
   * an entry past the end reads position 0, which always exists, and is
   * then set aside -- so that all loads of the row are in flight together
   * (one memory round trip, not one per entry) */
  uint32_t cp[PF], es[PF], pend = 0;
#pragma unroll
  for (int k = 0; k < PF; k++) {
    const uint32_t j = j0 + k < jend ? j0 + k : 0u;   /* an empty backward slot has j0 = 1, jend = 0 */
    cp[k] = d.col[j];
    es[k] = eslot[j];
    ld_row(a, j, c, mp[k]);
  }
  __builtin_amdgcn_sched_barrier(0);   /* every load issued before the first use */
#pragma unroll
  for (int k = 0; k < PF; k++) {
    const bool in = j0 + k < jend;
    if (!in) {
      cp[k] = 0;
      mp[k][0] = mp[k][1] = mp[k][2] = mp[k][3] = TV(0);
    }
    /* a column is below N < 2^31, so the second test always holds; it makes
     * the mask depend on the column, whose load then goes out with the
     * others instead of a second round trip inside the polls' branch */
    const bool ext = in && es[k] == 0xffu && cp[k] != 0xffffffffu;
    pend |= ext ? 1u << k : 0u;
    off[k] = !in ? (uint32_t)KF_TRSV_ZERO : ext ? (uint32_t)(KF_TRSV_EXT + 4 * (KF_TRSV_PF * qd + k)) : 4 * es[k];
  }
  TV xo[PF];
  bool ok = true;
  if (!aborted) ok = wait_own<PF, TV>(v, cp, pend, c, s, xo, d.err, d.spin_sleep);
#pragma unroll
  for (int k = 0; k < PF; k++)
    if (pend & (1u << k)) lds[off[k] + c] = xo[k];
  return ok;
}

/* x_r minus the products of its entries [j0, jend) in ascending order.  The
 * operands of the prefetched entries are read from LDS all together (an
 * entry past the end reads the zero block: its product is +0 and the
 * subtraction leaves x_r unchanged), the products are formed side by side,
 * and only the subtractions run in order; entries past the prefetch are
 * loaded (externals polled) here.  One straight-line body for every row:
 * a level costs about the same whatever its entries, so nothing is gained
 * by forming fewer products per level. */
template <int PF, typename TM, typename TV>
__device__ __forceinline__ TV row_sweep(const kf_dv &d, const uint8_t *__restrict__ eslot, const TM *__restrict__ a,
                                        const TV *v, typename sentinel<TV>::bits s, uint32_t j0, uint32_t jend,
                                        uint32_t c, const uint32_t off[PF], const TV mp[PF][4],
                                        const TV *lds, TV xr, bool &aborted)
{
  TV xin[PF][4], pr[PF];
#pragma unroll
  for (int k = 0; k < PF; k++) lds_get4(lds + off[k], xin[k]);
  __builtin_amdgcn_sched_barrier(0);   /* every LDS read issued before the first product */
  /* the 8 products row_dot(mp[k], xin[k]) stage by stage, so that each
   * stage's 8 independent operations hide one another's latency; every
   * product is summed in row_dot's order (0 + m0 x0, + m1 x1, ...) */
#pragma unroll
  for (int k = 0; k < PF; k++) pr[k] = mp[k][0] * xin[k][0];
#pragma unroll
  for (int k = 0; k < PF; k++) pr[k] = TV(0) + pr[k];
#pragma unroll
  for (int i = 1; i < 4; i++) {
    TV t[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) t[k] = mp[k][i] * xin[k][i];
#pragma unroll
    for (int k = 0; k < PF; k++) pr[k] = pr[k] + t[k];
  }
#pragma unroll
  for (int k = 0; k < PF; k++) xr = xr - pr[k];
  for (uint32_t j = j0 + PF; j < jend; j++) {
    TV m[4];
    ld_row(a, j, c, m);
    const uint32_t e = eslot[j];
    if (e == 0xffu) {
      TV xc[4] = {TV(0), TV(0), TV(0), TV(0)};
      if (!aborted) aborted = !wait_vec(v, d.col[j], s, xc, d.err, d.spin_sleep);
      xr = xr - row_dot(m, xc[0], xc[1], xc[2], xc[3]);
    } else {
      const TV *xc = lds + 4 * e;
      xr = xr - row_dot(m, xc[0], xc[1], xc[2], xc[3]);
    }
  }
  return xr;
}

/* forward sweep: y = L^-1 b; presets the backward output x of its rows to
 * the backward sentinel */
template <int PF, typename TM, typename TV, bool FD>
__global__ __launch_bounds__(64 * KF_TRSV_GMAX) void k_trsv_lower(kf_dv d, const uint32_t *__restrict__ trow,
                                                   const uint8_t *__restrict__ tlvl, const uint8_t *__restrict__ tdep,
                                                   const uint8_t *__restrict__ eslot, uint32_t ntask, uint32_t *ticket,
                                                   uint32_t base, typename sentinel<TV>::bits sy,
                                                   typename sentinel<TV>::bits sx, const TM *__restrict__ a,
                                                   const double *__restrict__ b, TV *y, TV *xnext, kf_fdb fd)
{
  __shared__ __align__(16) TV lds_all[KF_TRSV_GMAX][KF_TRSV_LDS];
  __shared__ uint32_t sbatch;
  const uint32_t G = blockDim.x >> 6, wv = threadIdx.x >> 6;
  TV *lds = lds_all[wv];
  const uint32_t lane = threadIdx.x & 63u, qd = lane >> 2, c = lane & 3;
  bool aborted = false;
  /* |x| / h of the finite-difference product: the same for every row (its
   * norm was formed before this launch) */
  const double am = FD ? *fd.xnorm / fd.h : 0.0;
  if (lane < 4) lds[KF_TRSV_ZERO + lane] = TV(0);
  if (threadIdx.x == 0) sbatch = atomicAdd(ticket, G) - base;
  __syncthreads();
  uint32_t bstart = sbatch;
  for (;;) {
    if (bstart >= ntask) break;   /* uniform over the workgroup */
    uint32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(ticket, G);   /* the next batch, read after the tasks */
    const uint32_t t = bstart + wv;
    if (t < ntask) {
      const uint32_t *rt = trow + 3 * ((size_t)t * KF_GROUP + qd);
      const uint32_t r = rt[0], j0 = rt[1], jd = rt[2];
      const uint32_t my = tlvl[(size_t)t * KF_GROUP + qd], dep = tdep[t];
      const bool active = r != ~0u;
      /* the right-hand side, loaded for every slot (an empty one reads row
       * 0 and keeps 0) beside the row's entries below */
      const size_t o = 4 * (size_t)(active ? r : 0u) + c;
      TV xr;
      if (FD) {
        /* b is R(w): the finite-difference product first, as
         * src/kernel.c:33-44 forms it, ((R(w) - r) + cdt (w - q)) |x| / h */
        const double cd = fd.cdt[active ? r : 0u];
        xr = (TV)((b[o] - fd.r[o] + cd * (fd.w[o] - fd.q0[o])) * am);
      } else {
        xr = (TV)b[o];
      }
      xr = active ? xr : TV(0);
      if (active) xnext[4 * (size_t)r + c] = sentinel<TV>::val(sx);
      uint32_t off[PF];
      TV mp[PF][4];
      if (!setup_row<PF, TM, TV>(d, eslot, a, y, sy, j0, jd, qd, c, lds, off, mp, aborted)) aborted = true;
      aborted = __any(aborted);   /* a wave gives up at most once per launch */
      wave_lds_sync();
      for (uint32_t l = 0; l < dep; l++) {
        if (my == l) {   /* the quad's four lanes take this branch together */
          xr = row_sweep<PF, TM, TV>(d, eslot, a, y, sy, j0, jd, c, off, mp, lds, xr, aborted);
          xr = canon(xr);
          lds[lane] = xr;
        }
        wave_lds_sync();
      }
      if (active) st_coherent(y + 4 * (size_t)r + c, xr);
    }
    __syncthreads();   /* every task of the batch done; sbatch free */
    if (threadIdx.x == 0) sbatch = nxt - base;
    __syncthreads();
    bstart = sbatch;
  }
}

/* backward sweep: x = U^-1 y (y from the forward sweep); resets y of its rows
 * to the forward sentinel of the next call */
template <int PF, typename TM, typename TV>
__global__ __launch_bounds__(64 * KF_TRSV_GMAX) void k_trsv_upper(kf_dv d, const uint32_t *__restrict__ trow,
                                                   const uint8_t *__restrict__ tlvl, const uint8_t *__restrict__ tdep,
                                                   const uint8_t *__restrict__ eslot, uint32_t ntask, uint32_t *ticket,
                                                   uint32_t base, typename sentinel<TV>::bits sx,
                                                   typename sentinel<TV>::bits synext, const TM *__restrict__ a,
                                                   TV *y, TV *x, double *__restrict__ xout)
{
  __shared__ __align__(16) TV lds_all[KF_TRSV_GMAX][KF_TRSV_LDS];
  __shared__ uint32_t sbatch;
  const uint32_t G = blockDim.x >> 6, wv = threadIdx.x >> 6;
  TV *lds = lds_all[wv];
  const uint32_t lane = threadIdx.x & 63u, qd = lane >> 2, c = lane & 3;
  bool aborted = false;
  if (lane < 4) lds[KF_TRSV_ZERO + lane] = TV(0);
  if (threadIdx.x == 0) sbatch = atomicAdd(ticket, G) - base;
  __syncthreads();
  uint32_t bstart = sbatch;
  for (;;) {
    if (bstart >= ntask) break;   /* uniform over the workgroup */
    uint32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(ticket, G);   /* the next batch, read after the tasks */
    const uint32_t t = bstart + wv;
    if (t < ntask) {
      const uint32_t *rt = trow + 3 * ((size_t)t * KF_GROUP + qd);
      const uint32_t rr = rt[0], jd = rt[1], j0 = jd + 1, j1 = rt[2];
      const uint32_t my = tlvl[(size_t)t * KF_GROUP + qd], dep = tdep[t];
      const bool active = rr != ~0u;
      const uint32_t r = active ? rr : 0;
      /* the diagonal inverse and y, loaded for every slot (an empty one
       * reads row 0 and discards it) beside the row's entries below */
      TV dg[4];
      ld_row(a, active ? jd : 0u, c, dg);
      TV x0 = y[4 * (size_t)r + c];
      x0 = active ? x0 : TV(0);
      if (active) y[4 * (size_t)r + c] = sentinel<TV>::val(synext);
      uint32_t off[PF];
      TV mp[PF][4];
      if (!setup_row<PF, TM, TV>(d, eslot, a, x, sx, j0, j1, qd, c, lds, off, mp, aborted)) aborted = true;
      aborted = __any(aborted);
      wave_lds_sync();
      TV res = TV(0);
      for (uint32_t l = 0; l < dep; l++) {
        if (my == l) {   /* the quad's four lanes take this branch together */
          const TV xr = row_sweep<PF, TM, TV>(d, eslot, a, x, sx, j0, j1, c, off, mp, lds, x0, aborted);
          /* x_r = inv(D_r) x_r: the quad exchanges its four components */
          const TV y0 = quad_get<0>(xr), y1 = quad_get<1>(xr), y2 = quad_get<2>(xr), y3 = quad_get<3>(xr);
          res = canon(row_dot(dg, y0, y1, y2, y3));
          lds[lane] = res;
        }
        wave_lds_sync();
      }
      if (active) {
        st_coherent(x + 4 * (size_t)r + c, res);
        if ((void *)xout != (void *)x) xout[4 * (size_t)r + c] = (double)res;
      }
    }
    __syncthreads();   /* every task of the batch done; sbatch free */
    if (threadIdx.x == 0) sbatch = nxt - base;
    __syncthreads();
    bstart = sbatch;
  }
}

/* forward output y (held between calls, preset to the next call's forward
 * sentinel by the previous backward sweep) and backward output x */
template <int PF, typename TM, typename TV>
static void launch_trsv(kf_ctx *c, const TM *a, const double *b, TV *y, int yslot, TV *x, double *xout,
                        const kf_fdb *fd)
{
  const uint32_t G = c->trsv_group;
  /* persistent workgroups of G wavefronts: one atomic per G tickets */
  /* waves: the context's setting, else the schedule's size capped by the
   * resident slots of these kernels -- the FP64 ones holding 8 entries per
   * row fit two waves per SIMD (184-186 VGPRs: 2048 on 256 CUs x 4 SIMDs),
   * those holding 7 three (160-168 VGPRs: 3072; 0.97 / 1.06 ms on the
   * 2.65M-vertex lattice against 1.02 / 1.10 ms with 2048), the FP32 ones
   * four, of which 3072 measured best on that lattice (0.86 / 0.88 ms
   * against 0.95 / 0.95 ms with 2048 and 0.88 / 0.91 ms with 4096) */
  const uint32_t cap = (sizeof(TV) == 8 && PF == KF_TRSV_PF) ? 2048u : 3072u;
  const uint32_t W = c->trsv_waves ? c->trsv_waves : (c->trsv_auto < cap ? c->trsv_auto : cap);
  const uint32_t wl = c->ntask_l < W ? c->ntask_l : W;
  const uint32_t wu = c->ntask_u < W ? c->ntask_u : W;
  const uint32_t nwl = (wl + G - 1) / G, nwu = (wu + G - 1) / G;
  const uint64_t k = c->tcall;
  const size_t n = 4 * (size_t)c->d.N;
  if (c->ystate[yslot] != k) {
    hipLaunchKernelGGL((k_fill_bits<TV>), dim3(kf_blocks_for(n, 256)), dim3(256), 0, c->stream, y, n,
                       sentinel<TV>::of(2 * k));
    KF_LAUNCH_CHECK();
  }
  int s = kf_prof_begin(c, KP_TRSV_L);
  const kf_fdb none = {NULL, NULL, NULL, NULL, NULL, 1.0};
  if (fd)
    hipLaunchKernelGGL((k_trsv_lower<PF, TM, TV, true>), dim3(nwl), dim3(64 * G), 0, c->stream, c->d, c->trow_l,
                       c->tlvl_l, c->tdep_l, c->eslot, c->ntask_l, c->ticket + 0, c->ticket_base[0],
                       sentinel<TV>::of(2 * k), sentinel<TV>::of(2 * k + 1), a, b, y, x, *fd);
  else
    hipLaunchKernelGGL((k_trsv_lower<PF, TM, TV, false>), dim3(nwl), dim3(64 * G), 0, c->stream, c->d, c->trow_l,
                       c->tlvl_l, c->tdep_l, c->eslot, c->ntask_l, c->ticket + 0, c->ticket_base[0],
                       sentinel<TV>::of(2 * k), sentinel<TV>::of(2 * k + 1), a, b, y, x, none);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  c->ticket_base[0] += kf_ticket_advance_groups(c->ntask_l, nwl, G);
  s = kf_prof_begin(c, KP_TRSV_U);
  hipLaunchKernelGGL((k_trsv_upper<PF, TM, TV>), dim3(nwu), dim3(64 * G), 0, c->stream, c->d, c->trow_u, c->tlvl_u,
                     c->tdep_u, c->eslot, c->ntask_u, c->ticket + 1, c->ticket_base[1], sentinel<TV>::of(2 * k + 1),
                     sentinel<TV>::of(2 * k + 2), a, y, x, xout);
  KF_LAUNCH_CHECK();
  kf_prof_end(c, s);
  c->ticket_base[1] += kf_ticket_advance_groups(c->ntask_u, nwu, G);
  c->tcall = k + 1;
  c->ystate[yslot] = k + 1;
}

/* whether the sweeps run at their resident-wave cap (the schedule offers
 * more tasks per level than fit: throughput-bound, as on the 2.65M-vertex
 * lattice) rather than below it (latency-bound along the task chain, as on
 * the 0.33M-vertex one) */
bool kf_trsv_capped(const kf_ctx *c)
{
  const uint32_t cap = (c->opt.prec == KF_PREC_FP64 && c->trsv_pf == KF_TRSV_PF) ? 2048u : 3072u;
  return c->trsv_waves ? c->trsv_waves >= cap : c->trsv_auto >= cap;
}

void kf_launch_sptrsv(kf_ctx *c, const double *b, double *x, const kf_fdb *fd)
{
  kf_check_aligned(b, "b"); kf_check_aligned(x, "x");
  const size_t n = 4 * (size_t)c->d.N;
  if (c->opt.prec == KF_PREC_FP64) {
    if (!c->ty64) KF_CHECK(hipMalloc((void **)&c->ty64, n * sizeof(double)));
    if (c->trsv_pf == 7) launch_trsv<7, double, double>(c, c->a64, b, c->ty64, 0, x, x, fd);
    else launch_trsv<KF_TRSV_PF, double, double>(c, c->a64, b, c->ty64, 0, x, x, fd);
    return;
  }
  if (!c->f32a) KF_CHECK(hipMalloc((void **)&c->f32a, n * sizeof(float)));
  if (!c->f32b) KF_CHECK(hipMalloc((void **)&c->f32b, n * sizeof(float)));
  const bool p7 = c->trsv_pf == 7;
  if (c->opt.prec == KF_PREC_FP32) {
    if (p7) launch_trsv<7, float, float>(c, c->a32, b, c->f32a, 1, c->f32b, x, fd);
    else launch_trsv<KF_TRSV_PF, float, float>(c, c->a32, b, c->f32a, 1, c->f32b, x, fd);
  } else {
    if (p7) launch_trsv<7, uint16_t, float>(c, c->abf, b, c->f32a, 1, c->f32b, x, fd);
    else launch_trsv<KF_TRSV_PF, uint16_t, float>(c, c->abf, b, c->f32a, 1, c->f32b, x, fd);
  }
}
