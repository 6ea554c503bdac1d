/*
 * ilu.hip -- block ILU(0) of the BSR Jacobian, in place: strictly lower
 * blocks become L (unit diagonal implied), upper blocks U, and the diagonal
 * holds inv(U(i,i)), the storage the reference's ilu_Edmond leaves in A
 * (matrix/ilu.cpp:49-138) and its triangular solve expects.
 *
 * The reference's one-thread sweep is sequential ILU(0) in natural row
 * order: for each row i and each column k of its pattern, ascending,
 *   a = A(i,k) - sum_{m < min(i,k)} L(i,m) U(m,k)   (m ascending)
 *   L(i,k) = a inv(U(k,k))  if k < i,   U(i,k) = a  otherwise.
 * Here row i runs on one wavefront as soon as the rows of its lower
 * pattern are final (level schedule of matrix/init.cpp:194-232, sync-free:
 * see sync.h).  The row is staged in LDS; for each lower block m ascending
 * 16 lanes form L(i,m) = a(i,m) inv(U(m,m)) and then 4 x 16 lanes subtract
 * L(i,m) U(m,k) from every matching block of the row, so each block sees
 * its updates in ascending m, in the reference's block-product order
 * (matrix/init.cpp:105-116).  16 lanes then form the cofactor inverse of the
 * diagonal block, one lane per entry, with the expressions of
 * matrix/inverse.cpp:195-352.
 */
#include "sync.h"
#include "handoff.h"

/* matrix/inverse.cpp:200-326: entry e of the adjugate is
 * s0*(a b c) s1*(...) ... six signed triple products summed left to right */
__constant__ uint8_t kf_cof_idx[16][6][3] = {
  {{5, 10, 15}, {5, 11, 14}, {9, 6, 15}, {9, 7, 14}, {13, 6, 11}, {13, 7, 10}},
  {{1, 10, 15}, {1, 11, 14}, {9, 2, 15}, {9, 3, 14}, {13, 2, 11}, {13, 3, 10}},
  {{1, 6, 15}, {1, 7, 14}, {5, 2, 15}, {5, 3, 14}, {13, 2, 7}, {13, 3, 6}},
  {{1, 6, 11}, {1, 7, 10}, {5, 2, 11}, {5, 3, 10}, {9, 2, 7}, {9, 3, 6}},
  {{4, 10, 15}, {4, 11, 14}, {8, 6, 15}, {8, 7, 14}, {12, 6, 11}, {12, 7, 10}},
  {{0, 10, 15}, {0, 11, 14}, {8, 2, 15}, {8, 3, 14}, {12, 2, 11}, {12, 3, 10}},
  {{0, 6, 15}, {0, 7, 14}, {4, 2, 15}, {4, 3, 14}, {12, 2, 7}, {12, 3, 6}},
  {{0, 6, 11}, {0, 7, 10}, {4, 2, 11}, {4, 3, 10}, {8, 2, 7}, {8, 3, 6}},
  {{4, 9, 15}, {4, 11, 13}, {8, 5, 15}, {8, 7, 13}, {12, 5, 11}, {12, 7, 9}},
  {{0, 9, 15}, {0, 11, 13}, {8, 1, 15}, {8, 3, 13}, {12, 1, 11}, {12, 3, 9}},
  {{0, 5, 15}, {0, 7, 13}, {4, 1, 15}, {4, 3, 13}, {12, 1, 7}, {12, 3, 5}},
  {{0, 5, 11}, {0, 7, 9}, {4, 1, 11}, {4, 3, 9}, {8, 1, 7}, {8, 3, 5}},
  {{4, 9, 14}, {4, 10, 13}, {8, 5, 14}, {8, 6, 13}, {12, 5, 10}, {12, 6, 9}},
  {{0, 9, 14}, {0, 10, 13}, {8, 1, 14}, {8, 2, 13}, {12, 1, 10}, {12, 2, 9}},
  {{0, 5, 14}, {0, 6, 13}, {4, 1, 14}, {4, 2, 13}, {12, 1, 6}, {12, 2, 5}},
  {{0, 5, 10}, {0, 6, 9}, {4, 1, 10}, {4, 2, 9}, {8, 1, 6}, {8, 2, 5}}};
/* bit t set: term t is subtracted (term 0: its first factor is negated) */
__constant__ uint8_t kf_cof_neg[16] = {0x26, 0x19, 0x26, 0x19, 0x19, 0x26, 0x19, 0x26,
                                      0x26, 0x19, 0x26, 0x19, 0x19, 0x26, 0x19, 0x26};

#define ILU_MAXROW 128

/* d = A B for 4x4 row-major blocks, entry (r, c) of lane e = 4 r + c, with the
 * reference's summation order (matrix/init.cpp:105-116) */
template <typename T>
__device__ __forceinline__ T blk_mul(const T *A, const T *B, uint32_t r, uint32_t c)
{
  T s = T(0);
#pragma unroll
  for (int k = 0; k < 4; k++) s += A[4 * r + k] * B[4 * k + c];
  return s;
}

/* EXACT: the reference's one-thread sweep bit for bit.  Each block it forms
 * ends with the product of the last common pair of its merge (matrix/ilu.cpp:
 * 71-86): subtracted, then added back, which rounds.  For a lower block (i,k)
 * that pair is L_old(i,k) U(k,k), where L_old is what the L copy held before
 * this sweep -- the previous sweep's L(i,k), or, for the first lower slot of
 * row i, the Jacobian block A(i, k_last) that matrix/matrix.cpp:145 copies
 * there during the fill (k_last = the row's largest lower column, the last
 * lower block the fill touches); for an upper block it is I A(i,k).  lold
 * keeps the previous sweep's L blocks (BSR positions), ud the non-inverted
 * U(k,k). */
template <typename T, bool EXACT>
__global__ __launch_bounds__(64) void k_ilu(kf_dv d, const uint32_t *__restrict__ rows, uint32_t nrows,
                                            uint32_t *ticket, uint32_t base, uint32_t *rdone, uint32_t epoch,
                                            T *a, T *lold, T *ud)
{
  __shared__ uint32_t scol[ILU_MAXROW];
  __shared__ T sa[ILU_MAXROW * 16];
  __shared__ T sq[16];
  const uint32_t lane = threadIdx.x;
  const uint32_t e = lane & 15, er = e >> 2, ec = e & 3, grp = lane >> 4;
  bool aborted = false;
  for (;;) {
    const uint32_t t = next_ticket(ticket, base);
    if (t >= nrows) break;
    const uint32_t i = rows[t];
    const uint32_t j0 = d.rowptr[i], j1 = d.rowptr[i + 1], jd = d.diag[i];
    const uint32_t L = j1 - j0, nlow = jd - j0, pd = jd - j0;
    for (uint32_t k = lane; k < L; k += 64) scol[k] = d.col[j0 + k];
    for (uint32_t k = lane; k < 16 * L; k += 64) sa[k] = a[16 * (size_t)j0 + k];
    __syncthreads();
    for (uint32_t k = lane; k < nlow; k += 64) wait_row(rdone, scol[k], epoch, d.err, aborted, d.spin_sleep);
    aborted = __any(aborted);
    consume_fence();
    __syncthreads();

    for (uint32_t p = 0; p < nlow; p++) {
      const uint32_t m = scol[p];
      const T *dm = a + 16 * (size_t)d.diag[m];
      if (EXACT) {
        /* a(i,m) = (a - L_old U(m,m)) + L_old U(m,m) */
        const T *lo = p == 0 ? a + 16 * ((size_t)j0 + nlow - 1) : lold + 16 * ((size_t)j0 + p);
        if (lane < 16) sq[e] = lo[e];
        __syncthreads();
        T dq = T(0);
        if (lane < 16) dq = blk_mul(sq, ud + 16 * (size_t)m, er, ec);
        __syncthreads();
        if (lane < 16) {
          T v = sa[16 * p + e];
          v = T(-1) * dq + v;
          v = T(1) * dq + v;
          sa[16 * p + e] = v;
        }
        __syncthreads();
      }
      /* L(i,m) = a(i,m) inv(U(m,m)) */
      T lv = T(0);
      if (lane < 16) lv = blk_mul(sa + 16 * p, dm, er, ec);
      __syncthreads();
      if (lane < 16) sa[16 * p + e] = lv;
      __syncthreads();
      /* a(i,k) -= L(i,m) U(m,k) for the upper blocks (m,k) of row m in row i */
      const uint32_t m1 = d.rowptr[m + 1];
      for (uint32_t jm = d.diag[m] + 1 + grp; jm < m1; jm += 4) {
        const uint32_t kc = d.col[jm];
        uint32_t lo = 0, hi = L;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (scol[mid] < kc) lo = mid + 1;
          else hi = mid;
        }
        if (lo < L && scol[lo] == kc) {
          const T s = blk_mul(sa + 16 * p, a + 16 * (size_t)jm, er, ec);
          sa[16 * lo + e] = sa[16 * lo + e] - s;
        }
      }
      __syncthreads();
    }

    if (EXACT) {
      /* upper blocks: a = (a - I A(i,k)) + I A(i,k), A = the fill's block */
      for (uint32_t p = pd + grp; p < L; p += 4) {
        const T *a0 = a + 16 * ((size_t)j0 + p);
        T dq = T(0);
#pragma unroll
        for (int k = 0; k < 4; k++) dq += (er == (uint32_t)k ? T(1) : T(0)) * a0[4 * k + ec];
        T v = sa[16 * p + e];
        v = T(-1) * dq + v;
        v = T(1) * dq + v;
        sa[16 * p + e] = v;
      }
      __syncthreads();
      if (lane < 16) ud[16 * (size_t)i + e] = sa[16 * pd + e];
    }

    /* diagonal: cofactor inverse, one lane per entry (lanes 16..63 mirror
     * lanes 0..15 so that the shuffles below read defined values) */
    {
      const T *s = sa + 16 * pd;
      const uint8_t neg = kf_cof_neg[e];
      T acc = s[kf_cof_idx[e][0][0]];
      if (neg & 1u) acc = -acc;
      acc = acc * s[kf_cof_idx[e][0][1]] * s[kf_cof_idx[e][0][2]];
#pragma unroll
      for (int tt = 1; tt < 6; tt++) {
        const T pr = s[kf_cof_idx[e][tt][0]] * s[kf_cof_idx[e][tt][1]] * s[kf_cof_idx[e][tt][2]];
        acc = (neg & (1u << tt)) ? acc - pr : acc + pr;
      }
      const T d0 = __shfl(acc, 0, 16), d4 = __shfl(acc, 4, 16), d8 = __shfl(acc, 8, 16), d12 = __shfl(acc, 12, 16);
      T det = s[0] * d0 + s[1] * d4 + s[2] * d8 + s[3] * d12;
      det = T(1) / det;
      acc *= det;
      __syncthreads();
      if (lane < 16) {
        if (!isfinite((double)acc)) atomicOr(d.err, KF_ERR_NONFINITE_PIVOT);
        sa[16 * pd + e] = acc;
      }
    }
    __syncthreads();
    for (uint32_t k = lane; k < 16 * L; k += 64) a[16 * (size_t)j0 + k] = sa[k];
    if (EXACT)
      for (uint32_t k = lane; k < 16 * nlow; k += 64) lold[16 * (size_t)j0 + k] = sa[k];
    publish_fence();
    __syncthreads();
    if (lane == 0) flag_store(rdone + i, epoch);
  }
}

/* Shared-memory layout of k_ilu_rows, sized on the host from the plan:
 * maxrow blocks per row, maxlow lower blocks per row, maxmatch update pairs
 * per row. */
struct ilu_layout {
  uint32_t scol, sa, sa0, sldg, smoff, sjm, slo, stage, bytes;
  __host__ __device__ ilu_layout(uint32_t maxrow, uint32_t maxlow, uint32_t maxmatch, uint32_t tsize, bool exact)
  {
    uint32_t o = 0;
    auto take = [&o](uint32_t n) { const uint32_t r = o; o += (n + 15u) & ~15u; return r; };
    scol = take(4 * maxrow);
    sa = take(16 * tsize * maxrow);
    sa0 = exact ? take(16 * tsize * maxrow) : 0;
    sldg = take(4 * maxlow);
    smoff = take(4 * (maxlow + 1));
    sjm = take(4 * maxmatch);
    slo = take(2 * maxmatch);
    stage = take(16 * tsize * ((exact ? 3 : 1) * maxlow + maxmatch));
    bytes = o;
  }
};

#define ILU_Q 4    /* values per lane per load round before the wait */
#define ILU_QB 12  /* values per lane per staging round after the wait */

/* k_ilu with the row's schedule precomputed on the host (kf_plan_ilu): a
 * record per ticket, the diagonal position of every row it reads from and
 * the list of update pairs (U(m,k) block, position of k in row i).  The wave
 * loads, in one round, its row, those tables and its L_old blocks, then
 * stages inv(U(m,m)), U(m,m) and the U(m,k) blocks in LDS from the
 * published copies (apub, udpub), retrying values that still hold the
 * call's sentinel (handoff.h): no flags, no fences.  The rest runs from
 * LDS; the row then stores its factors in place and publishes its diagonal
 * inverse, upper blocks and U(i,i).  Arithmetic and its order are those of
 * k_ilu. */
template <typename T, bool EXACT>
__global__ __launch_bounds__(256) void k_ilu_rows(kf_dv d, const uint32_t *__restrict__ rec, uint32_t nrows,
                                                 uint32_t *ticket, uint32_t base, T *apub, T *udpub,
                                                 typename sentinel<T>::bits snt, T *a, T *lold,
                                                 const uint32_t *__restrict__ ldiag,
                                                 const uint32_t *__restrict__ moff, const uint32_t *__restrict__ mjm,
                                                 const uint16_t *__restrict__ mlo, uint32_t maxrow, uint32_t maxlow,
                                                 uint32_t maxmatch)
{
  extern __shared__ __align__(16) unsigned char ilu_lds_all[];
  __shared__ uint32_t sbatch;
  const ilu_layout lay(maxrow, maxlow, maxmatch, sizeof(T), EXACT);
  const uint32_t G = blockDim.x >> 6, wv = threadIdx.x >> 6;
  unsigned char *ilu_lds = ilu_lds_all + (size_t)wv * lay.bytes;   /* this wave's region */
  uint32_t *scol = (uint32_t *)(ilu_lds + lay.scol);
  T *sa = (T *)(ilu_lds + lay.sa);
  T *sa0 = (T *)(ilu_lds + lay.sa0);
  uint32_t *sldg = (uint32_t *)(ilu_lds + lay.sldg);
  uint32_t *smoff = (uint32_t *)(ilu_lds + lay.smoff);
  uint32_t *sjm = (uint32_t *)(ilu_lds + lay.sjm);
  uint16_t *slo = (uint16_t *)(ilu_lds + lay.slo);
  T *stage = (T *)(ilu_lds + lay.stage);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t e = lane & 15, er = e >> 2, ec = e & 3, grp = lane >> 4;
  bool aborted = false;
  /* G consecutive tickets per atomic, one row per wavefront; the workgroup
   * takes the next G when all its rows are done (sync.h) */
  if (threadIdx.x == 0) sbatch = atomicAdd(ticket, G) - base;
  __syncthreads();
  uint32_t bstart = sbatch;
  for (;;) {
    if (bstart >= nrows) break;   /* uniform over the workgroup */
    uint32_t nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(ticket, G);
    const uint32_t t = bstart + wv;
    if (t < nrows) {
    const uint32_t rv = lane < 6 ? rec[8 * (size_t)t + lane] : 0;
    const uint32_t i = __shfl(rv, 0, 64), j0 = __shfl(rv, 1, 64), L = __shfl(rv, 2, 64);
    const uint32_t nlow = __shfl(rv, 3, 64), mbeg = __shfl(rv, 4, 64), mcnt = __shfl(rv, 5, 64);
    const uint32_t pd = nlow;
    /* staged slots: [0, nlow) inv(U(m,m)); EXACT: [nlow, 2 nlow) U(m,m),
     * [2 nlow, 3 nlow) L_old; then the U(m,k) blocks of the update pairs */
    const uint32_t mbase = EXACT ? 3 * nlow : nlow;
    /* one round of loads: row i, the tables, L_old (rows beyond the
     * register round take the loops below) */
    {
      T av[ILU_Q], lv[ILU_Q];
      uint32_t cv = 0, gv = 0, mo = 0, jv = 0;
      uint16_t ov = 0;
#pragma unroll
      for (int q = 0; q < ILU_Q; q++) {
        const uint32_t k = lane + 64 * q;
        if (k < 16 * L) av[q] = a[16 * (size_t)j0 + k];
        if (EXACT && k < 16 * nlow) {
          const uint32_t pp = k >> 4, ee = k & 15;
          lv[q] = pp == 0 ? a[16 * ((size_t)j0 + nlow - 1) + ee] : lold[16 * ((size_t)j0 + pp) + ee];
        }
      }
      if (lane < L) cv = d.col[j0 + lane];
      if (lane < nlow) { gv = ldiag[j0 + lane]; mo = moff[j0 + lane]; }
      if (lane < mcnt) { jv = mjm[mbeg + lane]; ov = mlo[mbeg + lane]; }
#pragma unroll
      for (int q = 0; q < ILU_Q; q++) {
        const uint32_t k = lane + 64 * q;
        if (k < 16 * L) {
          sa[k] = av[q];
          if (EXACT) sa0[k] = av[q];
        }
        if (EXACT && k < 16 * nlow) stage[32 * nlow + k] = lv[q];
      }
      if (lane < L) scol[lane] = cv;
      if (lane < nlow) { sldg[lane] = gv; smoff[lane] = mo; }
      if (lane < mcnt) { sjm[lane] = jv; slo[lane] = ov; }
      if (lane == 0) smoff[nlow] = mcnt;
      for (uint32_t k = lane + 64 * ILU_Q; k < 16 * L; k += 64) {
        const T v = a[16 * (size_t)j0 + k];
        sa[k] = v;
        if (EXACT) sa0[k] = v;
      }
      if (EXACT)
        for (uint32_t k = lane + 64 * ILU_Q; k < 16 * nlow; k += 64)
          stage[32 * nlow + k] = lold[16 * ((size_t)j0 + (k >> 4)) + (k & 15)];
      for (uint32_t k = lane + 64; k < L; k += 64) scol[k] = d.col[j0 + k];
      for (uint32_t k = lane + 64; k < nlow; k += 64) { sldg[k] = ldiag[j0 + k]; smoff[k] = moff[j0 + k]; }
      for (uint32_t k = lane + 64; k < mcnt; k += 64) { sjm[k] = mjm[mbeg + k]; slo[k] = mlo[mbeg + k]; }
    }
    wave_lds_sync();
    /* staging from the published copies, ILU_QB values per lane per round;
     * a value still holding the sentinel is loaded again */
    {
      const uint32_t npre = EXACT ? 2 * nlow : nlow;
      const uint32_t tot = 16 * (npre + mcnt);
      for (uint32_t b0 = 0; b0 < tot; b0 += 64 * ILU_QB) {
        uint32_t need = 0;
#pragma unroll
        for (int q = 0; q < ILU_QB; q++)
          if (b0 + lane + 64 * q < tot) need |= 1u << q;
        for (uint32_t n = 0; need && !aborted; n++) {
          if (n > 0) {
            if (n > KF_SPIN_LIMIT) {
              atomicOr(d.err, KF_ERR_SPIN_LIMIT);
              aborted = true;
              break;
            }
            for (uint32_t z = 0; z < d.spin_sleep; z++) __builtin_amdgcn_s_sleep(1);
          }
          T v[ILU_QB];
#pragma unroll
          for (int q = 0; q < ILU_QB; q++)
            if (need & (1u << q)) {
              const uint32_t k = b0 + lane + 64 * q, sl = k >> 4, ee = k & 15;
              const T *src;
              if (sl < nlow) src = apub + 16 * (size_t)sldg[sl];
              else if (EXACT && sl < 2 * nlow) src = udpub + 16 * (size_t)scol[sl - nlow];
              else src = apub + 16 * (size_t)sjm[sl - npre];
              v[q] = ld_coherent(src + ee);
            }
#pragma unroll
          for (int q = 0; q < ILU_QB; q++)
            if ((need & (1u << q)) && sentinel<T>::raw(v[q]) != snt) {
              const uint32_t k = b0 + lane + 64 * q, sl = k >> 4;
              stage[16 * (sl < npre ? sl : sl - npre + mbase) + (k & 15)] = v[q];
              need &= ~(1u << q);
            }
        }
      }
      aborted = __any(aborted);   /* a wave gives up at most once per launch */
    }
    wave_lds_sync();

    for (uint32_t p = 0; p < nlow; p++) {
      if (EXACT) {
        /* a(i,m) = (a - L_old U(m,m)) + L_old U(m,m) */
        if (lane < 16) {
          const T dq = blk_mul(stage + 16 * (2 * nlow + p), stage + 16 * (nlow + p), er, ec);
          T v = sa[16 * p + e];
          v = T(-1) * dq + v;
          v = T(1) * dq + v;
          sa[16 * p + e] = v;
        }
        wave_lds_sync();
      }
      /* L(i,m) = a(i,m) inv(U(m,m)) */
      T lv = T(0);
      if (lane < 16) lv = blk_mul(sa + 16 * p, stage + 16 * p, er, ec);
      wave_lds_sync();
      if (lane < 16) sa[16 * p + e] = lv;
      wave_lds_sync();
      /* a(i,k) -= L(i,m) U(m,k) for the update pairs of m (distinct k) */
      for (uint32_t idx = smoff[p] + grp; idx < smoff[p + 1]; idx += 4) {
        const uint32_t lo = slo[idx];
        const T s = blk_mul(sa + 16 * p, stage + 16 * (size_t)(mbase + idx), er, ec);
        sa[16 * lo + e] = sa[16 * lo + e] - s;
      }
      wave_lds_sync();
    }

    if (EXACT) {
      /* upper blocks: a = (a - I A(i,k)) + I A(i,k), A = the fill's block */
      for (uint32_t p = pd + grp; p < L; p += 4) {
        const T *a0 = sa0 + 16 * p;
        T dq = T(0);
#pragma unroll
        for (int k = 0; k < 4; k++) dq += (er == (uint32_t)k ? T(1) : T(0)) * a0[4 * k + ec];
        T v = sa[16 * p + e];
        v = T(-1) * dq + v;
        v = T(1) * dq + v;
        sa[16 * p + e] = v;
      }
      wave_lds_sync();
      if (lane < 16) st_coherent(udpub + 16 * (size_t)i + e, canon(sa[16 * pd + e]));
    }

    {
      const T *s = sa + 16 * pd;
      const uint8_t neg = kf_cof_neg[e];
      T acc = s[kf_cof_idx[e][0][0]];
      if (neg & 1u) acc = -acc;
      acc = acc * s[kf_cof_idx[e][0][1]] * s[kf_cof_idx[e][0][2]];
#pragma unroll
      for (int tt = 1; tt < 6; tt++) {
        const T pr = s[kf_cof_idx[e][tt][0]] * s[kf_cof_idx[e][tt][1]] * s[kf_cof_idx[e][tt][2]];
        acc = (neg & (1u << tt)) ? acc - pr : acc + pr;
      }
      const T d0 = __shfl(acc, 0, 16), d4 = __shfl(acc, 4, 16), d8 = __shfl(acc, 8, 16), d12 = __shfl(acc, 12, 16);
      T det = s[0] * d0 + s[1] * d4 + s[2] * d8 + s[3] * d12;
      det = T(1) / det;
      acc *= det;
      wave_lds_sync();
      if (lane < 16) {
        if (!isfinite((double)acc)) atomicOr(d.err, KF_ERR_NONFINITE_PIVOT);
        sa[16 * pd + e] = acc;
      }
    }
    wave_lds_sync();
    for (uint32_t k = lane; k < 16 * L; k += 64) a[16 * (size_t)j0 + k] = sa[k];
    if (EXACT)
      for (uint32_t k = lane; k < 16 * nlow; k += 64) lold[16 * (size_t)j0 + k] = sa[k];
    /* published copies of what later rows read: inv(U(i,i)) and U(i,k) */
    for (uint32_t k = lane; k < 16 * (L - pd); k += 64)
      st_coherent(apub + 16 * ((size_t)j0 + pd) + k, canon(sa[16 * pd + k]));
    }
    __syncthreads();   /* every row of the batch done; sbatch free */
    if (threadIdx.x == 0) sbatch = nxt - base;
    __syncthreads();
    bstart = sbatch;
  }
}

static uint32_t persistent_waves(uint32_t work, uint32_t cap)
{
  return work < cap ? (work ? work : 1) : cap;
}

template <typename T, bool EXACT>
static void launch_ilu(kf_ctx *c, T *a, T *lold, T *ud)
{
  /* the context's setting, else 2048 waves for FP64 and 3072 for the
   * single-precision factors (8.8 -> 8.1 ms per factorisation on the
   * 2.65M-vertex lattice; FP64 is flat from 2048 to 4096) */
  const uint32_t nw = persistent_waves(c->d.N, c->ilu_waves ? c->ilu_waves : (sizeof(T) == 8 ? 2048u : 3072u));
  const ilu_layout lay(c->ilu_maxrow, c->ilu_maxlow, c->ilu_maxmatch, sizeof(T), EXACT);
  /* wavefronts per workgroup: as asked, within 4 and 64 KB of LDS */
  uint32_t G = c->ilu_group;
  const size_t lds_cap = c->lds_max < 64 * 1024 ? c->lds_max : 64 * 1024;
  while (G > 1 && G * (size_t)lay.bytes > lds_cap) G--;
  if (lay.bytes <= lds_cap) {
    const uint32_t nb = (nw + G - 1) / G;
    /* published copies, preset to this call's sentinel */
    const size_t na = 16 * c->nnzb, nu = 16 * (size_t)c->d.N;
    if (!c->ilu_apub) KF_CHECK(hipMalloc(&c->ilu_apub, na * sizeof(double)));
    if (EXACT && !c->ilu_udpub) KF_CHECK(hipMalloc(&c->ilu_udpub, nu * sizeof(double)));
    const typename sentinel<T>::bits snt = sentinel<T>::of(++c->icall);
    T *apub = (T *)c->ilu_apub, *udpub = EXACT ? (T *)c->ilu_udpub : NULL;
    hipLaunchKernelGGL((k_fill_bits<T>), dim3(kf_blocks_for(na, 256)), dim3(256), 0, c->stream, apub, na, snt);
    KF_LAUNCH_CHECK();
    if (EXACT) {
      hipLaunchKernelGGL((k_fill_bits<T>), dim3(kf_blocks_for(nu, 256)), dim3(256), 0, c->stream, udpub, nu, snt);
      KF_LAUNCH_CHECK();
    }
    const int s = kf_prof_begin(c, KP_ILU);
    hipLaunchKernelGGL((k_ilu_rows<T, EXACT>), dim3(nb), dim3(64 * G), G * lay.bytes, c->stream, c->d, c->ilu_rec,
                       c->d.N, c->ticket + 2, c->ticket_base[2], apub, udpub, snt, a, lold, c->ilu_ldiag, c->ilu_moff,
                       c->ilu_mjm, c->ilu_mlo, c->ilu_maxrow, c->ilu_maxlow, c->ilu_maxmatch);
    KF_LAUNCH_CHECK();
    kf_prof_end(c, s);
    c->ticket_base[2] += kf_ticket_advance_groups(c->d.N, nb, G);
  } else {
    const uint32_t epoch = ++c->epoch;
    const int s = kf_prof_begin(c, KP_ILU);
    hipLaunchKernelGGL((k_ilu<T, EXACT>), dim3(nw), dim3(64), 0, c->stream, c->d, c->lrows, c->d.N, c->ticket + 2,
                       c->ticket_base[2], c->rdone, epoch, a, lold, ud);
    KF_LAUNCH_CHECK();
    kf_prof_end(c, s);
    c->ticket_base[2] += c->d.N + nw;
  }
}

__global__ __launch_bounds__(256) void k_to_bf16(size_t n, const float *__restrict__ src, uint16_t *__restrict__ dst)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  /* round to nearest even */
  uint32_t u = __float_as_uint(src[i]);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  dst[i] = (uint16_t)(u >> 16);
}

void kf_launch_ilu(kf_ctx *c)
{
  c->jsl_fresh = 0;   /* the Jacobi sweeps' sliced copy is now stale */
  if (c->maxrow > ILU_MAXROW) {
    fprintf(stderr, "kf: ILU supports at most %d blocks per row (mesh has %u)\n", ILU_MAXROW, c->maxrow);
    abort();
  }
  if (c->opt.prec == KF_PREC_FP64) {
    if (c->opt.ilu_exact) {
      if (!c->lold) {
        KF_CHECK(hipMalloc((void **)&c->lold, 16 * c->nnzb * sizeof(double)));
        KF_CHECK(hipMemset(c->lold, 0, 16 * c->nnzb * sizeof(double)));
        KF_CHECK(hipMalloc((void **)&c->ud, 16 * (size_t)c->d.N * sizeof(double)));
      }
      launch_ilu<double, true>(c, c->a64, c->lold, c->ud);
    } else {
      launch_ilu<double, false>(c, c->a64, NULL, NULL);
    }
  } else {
    launch_ilu<float, false>(c, c->a32, NULL, NULL);
    if (c->opt.prec == KF_PREC_BF16) {
      const size_t n = 16 * c->nnzb;
      if (!c->abf) KF_CHECK(hipMalloc((void **)&c->abf, n * sizeof(uint16_t)));
      hipLaunchKernelGGL(k_to_bf16, dim3(kf_blocks_for(n, 256)), dim3(256), 0, c->stream, n, c->a32, c->abf);
      KF_LAUNCH_CHECK();
    }
  }
}
