/*
 * solver.c -- the pseudo-transient Newton-Krylov driver of src/kernel.c on
 * the device context.  Host C: it issues the HIP entry points on the
 * context's stream and keeps only the scalar recurrences on the host (SER
 * CFL, Givens rotations of the (restart+1) x restart Hessenberg, back
 * substitution), exactly as the reference orders them.
 *
 *   iJacVec  src/kernel.c:24-45  matrix-free (J + diag(cdt)) x by forward
 *            difference: w = (h/|x|) x + q, y = ((R(w) - r) + cdt (w - q)) |x|/h
 *   Kernel   src/kernel.c:47-232 pseudo-time loop, left-preconditioned
 *            restarted GMRES with modified Gram-Schmidt (or batched CGS2);
 *            optionally right-preconditioned flexible GMRES (z_j = M(v_j),
 *            x += Z y), the form a varying preconditioner needs
 *
 * The only host/device round trips per Arnoldi step are the restart+2
 * Hessenberg scalars (one copy) needed by the Givens recurrence.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kf_host.h"

#define FD_EPS 1.490116119384766e-08 /* src/kernel.c:12 */

void iJacVec(const GEOMETRY *gc, const double *x, const double *r, const double norm, GRADIENT *grad, double *w,
             double *y)
{
  GEOMETRY *g = (GEOMETRY *)gc;
  const size_t sz = g->c->sz;
  kfs_nrm2(g, sz, x, KFS_SLOT_XNORM);
  const double h = FD_EPS * norm;
  kfs_fd_perturb(g, x, h, KFS_SLOT_XNORM, w);
  ComputeFlux(g, w, grad, y);
  kfs_fd_combine(g, r, w, h, KFS_SLOT_XNORM, y);
}

static double fetch1(GEOMETRY *g, int slot)
{
  double v;
  kfs_fetch(g, slot, 1, &v);
  return v;
}

int kf_kernel(GEOMETRY *g, kf_step_record *rec, int max_rec)
{
  const kf_options o = kf_get_options(g);
  const size_t sz = g->c->sz;
  const int restart = o.gmres_restart, nb = restart + 1;
  const int flexible = o.krylov == KF_KRYLOV_FGMRES;
  double *gxyz[3], *r, *w, *t, *x, *vecs;
  kfs_work(g, flexible ? nb + restart : nb, gxyz, &r, &w, &t, &x, &vecs);
  GRADIENT grad = {gxyz[0], gxyz[1], gxyz[2]};
#define VEC(k) (vecs + (size_t)(k) * sz)
#define ZVEC(k) (vecs + (size_t)(nb + (k)) * sz)   /* FGMRES: z_k = M(v_k) */
  double *ghh = (double *)calloc((size_t)nb * nb + nb, sizeof(double));
  double cc[64], ss[64], rs[64], hcol[64];
  double fnorm_init = 0.0, cfl = o.cfl_init;
  int its = 0, nrec = 0;
  while (its < o.ts_max) {
    /* the residual also leaves the time-step sums of this state */
    kfs_flux_spectral(g, g->q->q, &grad, r);
    kfs_nrm2(g, sz, r, KFS_SLOT_FNORM);
    const double fnorm = fetch1(g, KFS_SLOT_FNORM);
    if (!isfinite(fnorm)) {
      fprintf(stderr, "kf: non-finite residual norm at pseudo-step %d\n", its);
      break;
    }
    if (its == 0) fnorm_init = fnorm;
    else {
      double cfl0 = 1.1 * o.cfl_init * fnorm_init;
      cfl0 /= fnorm;
      cfl = (cfl0 < o.cfl_max) ? cfl0 : o.cfl_max;
    }
    kfs_timestep_spectral(g, cfl);   /* = ComputeTimeStep(cfl, g) on the same state */
    ComputeA(g);
    kfs_nrm2(g, sz, g->q->q, KFS_SLOT_QNORM);
    const double qnorm = sqrt(1.0 + fetch1(g, KFS_SLOT_QNORM));
    kfs_zero(g, sz, x);
    ComputeNumericalILU(g);
    if (!flexible) ComputeSparseTriangularSolve(g, r, VEC(0));
    else kfs_newaxpy(g, sz, 0.0, r, r, VEC(0));   /* v = r (x = 0) */
    double tol = 0.0;
    int total = 0, flag = 0;
    while (total <= o.gmres_max_it) {
      if (total > 0) {
        iJacVec(g, x, r, qnorm, &grad, w, t);
        if (!flexible) {
          kfs_newaxpy(g, sz, -1.0, t, r, w);
          ComputeSparseTriangularSolve(g, w, VEC(0));
        } else {
          kfs_newaxpy(g, sz, -1.0, t, r, VEC(0));   /* v = r - A x */
        }
      }
      kfs_nrm2(g, sz, VEC(0), KFS_SLOT_BETA);
      kfs_scale_recip(g, sz, VEC(0), KFS_SLOT_BETA);
      rs[0] = fetch1(g, KFS_SLOT_BETA);
      tol = (total == 0) ? o.gmres_rtol * rs[0] : tol;
      int it = 0;
      while (it < restart) {
        if (!flexible) {
          iJacVec(g, VEC(it), r, qnorm, &grad, w, t);
          ComputeSparseTriangularSolve(g, t, VEC(it + 1));
        } else {
          ComputeSparseTriangularSolve(g, VEC(it), ZVEC(it));   /* z = M(v) */
          iJacVec(g, ZVEC(it), r, qnorm, &grad, w, VEC(it + 1));
        }
        double *hh = ghh + (size_t)it * (nb + 1);
        if (o.ortho == KF_ORTHO_CGS2) {
          /* two passes of batched classical Gram-Schmidt */
          kfs_cgs_dots(g, sz, VEC(0), it + 1, VEC(it + 1), KFS_SLOT_H);
          kfs_cgs_update(g, sz, VEC(0), it + 1, KFS_SLOT_H, VEC(it + 1));
          kfs_fetch(g, KFS_SLOT_H, it + 1, hcol);
          kfs_cgs_dots(g, sz, VEC(0), it + 1, VEC(it + 1), KFS_SLOT_H);
          kfs_cgs_update(g, sz, VEC(0), it + 1, KFS_SLOT_H, VEC(it + 1));
          kfs_nrm2(g, sz, VEC(it + 1), KFS_SLOT_H + it + 1);
          kfs_scale_recip(g, sz, VEC(it + 1), KFS_SLOT_H + it + 1);
          double h2[64];
          kfs_fetch(g, KFS_SLOT_H, it + 2, h2);
          for (int v = 0; v <= it; v++) hh[v] = hcol[v] + h2[v];
          hh[it + 1] = h2[it + 1];
        } else {
          /* modified Gram-Schmidt, src/kernel.c:129-141 */
          kfs_mgs(g, sz, VEC(0), sz, it + 1, VEC(it + 1), KFS_SLOT_H);
          kfs_scale_recip(g, sz, VEC(it + 1), KFS_SLOT_H + it + 1);
          kfs_fetch(g, KFS_SLOT_H, it + 2, hh);
        }
        /* Givens, src/kernel.c:143-173 */
        for (int i = 0; i < it; i++) {
          const double tt = *hh;
          *hh = cc[i] * *hh + ss[i] * *(hh + 1);
          hh++;
          *hh = cc[i] * *hh - ss[i] * tt;
        }
        const double tt = sqrt(*hh * *hh + *(hh + 1) * *(hh + 1));
        const double ct = *hh / tt, st = *(hh + 1) / tt;
        cc[it] = ct;
        ss[it] = st;
        rs[it + 1] = -(st * rs[it]);
        rs[it] = ct * rs[it];
        *hh = ct * *hh + st * *(hh + 1);
        const double res = fabs(rs[it + 1]);
        it++;
        if (res <= tol) {
          flag = 1;
          break;
        }
      }
      total += it;
      /* back substitution, src/kernel.c:178-188 */
      rs[it - 1] /= ghh[(size_t)(it - 1) * (nb + 1) + (it - 1)];
      for (int i = 1; i <= it - 1; i++) {
        const int k = (it - 1) - i;
        double tt = rs[k];
        for (int j = k + 1; j <= it - 1; j++) tt -= ghh[(size_t)j * (nb + 1) + k] * rs[j];
        rs[k] = tt / ghh[(size_t)k * (nb + 1) + k];
      }
      for (int v = 0; v < it; v++) kfs_axpy(g, sz, rs[v], flexible ? ZVEC(v) : VEC(v), x);
      if (flag) break;
    }
    kfs_axpy(g, sz, -1.0, x, g->q->q);
    double forces[3];
    ComputeForces(g, forces);
    if (o.verbose) {
      printf(">> GMRES: TOTAL ITERATIONS %d\n", total);
      printf("STEP: %d CFL: %g NORM: %g\n", its, cfl, fnorm);
      printf("LIFT: %g DRAG: %g MOMN: %g\n", forces[0], forces[1], forces[2]);
      fflush(stdout);
    }
    if (rec && nrec < max_rec) {
      rec[nrec].step = its;
      rec[nrec].cfl = cfl;
      rec[nrec].fnorm = fnorm;
      rec[nrec].krylov = total;
      memcpy(rec[nrec].forces, forces, sizeof(forces));
    }
    nrec++;
    /* src/kernel.c:214 rounds the ratio; compared in double (no uint32 wrap) */
    if (round(fnorm_init / fnorm) >= o.fnorm_ratio) break;
    its++;
  }
#undef VEC
#undef ZVEC
  free(ghh);
  return nrec;
}

void Kernel(GEOMETRY *g) { kf_kernel(g, NULL, 0); }
