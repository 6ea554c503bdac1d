/*
 * kf_main.c -- a C caller of the library built only against
 * include/kfun3d.h, the way the reference's driver calls its own path:
 *
 *   src/main.c:51-59   GEOMETRY allocated by the caller, Setup(fpath, g),
 *                      Kernel(g);
 *   src/main.c:65-122  cleanup (here kf_geometry_free: the state, the time
 *                      step and g->matrix live on the device);
 *   inc/utils.h:18-28  Compute2ndNorm, ComputeAXPY, ComputeNewAXPY and
 *                      Normalize on device vectors, as src/kernel.c calls
 *                      them.
 *
 *   kf_main <mesh.msh> <seq|tree> <out.bin>
 *
 * seq: norms and dot products summed in index order, as the reference's
 * one-thread loops (the history then equals the reference's bit for bit);
 * tree: the default fixed-order reduction tree.  Kernel prints the
 * reference's per-step lines (src/kernel.c:198-212).  After the solve the
 * program prints the final forces, then runs y = 0.37 q + y, w = -1.25 q + y,
 * |w| and Normalize(w) with q the converged state, and writes q, the initial
 * y, the final y and the final w (4N doubles each) to out.bin.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kfun3d.h"

int main(int argc, char **argv)
{
  if (argc != 4 || (strcmp(argv[2], "seq") && strcmp(argv[2], "tree"))) {
    fprintf(stderr, "usage: kf_main <mesh.msh> <seq|tree> <out.bin>\n");
    return 2;
  }
  GEOMETRY *g = (GEOMETRY *)calloc(1, sizeof(GEOMETRY));
  if (!g) return 3;

  Setup(argv[1], g); /* src/main.c:53 */
  kf_options o = kf_get_options(g);
  o.verbose = 1;
  o.reduce_seq = strcmp(argv[2], "seq") == 0;
  kf_set_options(g, &o);
  Kernel(g); /* src/main.c:59 */

  double f[3];
  ComputeForces(g, f);
  printf("FORCES %.17g %.17g %.17g\n", f[0], f[1], f[2]);

  const size_t n = g->c->sz;
  double *q = (double *)malloc(n * sizeof(double));
  double *y0 = (double *)malloc(n * sizeof(double));
  double *y1 = (double *)malloc(n * sizeof(double));
  double *w1 = (double *)malloc(n * sizeof(double));
  if (!q || !y0 || !y1 || !w1) return 3;
  kf_get_state(g, q);
  for (size_t i = 0; i < n; i++) y0[i] = 1.0 / (double)(i + 1) - 0.5 * (double)(i % 3);
  double *y = (double *)kf_device_alloc(n * sizeof(double));
  double *w = (double *)kf_device_alloc(n * sizeof(double));
  kf_memcpy_h2d(y, y0, n * sizeof(double));
  ComputeAXPY(n, 0.37, g->q->q, y);         /* y = 0.37 q + y   (src/utils.c:67-83) */
  ComputeNewAXPY(n, -1.25, g->q->q, y, w);  /* w = -1.25 q + y  (src/utils.c:86-103) */
  const double nw = Compute2ndNorm(n, w);   /* src/utils.c:53-64 */
  const double nn = Normalize(n, w);        /* src/utils.c:106-123 */
  printf("NORM2 %.17g\nNORMALIZE %.17g\n", nw, nn);
  kf_memcpy_d2h(y1, y, n * sizeof(double));
  kf_memcpy_d2h(w1, w, n * sizeof(double));
  const int err = kf_last_error(g);
  printf("ERROR %d\n", err);

  FILE *out = fopen(argv[3], "wb");
  if (!out) return 4;
  const double *arrs[4] = {q, y0, y1, w1};
  for (int k = 0; k < 4; k++)
    if (fwrite(arrs[k], sizeof(double), n, out) != n) return 4;
  fclose(out);

  kf_device_free(y);
  kf_device_free(w);
  kf_geometry_free(g); /* replaces src/main.c:65-121 */
  free(g);
  free(q); free(y0); free(y1); free(w1);
  return err ? 1 : 0;
}
