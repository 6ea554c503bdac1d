/*
 * ref_peek.cpp -- read access to the reference's fun3d::Matrix storage for
 * the golden-vector harness (ref_driver.c).  TEST INFRASTRUCTURE ONLY.
 * The class keeps A/L/U/Ut private (matrix/matrix.h:20-41); the harness
 * opens them by compiling this one translation unit with private -> public.
 */
#include <cstddef>
#include <cstdio>
#include <cstring>

#define private public
#include "matrix.h"
#undef private

extern "C" size_t kref_nnzb(void *m)
{
  fun3d::Matrix *M = (fun3d::Matrix *)m;
  return (size_t)M->A->i[M->nrows];
}

extern "C" void kref_blocks(void *m, double *out)
{
  fun3d::Matrix *M = (fun3d::Matrix *)m;
  std::memcpy(out, M->A->a, kref_nnzb(m) * M->bs2 * sizeof(double));
}

extern "C" void kref_pattern(void *m, unsigned int *rowptr, unsigned int *col)
{
  fun3d::Matrix *M = (fun3d::Matrix *)m;
  std::memcpy(rowptr, M->A->i, (M->nrows + 1) * sizeof(unsigned int));
  std::memcpy(col, M->A->j, kref_nnzb(m) * sizeof(unsigned int));
}
