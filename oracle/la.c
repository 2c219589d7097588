/*
 * Oracle: SpMV, dot products, CG and Jacobi-PCG.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* la/matrix_csr_impl.h:259-286 (spmv): y += A x, row-serial accumulation,
 * block entries row-major. */
void orc_spmv(int32_t nrows, int bs, const int64_t* row_ptr,
              const int32_t* cols, const double* values, const double* x,
              double* y)
{
  for (int k0 = 0; k0 < bs; ++k0)
    for (int32_t i = 0; i < nrows; ++i)
    {
      double vi = 0;
      for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
        for (int k1 = 0; k1 < bs; ++k1)
          vi += values[j * bs * bs + k0 * bs + k1] * x[cols[j] * bs + k1];
      y[i * bs + k0] += vi;
    }
}

/* la/Vector.h:434-460 (inner_product: std::transform_reduce on the owned
 * range; summation order of libstdc++ is left to right). */
double orc_dot(int32_t n, const double* a, const double* b)
{
  double s = 0;
  for (int32_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

/* cpp/demo/poisson_matrix_free/main.cpp:84-132 (cg) with an assembled
 * operator; stops when ||r||^2/||r0||^2 < rtol^2 (recursive residual).
 * Returns iterations, -1 when kmax is reached without convergence
 * (python/test/conftest.py:200-202 raises in that case). */
int orc_cg(int32_t nrows, int bs, const int64_t* row_ptr, const int32_t* cols,
           const double* values, const double* b, double* x, int kmax,
           double rtol)
{
  const int32_t n = nrows * bs;
  double* r = (double*)malloc(sizeof(double) * n);
  double* y = (double*)calloc(n, sizeof(double));
  double* p = (double*)malloc(sizeof(double) * n);
  orc_spmv(nrows, bs, row_ptr, cols, values, x, y);
  for (int32_t i = 0; i < n; ++i)
    r[i] = b[i] - y[i];
  memcpy(p, r, sizeof(double) * n);
  const double rnorm0 = orc_dot(n, r, r);
  double rnorm = rnorm0;
  const double rtol2 = rtol * rtol;
  int k = 0, conv = 0;
  while (k < kmax)
  {
    ++k;
    memset(y, 0, sizeof(double) * n);
    orc_spmv(nrows, bs, row_ptr, cols, values, p, y);
    const double alpha = rnorm / orc_dot(n, p, y);
    for (int32_t i = 0; i < n; ++i)
      x[i] += alpha * p[i];
    for (int32_t i = 0; i < n; ++i)
      r[i] -= alpha * y[i];
    const double rnew = orc_dot(n, r, r);
    const double beta = rnew / rnorm;
    rnorm = rnew;
    if (rnorm / rnorm0 < rtol2)
    {
      conv = 1;
      break;
    }
    for (int32_t i = 0; i < n; ++i)
      p[i] = beta * p[i] + r[i];
  }
  free(r);
  free(y);
  free(p);
  return conv ? k : -1;
}

/* Jacobi-preconditioned CG. Not in the reference tree (PETSc's
 * pc_type=jacobi in python/demo/demo_elasticity.py:203-209); restated
 * from Saad (2003), Algorithm 9.1, with the reference CG's stopping test
 * on the unpreconditioned recursive residual ||r||^2/||r0||^2 < rtol^2. */
int orc_pcg_jacobi(int32_t nrows, int bs, const int64_t* row_ptr,
                   const int32_t* cols, const double* values, const double* b,
                   double* x, int kmax, double rtol)
{
  const int32_t n = nrows * bs;
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* y = (double*)calloc(n, sizeof(double));
  double* p = (double*)malloc(sizeof(double) * n);
  double* dinv = (double*)malloc(sizeof(double) * n);
  for (int32_t i = 0; i < nrows; ++i)
    for (int a = 0; a < bs; ++a)
    {
      double d = 0;
      for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j)
        if (cols[j] == i)
          d = values[j * bs * bs + a * bs + a];
      dinv[i * bs + a] = 1.0 / d;
    }
  orc_spmv(nrows, bs, row_ptr, cols, values, x, y);
  for (int32_t i = 0; i < n; ++i)
  {
    r[i] = b[i] - y[i];
    z[i] = dinv[i] * r[i];
  }
  memcpy(p, z, sizeof(double) * n);
  const double rnorm0 = orc_dot(n, r, r);
  double rz = orc_dot(n, r, z);
  const double rtol2 = rtol * rtol;
  int k = 0, conv = rnorm0 == 0.0;
  while (k < kmax && !conv)
  {
    ++k;
    memset(y, 0, sizeof(double) * n);
    orc_spmv(nrows, bs, row_ptr, cols, values, p, y);
    const double alpha = rz / orc_dot(n, p, y);
    for (int32_t i = 0; i < n; ++i)
    {
      x[i] += alpha * p[i];
      r[i] -= alpha * y[i];
      z[i] = dinv[i] * r[i];
    }
    const double rnorm = orc_dot(n, r, r);
    if (rnorm / rnorm0 < rtol2)
    {
      conv = 1;
      break;
    }
    const double rznew = orc_dot(n, r, z);
    const double beta = rznew / rz;
    rz = rznew;
    for (int32_t i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  free(r);
  free(z);
  free(y);
  free(p);
  free(dinv);
  return conv ? k : -1;
}
