/* The CPU oracle (oracle/*.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: meshes, entities, dofmaps P1-P3, boundary
 * nodes, sparsity, matrix and vector assembly with BCs, lifting, facet and
 * interior-facet integrals, SpMV and both CG solvers on small boxes. Built
 * and run by tests/test_sanitizers.py; exits non-zero on a failed check. */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(c)                                                               \
  do                                                                           \
  {                                                                            \
    if (!(c))                                                                  \
    {                                                                          \
      fprintf(stderr, "failed: %s\n", #c);                                     \
      return 1;                                                                \
    }                                                                          \
  } while (0)

static int run(int celltype, const int64_t n[3], int degree)
{
  const double p0[3] = {0, 0, 0}, p1[3] = {1, 1, 1};
  orc_mesh* m = orc_mesh_create_box(p0, p1, n, celltype);
  CHECK(m);
  int32_t *c2e = NULL, *e2v = NULL;
  CHECK(orc_entities(m, 1, &c2e, &e2v) > 0);
  orc_free(c2e);
  orc_free(e2v);
  const int k = orc_element_ndofs(celltype, degree);
  int32_t* dm = malloc(sizeof(int32_t) * (size_t)m->num_cells * k);
  const int32_t nd = orc_dofmap(m, degree, dm);
  CHECK(nd > 0);
  int32_t* nodes = NULL;
  const int32_t nb = orc_boundary_nodes(m, degree, dm, 0, 0, 0.0, &nodes);
  CHECK(nb > 0);
  int64_t* rp = malloc(sizeof(int64_t) * ((size_t)nd + 1));
  int32_t* cols = NULL;
  const int64_t nnz = orc_sparsity(dm, k, dm, k, m->num_cells, nd, rp, &cols);
  CHECK(nnz > 0);
  int8_t* mark = calloc((size_t)nd, 1);
  for (int32_t i = 0; i < nb; ++i)
    mark[nodes[i]] = 1;
  if (degree == 2)
  {
    /* no scalar P2 form in the oracle: setup only */
    free(mark);
    orc_free(cols);
    free(rp);
    orc_free(nodes);
    free(dm);
    orc_mesh_free(m);
    return 0;
  }
  const int kA = celltype == ORC_HEX ? ORC_K_Q1HEX_A : (degree == 3 ? ORC_K_P3TET_A : ORC_K_P1TET_A);
  const int kL = celltype == ORC_HEX ? ORC_K_Q1HEX_L : (degree == 3 ? ORC_K_P3TET_L : ORC_K_P1TET_L);
  double* kap = malloc(sizeof(double) * (size_t)m->num_cells);
  for (int32_t c = 0; c < m->num_cells; ++c)
    kap[c] = 1.0 + 0.5 * (c % 3);
  const double cst[1] = {0.5};
  double* A = calloc((size_t)nnz, sizeof(double));
  CHECK(orc_assemble_matrix(kA, m->x, m->cells, m->nvpc, m->num_cells, dm, k, 1, kap, 1, cst,
                            mark, mark, rp, cols, A) == 0);
  CHECK(orc_set_diagonal(nodes, nb, 1, 1.0, rp, cols, A) == 0);
  const int nvl = degree == 3 ? 1 : m->nvpc;
  double* f = malloc(sizeof(double) * (size_t)m->num_cells * nvl);
  for (int64_t i = 0; i < (int64_t)m->num_cells * nvl; ++i)
    f[i] = 1.0;
  double* b = calloc((size_t)nd, sizeof(double));
  if (degree == 3)
    orc_assemble_vector(kL, m->x, m->cells, m->nvpc, m->num_cells, dm, k, 1, NULL, 0, cst, b);
  else
    orc_assemble_vector(kL, m->x, m->cells, m->nvpc, m->num_cells, dm, k, 1, f, nvl, cst, b);
  double* g = malloc(sizeof(double) * (size_t)nd);
  for (int32_t i = 0; i < nd; ++i)
    g[i] = 0.25 * (i % 5);
  orc_lift_bc(kA, m->x, m->cells, m->nvpc, m->num_cells, dm, k, 1, kap, 1, cst, g, mark, NULL,
              1.0, b);
  orc_bc_set(b, nodes, nb, g, NULL, 1.0);
  double* y = calloc((size_t)nd, sizeof(double));
  orc_spmv(nd, 1, rp, cols, A, b, y);
  double* x = calloc((size_t)nd, sizeof(double));
  CHECK(orc_pcg_jacobi(nd, 1, rp, cols, A, b, x, 5000, 1e-10) >= 0);
  memset(x, 0, sizeof(double) * (size_t)nd);
  CHECK(orc_cg(nd, 1, rp, cols, A, b, x, 5000, 1e-10) >= 0);
  if (degree == 1)
  {
    int32_t* pairs = NULL;
    const int32_t np = orc_facet_domain(m, 0, 0, 0.0, &pairs);
    CHECK(np > 0);
    const int kR = celltype == ORC_HEX ? ORC_K_Q1HEX_ROBIN_A : ORC_K_P1TET_ROBIN_A;
    double* beta = malloc(sizeof(double) * (size_t)m->num_cells);
    for (int32_t c = 0; c < m->num_cells; ++c)
      beta[c] = 2.0;
    CHECK(orc_assemble_matrix_facets(kR, m->x, m->cells, m->nvpc, dm, k, 1, beta, 1, cst, pairs,
                                     np, mark, mark, rp, cols, A) == 0);
    free(beta);
    orc_free(pairs);
    if (celltype == ORC_TET)
    {
      int32_t* quads = NULL;
      const int32_t nf = orc_interior_facet_domain(m, &quads);
      CHECK(nf > 0);
      int64_t* rp2 = malloc(sizeof(int64_t) * ((size_t)nd + 1));
      int32_t* cols2 = NULL;
      (void)orc_sparsity(dm, k, dm, k, m->num_cells, nd, rp2, &cols2);
      orc_free(cols2);
      free(rp2);
      orc_free(quads);
    }
  }
  free(x);
  free(y);
  free(g);
  free(b);
  free(f);
  free(A);
  free(kap);
  free(mark);
  orc_free(cols);
  free(rp);
  orc_free(nodes);
  free(dm);
  orc_mesh_free(m);
  return 0;
}

int main(void)
{
  const int64_t a[3] = {3, 2, 2}, b[3] = {4, 3, 3}, c[3] = {2, 2, 2};
  if (run(ORC_TET, a, 1) || run(ORC_TET, a, 2) || run(ORC_TET, c, 3) || run(ORC_HEX, b, 1))
    return 1;
  printf("oracle ok\n");
  return 0;
}
