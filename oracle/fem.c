/*
 * Oracle: dofmaps, Dirichlet dofs, sparsity, element kernels, assembly.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const int TET_FACETS[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int TET_EDGES[6][2]
    = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
static const int HEX_FACETS[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                     {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};

/* Lagrange element dof layout (Basix): vertices first, then edge interiors
 * (degree-1 dofs each, in edge order), then face interiors. */
int orc_element_ndofs(int celltype, int degree)
{
  if (celltype == ORC_HEX)
    return 8;
  return (degree + 1) * (degree + 2) * (degree + 3) / 6;
}

/* fem/dofmapbuilder.cpp:153-397 (build_basic_dofmaps), :421-490
 * (compute_reordering_map with reorder_fn == nullptr: first-traversal
 * order over cells), fem/utils.cpp:32-76 (create_dofmap incl. the P3
 * permute_inv of reflected edges). */
int32_t orc_dofmap(const orc_mesh* m, int degree, int32_t* dofmap)
{
  const int k = orc_element_ndofs(m->celltype, degree);
  const int32_t nc = m->num_cells;
  const int32_t nv = m->num_vertices;
  int32_t *c2e = NULL, *e2v = NULL, *c2f = NULL, *f2v = NULL;
  int32_t ne = 0, nf = 0;
  int32_t* basic = (int32_t*)malloc(sizeof(int32_t) * (int64_t)nc * k);
  const int ede = degree - 1;                           /* dofs per edge */
  const int fde = degree >= 3 ? (degree - 1) * (degree - 2) / 2 : 0;
  if (m->celltype == ORC_TET && ede > 0)
    ne = orc_entities(m, 1, &c2e, &e2v);
  if (m->celltype == ORC_TET && fde > 0)
    nf = orc_entities(m, 2, &c2f, &f2v);
  for (int32_t c = 0; c < nc; ++c)
  {
    int32_t* d = basic + (int64_t)c * k;
    for (int i = 0; i < m->nvpc; ++i)
      d[i] = m->cells[(int64_t)c * m->nvpc + i];
    if (m->celltype == ORC_TET)
    {
      for (int e = 0; e < 6 && ede > 0; ++e)
        for (int j = 0; j < ede; ++j)
          d[4 + ede * e + j] = nv + ede * c2e[(int64_t)c * 6 + e] + j;
      for (int f = 0; f < 4 && fde > 0; ++f)
        for (int j = 0; j < fde; ++j)
          d[4 + 6 * ede + fde * f + j]
              = nv + ede * ne + fde * c2f[(int64_t)c * 4 + f] + j;
    }
  }
  const int32_t ndofs = nv + ede * ne + fde * nf;
  int32_t* o2n = (int32_t*)malloc(sizeof(int32_t) * (ndofs > 0 ? ndofs : 1));
  for (int32_t i = 0; i < ndofs; ++i)
    o2n[i] = -1;
  int32_t cnt = 0;
  for (int64_t i = 0; i < (int64_t)nc * k; ++i)
    if (o2n[basic[i]] == -1)
      o2n[basic[i]] = cnt++;
  for (int32_t i = 0; i < ndofs; ++i)
    if (o2n[i] == -1)
      o2n[i] = cnt++;
  for (int64_t i = 0; i < (int64_t)nc * k; ++i)
    dofmap[i] = o2n[basic[i]];

  if (m->celltype == ORC_TET && ede == 2)
  {
    /* P3: dof transformations are permutations; a reflected edge swaps its
     * two interior dofs (fem/utils.cpp:61-73, FiniteElement.cpp:111-129). */
    uint32_t* bits = (uint32_t*)malloc(sizeof(uint32_t) * nc);
    orc_edge_reflections(m, c2e, e2v, bits);
    for (int32_t c = 0; c < nc; ++c)
      for (int e = 0; e < 6; ++e)
        if (bits[c] & (1u << e))
        {
          int32_t* d = dofmap + (int64_t)c * k;
          int32_t t = d[4 + 2 * e];
          d[4 + 2 * e] = d[5 + 2 * e];
          d[5 + 2 * e] = t;
        }
    free(bits);
  }
  free(o2n);
  free(basic);
  free(c2e);
  free(e2v);
  free(c2f);
  free(f2v);
  return ndofs;
}

/* Local dofs in the closure of local facet f (Basix entity closure dofs). */
static int facet_closure_dofs(int celltype, int degree, int f, int* out)
{
  int n = 0;
  if (celltype == ORC_HEX)
  {
    for (int i = 0; i < 4; ++i)
      out[n++] = HEX_FACETS[f][i];
    return n;
  }
  int in[4] = {0, 0, 0, 0};
  for (int i = 0; i < 3; ++i)
  {
    in[TET_FACETS[f][i]] = 1;
    out[n++] = TET_FACETS[f][i];
  }
  const int ede = degree - 1;
  for (int e = 0; e < 6 && ede > 0; ++e)
    if (in[TET_EDGES[e][0]] && in[TET_EDGES[e][1]])
      for (int j = 0; j < ede; ++j)
        out[n++] = 4 + ede * e + j;
  if (degree >= 3)
    out[n++] = 4 + 6 * ede + f;
  return n;
}

static int cmp_i32(const void* a, const void* b)
{
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* fem/DirichletBC.cpp:201-310 (locate_dofs_topological on exterior facets,
 * mesh/utils.cpp:61-89), with the facet marker of
 * mesh::locate_entities_boundary for marker == 1. Sorted unique nodes. */
int32_t orc_boundary_nodes(const orc_mesh* m, int degree,
                           const int32_t* dofmap, int marker, int axis,
                           double value, int32_t** nodes)
{
  int32_t* cf;
  const int32_t nb = orc_exterior_facets(m, &cf);
  const int k = orc_element_ndofs(m->celltype, degree);
  int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * ((int64_t)nb * 16 + 1));
  int64_t n = 0;
  for (int32_t i = 0; i < nb; ++i)
  {
    const int32_t c = cf[2 * i];
    const int f = cf[2 * i + 1];
    if (marker == 1)
    {
      const int nfv = m->celltype == ORC_TET ? 3 : 4;
      int ok = 1;
      for (int j = 0; j < nfv; ++j)
      {
        const int lv = m->celltype == ORC_TET ? TET_FACETS[f][j] : HEX_FACETS[f][j];
        const int32_t v = m->cells[(int64_t)c * m->nvpc + lv];
        if (fabs(m->x[3 * (int64_t)v + axis] - value) > 1e-12)
          ok = 0;
      }
      if (!ok)
        continue;
    }
    int loc[16];
    const int nl = facet_closure_dofs(m->celltype, degree, f, loc);
    for (int j = 0; j < nl; ++j)
      buf[n++] = dofmap[(int64_t)c * k + loc[j]];
  }
  qsort(buf, n, sizeof(int32_t), cmp_i32);
  int64_t u = 0;
  for (int64_t i = 0; i < n; ++i)
    if (u == 0 || buf[u - 1] != buf[i])
      buf[u++] = buf[i];
  free(cf);
  *nodes = buf;
  return (int32_t)u;
}

/* fem/utils.h:197-314 + fem/sparsitybuild.h:36-49 (insert cell outer
 * products) + la/SparsityPattern.cpp:264-491 (finalize: per-row sort and
 * deduplicate). Rows are bucketed as in the reference row cache. */
int64_t orc_sparsity(const int32_t* dofmap0, int k0, const int32_t* dofmap1,
                     int k1, int32_t num_cells, int32_t nrows,
                     int64_t* row_ptr, int32_t** cols_out)
{
  int64_t* cnt = (int64_t*)calloc((size_t)nrows + 1, sizeof(int64_t));
  for (int64_t c = 0; c < num_cells; ++c)
    for (int i = 0; i < k0; ++i)
      cnt[dofmap0[c * k0 + i] + 1] += k1;
  for (int32_t r = 0; r < nrows; ++r)
    cnt[r + 1] += cnt[r];
  int32_t* cache = (int32_t*)malloc(sizeof(int32_t) * (cnt[nrows] + 1));
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nrows + 1));
  memcpy(pos, cnt, sizeof(int64_t) * ((size_t)nrows + 1));
  for (int64_t c = 0; c < num_cells; ++c)
    for (int i = 0; i < k0; ++i)
    {
      const int32_t r = dofmap0[c * k0 + i];
      for (int j = 0; j < k1; ++j)
        cache[pos[r]++] = dofmap1[c * k1 + j];
    }
  int64_t nnz = 0;
  row_ptr[0] = 0;
  for (int32_t r = 0; r < nrows; ++r)
  {
    int32_t* seg = cache + cnt[r];
    const int64_t len = cnt[r + 1] - cnt[r];
    qsort(seg, len, sizeof(int32_t), cmp_i32);
    int64_t u = 0;
    for (int64_t i = 0; i < len; ++i)
      if (u == 0 || seg[u - 1] != seg[i])
        seg[u++] = seg[i];
    memmove(cache + nnz, seg, sizeof(int32_t) * u);
    nnz += u;
    row_ptr[r + 1] = nnz;
  }
  free(pos);
  free(cnt);
  *cols_out = (int32_t*)realloc(cache, sizeof(int32_t) * (nnz + 1));
  return nnz;
}

/* ====================================================================== */
/* Element kernels. The arithmetic of the FFCx-generated kernels is not in
 * the reference tree (fem/kernel.h:18-60 only fixes the ABI); these are
 * written from the mathematics: affine maps, exact quadrature. */

static double det3(const double J[9])
{
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6])
         + J[2] * (J[3] * J[7] - J[4] * J[6]);
}

/* K = J^{-1} (row-major), returns det J */
static double inv3(const double J[9], double K[9])
{
  const double d = det3(J);
  K[0] = (J[4] * J[8] - J[5] * J[7]) / d;
  K[1] = (J[2] * J[7] - J[1] * J[8]) / d;
  K[2] = (J[1] * J[5] - J[2] * J[4]) / d;
  K[3] = (J[5] * J[6] - J[3] * J[8]) / d;
  K[4] = (J[0] * J[8] - J[2] * J[6]) / d;
  K[5] = (J[2] * J[3] - J[0] * J[5]) / d;
  K[6] = (J[3] * J[7] - J[4] * J[6]) / d;
  K[7] = (J[1] * J[6] - J[0] * J[7]) / d;
  K[8] = (J[0] * J[4] - J[1] * J[3]) / d;
  return d;
}

/* Affine tet Jacobian J[i][a] = x_{a+1,i} - x_{0,i} */
static double tet_jacobian(const double* cd, double J[9], double K[9])
{
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 3; ++a)
      J[3 * i + a] = cd[3 * (a + 1) + i] - cd[i];
  return inv3(J, K);
}

/* P1 tet: A_ij = kappa |K| grad phi_i . grad phi_j + c |K| (1+d_ij)/20 */
static void k_p1tet_a(double* A, const double* w, const double* c,
                      const double* cd, const int* e, const uint8_t* q, void* u)
{
  (void)e;
  (void)q;
  (void)u;
  double J[9], K[9];
  const double vol = fabs(tet_jacobian(cd, J, K)) / 6.0;
  /* reference gradients of phi_0..3 */
  const double dr[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double g[4][3];
  for (int i = 0; i < 4; ++i)
    for (int d = 0; d < 3; ++d)
      g[i][d] = dr[i][0] * K[0 * 3 + d] + dr[i][1] * K[1 * 3 + d]
                + dr[i][2] * K[2 * 3 + d];
  const double kap = w[0], rc = c[0];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      A[4 * i + j] += kap * vol
                          * (g[i][0] * g[j][0] + g[i][1] * g[j][1]
                             + g[i][2] * g[j][2])
                      + rc * vol * (i == j ? 2.0 : 1.0) / 20.0;
}

/* P1 tet load with P1 coefficient f: b_i = sum_j M_ij f_j */
static void k_p1tet_l(double* A, const double* w, const double* c,
                      const double* cd, const int* e, const uint8_t* q, void* u)
{
  (void)c;
  (void)e;
  (void)q;
  (void)u;
  double J[9], K[9];
  const double vol = fabs(tet_jacobian(cd, J, K)) / 6.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      A[i] += vol * (i == j ? 2.0 : 1.0) / 20.0 * w[j];
}

/* Q1 hex, 2x2x2 Gauss-Legendre (exact on parallelepipeds). */
static void q1_tab(const double xi[3], double phi[8], double dphi[8][3])
{
  for (int i = 0; i < 8; ++i)
  {
    const int b[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
    double f[3], df[3];
    for (int d = 0; d < 3; ++d)
    {
      f[d] = b[d] ? xi[d] : 1.0 - xi[d];
      df[d] = b[d] ? 1.0 : -1.0;
    }
    phi[i] = f[0] * f[1] * f[2];
    dphi[i][0] = df[0] * f[1] * f[2];
    dphi[i][1] = f[0] * df[1] * f[2];
    dphi[i][2] = f[0] * f[1] * df[2];
  }
}

static void hex_qp(int q, double xi[3])
{
  const double a = 0.5 - 0.5 / sqrt(3.0), b = 0.5 + 0.5 / sqrt(3.0);
  xi[0] = (q & 1) ? b : a;
  xi[1] = ((q >> 1) & 1) ? b : a;
  xi[2] = ((q >> 2) & 1) ? b : a;
}

static void k_q1hex_a(double* A, const double* w, const double* c,
                      const double* cd, const int* e, const uint8_t* qq,
                      void* u)
{
  (void)e;
  (void)qq;
  (void)u;
  for (int q = 0; q < 8; ++q)
  {
    double xi[3], phi[8], dphi[8][3], J[9] = {0}, K[9];
    hex_qp(q, xi);
    q1_tab(xi, phi, dphi);
    for (int i = 0; i < 8; ++i)
      for (int r = 0; r < 3; ++r)
        for (int a = 0; a < 3; ++a)
          J[3 * r + a] += cd[3 * i + r] * dphi[i][a];
    const double wdet = fabs(inv3(J, K)) / 8.0;
    double g[8][3];
    for (int i = 0; i < 8; ++i)
      for (int d = 0; d < 3; ++d)
        g[i][d] = dphi[i][0] * K[d] + dphi[i][1] * K[3 + d]
                  + dphi[i][2] * K[6 + d];
    for (int i = 0; i < 8; ++i)
      for (int j = 0; j < 8; ++j)
        A[8 * i + j] += wdet
                        * (w[0]
                               * (g[i][0] * g[j][0] + g[i][1] * g[j][1]
                                  + g[i][2] * g[j][2])
                           + c[0] * phi[i] * phi[j]);
  }
}

static void k_q1hex_l(double* A, const double* w, const double* c,
                      const double* cd, const int* e, const uint8_t* qq,
                      void* u)
{
  (void)c;
  (void)e;
  (void)qq;
  (void)u;
  for (int q = 0; q < 8; ++q)
  {
    double xi[3], phi[8], dphi[8][3], J[9] = {0}, K[9];
    hex_qp(q, xi);
    q1_tab(xi, phi, dphi);
    for (int i = 0; i < 8; ++i)
      for (int r = 0; r < 3; ++r)
        for (int a = 0; a < 3; ++a)
          J[3 * r + a] += cd[3 * i + r] * dphi[i][a];
    const double wdet = fabs(inv3(J, K)) / 8.0;
    double f = 0;
    for (int j = 0; j < 8; ++j)
      f += w[j] * phi[j];
    for (int i = 0; i < 8; ++i)
      A[i] += wdet * f * phi[i];
  }
}

/* ---- exterior facet loads, g v ds with g at the cell's vertices ---- */
/* Facet vertices of the reference cells in DOLFINx sub-entity order
 * (cell_types.cpp facet tables; quadrilateral facets in tensor order). */
static const int TET_FV[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int HEX_FV[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                 {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};

static void cross3(const double a[3], const double b[3], double n[3])
{
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
}

/* P1 tet, facet e[0]: b_i = sum_j |F| (1 + d_ij) / 12 g_j over the facet's
 * vertices (the P1 mass matrix of the facet triangle). */
static void k_p1tet_neumann(double* A, const double* w, const double* c,
                            const double* cd, const int* e, const uint8_t* q,
                            void* u)
{
  (void)c;
  (void)q;
  (void)u;
  const int* fv = TET_FV[e[0]];
  double t0[3], t1[3], n[3];
  for (int d = 0; d < 3; ++d)
  {
    t0[d] = cd[3 * fv[1] + d] - cd[3 * fv[0] + d];
    t1[d] = cd[3 * fv[2] + d] - cd[3 * fv[0] + d];
  }
  cross3(t0, t1, n);
  const double area = 0.5 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      A[fv[a]] += area * (a == b ? 2.0 : 1.0) / 12.0 * w[fv[b]];
}

/* Q1 hex, facet e[0]: 2 x 2 Gauss-Legendre on the bilinear facet map
 * through its four vertices (exact for g v on parallelogram facets). */
static void k_q1hex_neumann(double* A, const double* w, const double* c,
                            const double* cd, const int* e, const uint8_t* qq,
                            void* u)
{
  (void)c;
  (void)qq;
  (void)u;
  const int* fv = HEX_FV[e[0]];
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  const double pts[2] = {g0, g1};
  for (int qs = 0; qs < 2; ++qs)
    for (int qt = 0; qt < 2; ++qt)
    {
      const double s = pts[qs], t = pts[qt];
      const double phi[4] = {(1 - s) * (1 - t), s * (1 - t), (1 - s) * t, s * t};
      const double ds[4] = {-(1 - t), 1 - t, -t, t};
      const double dt[4] = {-(1 - s), -s, 1 - s, s};
      double xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0}, n[3];
      for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d)
        {
          xs[d] += ds[a] * cd[3 * fv[a] + d];
          xt[d] += dt[a] * cd[3 * fv[a] + d];
        }
      cross3(xs, xt, n);
      const double dA = 0.25 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      double g = 0;
      for (int a = 0; a < 4; ++a)
        g += phi[a] * w[fv[a]];
      for (int a = 0; a < 4; ++a)
        A[fv[a]] += dA * g * phi[a];
    }
}

/* P1 tet, facet e[0]: A_ab = beta |F| (1 + d_ab) / 12 over the facet's
 * vertices (beta u v ds, beta = w[0] constant on the cell), written into the
 * 4 x 4 cell matrix. */
static void k_p1tet_robin(double* A, const double* w, const double* c,
                          const double* cd, const int* e, const uint8_t* q,
                          void* u)
{
  (void)c;
  (void)q;
  (void)u;
  const int* fv = TET_FV[e[0]];
  double t0[3], t1[3], n[3];
  for (int d = 0; d < 3; ++d)
  {
    t0[d] = cd[3 * fv[1] + d] - cd[3 * fv[0] + d];
    t1[d] = cd[3 * fv[2] + d] - cd[3 * fv[0] + d];
  }
  cross3(t0, t1, n);
  const double area = 0.5 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      A[4 * fv[a] + fv[b]] += w[0] * area * (a == b ? 2.0 : 1.0) / 12.0;
}

/* Q1 hex, facet e[0]: beta u v ds by 2 x 2 Gauss-Legendre on the bilinear
 * facet map (exact on parallelogram facets), into the 8 x 8 cell matrix. */
static void k_q1hex_robin(double* A, const double* w, const double* c,
                          const double* cd, const int* e, const uint8_t* qq,
                          void* u)
{
  (void)c;
  (void)qq;
  (void)u;
  const int* fv = HEX_FV[e[0]];
  const double g0 = 0.5 - 0.5 / sqrt(3.0), g1 = 0.5 + 0.5 / sqrt(3.0);
  const double pts[2] = {g0, g1};
  for (int qs = 0; qs < 2; ++qs)
    for (int qt = 0; qt < 2; ++qt)
    {
      const double s = pts[qs], t = pts[qt];
      const double phi[4] = {(1 - s) * (1 - t), s * (1 - t), (1 - s) * t, s * t};
      const double ds[4] = {-(1 - t), 1 - t, -t, t};
      const double dt[4] = {-(1 - s), -s, 1 - s, s};
      double xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0}, n[3];
      for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d)
        {
          xs[d] += ds[a] * cd[3 * fv[a] + d];
          xt[d] += dt[a] * cd[3 * fv[a] + d];
        }
      cross3(xs, xt, n);
      const double dA = 0.25 * sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          A[8 * fv[a] + fv[b]] += dA * w[0] * phi[a] * phi[b];
    }
}

/* P1 tet pair across an interior facet: gamma [grad u.n][grad v.n] dS on
 * the 8 x 8 macro element (cell 0's dofs, then cell 1's; cd holds both
 * cells' coordinates), n the unit normal of cell 0's facet e[0],
 * j_a = grad phi_a.n on cell 0 and -grad phi_a.n on cell 1, A_ab =
 * gamma |F| j_a j_b (the gradients are constant: exact). */
static void k_p1tet_gradjump(double* A, const double* w, const double* c,
                             const double* cd, const int* e, const uint8_t* q,
                             void* u)
{
  (void)w;
  (void)q;
  (void)u;
  const int* fv = TET_FV[e[0]];
  double t0[3], t1[3], n[3];
  for (int d = 0; d < 3; ++d)
  {
    t0[d] = cd[3 * fv[1] + d] - cd[3 * fv[0] + d];
    t1[d] = cd[3 * fv[2] + d] - cd[3 * fv[0] + d];
  }
  cross3(t0, t1, n);
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int d = 0; d < 3; ++d)
    n[d] /= nn;
  const double area = 0.5 * nn;
  double jmp[8];
  for (int s = 0; s < 2; ++s)
  {
    double J[9], K[9];
    const double* X = cd + 12 * s;
    for (int r = 0; r < 3; ++r)
      for (int a = 0; a < 3; ++a)
        J[3 * r + a] = X[3 * (a + 1) + r] - X[r];
    inv3(J, K);
    for (int i = 0; i < 4; ++i)
    {
      /* grad lambda_i: row i-1 of K, lambda_0 = 1 - sum */
      double g[3];
      for (int d = 0; d < 3; ++d)
        g[d] = i == 0 ? -(K[d] + K[3 + d] + K[6 + d]) : K[3 * (i - 1) + d];
      const double dn = g[0] * n[0] + g[1] * n[1] + g[2] * n[2];
      jmp[4 * s + i] = s == 0 ? dn : -dn;
    }
  }
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      A[8 * a + b] += c[0] * area * jmp[a] * jmp[b];
}

/* ---- P2/P3 tetrahedra: nodal basis from the monomial Vandermonde ---- */
typedef struct
{
  int degree, ndofs, nq;
  double* qw;     /* nq */
  double* phi;    /* nq x ndofs */
  double* dphi;   /* nq x ndofs x 3 (reference derivatives) */
} tet_tab;

static tet_tab TAB[4];

/* Basix Lagrange points: vertices, then edge interiors from the edge's first
 * to its second vertex (GLL points for degree 3, the default gll_warped
 * variant), then face centroids (degree 3). */
static int lagrange_points(int degree, double pts[][3])
{
  const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  int n = 0;
  for (int i = 0; i < 4; ++i, ++n)
    memcpy(pts[n], V[i], sizeof(V[i]));
  double t[2];
  int nt = degree - 1;
  if (degree == 2)
    t[0] = 0.5;
  if (degree == 3)
  {
    t[0] = 0.5 - 0.5 / sqrt(5.0);
    t[1] = 0.5 + 0.5 / sqrt(5.0);
  }
  for (int e = 0; e < 6; ++e)
    for (int j = 0; j < nt; ++j, ++n)
      for (int d = 0; d < 3; ++d)
        pts[n][d] = V[TET_EDGES[e][0]][d]
                    + t[j] * (V[TET_EDGES[e][1]][d] - V[TET_EDGES[e][0]][d]);
  if (degree == 3)
    for (int f = 0; f < 4; ++f, ++n)
      for (int d = 0; d < 3; ++d)
        pts[n][d] = (V[TET_FACETS[f][0]][d] + V[TET_FACETS[f][1]][d]
                     + V[TET_FACETS[f][2]][d])
                    / 3.0;
  return n;
}

static int monomials(int degree, int mono[][3])
{
  int n = 0;
  for (int s = 0; s <= degree; ++s)
    for (int a = s; a >= 0; --a)
      for (int b = s - a; b >= 0; --b)
      {
        mono[n][0] = a;
        mono[n][1] = b;
        mono[n][2] = s - a - b;
        ++n;
      }
  return n;
}

static double ipow(double x, int p)
{
  double r = 1;
  for (int i = 0; i < p; ++i)
    r *= x;
  return r;
}

/* Gauss-Legendre on [0,1], m points (Newton on P_m). */
static void gauss01(int m, double* x, double* w)
{
  for (int i = 0; i < m; ++i)
  {
    double z = cos(M_PI * (i + 0.75) / (m + 0.5)), pp = 0;
    for (int it = 0; it < 100; ++it)
    {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= m; ++j)
      {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = m * (z * p1 - p2) / (z * z - 1);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16)
        break;
    }
    x[i] = 0.5 * (1 - z);
    w[i] = 1.0 / ((1 - z * z) * pp * pp);
  }
}

/* Solve A X = B (n x n, n x m) by Gaussian elimination with pivoting. */
static void solve(int n, double* A, int m, double* B)
{
  for (int c = 0; c < n; ++c)
  {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(A[r * n + c]) > fabs(A[p * n + c]))
        p = r;
    for (int j = 0; j < n; ++j)
    {
      double t = A[c * n + j];
      A[c * n + j] = A[p * n + j];
      A[p * n + j] = t;
    }
    for (int j = 0; j < m; ++j)
    {
      double t = B[c * m + j];
      B[c * m + j] = B[p * m + j];
      B[p * m + j] = t;
    }
    for (int r = 0; r < n; ++r)
    {
      if (r == c)
        continue;
      const double f = A[r * n + c] / A[c * n + c];
      for (int j = 0; j < n; ++j)
        A[r * n + j] -= f * A[c * n + j];
      for (int j = 0; j < m; ++j)
        B[r * m + j] -= f * B[c * m + j];
    }
  }
  for (int r = 0; r < n; ++r)
    for (int j = 0; j < m; ++j)
      B[r * m + j] /= A[r * n + r];
}

static const tet_tab* tet_table(int degree)
{
  tet_tab* T = &TAB[degree];
  if (T->qw)
    return T;
  double pts[20][3];
  int mono[20][3];
  const int n = lagrange_points(degree, pts);
  monomials(degree, mono);
  /* V[i][j] = m_j(p_i); basis coefficients C = V^{-1}: phi_i = sum_j C[j][i] m_j */
  double* Vm = (double*)malloc(sizeof(double) * n * n);
  double* C = (double*)calloc(n * n, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      Vm[i * n + j] = ipow(pts[i][0], mono[j][0]) * ipow(pts[i][1], mono[j][1])
                      * ipow(pts[i][2], mono[j][2]);
  for (int i = 0; i < n; ++i)
    C[i * n + i] = 1.0;
  solve(n, Vm, n, C); /* C = V^{-1} */
  /* collapsed Gauss-Legendre, 5 points per direction (exact to degree 7) */
  const int m = 5;
  double gx[5], gw[5];
  gauss01(m, gx, gw);
  T->degree = degree;
  T->ndofs = n;
  T->nq = m * m * m;
  T->qw = (double*)malloc(sizeof(double) * T->nq);
  T->phi = (double*)malloc(sizeof(double) * T->nq * n);
  T->dphi = (double*)malloc(sizeof(double) * T->nq * n * 3);
  int q = 0;
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b)
      for (int cc = 0; cc < m; ++cc, ++q)
      {
        const double u = gx[a], v = gx[b], w = gx[cc];
        const double X[3] = {u * (1 - v) * (1 - w), v * (1 - w), w};
        T->qw[q] = gw[a] * gw[b] * gw[cc] * (1 - v) * (1 - w) * (1 - w);
        for (int i = 0; i < n; ++i)
        {
          double p = 0, d0 = 0, d1 = 0, d2 = 0;
          for (int j = 0; j < n; ++j)
          {
            const int* e = mono[j];
            const double cj = C[j * n + i];
            p += cj * ipow(X[0], e[0]) * ipow(X[1], e[1]) * ipow(X[2], e[2]);
            if (e[0])
              d0 += cj * e[0] * ipow(X[0], e[0] - 1) * ipow(X[1], e[1])
                    * ipow(X[2], e[2]);
            if (e[1])
              d1 += cj * e[1] * ipow(X[0], e[0]) * ipow(X[1], e[1] - 1)
                    * ipow(X[2], e[2]);
            if (e[2])
              d2 += cj * e[2] * ipow(X[0], e[0]) * ipow(X[1], e[1])
                    * ipow(X[2], e[2] - 1);
          }
          T->phi[q * n + i] = p;
          T->dphi[(q * n + i) * 3 + 0] = d0;
          T->dphi[(q * n + i) * 3 + 1] = d1;
          T->dphi[(q * n + i) * 3 + 2] = d2;
        }
      }
  free(Vm);
  free(C);
  return T;
}

/* P2 vector elasticity: K[(i,c),(j,d)] =
 *   int mu (d_cd grad phi_i.grad phi_j + d_d phi_i d_c phi_j)
 *     + lambda d_c phi_i d_d phi_j,  row = 3 i + c (interleaved blocks) */
static void k_p2_elast_a(double* A, const double* w, const double* c,
                         const double* cd, const int* e, const uint8_t* qq,
                         void* u)
{
  (void)c;
  (void)e;
  (void)qq;
  (void)u;
  const tet_tab* T = tet_table(2);
  double J[9], K[9];
  const double det = fabs(tet_jacobian(cd, J, K));
  const double lam = w[0], mu = w[1];
  const int n = 10;
  for (int q = 0; q < T->nq; ++q)
  {
    double g[10][3];
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d)
      {
        const double* r = T->dphi + (q * n + i) * 3;
        g[i][d] = r[0] * K[d] + r[1] * K[3 + d] + r[2] * K[6 + d];
      }
    const double s = T->qw[q] * det;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
      {
        const double gg = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
        for (int cc = 0; cc < 3; ++cc)
          for (int dd = 0; dd < 3; ++dd)
            A[(3 * i + cc) * 30 + 3 * j + dd]
                += s * (mu * ((cc == dd ? gg : 0.0) + g[i][dd] * g[j][cc])
                        + lam * g[i][cc] * g[j][dd]);
      }
  }
}

static void k_p2_elast_l(double* A, const double* w, const double* c,
                         const double* cd, const int* e, const uint8_t* qq,
                         void* u)
{
  (void)w;
  (void)e;
  (void)qq;
  (void)u;
  const tet_tab* T = tet_table(2);
  double J[9], K[9];
  const double det = fabs(tet_jacobian(cd, J, K));
  for (int q = 0; q < T->nq; ++q)
    for (int i = 0; i < 10; ++i)
      for (int cc = 0; cc < 3; ++cc)
        A[3 * i + cc] += T->qw[q] * det * T->phi[q * 10 + i] * c[cc];
}

static void k_p3_a(double* A, const double* w, const double* c,
                   const double* cd, const int* e, const uint8_t* qq, void* u)
{
  (void)e;
  (void)qq;
  (void)u;
  const tet_tab* T = tet_table(3);
  double J[9], K[9];
  const double det = fabs(tet_jacobian(cd, J, K));
  const int n = 20;
  for (int q = 0; q < T->nq; ++q)
  {
    double g[20][3];
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d)
      {
        const double* r = T->dphi + (q * n + i) * 3;
        g[i][d] = r[0] * K[d] + r[1] * K[3 + d] + r[2] * K[6 + d];
      }
    const double s = T->qw[q] * det;
    const double* ph = T->phi + q * n;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        A[n * i + j] += s
                        * (w[0]
                               * (g[i][0] * g[j][0] + g[i][1] * g[j][1]
                                  + g[i][2] * g[j][2])
                           + c[0] * ph[i] * ph[j]);
  }
}

static void k_p3_l(double* A, const double* w, const double* c,
                   const double* cd, const int* e, const uint8_t* qq, void* u)
{
  (void)w;
  (void)e;
  (void)qq;
  (void)u;
  const tet_tab* T = tet_table(3);
  double J[9], K[9];
  const double det = fabs(tet_jacobian(cd, J, K));
  for (int q = 0; q < T->nq; ++q)
    for (int i = 0; i < 20; ++i)
      A[i] += T->qw[q] * det * T->phi[q * 20 + i] * c[0];
}

orc_kernel_fn orc_get_kernel(int id)
{
  switch (id)
  {
  case ORC_K_P1TET_A:
    return k_p1tet_a;
  case ORC_K_P1TET_L:
    return k_p1tet_l;
  case ORC_K_Q1HEX_A:
    return k_q1hex_a;
  case ORC_K_Q1HEX_L:
    return k_q1hex_l;
  case ORC_K_P2TET_ELAST_A:
    return k_p2_elast_a;
  case ORC_K_P2TET_ELAST_L:
    return k_p2_elast_l;
  case ORC_K_P3TET_A:
    return k_p3_a;
  case ORC_K_P3TET_L:
    return k_p3_l;
  case ORC_K_P1TET_NEUMANN:
    return k_p1tet_neumann;
  case ORC_K_Q1HEX_NEUMANN:
    return k_q1hex_neumann;
  case ORC_K_P1TET_ROBIN_A:
    return k_p1tet_robin;
  case ORC_K_Q1HEX_ROBIN_A:
    return k_q1hex_robin;
  case ORC_K_P1TET_GRADJUMP_A:
    return k_p1tet_gradjump;
  default:
    return NULL;
  }
}

void orc_tabulate(int id, double* A, const double* w, const double* c,
                  const double* coordinate_dofs)
{
  orc_get_kernel(id)(A, w, c, coordinate_dofs, NULL, NULL, NULL);
}

/* ====================================================================== */
/* One cell of fem/assemble_matrix_impl.h:92-200 after the kernel call: zero
 * the BC rows (bc0) and columns (bc1) of Ae (:156-180), then mat_set =
 * MatrixCSR::add -> la/matrix_csr_impl.h:67-109 (insert_csr<BS,BS>,
 * lower_bound per column, "+="). Ae is modified. Returns -1 if an entry is
 * not in the sparsity. */
static int insert_cell(const int32_t* dofs, int k, int bs, double* Ae,
                       const int8_t* bc0, const int8_t* bc1,
                       const int64_t* row_ptr, const int32_t* cols,
                       double* values)
{
  const int nd = bs * k;
  if (bc0)
    for (int i = 0; i < k; ++i)
      for (int b = 0; b < bs; ++b)
        if (bc0[bs * dofs[i] + b])
          memset(Ae + (bs * i + b) * nd, 0, sizeof(double) * nd);
  if (bc1)
    for (int j = 0; j < k; ++j)
      for (int b = 0; b < bs; ++b)
        if (bc1[bs * dofs[j] + b])
          for (int r = 0; r < nd; ++r)
            Ae[r * nd + bs * j + b] = 0;
  for (int i = 0; i < k; ++i)
  {
    const int32_t row = dofs[i];
    const int32_t* c0 = cols + row_ptr[row];
    const int64_t len = row_ptr[row + 1] - row_ptr[row];
    for (int j = 0; j < k; ++j)
    {
      const int32_t col = dofs[j];
      int64_t lo = 0, hi = len;
      while (lo < hi)
      {
        const int64_t mid = (lo + hi) / 2;
        if (c0[mid] < col)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo == len || c0[lo] != col)
        return -1;
      double* dst = values + (row_ptr[row] + lo) * bs * bs;
      for (int a = 0; a < bs; ++a)
        for (int b = 0; b < bs; ++b)
          dst[a * bs + b] += Ae[(bs * i + a) * nd + bs * j + b];
    }
  }
  return 0;
}

/* fem/assemble_matrix_impl.h:92-200 (assemble_cells_matrix, non-lifting):
 * per cell gather the coordinates, zero Ae, call the kernel, insert. */
int orc_assemble_matrix(int kernel, const double* x, const int32_t* x_dofmap,
                        int npc, int32_t num_cells, const int32_t* dofmap,
                        int k, int bs, const double* coeffs, int cstride,
                        const double* constants, const int8_t* bc0,
                        const int8_t* bc1, const int64_t* row_ptr,
                        const int32_t* cols, double* values)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  const int nd = bs * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  double cdofs[3 * 20];
  const double zero = 0.0;
  for (int32_t c = 0; c < num_cells; ++c)
  {
    const int32_t* dofs = dofmap + (int64_t)c * k;
    for (int i = 0; i < npc; ++i)
      memcpy(cdofs + 3 * i, x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i],
             3 * sizeof(double));
    memset(Ae, 0, sizeof(double) * nd * nd);
    fn(Ae, coeffs ? coeffs + (int64_t)c * cstride : &zero,
       constants ? constants : &zero, cdofs, NULL, NULL, NULL);
    if (insert_cell(dofs, k, bs, Ae, bc0, bc1, row_ptr, cols, values))
    {
      free(Ae);
      return -1;
    }
  }
  free(Ae);
  return 0;
}

/* The same cell loop with the element tensors supplied by the caller (the
 * outputs of a user kernel, as in cpp/demo/custom_kernel/main.cpp:73-91):
 * E holds num_cells row-major (bs k) x (bs k) matrices. */
int orc_add_element_matrices(int32_t num_cells, const int32_t* dofmap, int k,
                             int bs, const double* E, const int8_t* bc0,
                             const int8_t* bc1, const int64_t* row_ptr,
                             const int32_t* cols, double* values)
{
  const int nd = bs * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  for (int32_t c = 0; c < num_cells; ++c)
  {
    memcpy(Ae, E + (int64_t)c * nd * nd, sizeof(double) * nd * nd);
    if (insert_cell(dofmap + (int64_t)c * k, k, bs, Ae, bc0, bc1, row_ptr, cols,
                    values))
    {
      free(Ae);
      return -1;
    }
  }
  free(Ae);
  return 0;
}

/* fem/assemble_vector_impl.h:72-116 (assemble_cells) with the element
 * vectors supplied by the caller: b[bs dofs_i + a] += e_c[bs i + a]. */
void orc_add_element_vectors(int32_t num_cells, const int32_t* dofmap, int k,
                             int bs, const double* e, double* b)
{
  for (int32_t c = 0; c < num_cells; ++c)
  {
    const int32_t* dofs = dofmap + (int64_t)c * k;
    const double* be = e + (int64_t)c * k * bs;
    for (int i = 0; i < k; ++i)
      for (int a = 0; a < bs; ++a)
        b[bs * dofs[i] + a] += be[bs * i + a];
  }
}

/* fem/assemble_vector_impl.h:72-116 (assemble_cells) */
void orc_assemble_vector(int kernel, const double* x, const int32_t* x_dofmap,
                         int npc, int32_t num_cells, const int32_t* dofmap,
                         int k, int bs, const double* coeffs, int cstride,
                         const double* constants, double* b)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  double be[60];
  double cdofs[3 * 20];
  const double zero = 0.0;
  for (int32_t c = 0; c < num_cells; ++c)
  {
    for (int i = 0; i < npc; ++i)
      memcpy(cdofs + 3 * i, x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i],
             3 * sizeof(double));
    memset(be, 0, sizeof(be));
    fn(be, coeffs ? coeffs + (int64_t)c * cstride : &zero,
       constants ? constants : &zero, cdofs, NULL, NULL, NULL);
    const int32_t* dofs = dofmap + (int64_t)c * k;
    for (int i = 0; i < k; ++i)
      for (int a = 0; a < bs; ++a)
        b[bs * dofs[i] + a] += be[bs * i + a];
  }
}

/* fem/assemble_vector_impl.h:118-174 (assemble_exterior_facets): for each
 * (cell, local facet) pair in order, the kernel with the local facet as
 * entity_local_index, then b[bs dofs_i + a] += be[bs i + a]. */
void orc_assemble_vector_facets(int kernel, const double* x,
                                const int32_t* x_dofmap, int npc,
                                const int32_t* dofmap, int k, int bs,
                                const double* coeffs, int cstride,
                                const double* constants, const int32_t* pairs,
                                int32_t npairs, double* b)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  double be[60];
  double cdofs[3 * 20];
  const double zero = 0.0;
  for (int32_t p = 0; p < npairs; ++p)
  {
    const int32_t c = pairs[2 * p];
    const int lf = pairs[2 * p + 1];
    for (int i = 0; i < npc; ++i)
      memcpy(cdofs + 3 * i, x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i],
             3 * sizeof(double));
    memset(be, 0, sizeof(be));
    fn(be, coeffs ? coeffs + (int64_t)c * cstride : &zero,
       constants ? constants : &zero, cdofs, &lf, NULL, NULL);
    const int32_t* dofs = dofmap + (int64_t)c * k;
    for (int i = 0; i < k; ++i)
      for (int a = 0; a < bs; ++a)
        b[bs * dofs[i] + a] += be[bs * i + a];
  }
}

/* fem/assemble_matrix_impl.h:264-383 (assemble_entities, non-lifting) on
 * exterior facets: for each (cell, local facet) pair in order, zero Ae, the
 * kernel with the local facet as entity_local_index, zero BC rows and
 * columns, insert (+=) into the CSR values. Returns -1 if an entry is not
 * in the sparsity. */
int orc_assemble_matrix_facets(int kernel, const double* x,
                               const int32_t* x_dofmap, int npc,
                               const int32_t* dofmap, int k, int bs,
                               const double* coeffs, int cstride,
                               const double* constants, const int32_t* pairs,
                               int32_t npairs, const int8_t* bc0,
                               const int8_t* bc1, const int64_t* row_ptr,
                               const int32_t* cols, double* values)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  const int nd = bs * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  double cdofs[3 * 20];
  const double zero = 0.0;
  for (int32_t p = 0; p < npairs; ++p)
  {
    const int32_t c = pairs[2 * p];
    const int lf = pairs[2 * p + 1];
    for (int i = 0; i < npc; ++i)
      memcpy(cdofs + 3 * i, x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i],
             3 * sizeof(double));
    memset(Ae, 0, sizeof(double) * nd * nd);
    fn(Ae, coeffs ? coeffs + (int64_t)c * cstride : &zero,
       constants ? constants : &zero, cdofs, &lf, NULL, NULL);
    if (insert_cell(dofmap + (int64_t)c * k, k, bs, Ae, bc0, bc1, row_ptr, cols,
                    values))
    {
      free(Ae);
      return -1;
    }
  }
  free(Ae);
  return 0;
}

/* fem/assemble_matrix_impl.h:442-667 (assemble_interior_facets,
 * non-lifting): per facet (cell0, lf0, cell1, lf1) in order, both cells'
 * coordinates, the kernel with local facets {lf0, lf1}, zero the rows and
 * columns of BC dofs of the joint dofmap (dofs of cell0 then cell1), then
 * mat_set(joint, joint, Ae): entry by entry, shared dofs twice. Scalar
 * spaces. Returns -1 if an entry is not in the sparsity. */
int orc_assemble_matrix_interior(int kernel, const double* x,
                                 const int32_t* x_dofmap, int npc,
                                 const int32_t* dofmap, int k,
                                 const double* constants, const int32_t* quads,
                                 int32_t nfacets, const int8_t* bc0,
                                 const int8_t* bc1, const int64_t* row_ptr,
                                 const int32_t* cols, double* values)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  const int nd = 2 * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  int32_t joint[40];
  double cdofs[2 * 3 * 20];
  const double zero = 0.0;
  for (int32_t f = 0; f < nfacets; ++f)
  {
    const int lf[2] = {quads[4 * f + 1], quads[4 * f + 3]};
    for (int s = 0; s < 2; ++s)
    {
      const int32_t c = quads[4 * f + 2 * s];
      for (int i = 0; i < npc; ++i)
        memcpy(cdofs + 3 * (s * npc + i),
               x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i], 3 * sizeof(double));
      for (int i = 0; i < k; ++i)
        joint[s * k + i] = dofmap[(int64_t)c * k + i];
    }
    memset(Ae, 0, sizeof(double) * nd * nd);
    fn(Ae, &zero, constants ? constants : &zero, cdofs, lf, NULL, NULL);
    if (bc0)
      for (int i = 0; i < nd; ++i)
        if (bc0[joint[i]])
          memset(Ae + i * nd, 0, sizeof(double) * nd);
    if (bc1)
      for (int j = 0; j < nd; ++j)
        if (bc1[joint[j]])
          for (int i = 0; i < nd; ++i)
            Ae[i * nd + j] = 0;
    for (int i = 0; i < nd; ++i)
    {
      const int32_t row = joint[i];
      for (int j = 0; j < nd; ++j)
      {
        int64_t p = -1;
        for (int64_t q = row_ptr[row]; q < row_ptr[row + 1]; ++q)
          if (cols[q] == joint[j])
            p = q;
        if (p < 0)
        {
          free(Ae);
          return -1;
        }
        values[p] += Ae[i * nd + j];
      }
    }
  }
  free(Ae);
  return 0;
}

/* lift_bc (fem/assemble_vector_impl.h:361-414) over an interior facet
 * integral: facets whose joint dofmap has a BC dof; b[joint_i] -=
 * Ae[i, j] alpha (g_j - x0_j) over the BC columns j, column by column. */
void orc_lift_bc_interior(int kernel, const double* x, const int32_t* x_dofmap,
                          int npc, const int32_t* dofmap, int k,
                          const double* constants, const int32_t* quads,
                          int32_t nfacets, const double* bc_values,
                          const int8_t* bc_markers, const double* x0,
                          double alpha, double* b)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  const int nd = 2 * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  int32_t joint[40];
  double cdofs[2 * 3 * 20];
  const double zero = 0.0;
  for (int32_t f = 0; f < nfacets; ++f)
  {
    const int lf[2] = {quads[4 * f + 1], quads[4 * f + 3]};
    int has = 0;
    for (int s = 0; s < 2; ++s)
    {
      const int32_t c = quads[4 * f + 2 * s];
      for (int i = 0; i < npc; ++i)
        memcpy(cdofs + 3 * (s * npc + i),
               x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i], 3 * sizeof(double));
      for (int i = 0; i < k; ++i)
      {
        joint[s * k + i] = dofmap[(int64_t)c * k + i];
        has |= bc_markers[joint[s * k + i]];
      }
    }
    if (!has)
      continue;
    memset(Ae, 0, sizeof(double) * nd * nd);
    fn(Ae, &zero, constants ? constants : &zero, cdofs, lf, NULL, NULL);
    for (int j = 0; j < nd; ++j)
    {
      if (!bc_markers[joint[j]])
        continue;
      const double v = alpha * (bc_values[joint[j]] - (x0 ? x0[joint[j]] : 0.0));
      for (int i = 0; i < nd; ++i)
        b[joint[i]] -= Ae[i * nd + j] * v;
    }
  }
  free(Ae);
}

/* fem/assemble_vector_impl.h:361-414 (lift_bc) on exterior facets
 * (assemble_entities<true>, fem/assemble_matrix_impl.h:316-322: facets whose
 * cell has a BC column dof): b[row] -= Ae[row, col] alpha (g - x0) over the
 * BC columns, facet by facet in order. */
void orc_lift_bc_facets(int kernel, const double* x, const int32_t* x_dofmap,
                        int npc, const int32_t* dofmap, int k, int bs,
                        const double* coeffs, int cstride,
                        const double* constants, const int32_t* pairs,
                        int32_t npairs, const double* bc_values,
                        const int8_t* bc_markers, const double* x0,
                        double alpha, double* b)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  const int nd = bs * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  double cdofs[3 * 20];
  const double zero = 0.0;
  for (int32_t p = 0; p < npairs; ++p)
  {
    const int32_t c = pairs[2 * p];
    const int lf = pairs[2 * p + 1];
    const int32_t* dofs = dofmap + (int64_t)c * k;
    int has = 0;
    for (int j = 0; j < k && !has; ++j)
      for (int a = 0; a < bs; ++a)
        if (bc_markers[bs * dofs[j] + a])
          has = 1;
    if (!has)
      continue;
    for (int i = 0; i < npc; ++i)
      memcpy(cdofs + 3 * i, x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i],
             3 * sizeof(double));
    memset(Ae, 0, sizeof(double) * nd * nd);
    fn(Ae, coeffs ? coeffs + (int64_t)c * cstride : &zero,
       constants ? constants : &zero, cdofs, &lf, NULL, NULL);
    for (int j = 0; j < k; ++j)
      for (int a = 0; a < bs; ++a)
      {
        const int32_t jj = bs * dofs[j] + a;
        if (!bc_markers[jj])
          continue;
        const double xb = bc_values[jj];
        const double x0v = x0 ? x0[jj] : 0.0;
        for (int i = 0; i < k; ++i)
          for (int m = 0; m < bs; ++m)
            b[bs * dofs[i] + m]
                -= Ae[(bs * i + m) * nd + bs * j + a] * alpha * (xb - x0v);
      }
  }
  free(Ae);
}

/* fem/assemble_vector_impl.h:361-414 (lift_bc) through
 * assemble_matrix<true> (cells with a BC column dof only, :139-143). */
void orc_lift_bc(int kernel, const double* x, const int32_t* x_dofmap,
                 int npc, int32_t num_cells, const int32_t* dofmap, int k,
                 int bs, const double* coeffs, int cstride,
                 const double* constants, const double* bc_values,
                 const int8_t* bc_markers, const double* x0, double alpha,
                 double* b)
{
  orc_kernel_fn fn = orc_get_kernel(kernel);
  const int nd = bs * k;
  double* Ae = (double*)malloc(sizeof(double) * nd * nd);
  double cdofs[3 * 20];
  const double zero = 0.0;
  for (int32_t c = 0; c < num_cells; ++c)
  {
    const int32_t* dofs = dofmap + (int64_t)c * k;
    int has = 0;
    for (int j = 0; j < k && !has; ++j)
      for (int a = 0; a < bs; ++a)
        if (bc_markers[bs * dofs[j] + a])
          has = 1;
    if (!has)
      continue;
    for (int i = 0; i < npc; ++i)
      memcpy(cdofs + 3 * i, x + 3 * (int64_t)x_dofmap[(int64_t)c * npc + i],
             3 * sizeof(double));
    memset(Ae, 0, sizeof(double) * nd * nd);
    fn(Ae, coeffs ? coeffs + (int64_t)c * cstride : &zero,
       constants ? constants : &zero, cdofs, NULL, NULL, NULL);
    for (int j = 0; j < k; ++j)
      for (int a = 0; a < bs; ++a)
      {
        const int32_t jj = bs * dofs[j] + a;
        if (!bc_markers[jj])
          continue;
        const double xb = bc_values[jj];
        const double x0v = x0 ? x0[jj] : 0.0;
        for (int i = 0; i < k; ++i)
          for (int m = 0; m < bs; ++m)
            b[bs * dofs[i] + m]
                -= Ae[(bs * i + m) * nd + bs * j + a] * alpha * (xb - x0v);
      }
  }
  free(Ae);
}

/* fem/DirichletBC.h:495-578 (DirichletBC::set): x[d] = alpha (g - x0) */
void orc_bc_set(double* b, const int32_t* dofs, int32_t ndofs,
                const double* g, const double* x0, double alpha)
{
  for (int32_t i = 0; i < ndofs; ++i)
  {
    const int32_t d = dofs[i];
    b[d] = alpha * (g[d] - (x0 ? x0[d] : 0.0));
  }
}

/* fem/assembler.h:644-686 (set_diagonal through mat_set_values with bs=1
 * data, i.e. insert_nonblocked_csr for blocked matrices). rows are
 * unrolled dof indices. */
int orc_set_diagonal(const int32_t* rows, int32_t nrows, int bs, double diag,
                     const int64_t* row_ptr, const int32_t* cols,
                     double* values)
{
  for (int32_t i = 0; i < nrows; ++i)
  {
    const int32_t r = rows[i] / bs, a = rows[i] % bs;
    int64_t p = -1;
    for (int64_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q)
      if (cols[q] == r)
        p = q;
    if (p < 0)
      return -1;
    values[p * bs * bs + a * bs + a] = diag;
  }
  return 0;
}
