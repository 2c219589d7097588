/*
 * Oracle: box mesh, dual graph, RCM, topology and entities.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Reference-cell sub-entity tables (Basix convention; reached from
 * mesh/cell_types.cpp:66-70 via basix::cell::topology). */
static const int TET_FACETS[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int TET_EDGES[6][2]
    = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
static const int HEX_FACETS[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                     {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};
static const int HEX_EDGES[12][2]
    = {{0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 5}, {2, 3},
       {2, 6}, {3, 7}, {4, 5}, {4, 6}, {5, 7}, {6, 7}};

void orc_free(void* p) { free(p); }

int orc_cell_num_entities(int celltype, int dim)
{
  static const int tet[4] = {4, 6, 4, 1};
  static const int hex[4] = {8, 12, 6, 1};
  return celltype == ORC_TET ? tet[dim] : hex[dim];
}

static int entity_nverts(int celltype, int dim)
{
  if (dim == 0)
    return 1;
  if (dim == 1)
    return 2;
  if (dim == 2)
    return celltype == ORC_TET ? 3 : 4;
  return celltype == ORC_TET ? 4 : 8;
}

static const int* entity_local_vertices(int celltype, int dim, int e)
{
  if (dim == 1)
    return celltype == ORC_TET ? TET_EDGES[e] : HEX_EDGES[e];
  return celltype == ORC_TET ? TET_FACETS[e] : HEX_FACETS[e];
}

/* ---------------------------------------------------------------------- */
/* generation.h:332-377 (create_geom) */
int64_t orc_box_num_vertices(const int64_t n[3])
{
  return (n[0] + 1) * (n[1] + 1) * (n[2] + 1);
}

int64_t orc_box_num_cells(const int64_t n[3], int celltype)
{
  return (celltype == ORC_TET ? 6 : 1) * n[0] * n[1] * n[2];
}

void orc_box_geometry(const double p0[3], const double p1[3],
                      const int64_t n[3], double* x)
{
  double ext[3];
  for (int i = 0; i < 3; ++i)
    ext[i] = (p1[i] - p0[i]) / (double)n[i];
  const int64_t npts = orc_box_num_vertices(n);
  const int64_t sqxy = (n[0] + 1) * (n[1] + 1);
  for (int64_t v = 0; v < npts; ++v)
  {
    const int64_t p = v % sqxy;
    const int64_t idx[3] = {p % (n[0] + 1), p / (n[0] + 1), v / sqxy};
    for (int i = 0; i < 3; ++i)
      x[3 * v + i] = p0[i] + (double)idx[i] * ext[i];
  }
}

/* generation.h:379-426 (build_tet) and :428-471 (build_hex) */
void orc_box_cells(const int64_t n[3], int celltype, int64_t* cells)
{
  const int64_t nx = n[0], ny = n[1], nz = n[2];
  int64_t pos = 0;
  for (int64_t i = 0; i < nx * ny * nz; ++i)
  {
    const int64_t iz = i / (nx * ny);
    const int64_t j = i % (nx * ny);
    const int64_t iy = j / nx;
    const int64_t ix = j % nx;
    const int64_t v0 = (iz * (ny + 1) + iy) * (nx + 1) + ix;
    const int64_t v1 = v0 + 1;
    const int64_t v2 = v0 + (nx + 1);
    const int64_t v3 = v1 + (nx + 1);
    const int64_t v4 = v0 + (nx + 1) * (ny + 1);
    const int64_t v5 = v1 + (nx + 1) * (ny + 1);
    const int64_t v6 = v2 + (nx + 1) * (ny + 1);
    const int64_t v7 = v3 + (nx + 1) * (ny + 1);
    if (celltype == ORC_TET)
    {
      const int64_t t[24] = {v0, v1, v3, v7, v0, v1, v7, v5, v0, v5, v7, v4,
                             v0, v3, v2, v7, v0, v6, v4, v7, v0, v2, v6, v7};
      memcpy(cells + pos, t, sizeof(t));
      pos += 24;
    }
    else
    {
      const int64_t h[8] = {v0, v1, v2, v3, v4, v5, v6, v7};
      memcpy(cells + pos, h, sizeof(h));
      pos += 8;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Stable LSD radix sort of a permutation by a 64-bit non-negative key,
 * 16 bits per pass (common/sort.h:81-177). */
static void radix_sort_perm(int32_t* perm, int64_t n, const int64_t* key)
{
  if (n <= 1)
    return;
  uint64_t maxv = 0;
  for (int64_t i = 0; i < n; ++i)
    if ((uint64_t)key[perm[i]] > maxv)
      maxv = (uint64_t)key[perm[i]];
  int passes = 0;
  while (maxv)
  {
    maxv >>= 16;
    ++passes;
  }
  int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * n);
  int64_t* count = (int64_t*)malloc(sizeof(int64_t) * 65536);
  int32_t* cur = perm;
  int32_t* nxt = buf;
  for (int pass = 0; pass < passes; ++pass)
  {
    const int shift = 16 * pass;
    memset(count, 0, sizeof(int64_t) * 65536);
    for (int64_t i = 0; i < n; ++i)
      count[((uint64_t)key[cur[i]] >> shift) & 0xFFFF]++;
    int64_t s = 0;
    for (int b = 0; b < 65536; ++b)
    {
      int64_t c = count[b];
      count[b] = s;
      s += c;
    }
    for (int64_t i = 0; i < n; ++i)
      nxt[count[((uint64_t)key[cur[i]] >> shift) & 0xFFFF]++] = cur[i];
    int32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (cur != perm)
    memcpy(perm, cur, sizeof(int32_t) * n);
  free(buf);
  free(count);
}

/* Row permutation sorting a column-major table lexicographically, column 0
 * most significant; stable (common/sort.h:200-270, sort_by_perm). */
static int32_t* sort_by_perm(int64_t* const* cols, int ncols, int64_t n)
{
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i)
    perm[i] = (int32_t)i;
  for (int c = ncols - 1; c >= 0; --c)
    radix_sort_perm(perm, n, cols[c]);
  return perm;
}

static void sort_small(int64_t* v, int n)
{
  for (int i = 1; i < n; ++i)
  {
    int64_t t = v[i];
    int j = i - 1;
    while (j >= 0 && v[j] > t)
    {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = t;
  }
}

/* ---------------------------------------------------------------------- */
/* mesh/graphbuild.cpp:564-900 (build_local_dual_graph), serial branch. */
int orc_dual_graph(int celltype, const int64_t* cells, int32_t num_cells,
                   int32_t** offsets_out, int32_t** data_out)
{
  const int nvpc = celltype == ORC_TET ? 4 : 8;
  const int nf = celltype == ORC_TET ? 4 : 6;
  const int fv = celltype == ORC_TET ? 3 : 4;
  const int64_t nfacets = (int64_t)num_cells * nf;

  /* Column-major facet table: fv vertex columns, then the cell column */
  int64_t* cols[5];
  for (int k = 0; k <= fv; ++k)
    cols[k] = (int64_t*)malloc(sizeof(int64_t) * (nfacets > 0 ? nfacets : 1));
  for (int32_t c = 0; c < num_cells; ++c)
  {
    for (int f = 0; f < nf; ++f)
    {
      const int* lv = entity_local_vertices(celltype, 2, f);
      int64_t row[4];
      for (int k = 0; k < fv; ++k)
        row[k] = cells[(int64_t)c * nvpc + lv[k]];
      sort_small(row, fv);
      const int64_t idx = (int64_t)c * nf + f;
      for (int k = 0; k < fv; ++k)
        cols[k][idx] = row[k];
      cols[fv][idx] = c;
    }
  }

  int32_t* perm = sort_by_perm(cols, fv, nfacets);

  /* Walk runs of equal facets; add one edge per pair (cell_a, cell_b). */
  int64_t cap = nfacets;
  int32_t(*edges)[2] = malloc(sizeof(int32_t[2]) * (cap > 0 ? cap : 1));
  int64_t nedges = 0;
  int64_t it = 0;
  while (it < nfacets)
  {
    int64_t end = it + 1;
    while (end < nfacets)
    {
      int eq = 1;
      for (int k = 0; k < fv; ++k)
        if (cols[k][perm[end]] != cols[k][perm[it]])
        {
          eq = 0;
          break;
        }
      if (!eq)
        break;
      ++end;
    }
    for (int64_t a = it; a < end; ++a)
      for (int64_t b = a + 1; b < end; ++b)
      {
        edges[nedges][0] = (int32_t)cols[fv][perm[a]];
        edges[nedges][1] = (int32_t)cols[fv][perm[b]];
        ++nedges;
      }
    it = end;
  }

  int32_t* offsets = (int32_t*)calloc(num_cells + 1, sizeof(int32_t));
  for (int64_t e = 0; e < nedges; ++e)
  {
    offsets[edges[e][0] + 1]++;
    offsets[edges[e][1] + 1]++;
  }
  for (int32_t i = 0; i < num_cells; ++i)
    offsets[i + 1] += offsets[i];
  int32_t* data = (int32_t*)malloc(sizeof(int32_t) * (2 * nedges + 1));
  int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (num_cells + 1));
  memcpy(pos, offsets, sizeof(int32_t) * (num_cells + 1));
  for (int64_t e = 0; e < nedges; ++e)
  {
    data[pos[edges[e][0]]++] = edges[e][1];
    data[pos[edges[e][1]]++] = edges[e][0];
  }
  free(pos);
  free(edges);
  free(perm);
  for (int k = 0; k <= fv; ++k)
    free(cols[k]);
  *offsets_out = offsets;
  *data_out = data;
  return (int)nedges;
}

/* ---------------------------------------------------------------------- */
/* graph/ordering.cpp:22-59 (create_level_structure). Returns number of
 * levels; fills order[] (BFS order) and level_off[] (levels+1). */
static int level_structure(const int32_t* off, const int32_t* adj, int32_t n,
                           int32_t s, int32_t* order, int32_t* level_off,
                           int8_t* lab)
{
  memset(lab, 0, n);
  lab[s] = 1;
  int32_t size = 1;
  order[0] = s;
  int nl = 0;
  level_off[0] = 0;
  while (size > level_off[nl])
  {
    level_off[nl + 1] = size;
    for (int32_t i = level_off[nl]; i < level_off[nl + 1]; ++i)
    {
      const int32_t node = order[i];
      for (int32_t k = off[node]; k < off[node + 1]; ++k)
      {
        const int32_t w = adj[k];
        if (lab[w])
          continue;
        order[size++] = w;
        lab[w] = 1;
      }
    }
    ++nl;
  }
  return nl;
}

/* graph/ordering.cpp:65-144 (rcm_reorder_unlabelled); returns count of rv */
static int32_t rcm_unlabelled(const int32_t* off, const int32_t* adj,
                              int32_t n, const int32_t* rlabel, int32_t* rv)
{
#define DEG(i) (off[(i) + 1] - off[(i)])
  int32_t v = 0;
  int32_t dmin = 0x7fffffff;
  for (int32_t i = 0; i < n; ++i)
  {
    const int32_t d = DEG(i);
    if (rlabel[i] == -1 && d < dmin)
    {
      v = i;
      dmin = d;
    }
  }
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  int32_t* loff = (int32_t*)malloc(sizeof(int32_t) * (n + 2));
  int32_t* order2 = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  int32_t* loff2 = (int32_t*)malloc(sizeof(int32_t) * (n + 2));
  int8_t* lab = (int8_t*)malloc(n > 0 ? n : 1);
  int nl = level_structure(off, adj, n, v, order, loff, lab);
  for (;;)
  {
    /* first vertex of minimal degree in the last level */
    int32_t s = order[loff[nl - 1]];
    for (int32_t i = loff[nl - 1] + 1; i < loff[nl]; ++i)
      if (DEG(order[i]) < DEG(s))
        s = order[i];
    int nl2 = level_structure(off, adj, n, s, order2, loff2, lab);
    if (nl2 > nl)
    {
      v = s;
      int32_t* t = order;
      order = order2;
      order2 = t;
      t = loff;
      loff = loff2;
      loff2 = t;
      nl = nl2;
    }
    else
      break;
  }

  /* Cuthill-McKee from v; each vertex's new neighbours sorted stably by
   * degree (std::ranges::sort on <= 16 elements is an insertion sort). */
  memset(lab, 0, n);
  int32_t cnt = 0;
  rv[cnt++] = v;
  lab[v] = 1;
  int32_t nbr[64];
  for (int32_t cur = 0; cur < cnt; ++cur)
  {
    int nn = 0;
    const int32_t u = rv[cur];
    for (int32_t k = off[u]; k < off[u + 1]; ++k)
    {
      const int32_t w = adj[k];
      if (!lab[w])
      {
        nbr[nn++] = w;
        lab[w] = 1;
      }
    }
    for (int i = 1; i < nn; ++i)
    {
      int32_t t = nbr[i];
      int j = i - 1;
      while (j >= 0 && DEG(nbr[j]) > DEG(t))
      {
        nbr[j + 1] = nbr[j];
        --j;
      }
      nbr[j + 1] = t;
    }
    for (int i = 0; i < nn; ++i)
      rv[cnt++] = nbr[i];
  }
  /* reverse */
  for (int32_t i = 0; i < cnt / 2; ++i)
  {
    int32_t t = rv[i];
    rv[i] = rv[cnt - 1 - i];
    rv[cnt - 1 - i] = t;
  }
  free(order);
  free(loff);
  free(order2);
  free(loff2);
  free(lab);
  return cnt;
#undef DEG
}

/* graph/ordering.cpp:148-170 (reorder_rcm): remap[old] = new */
void orc_rcm(const int32_t* offsets, const int32_t* data, int32_t n,
             int32_t* remap)
{
  for (int32_t i = 0; i < n; ++i)
    remap[i] = -1;
  int32_t* rv = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
  int32_t count = 0;
  while (count < n)
  {
    int32_t m = rcm_unlabelled(offsets, data, n, remap, rv);
    for (int32_t i = 0; i < m; ++i)
      remap[rv[i]] = count++;
  }
  free(rv);
}

/* ---------------------------------------------------------------------- */
/* mesh/generation.h:104-134 -> mesh/utils.h:1073-1259 (create_mesh, no
 * partitioner) -> utils.h:232-309 (dual graph + reorder_fn + reorder_list)
 * -> Topology.cpp:1222-1286 (vertex numbering by first appearance) ->
 * Geometry (P1 geometry dofmap equals the cell-vertex map). */
/* mesh::create_mesh on one rank (mesh/utils.h:1073-1259): dual graph,
 * RCM cell reordering, vertices numbered by first appearance. cells_in
 * (nc x nvpc, input vertex indices) and x_in (nv x 3) are input order. */
orc_mesh* orc_mesh_create(int celltype, const int64_t* cells_in, int32_t nc,
                          const double* x_in, int32_t nv)
{
  orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
  m->celltype = celltype;
  m->nvpc = celltype == ORC_TET ? 4 : 8;
  m->num_cells = nc;
  m->num_vertices = nv;
  const int nvpc = m->nvpc;

  int32_t *goff, *gdata;
  orc_dual_graph(celltype, cells_in, (int32_t)nc, &goff, &gdata);
  int32_t* remap = (int32_t*)malloc(sizeof(int32_t) * (nc + 1));
  orc_rcm(goff, gdata, (int32_t)nc, remap);
  free(goff);
  free(gdata);

  /* reorder_list (mesh/utils.h:57-71) and original index update */
  int64_t* cells_r = (int64_t*)malloc(sizeof(int64_t) * nvpc * nc);
  m->original_cell_index = (int64_t*)malloc(sizeof(int64_t) * nc);
  for (int64_t c = 0; c < nc; ++c)
  {
    memcpy(cells_r + (int64_t)remap[c] * nvpc, cells_in + c * nvpc,
           sizeof(int64_t) * nvpc);
    m->original_cell_index[remap[c]] = c;
  }
  free(remap);

  /* Vertex numbering by first appearance (single rank: all owned,
   * owned_vertices == 0..nv-1, Topology.cpp:1224-1242) */
  int32_t* local = (int32_t*)malloc(sizeof(int32_t) * nv);
  for (int64_t v = 0; v < nv; ++v)
    local[v] = -1;
  int32_t cnt = 0;
  m->cells = (int32_t*)malloc(sizeof(int32_t) * nvpc * nc);
  for (int64_t i = 0; i < nc * nvpc; ++i)
  {
    const int64_t v = cells_r[i];
    if (local[v] < 0)
      local[v] = cnt++;
    m->cells[i] = local[v];
  }
  m->x = (double*)malloc(sizeof(double) * 3 * nv);
  m->input_vertex_index = (int64_t*)malloc(sizeof(int64_t) * nv);
  for (int64_t v = 0; v < nv; ++v)
  {
    memcpy(m->x + 3 * (int64_t)local[v], x_in + 3 * v, 3 * sizeof(double));
    m->input_vertex_index[local[v]] = v;
  }
  free(local);
  free(cells_r);
  return m;
}

/* mesh::create_box (mesh/generation.h:104-134): the box's cells and
 * geometry through create_mesh. */
orc_mesh* orc_mesh_create_box(const double p0[3], const double p1[3],
                              const int64_t n[3], int celltype)
{
  const int nvpc = celltype == ORC_TET ? 4 : 8;
  const int64_t nv = orc_box_num_vertices(n);
  const int64_t nc = orc_box_num_cells(n, celltype);
  double* x_in = (double*)malloc(sizeof(double) * 3 * nv);
  int64_t* cells_in = (int64_t*)malloc(sizeof(int64_t) * nvpc * nc);
  orc_box_geometry(p0, p1, n, x_in);
  orc_box_cells(n, celltype, cells_in);
  orc_mesh* m = orc_mesh_create(celltype, cells_in, (int32_t)nc, x_in, (int32_t)nv);
  free(cells_in);
  free(x_in);
  return m;
}

void orc_mesh_free(orc_mesh* m)
{
  if (!m)
    return;
  free(m->x);
  free(m->cells);
  free(m->original_cell_index);
  free(m->input_vertex_index);
  free(m);
}

/* ---------------------------------------------------------------------- */
/* mesh/topologycomputation.cpp:709-952 (compute_entities_by_key_matching)
 * and :57-248 (build_entity_list), single rank: entity numbers are the
 * ranks of the sorted vertex keys in a stable lexicographic sort;
 * get_local_indexing (:586-615) is the identity. */
int32_t orc_entities(const orc_mesh* m, int dim, int32_t** c2e_out,
                     int32_t** e2v_out)
{
  const int ct = m->celltype;
  const int ne = orc_cell_num_entities(ct, dim);
  const int nve = entity_nverts(ct, dim);
  const int64_t n = (int64_t)m->num_cells * ne;
  int64_t* cols[4];
  for (int k = 0; k < nve; ++k)
    cols[k] = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  for (int32_t c = 0; c < m->num_cells; ++c)
    for (int e = 0; e < ne; ++e)
    {
      const int* lv = entity_local_vertices(ct, dim, e);
      int64_t row[4];
      for (int k = 0; k < nve; ++k)
        row[k] = m->cells[(int64_t)c * m->nvpc + lv[k]];
      sort_small(row, nve);
      for (int k = 0; k < nve; ++k)
        cols[k][(int64_t)c * ne + e] = row[k];
    }
  int32_t* perm = sort_by_perm(cols, nve, n);
  int32_t* c2e = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  int32_t* e2v = (int32_t*)malloc(sizeof(int32_t) * (n * nve > 0 ? n * nve : 1));
  int32_t count = 0;
  int64_t it = 0;
  while (it < n)
  {
    int64_t end = it + 1;
    while (end < n)
    {
      int eq = 1;
      for (int k = 0; k < nve; ++k)
        if (cols[k][perm[end]] != cols[k][perm[it]])
        {
          eq = 0;
          break;
        }
      if (!eq)
        break;
      ++end;
    }
    for (int64_t a = it; a < end; ++a)
      c2e[perm[a]] = count;
    for (int k = 0; k < nve; ++k)
      e2v[(int64_t)count * nve + k] = (int32_t)cols[k][perm[it]];
    ++count;
    it = end;
  }
  free(perm);
  for (int k = 0; k < nve; ++k)
    free(cols[k]);
  *c2e_out = c2e;
  *e2v_out = e2v;
  return count;
}

/* mesh/permutationcomputation.cpp:241-310 (compute_edge_reflections):
 * reflected iff (pos(v_hi) < pos(v_lo)) in the cell's vertex list. */
void orc_edge_reflections(const orc_mesh* m, const int32_t* c2e,
                          const int32_t* e2v, uint32_t* bits)
{
  const int ne = orc_cell_num_entities(m->celltype, 1);
  for (int32_t c = 0; c < m->num_cells; ++c)
  {
    const int32_t* cv = m->cells + (int64_t)c * m->nvpc;
    uint32_t b = 0;
    for (int e = 0; e < ne; ++e)
    {
      const int32_t* ev = e2v + 2 * (int64_t)c2e[(int64_t)c * ne + e];
      int it0 = -1, it1 = -1;
      for (int k = 0; k < m->nvpc; ++k)
      {
        if (cv[k] == ev[0] && it0 < 0)
          it0 = k;
        if (cv[k] == ev[1] && it1 < 0)
          it1 = k;
      }
      if ((it1 < it0) == (ev[1] > ev[0]))
        b |= (1u << e);
    }
    bits[c] = b;
  }
}

/* mesh/utils.cpp:61-89 (exterior_facet_indices): facets with one cell. */
/* fem::compute_integration_domains(IntegralType::exterior_facet, ...)
 * (fem/utils.cpp:154-273) over the exterior facets (mesh/utils.h:180-194)
 * in facet-index order: for each, its cell and local index in that cell. */
int32_t orc_facet_domain(const orc_mesh* m, int marker, int axis, double value,
                         int32_t** pairs)
{
  const int tdim = 3;
  int32_t *c2f, *f2v;
  const int32_t nfac = orc_entities(m, tdim - 1, &c2f, &f2v);
  const int nf = orc_cell_num_entities(m->celltype, tdim - 1);
  const int nvf = m->celltype == ORC_TET ? 3 : 4;
  int32_t* count = (int32_t*)calloc(nfac, sizeof(int32_t));
  int32_t* first = (int32_t*)malloc(sizeof(int32_t) * (nfac > 0 ? nfac : 1));
  for (int64_t i = 0; i < (int64_t)m->num_cells * nf; ++i)
  {
    if (count[c2f[i]] == 0)
      first[c2f[i]] = (int32_t)i;
    count[c2f[i]]++;
  }
  int32_t* out = (int32_t*)malloc(sizeof(int32_t) * 2 * (nfac > 0 ? nfac : 1));
  int32_t n = 0;
  for (int32_t f = 0; f < nfac; ++f)
  {
    if (count[f] != 1)
      continue;
    if (marker)
    {
      int on = 1;
      for (int a = 0; a < nvf; ++a)
        if (m->x[3 * (int64_t)f2v[(int64_t)f * nvf + a] + axis] != value)
          on = 0;
      if (!on)
        continue;
    }
    out[2 * n] = first[f] / nf;
    out[2 * n + 1] = first[f] % nf;
    ++n;
  }
  free(count);
  free(first);
  free(c2f);
  free(f2v);
  *pairs = out;
  return n;
}

/* fem::compute_integration_domains(interior_facet) (fem/utils.cpp:223-253,
 * impl::get_cell_facet_pairs<2>): every facet with two cells, in facet-index
 * order, as (cell0, lf0, cell1, lf1) in the order of the facet-to-cell
 * links (the transpose of cell-to-facet: ascending cells). */
int32_t orc_interior_facet_domain(const orc_mesh* m, int32_t** quads)
{
  const int tdim = 3;
  int32_t *c2f, *f2v;
  const int32_t nfac = orc_entities(m, tdim - 1, &c2f, &f2v);
  const int nf = orc_cell_num_entities(m->celltype, tdim - 1);
  int32_t* count = (int32_t*)calloc(nfac > 0 ? nfac : 1, sizeof(int32_t));
  int32_t* at = (int32_t*)malloc(sizeof(int32_t) * 2 * (nfac > 0 ? nfac : 1));
  for (int64_t i = 0; i < (int64_t)m->num_cells * nf; ++i)
  {
    const int32_t f = c2f[i];
    if (count[f] < 2)
      at[2 * f + count[f]] = (int32_t)i;
    count[f]++;
  }
  int32_t* out = (int32_t*)malloc(sizeof(int32_t) * 4 * (nfac > 0 ? nfac : 1));
  int32_t n = 0;
  for (int32_t f = 0; f < nfac; ++f)
  {
    if (count[f] != 2)
      continue;
    for (int s = 0; s < 2; ++s)
    {
      out[4 * n + 2 * s] = at[2 * f + s] / nf;
      out[4 * n + 2 * s + 1] = at[2 * f + s] % nf;
    }
    ++n;
  }
  free(count);
  free(at);
  free(c2f);
  free(f2v);
  *quads = out;
  return n;
}

int32_t orc_exterior_facets(const orc_mesh* m, int32_t** cell_facet)
{
  const int tdim = 3;
  int32_t *c2f, *f2v;
  const int32_t nfac = orc_entities(m, tdim - 1, &c2f, &f2v);
  const int nf = orc_cell_num_entities(m->celltype, tdim - 1);
  int32_t* count = (int32_t*)calloc(nfac, sizeof(int32_t));
  for (int64_t i = 0; i < (int64_t)m->num_cells * nf; ++i)
    count[c2f[i]]++;
  int32_t nb = 0;
  for (int64_t i = 0; i < (int64_t)m->num_cells * nf; ++i)
    if (count[c2f[i]] == 1)
      ++nb;
  int32_t* out = (int32_t*)malloc(sizeof(int32_t) * 2 * (nb > 0 ? nb : 1));
  nb = 0;
  for (int32_t c = 0; c < m->num_cells; ++c)
    for (int f = 0; f < nf; ++f)
      if (count[c2f[(int64_t)c * nf + f]] == 1)
      {
        out[2 * nb] = c;
        out[2 * nb + 1] = f;
        ++nb;
      }
  free(count);
  free(c2f);
  free(f2v);
  *cell_facet = out;
  return nb;
}
