// Host-side setup of the assemble-and-solve path: box mesh, cell
// reordering, vertex/entity numbering, dofmaps, Dirichlet dofs and the
// sparsity pattern. Integer-only work, done once per mesh; its outputs must
// equal the reference's bit for bit, so every ordering rule below follows
// the reference step it names.
#include "common.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

namespace
{

// Reference-cell sub-entity tables in the Basix convention (reached from
// mesh/cell_types.cpp:66-70 via basix::cell::topology).
constexpr int kTetFacets[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
constexpr int kTetEdges[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
constexpr int kHexFacets[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {0, 2, 4, 6},
                                  {1, 3, 5, 7}, {2, 3, 6, 7}, {4, 5, 6, 7}};
constexpr int kHexEdges[12][2] = {{0, 1}, {0, 2}, {0, 4}, {1, 3}, {1, 5}, {2, 3},
                                  {2, 6}, {3, 7}, {4, 5}, {4, 6}, {5, 7}, {6, 7}};

int num_entities_per_cell(int cell, int dim)
{
  if (dim == 1)
    return cell == DFX_TET ? 6 : 12;
  return cell == DFX_TET ? 4 : 6;
}

int entity_vertex_count(int cell, int dim)
{
  return dim == 1 ? 2 : (cell == DFX_TET ? 3 : 4);
}

const int* entity_vertices(int cell, int dim, int e)
{
  if (dim == 1)
    return cell == DFX_TET ? kTetEdges[e] : kHexEdges[e];
  return cell == DFX_TET ? kTetFacets[e] : kHexFacets[e];
}

unsigned hw_threads()
{
  unsigned n = std::thread::hardware_concurrency();
  return std::clamp(n, 1u, 16u);
}

template <typename F>
void parallel_for(std::int64_t n, F&& f)
{
  const unsigned nt = n > 200000 ? hw_threads() : 1;
  if (nt == 1)
  {
    f(std::int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
  {
    std::int64_t a = n * t / nt, b = n * (t + 1) / nt;
    th.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto& t : th)
    t.join();
}

/// Order of items whose keys are sorted vertex tuples of width W (<= 4):
/// lexicographic by key, ties kept in item order. Implemented as a
/// counting sort on the leading vertex followed by a comparison sort of
/// each (small) bucket on the remaining vertices and the item index. This
/// is the order the reference's stable LSD radix sort_by_perm
/// (common/sort.h:200-270) produces.
std::vector<std::int32_t> lexicographic_order(const std::vector<std::int32_t>& keys,
                                              int W, std::int32_t vmax)
{
  const std::int64_t n = static_cast<std::int64_t>(keys.size()) / W;
  std::vector<std::int64_t> start(static_cast<std::size_t>(vmax) + 2, 0);
  for (std::int64_t i = 0; i < n; ++i)
    ++start[keys[i * W] + 1];
  std::partial_sum(start.begin(), start.end(), start.begin());
  std::vector<std::int32_t> order(n);
  {
    std::vector<std::int64_t> pos(start.begin(), start.end() - 1);
    for (std::int64_t i = 0; i < n; ++i)
      order[pos[keys[i * W]]++] = static_cast<std::int32_t>(i);
  }
  parallel_for(vmax + 1,
               [&](std::int64_t a, std::int64_t b)
               {
                 for (std::int64_t v = a; v < b; ++v)
                 {
                   auto first = order.begin() + start[v];
                   auto last = order.begin() + start[v + 1];
                   if (last - first < 2)
                     continue;
                   std::sort(first, last,
                             [&keys, W](std::int32_t p, std::int32_t q)
                             {
                               for (int k = 1; k < W; ++k)
                               {
                                 std::int32_t kp = keys[std::int64_t(p) * W + k];
                                 std::int32_t kq = keys[std::int64_t(q) * W + k];
                                 if (kp != kq)
                                   return kp < kq;
                               }
                               return p < q;
                             });
                 }
               });
  return order;
}

bool same_key(const std::vector<std::int32_t>& keys, int W, std::int32_t a,
              std::int32_t b)
{
  for (int k = 0; k < W; ++k)
    if (keys[std::int64_t(a) * W + k] != keys[std::int64_t(b) * W + k])
      return false;
  return true;
}

template <typename T>
void sort_small(T* v, int n)
{
  for (int i = 1; i < n; ++i)
    for (int j = i; j > 0 && v[j - 1] > v[j]; --j)
      std::swap(v[j - 1], v[j]);
}

struct Graph
{
  std::vector<std::int32_t> offsets, adj;
  std::int32_t degree(std::int32_t i) const { return offsets[i + 1] - offsets[i]; }
};

/// Cell dual graph (mesh/graphbuild.cpp:564-900): cells sharing a facet
/// are linked; each cell's links appear in the order of the shared facet's
/// sorted vertex key (edges are emitted while walking facets in key order).
Graph dual_graph(int cell, const std::vector<std::int32_t>& cells,
                 std::int32_t nc, std::int32_t nv)
{
  const int nvpc = cell == DFX_TET ? 4 : 8;
  const int nf = num_entities_per_cell(cell, 2);
  const int W = entity_vertex_count(cell, 2);
  std::vector<std::int32_t> keys(std::int64_t(nc) * nf * W);
  parallel_for(nc,
               [&](std::int64_t a, std::int64_t b)
               {
                 for (std::int64_t c = a; c < b; ++c)
                   for (int f = 0; f < nf; ++f)
                   {
                     std::int32_t* k = keys.data() + (c * nf + f) * W;
                     const int* lv = entity_vertices(cell, 2, f);
                     for (int j = 0; j < W; ++j)
                       k[j] = cells[c * nvpc + lv[j]];
                     sort_small(k, W);
                   }
               });
  const std::vector<std::int32_t> order = lexicographic_order(keys, W, nv);
  std::vector<std::array<std::int32_t, 2>> edges;
  edges.reserve(std::size_t(nc) * nf / 2 + 1);
  const std::int64_t n = static_cast<std::int64_t>(order.size());
  for (std::int64_t i = 0; i < n;)
  {
    std::int64_t j = i + 1;
    while (j < n && same_key(keys, W, order[i], order[j]))
      ++j;
    for (std::int64_t a = i; a < j; ++a)
      for (std::int64_t b = a + 1; b < j; ++b)
        edges.push_back({order[a] / nf, order[b] / nf});
    i = j;
  }
  Graph g;
  g.offsets.assign(std::size_t(nc) + 1, 0);
  for (auto& e : edges)
  {
    ++g.offsets[e[0] + 1];
    ++g.offsets[e[1] + 1];
  }
  std::partial_sum(g.offsets.begin(), g.offsets.end(), g.offsets.begin());
  g.adj.resize(g.offsets.back());
  std::vector<std::int32_t> pos(g.offsets.begin(), g.offsets.end() - 1);
  for (auto& e : edges)
  {
    g.adj[pos[e[0]]++] = e[1];
    g.adj[pos[e[1]]++] = e[0];
  }
  return g;
}

/// Breadth-first level structure rooted at s (graph/ordering.cpp:22-59).
/// Returns the number of levels and fills `order`/`level_start`.
int levels_from(const Graph& g, std::int32_t s, std::vector<std::int32_t>& order,
                std::vector<std::int32_t>& level_start, std::vector<char>& seen)
{
  std::fill(seen.begin(), seen.end(), 0);
  order.clear();
  level_start.clear();
  order.push_back(s);
  seen[s] = 1;
  level_start.push_back(0);
  std::size_t head = 0;
  while (order.size() > static_cast<std::size_t>(level_start.back()))
  {
    const std::size_t end = order.size();
    level_start.push_back(static_cast<std::int32_t>(end));
    for (; head < end; ++head)
    {
      const std::int32_t u = order[head];
      for (std::int32_t k = g.offsets[u]; k < g.offsets[u + 1]; ++k)
      {
        const std::int32_t w = g.adj[k];
        if (!seen[w])
        {
          seen[w] = 1;
          order.push_back(w);
        }
      }
    }
  }
  return static_cast<int>(level_start.size()) - 1;
}

/// graph::reorder_rcm (graph/ordering.cpp:65-170): returns new index of
/// each node. Root: first node of minimum degree, then George-Liu
/// pseudo-peripheral search (first minimum-degree node of the last level,
/// accept while the depth grows); Cuthill-McKee appends each node's newly
/// found neighbours stably sorted by degree; the order is reversed.
std::vector<std::int32_t> reorder_rcm(const Graph& g)
{
  const std::int32_t n = static_cast<std::int32_t>(g.offsets.size()) - 1;
  std::vector<std::int32_t> remap(n, -1);
  std::vector<char> seen(n);
  std::vector<std::int32_t> ord, lev, ord2, lev2, rv, nbr;
  std::int32_t count = 0;
  while (count < n)
  {
    std::int32_t v = 0, dmin = INT32_MAX;
    for (std::int32_t i = 0; i < n; ++i)
      if (remap[i] == -1 && g.degree(i) < dmin)
      {
        v = i;
        dmin = g.degree(i);
      }
    int depth = levels_from(g, v, ord, lev, seen);
    for (;;)
    {
      const std::int32_t* last = ord.data() + lev[depth - 1];
      const std::int32_t* last_end = ord.data() + lev[depth];
      std::int32_t s = *std::min_element(last, last_end,
                                         [&g](std::int32_t a, std::int32_t b)
                                         { return g.degree(a) < g.degree(b); });
      int depth2 = levels_from(g, s, ord2, lev2, seen);
      if (depth2 <= depth)
        break;
      v = s;
      std::swap(ord, ord2);
      std::swap(lev, lev2);
      depth = depth2;
    }
    std::fill(seen.begin(), seen.end(), 0);
    rv.clear();
    rv.push_back(v);
    seen[v] = 1;
    for (std::size_t cur = 0; cur < rv.size(); ++cur)
    {
      nbr.clear();
      const std::int32_t u = rv[cur];
      for (std::int32_t k = g.offsets[u]; k < g.offsets[u + 1]; ++k)
        if (!seen[g.adj[k]])
        {
          seen[g.adj[k]] = 1;
          nbr.push_back(g.adj[k]);
        }
      std::stable_sort(nbr.begin(), nbr.end(), [&g](std::int32_t a, std::int32_t b)
                       { return g.degree(a) < g.degree(b); });
      rv.insert(rv.end(), nbr.begin(), nbr.end());
    }
    for (auto it = rv.rbegin(); it != rv.rend(); ++it)
      remap[*it] = count++;
  }
  return remap;
}

/// Entities of dimension dim (topologycomputation.cpp:709-952): numbered by
/// the rank of their sorted vertex key; the identity local indexing of a
/// single rank (:586-615).
void compute_entities(dfx_mesh& m, int dim)
{
  if (m.have_entities[dim])
    return;
  const int ne = num_entities_per_cell(m.cell, dim);
  const int W = entity_vertex_count(m.cell, dim);
  const std::int64_t n = std::int64_t(m.num_cells) * ne;
  std::vector<std::int32_t> keys(n * W);
  parallel_for(m.num_cells,
               [&](std::int64_t a, std::int64_t b)
               {
                 for (std::int64_t c = a; c < b; ++c)
                   for (int e = 0; e < ne; ++e)
                   {
                     std::int32_t* k = keys.data() + (c * ne + e) * W;
                     const int* lv = entity_vertices(m.cell, dim, e);
                     for (int j = 0; j < W; ++j)
                       k[j] = m.cells[c * m.nvpc + lv[j]];
                     sort_small(k, W);
                   }
               });
  const std::vector<std::int32_t> order = lexicographic_order(keys, W, m.num_vertices);
  std::vector<std::int32_t>& c2e = m.c2e[dim];
  std::vector<std::int32_t>& e2v = m.e2v[dim];
  c2e.assign(n, -1);
  e2v.clear();
  std::int32_t count = 0;
  for (std::int64_t i = 0; i < n;)
  {
    std::int64_t j = i + 1;
    while (j < n && same_key(keys, W, order[i], order[j]))
      ++j;
    for (std::int64_t a = i; a < j; ++a)
      c2e[order[a]] = count;
    e2v.insert(e2v.end(), keys.begin() + std::int64_t(order[i]) * W,
               keys.begin() + std::int64_t(order[i]) * W + W);
    ++count;
    i = j;
  }
  m.num_entities[dim] = count;
  m.have_entities[dim] = true;
}

int lagrange_dofs(int cell, int degree)
{
  if (cell == DFX_HEX)
    return 8;
  return (degree + 1) * (degree + 2) * (degree + 3) / 6;
}

/// Local dofs in the closure of local facet f (Basix entity closure dofs).
int facet_closure(int cell, int degree, int f, int* out)
{
  int n = 0;
  if (cell == DFX_HEX)
  {
    for (int i = 0; i < 4; ++i)
      out[n++] = kHexFacets[f][i];
    return n;
  }
  bool on[4] = {false, false, false, false};
  for (int i = 0; i < 3; ++i)
  {
    on[kTetFacets[f][i]] = true;
    out[n++] = kTetFacets[f][i];
  }
  const int ede = degree - 1;
  for (int e = 0; e < 6 && ede > 0; ++e)
    if (on[kTetEdges[e][0]] && on[kTetEdges[e][1]])
      for (int j = 0; j < ede; ++j)
        out[n++] = 4 + ede * e + j;
  if (degree >= 3)
    out[n++] = 4 + 6 * ede + f;
  return n;
}

/// mesh::create_mesh on one rank (mesh/utils.h:1073-1259) from input cells
/// (input vertex indices) and input geometry: cells reordered by RCM on the
/// dual graph, vertices numbered by first appearance in the reordered cells.
void build_mesh(dfx_mesh& mesh, const std::vector<std::int32_t>& cin,
                const std::vector<double>& xin)
{
  dfx_mesh* m = &mesh;
  const int cell = m->cell;
  const int nvpc = m->nvpc;
  const std::int64_t nc = m->num_cells;
  const std::int64_t nv = m->num_vertices;
  // Reorder cells by RCM on the dual graph (mesh/utils.h:232-309)
  const Graph g = dual_graph(cell, cin, m->num_cells, m->num_vertices);
  const std::vector<std::int32_t> remap = reorder_rcm(g);
  std::vector<std::int32_t> cr(nc * nvpc);
  m->original_cell.resize(nc);
  for (std::int64_t c = 0; c < nc; ++c)
  {
    std::copy_n(cin.begin() + c * nvpc, nvpc, cr.begin() + std::int64_t(remap[c]) * nvpc);
    m->original_cell[remap[c]] = c;
  }
  // Vertices numbered by first appearance (Topology.cpp:1222-1258)
  std::vector<std::int32_t> local(nv, -1);
  std::int32_t next = 0;
  m->cells.resize(nc * nvpc);
  for (std::int64_t i = 0; i < nc * nvpc; ++i)
  {
    std::int32_t& l = local[cr[i]];
    if (l < 0)
      l = next++;
    m->cells[i] = l;
  }
  m->x.resize(3 * nv);
  m->input_vertex.resize(nv);
  for (std::int64_t v = 0; v < nv; ++v)
  {
    std::copy_n(xin.begin() + 3 * v, 3, m->x.begin() + 3 * std::int64_t(local[v]));
    m->input_vertex[local[v]] = v;
  }
}

} // namespace

extern "C" {

dfx_status dfx_mesh_create_box(const double p0[3], const double p1[3],
                               const int64_t n[3], int cell, dfx_mesh** out)
{
  return dfx::guard(
      [&]
      {
        dfx::require(out && p0 && p1 && n, "null argument");
        dfx::require(cell == DFX_TET || cell == DFX_HEX, "unsupported cell type");
        for (int i = 0; i < 3; ++i)
        {
          dfx::require(n[i] >= 1, "At least one cell is required.");
          dfx::require(p0[i] < p1[i], "It must hold p[0] < p[1].");
        }
        auto m = std::make_unique<dfx_mesh>();
        m->cell = cell;
        m->nvpc = cell == DFX_TET ? 4 : 8;
        const std::int64_t nx = n[0], ny = n[1], nz = n[2];
        const std::int64_t nv = (nx + 1) * (ny + 1) * (nz + 1);
        const std::int64_t nc = (cell == DFX_TET ? 6 : 1) * nx * ny * nz;
        dfx::require(nv < INT32_MAX && nc * m->nvpc < INT32_MAX, "mesh too large");
        m->num_vertices = static_cast<std::int32_t>(nv);
        m->num_cells = static_cast<std::int32_t>(nc);

        // Geometry: p0 + idx * extent, x fastest (generation.h:332-377)
        std::vector<double> xin(3 * nv);
        const double h[3] = {(p1[0] - p0[0]) / double(nx), (p1[1] - p0[1]) / double(ny),
                             (p1[2] - p0[2]) / double(nz)};
        for (std::int64_t v = 0; v < nv; ++v)
        {
          const std::int64_t idx[3]
              = {v % (nx + 1), (v / (nx + 1)) % (ny + 1), v / ((nx + 1) * (ny + 1))};
          for (int i = 0; i < 3; ++i)
            xin[3 * v + i] = p0[i] + double(idx[i]) * h[i];
        }
        // Cells (generation.h:379-471): six tets around the v0-v7 diagonal,
        // or the lexicographic hexahedron
        const int nvpc = m->nvpc;
        std::vector<std::int32_t> cin(nc * nvpc);
        for (std::int64_t iz = 0, q = 0; iz < nz; ++iz)
          for (std::int64_t iy = 0; iy < ny; ++iy)
            for (std::int64_t ix = 0; ix < nx; ++ix)
            {
              std::int32_t v[8];
              v[0] = static_cast<std::int32_t>((iz * (ny + 1) + iy) * (nx + 1) + ix);
              v[1] = v[0] + 1;
              v[2] = v[0] + static_cast<std::int32_t>(nx + 1);
              v[3] = v[1] + static_cast<std::int32_t>(nx + 1);
              for (int k = 0; k < 4; ++k)
                v[4 + k] = v[k] + static_cast<std::int32_t>((nx + 1) * (ny + 1));
              if (cell == DFX_TET)
              {
                static constexpr int t[24] = {0, 1, 3, 7, 0, 1, 7, 5, 0, 5, 7, 4,
                                              0, 3, 2, 7, 0, 6, 4, 7, 0, 2, 6, 7};
                for (int k = 0; k < 24; ++k)
                  cin[q++] = v[t[k]];
              }
              else
                for (int k = 0; k < 8; ++k)
                  cin[q++] = v[k];
            }

        build_mesh(*m, cin, xin);
        *out = m.release();
      });
}

dfx_status dfx_mesh_create(int cell, int32_t num_cells, const int64_t* cells,
                           int32_t num_vertices, const double* x, dfx_mesh** out)
{
  return dfx::guard(
      [&]
      {
        dfx::require(out && (num_cells == 0 || cells) && (num_vertices == 0 || x),
                     "null argument");
        dfx::require(cell == DFX_TET || cell == DFX_HEX, "unsupported cell type");
        dfx::require(num_cells >= 0 && num_vertices >= 0, "negative size");
        auto m = std::make_unique<dfx_mesh>();
        m->cell = cell;
        m->nvpc = cell == DFX_TET ? 4 : 8;
        dfx::require(std::int64_t(num_cells) * m->nvpc < INT32_MAX, "mesh too large");
        m->num_vertices = num_vertices;
        m->num_cells = num_cells;
        const std::int64_t n = std::int64_t(num_cells) * m->nvpc;
        std::vector<std::int32_t> cin(n);
        std::vector<char> used(num_vertices, 0);
        for (std::int64_t i = 0; i < n; ++i)
        {
          dfx::require(cells[i] >= 0 && cells[i] < num_vertices, "cell vertex out of range");
          cin[i] = static_cast<std::int32_t>(cells[i]);
          used[cin[i]] = 1;
        }
        for (std::int32_t v = 0; v < num_vertices; ++v)
          dfx::require(used[v], "every vertex must belong to a cell");
        for (std::int32_t c = 0; c < num_cells; ++c)
          for (int a = 0; a < m->nvpc; ++a)
            for (int b = a + 1; b < m->nvpc; ++b)
              dfx::require(cin[std::int64_t(c) * m->nvpc + a] != cin[std::int64_t(c) * m->nvpc + b],
                           "a cell repeats a vertex");
        std::vector<double> xin(x, x + 3 * std::int64_t(num_vertices));
        build_mesh(*m, cin, xin);
        *out = m.release();
      });
}

dfx_status dfx_mesh_destroy(dfx_mesh* m)
{
  delete m;
  return DFX_OK;
}

dfx_status dfx_mesh_sizes(const dfx_mesh* m, int32_t* nc, int32_t* nv, int32_t* npc)
{
  return dfx::guard(
      [&]
      {
        dfx::require(m, "null mesh");
        if (nc)
          *nc = m->num_cells;
        if (nv)
          *nv = m->num_vertices;
        if (npc)
          *npc = m->nvpc;
      });
}

dfx_status dfx_mesh_copy(const dfx_mesh* m, double* x, int32_t* cells,
                         int64_t* original_cell_index, int64_t* input_vertex_index)
{
  return dfx::guard(
      [&]
      {
        dfx::require(m, "null mesh");
        if (x)
          std::copy(m->x.begin(), m->x.end(), x);
        if (cells)
          std::copy(m->cells.begin(), m->cells.end(), cells);
        if (original_cell_index)
          std::copy(m->original_cell.begin(), m->original_cell.end(), original_cell_index);
        if (input_vertex_index)
          std::copy(m->input_vertex.begin(), m->input_vertex.end(), input_vertex_index);
      });
}

dfx_status dfx_space_create(dfx_mesh* m, int degree, int bs, dfx_space** out)
{
  return dfx::guard(
      [&]
      {
        dfx::require(m && out, "null argument");
        dfx::require(bs >= 1 && bs <= 3, "block size must be 1..3");
        if (m->cell == DFX_HEX)
          dfx::require(degree == 1, "hexahedra: only Q1 is supported");
        else
          dfx::require(degree >= 1 && degree <= 3, "tetrahedra: degree 1..3");
        auto V = std::make_unique<dfx_space>();
        V->mesh = m;
        V->degree = degree;
        V->bs = bs;
        const int k = lagrange_dofs(m->cell, degree);
        V->dofs_per_cell = k;
        const int ede = m->cell == DFX_TET ? degree - 1 : 0;
        const int fde = (m->cell == DFX_TET && degree >= 3) ? 1 : 0;
        if (ede > 0)
          compute_entities(*m, 1);
        if (fde > 0)
          compute_entities(*m, 2);
        const std::int32_t nv = m->num_vertices;
        const std::int32_t ne = ede > 0 ? m->num_entities[1] : 0;
        const std::int32_t nf = fde > 0 ? m->num_entities[2] : 0;
        const std::int64_t nc = m->num_cells;

        // Entity-based numbering (dofmapbuilder.cpp:153-397): vertices,
        // then edge interiors, then face interiors.
        std::vector<std::int32_t> basic(nc * k);
        for (std::int64_t c = 0; c < nc; ++c)
        {
          std::int32_t* d = basic.data() + c * k;
          for (int i = 0; i < m->nvpc; ++i)
            d[i] = m->cells[c * m->nvpc + i];
          for (int e = 0; e < 6 && ede > 0; ++e)
            for (int j = 0; j < ede; ++j)
              d[4 + ede * e + j] = nv + ede * m->c2e[1][c * 6 + e] + j;
          for (int f = 0; f < 4 && fde > 0; ++f)
            d[4 + 6 * ede + f] = nv + ede * ne + m->c2e[2][c * 4 + f];
        }
        // Renumber by first traversal over cells; no graph reordering
        // (reorder_fn == nullptr, dofmapbuilder.cpp:421-490)
        const std::int32_t ndofs = nv + ede * ne + fde * nf;
        std::vector<std::int32_t> o2n(ndofs, -1);
        std::int32_t next = 0;
        for (std::int32_t d : basic)
          if (o2n[d] < 0)
            o2n[d] = next++;
        for (auto& d : o2n)
          if (d < 0)
            d = next++;
        V->dofmap.resize(nc * k);
        for (std::int64_t i = 0; i < nc * k; ++i)
          V->dofmap[i] = o2n[basic[i]];
        // P3: a reflected edge swaps its two interior dofs (fem/utils.cpp:
        // 61-73; reflection test permutationcomputation.cpp:241-310)
        if (ede == 2)
        {
          for (std::int64_t c = 0; c < nc; ++c)
          {
            const std::int32_t* cv = m->cells.data() + c * 4;
            std::int32_t* d = V->dofmap.data() + c * k;
            for (int e = 0; e < 6; ++e)
            {
              // edge vertices in cell order are (cv[a], cv[b]) with a < b;
              // the global orientation runs from the smaller vertex index
              if (cv[kTetEdges[e][0]] > cv[kTetEdges[e][1]])
                std::swap(d[4 + 2 * e], d[5 + 2 * e]);
            }
          }
        }
        V->num_nodes = ndofs;
        *out = V.release();
      });
}

dfx_status dfx_space_destroy(dfx_space* V)
{
  delete V;
  return DFX_OK;
}

dfx_status dfx_space_sizes(const dfx_space* V, int32_t* nn, int32_t* k, int32_t* bs)
{
  return dfx::guard(
      [&]
      {
        dfx::require(V, "null space");
        if (nn)
          *nn = V->num_nodes;
        if (k)
          *k = V->dofs_per_cell;
        if (bs)
          *bs = V->bs;
      });
}

dfx_status dfx_space_dofmap(const dfx_space* V, int32_t* dofmap)
{
  return dfx::guard(
      [&]
      {
        dfx::require(V && dofmap, "null argument");
        std::copy(V->dofmap.begin(), V->dofmap.end(), dofmap);
      });
}

dfx_status dfx_locate_boundary_dofs(const dfx_space* V, int marker, int axis,
                                    double value, int32_t* count, int32_t* dofs)
{
  return dfx::guard(
      [&]
      {
        dfx::require(V && count, "null argument");
        dfx::require(marker == 0 || (marker == 1 && axis >= 0 && axis < 3), "bad marker");
        dfx_mesh& m = *V->mesh;
        compute_entities(m, 2);
        const int nf = num_entities_per_cell(m.cell, 2);
        const int fv = entity_vertex_count(m.cell, 2);
        std::vector<std::int32_t> ncell(m.num_entities[2], 0);
        for (std::int32_t f : m.c2e[2])
          ++ncell[f];
        std::vector<std::int32_t> nodes;
        int loc[16];
        for (std::int64_t c = 0; c < m.num_cells; ++c)
          for (int f = 0; f < nf; ++f)
          {
            if (ncell[m.c2e[2][c * nf + f]] != 1)
              continue; // interior facet
            if (marker == 1)
            {
              bool all = true;
              const int* lv = entity_vertices(m.cell, 2, f);
              for (int j = 0; j < fv; ++j)
              {
                const std::int32_t v = m.cells[c * m.nvpc + lv[j]];
                all = all && std::abs(m.x[3 * std::int64_t(v) + axis] - value) <= 1e-12;
              }
              if (!all)
                continue;
            }
            const int nl = facet_closure(m.cell, V->degree, f, loc);
            for (int j = 0; j < nl; ++j)
              nodes.push_back(V->dofmap[c * V->dofs_per_cell + loc[j]]);
          }
        std::sort(nodes.begin(), nodes.end());
        nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
        const std::int64_t n = std::int64_t(nodes.size()) * V->bs;
        if (dofs)
        {
          dfx::require(*count >= n, "dofs buffer too small");
          for (std::size_t i = 0; i < nodes.size(); ++i)
            for (int b = 0; b < V->bs; ++b)
              dofs[i * V->bs + b] = nodes[i] * V->bs + b;
        }
        *count = static_cast<std::int32_t>(n);
      });
}

dfx_status dfx_facet_domain(const dfx_mesh* mesh, int marker, int axis, double value,
                            int32_t* count, int32_t* pairs)
{
  return dfx::guard(
      [&]
      {
        dfx::require(mesh && count, "null argument");
        dfx::require(marker == 0 || (marker == 1 && axis >= 0 && axis < 3), "bad marker");
        dfx_mesh& m = const_cast<dfx_mesh&>(*mesh);
        compute_entities(m, 2);
        const int nf = num_entities_per_cell(m.cell, 2);
        const int fv = entity_vertex_count(m.cell, 2);
        // cells per facet and the first (cell, local facet) of each
        std::vector<std::int32_t> ncell(m.num_entities[2], 0), first(m.num_entities[2], 0);
        for (std::int64_t i = 0; i < std::int64_t(m.num_cells) * nf; ++i)
          if (ncell[m.c2e[2][i]]++ == 0)
            first[m.c2e[2][i]] = static_cast<std::int32_t>(i);
        std::int32_t n = 0;
        for (std::int32_t f = 0; f < m.num_entities[2]; ++f)
        {
          if (ncell[f] != 1)
            continue; // interior facet
          const std::int32_t c = first[f] / nf, lf = first[f] % nf;
          if (marker == 1)
          {
            bool all = true;
            const int* lv = entity_vertices(m.cell, 2, lf);
            for (int j = 0; j < fv; ++j)
            {
              const std::int32_t v = m.cells[std::int64_t(c) * m.nvpc + lv[j]];
              all = all && std::abs(m.x[3 * std::int64_t(v) + axis] - value) <= 1e-12;
            }
            if (!all)
              continue;
          }
          if (pairs)
          {
            dfx::require(n < *count, "pairs buffer too small");
            pairs[2 * n] = c;
            pairs[2 * n + 1] = lf;
          }
          ++n;
        }
        *count = n;
      });
}

dfx_status dfx_interior_facet_domain(const dfx_mesh* mesh, int32_t* count, int32_t* quads)
{
  return dfx::guard(
      [&]
      {
        dfx::require(mesh && count, "null argument");
        dfx_mesh& m = const_cast<dfx_mesh&>(*mesh);
        compute_entities(m, 2);
        const int nf = num_entities_per_cell(m.cell, 2);
        // the (cell, local facet) positions of each facet in cell order:
        // the facet-to-cell links of the transposed connectivity
        std::vector<std::int32_t> ncell(m.num_entities[2], 0), at(2 * std::size_t(m.num_entities[2]), 0);
        for (std::int64_t i = 0; i < std::int64_t(m.num_cells) * nf; ++i)
        {
          const std::int32_t f = m.c2e[2][i];
          if (ncell[f] < 2)
            at[2 * std::size_t(f) + ncell[f]] = static_cast<std::int32_t>(i);
          ++ncell[f];
        }
        std::int32_t n = 0;
        for (std::int32_t f = 0; f < m.num_entities[2]; ++f)
        {
          if (ncell[f] != 2)
            continue;
          if (quads)
          {
            dfx::require(n < *count, "quads buffer too small");
            for (int s = 0; s < 2; ++s)
            {
              quads[4 * n + 2 * s] = at[2 * std::size_t(f) + s] / nf;
              quads[4 * n + 2 * s + 1] = at[2 * std::size_t(f) + s] % nf;
            }
          }
          ++n;
        }
        *count = n;
      });
}

dfx_status dfx_sparsity_create(const dfx_space* V0, const dfx_space* V1,
                               dfx_sparsity** out)
{
  return dfx_sparsity_create_facets(V0, V1, 0, nullptr, out);
}

dfx_status dfx_sparsity_create_facets(const dfx_space* V0, const dfx_space* V1, int32_t nfacets,
                                      const int32_t* quads, dfx_sparsity** out)
{
  return dfx::guard(
      [&]
      {
        dfx::require(V0 && V1 && out, "null argument");
        dfx::require(V0->mesh == V1->mesh, "spaces on different meshes");
        dfx::require(V0->bs == V1->bs, "rectangular blocks are not supported");
        dfx::require(nfacets >= 0 && (nfacets == 0 || quads), "bad interior facets");
        const std::int64_t nc = V0->mesh->num_cells;
        const int k0 = V0->dofs_per_cell, k1 = V1->dofs_per_cell;
        const std::int32_t nrows = V0->num_nodes;
        const int nfc = num_entities_per_cell(V0->mesh->cell, 2);
        for (std::int64_t e = 0; e < 2 * std::int64_t(nfacets); ++e)
          dfx::require(quads[2 * e] >= 0 && quads[2 * e] < nc && quads[2 * e + 1] >= 0
                           && quads[2 * e + 1] < nfc,
                       "interior facet out of range");
        // Interior facets couple the dofs of both their cells
        // (sparsitybuild::interior_facets, fem/sparsitybuild.cpp:16-): row
        // -> incident facets
        std::vector<std::int64_t> finc_off(std::size_t(nrows) + 1, 0);
        std::vector<std::int32_t> finc;
        if (nfacets > 0)
        {
          for (std::int64_t e = 0; e < nfacets; ++e)
            for (int s = 0; s < 2; ++s)
              for (int i = 0; i < k0; ++i)
                ++finc_off[V0->dofmap[std::int64_t(quads[4 * e + 2 * s]) * k0 + i] + 1];
          std::partial_sum(finc_off.begin(), finc_off.end(), finc_off.begin());
          finc.resize(finc_off.back());
          std::vector<std::int64_t> pos(finc_off.begin(), finc_off.end() - 1);
          for (std::int64_t e = 0; e < nfacets; ++e)
            for (int s = 0; s < 2; ++s)
              for (int i = 0; i < k0; ++i)
                finc[pos[V0->dofmap[std::int64_t(quads[4 * e + 2 * s]) * k0 + i]]++]
                    = static_cast<std::int32_t>(e);
        }
        // Transpose of the row dofmap: row -> incident (cell) list
        std::vector<std::int64_t> inc_off(std::size_t(nrows) + 1, 0);
        for (std::int32_t d : V0->dofmap)
          ++inc_off[d + 1];
        std::partial_sum(inc_off.begin(), inc_off.end(), inc_off.begin());
        std::vector<std::int32_t> inc(inc_off.back());
        {
          std::vector<std::int64_t> pos(inc_off.begin(), inc_off.end() - 1);
          for (std::int64_t c = 0; c < nc; ++c)
            for (int i = 0; i < k0; ++i)
              inc[pos[V0->dofmap[c * k0 + i]]++] = static_cast<std::int32_t>(c);
        }
        // Each row: union of the column dofs of its cells, sorted and
        // deduplicated (the finalised row of la/SparsityPattern.cpp:448-479)
        std::vector<std::int32_t> rowlen(nrows);
        std::vector<std::vector<std::int32_t>> chunk_cols(hw_threads() + 1);
        std::vector<std::int64_t> chunk_start;
        auto build_rows = [&](std::int64_t a, std::int64_t b, std::vector<std::int32_t>& dst)
        {
          std::vector<std::int32_t> buf;
          for (std::int64_t r = a; r < b; ++r)
          {
            buf.clear();
            for (std::int64_t p = inc_off[r]; p < inc_off[r + 1]; ++p)
            {
              const std::int32_t* d = V1->dofmap.data() + std::int64_t(inc[p]) * k1;
              buf.insert(buf.end(), d, d + k1);
            }
            if (nfacets > 0)
              for (std::int64_t p = finc_off[r]; p < finc_off[r + 1]; ++p)
                for (int s = 0; s < 2; ++s)
                {
                  const std::int32_t* d
                      = V1->dofmap.data() + std::int64_t(quads[4 * std::int64_t(finc[p]) + 2 * s]) * k1;
                  buf.insert(buf.end(), d, d + k1);
                }
            std::sort(buf.begin(), buf.end());
            buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
            rowlen[r] = static_cast<std::int32_t>(buf.size());
            dst.insert(dst.end(), buf.begin(), buf.end());
          }
        };
        const unsigned nt = nrows > 200000 ? hw_threads() : 1;
        {
          std::vector<std::thread> th;
          for (unsigned t = 0; t < nt; ++t)
          {
            std::int64_t a = std::int64_t(nrows) * t / nt, b = std::int64_t(nrows) * (t + 1) / nt;
            th.emplace_back([&, a, b, t] { build_rows(a, b, chunk_cols[t]); });
          }
          for (auto& t : th)
            t.join();
        }
        auto sp = std::make_unique<dfx_sparsity>();
        sp->nrows = nrows;
        sp->ncols = V1->num_nodes;
        sp->bs = V0->bs;
        sp->row_ptr.resize(std::size_t(nrows) + 1);
        sp->row_ptr[0] = 0;
        for (std::int32_t r = 0; r < nrows; ++r)
          sp->row_ptr[r + 1] = sp->row_ptr[r] + rowlen[r];
        sp->cols.reserve(sp->row_ptr.back());
        for (unsigned t = 0; t < nt; ++t)
          sp->cols.insert(sp->cols.end(), chunk_cols[t].begin(), chunk_cols[t].end());
        *out = sp.release();
      });
}

dfx_status dfx_sparsity_destroy(dfx_sparsity* sp)
{
  delete sp;
  return DFX_OK;
}

dfx_status dfx_sparsity_sizes(const dfx_sparsity* sp, int32_t* nrows, int64_t* nnz,
                              int32_t* bs)
{
  return dfx::guard(
      [&]
      {
        dfx::require(sp, "null sparsity");
        if (nrows)
          *nrows = sp->nrows;
        if (nnz)
          *nnz = sp->row_ptr.back();
        if (bs)
          *bs = sp->bs;
      });
}

dfx_status dfx_sparsity_copy(const dfx_sparsity* sp, int64_t* row_ptr, int32_t* cols)
{
  return dfx::guard(
      [&]
      {
        dfx::require(sp, "null sparsity");
        if (row_ptr)
          std::copy(sp->row_ptr.begin(), sp->row_ptr.end(), row_ptr);
        if (cols)
          std::copy(sp->cols.begin(), sp->cols.end(), cols);
      });
}

} // extern "C"
