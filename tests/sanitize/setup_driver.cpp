// The host setup of libdfx (dolfinx_amd/csrc/setup.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer: boxes of tetrahedra and
// hexahedra (one above 200,000 cells, which takes the threaded paths),
// create_mesh on a shuffled box, Lagrange spaces of degree 1-3 and block
// size 3, boundary dofs, exterior and interior facet domains, and the
// sparsity patterns with and without interior-facet couplings. Built and
// run by tests/test_sanitizers.py; exits non-zero when a call fails.
#include "dfx.h"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace dfx
{
// the library keeps this message for dfx_last_error; here it is printed
void set_last_error(const std::string& msg) { std::fprintf(stderr, "error: %s\n", msg.c_str()); }
} // namespace dfx

#define CHECK(call)                                                            \
  do                                                                           \
  {                                                                            \
    if ((call) != DFX_OK)                                                      \
    {                                                                          \
      std::fprintf(stderr, "failed: %s\n", #call);                             \
      return 1;                                                                \
    }                                                                          \
  } while (0)

static int spaces_and_patterns(dfx_mesh* m, int maxdeg)
{
  for (int deg = 1; deg <= maxdeg; ++deg)
    for (int bs : {1, 3})
    {
      if (bs == 3 && deg != 2)
        continue;
      dfx_space* V = nullptr;
      CHECK(dfx_space_create(m, deg, bs, &V));
      int32_t nn = 0, k = 0, b = 0;
      CHECK(dfx_space_sizes(V, &nn, &k, &b));
      int32_t nc = 0, nv = 0, npc = 0;
      CHECK(dfx_mesh_sizes(m, &nc, &nv, &npc));
      std::vector<int32_t> dm(std::size_t(nc) * k);
      CHECK(dfx_space_dofmap(V, dm.data()));
      int32_t count = 0;
      CHECK(dfx_locate_boundary_dofs(V, 0, 0, 0.0, &count, nullptr));
      std::vector<int32_t> dofs(std::max(count, 1));
      CHECK(dfx_locate_boundary_dofs(V, 0, 0, 0.0, &count, dofs.data()));
      CHECK(dfx_locate_boundary_dofs(V, 1, 0, 0.0, &count, nullptr));
      dfx_sparsity* sp = nullptr;
      CHECK(dfx_sparsity_create(V, V, &sp));
      int32_t nrows = 0, sbs = 0;
      int64_t nnz = 0;
      CHECK(dfx_sparsity_sizes(sp, &nrows, &nnz, &sbs));
      std::vector<int64_t> rp(std::size_t(nrows) + 1);
      std::vector<int32_t> cols(std::max<int64_t>(nnz, 1));
      CHECK(dfx_sparsity_copy(sp, rp.data(), cols.data()));
      CHECK(dfx_sparsity_destroy(sp));
      if (deg == 1 && bs == 1 && npc == 4)
      {
        int32_t nf = 0;
        CHECK(dfx_interior_facet_domain(m, &nf, nullptr));
        std::vector<int32_t> quads(4 * std::size_t(std::max(nf, 1)));
        CHECK(dfx_interior_facet_domain(m, &nf, quads.data()));
        CHECK(dfx_sparsity_create_facets(V, V, nf, quads.data(), &sp));
        CHECK(dfx_sparsity_destroy(sp));
      }
      CHECK(dfx_space_destroy(V));
    }
  int32_t np = 0;
  CHECK(dfx_facet_domain(m, 0, 0, 0.0, &np, nullptr));
  std::vector<int32_t> pairs(2 * std::size_t(std::max(np, 1)));
  CHECK(dfx_facet_domain(m, 0, 0, 0.0, &np, pairs.data()));
  CHECK(dfx_facet_domain(m, 1, 2, 1.0, &np, nullptr));
  return 0;
}

int main()
{
  const double p0[3] = {0, 0, 0}, p1[3] = {1, 2, 1};
  struct Case
  {
    int cell;
    int64_t n[3];
    int maxdeg;
  };
  // the 40 x 40 x 25 tetrahedra (240,000 cells) take the threaded paths
  const Case cases[] = {{DFX_TET, {3, 2, 2}, 3}, {DFX_HEX, {4, 3, 2}, 1},
                        {DFX_TET, {40, 40, 25}, 2}, {DFX_HEX, {12, 12, 12}, 1}};
  for (const Case& c : cases)
  {
    dfx_mesh* m = nullptr;
    CHECK(dfx_mesh_create_box(p0, p1, c.n, c.cell, &m));
    if (spaces_and_patterns(m, c.maxdeg))
      return 1;
    // create_mesh on the same cells, shuffled, from a copy of the box
    int32_t nc = 0, nv = 0, npc = 0;
    CHECK(dfx_mesh_sizes(m, &nc, &nv, &npc));
    std::vector<double> x(3 * std::size_t(nv));
    std::vector<int32_t> cells(std::size_t(nc) * npc);
    CHECK(dfx_mesh_copy(m, x.data(), cells.data(), nullptr, nullptr));
    std::vector<int32_t> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::shuffle(order.begin(), order.end(), std::mt19937(7));
    std::vector<int64_t> shuffled(cells.size());
    for (int32_t q = 0; q < nc; ++q)
      for (int j = 0; j < npc; ++j)
        shuffled[std::size_t(q) * npc + j] = cells[std::size_t(order[q]) * npc + j];
    dfx_mesh* m2 = nullptr;
    CHECK(dfx_mesh_create(c.cell, nc, shuffled.data(), nv, x.data(), &m2));
    if (spaces_and_patterns(m2, 1))
      return 1;
    CHECK(dfx_mesh_destroy(m2));
    CHECK(dfx_mesh_destroy(m));
  }
  // an invalid input is reported, not a crash
  const int64_t bad[3] = {0, 1, 1};
  dfx_mesh* mb = nullptr;
  if (dfx_mesh_create_box(p0, p1, bad, DFX_TET, &mb) == DFX_OK)
    dfx_mesh_destroy(mb);
  std::printf("setup ok\n");
  return 0;
}
