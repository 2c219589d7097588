// The custom-kernel pattern of cpp/demo/custom_kernel/main.cpp: forms
// whose cell integrals are element kernels with the reference's kernel ABI
// (kernelptr_t), here written by hand -- the P1 tetrahedron stiffness
// kappa grad u . grad v + c u v (DG0 kappa, constant c) and the load f v (f
// at the cell's vertices) -- assembled with the drop-in's
// fem::assemble_matrix / assemble_vector on the GPU, beside the library's
// own device kernels for the same forms; and the same stiffness kernel
// called in a host loop over the cells whose element matrices go through
// the matrix's MatSet functor (A.mat_add_values(), then scatter_rev), as
// the reference's custom assembly loop does. An exterior facet integral,
// beta u v ds on every boundary facet, is given the same way: a
// hand-written facet kernel called per (cell, local facet) pair.
// Usage: custom_kernel_demo <n> <out.bin>
// Writes the matrix and vector values (FP64) to out.bin: kernel-ABI form,
// its vector, device kernels, their vector, the host-loop matrix, the
// kernel-ABI form without BCs, then the facet matrix from the facet kernel
// and from the library's device Robin kernel.
#include <dolfinx_amd/dolfinx_amd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>

using namespace dolfinx_amd;

namespace
{
/// Gradients of the barycentric coordinates of a tetrahedron and its
/// volume (from the inverse of the Jacobian x1 - x0, x2 - x0, x3 - x0).
double p1_gradients(const double* X, double (&grad)[4][3])
{
  double J[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      J[r][c] = X[3 * (c + 1) + r] - X[r];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                     - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                     + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  // rows of J^{-1} (cofactors / det): gradient of lambda_{c+1}
  double inv[3][3];
  inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  for (int k = 0; k < 3; ++k)
  {
    grad[k + 1][0] = inv[k][0];
    grad[k + 1][1] = inv[k][1];
    grad[k + 1][2] = inv[k][2];
    grad[0][k] = -(inv[0][k] + inv[1][k] + inv[2][k]);
  }
  return std::abs(det) / 6.0;
}

/// kappa grad u . grad v + c u v on a P1 tetrahedron; w = [kappa], c = [c].
void stiffness_kernel(double* A, const double* w, const double* c, const double* X,
                      const int*, const std::uint8_t*, void*)
{
  double g[4][3];
  const double vol = p1_gradients(X, g);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      A[4 * i + j] += w[0] * vol * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2])
                      + c[0] * vol * (i == j ? 2.0 : 1.0) / 20.0;
}

/// f v on a P1 tetrahedron with f given at the vertices: b_i = |K|/20 (f_i + sum f).
void load_kernel(double* b, const double* w, const double*, const double* X, const int*,
                 const std::uint8_t*, void*)
{
  double g[4][3];
  const double vol = p1_gradients(X, g);
  const double sum = w[0] + w[1] + w[2] + w[3];
  for (int i = 0; i < 4; ++i)
    b[i] += vol / 20.0 * (w[i] + sum);
}
/// beta u v ds on local facet lf of a P1 tetrahedron (the facet opposite
/// vertex lf): |F|/12 (1 + delta_ij) beta for i, j on the facet; w = [beta].
void robin_kernel(double* A, const double* w, const double*, const double* X, const int* lf,
                  const std::uint8_t*, void*)
{
  int v[3], k = 0;
  for (int i = 0; i < 4; ++i)
    if (i != lf[0])
      v[k++] = i;
  double e1[3], e2[3];
  for (int d = 0; d < 3; ++d)
  {
    e1[d] = X[3 * v[1] + d] - X[3 * v[0] + d];
    e2[d] = X[3 * v[2] + d] - X[3 * v[0] + d];
  }
  const double n0 = e1[1] * e2[2] - e1[2] * e2[1];
  const double n1 = e1[2] * e2[0] - e1[0] * e2[2];
  const double n2 = e1[0] * e2[1] - e1[1] * e2[0];
  const double area = 0.5 * std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      A[4 * v[i] + v[j]] += w[0] * area / 12.0 * (i == j ? 2.0 : 1.0);
}
} // namespace

int main(int argc, char* argv[])
{
  if (argc < 3)
  {
    std::cerr << "usage: custom_kernel_demo <n> <out.bin>\n";
    return 2;
  }
  const std::int64_t n = std::atoll(argv[1]);
  try
  {
    auto msh = mesh::create_box({{{0.0, 0.0, 0.0}, {1.0, 1.0, 1.0}}}, {n, n, n},
                                mesh::CellType::tetrahedron);
    auto V = fem::create_functionspace(msh, 1);
    const std::int32_t nc = msh->num_cells();
    std::vector<double> kappa(nc), f(std::size_t(nc) * 4);
    const auto orig = msh->original_cell_index();
    for (std::int32_t c = 0; c < nc; ++c)
    {
      kappa[c] = 1.0 + 0.5 * std::sin(double(orig[c]));
      for (int i = 0; i < 4; ++i)
        f[4 * c + i] = std::cos(0.1 * double(orig[c]) + i);
    }
    const double creac = 0.75;
    fem::Form<double> ak(&stiffness_kernel, 2, V, kappa, 1, {creac});
    fem::Form<double> Lk(&load_kernel, 1, V, f, 4, {});
    fem::Form<double> ad(DFX_K_P1TET_A, V, kappa, 1, {creac});
    fem::Form<double> Ld(DFX_K_P1TET_L, V, f, 4, {});
    fem::DirichletBC<double> bc(V, fem::locate_boundary_dofs(*V));

    la::MatrixCSR<double> Ak(fem::create_sparsity_pattern(ak));
    fem::assemble_matrix(Ak.mat_add_values(), ak, {bc});
    fem::set_diagonal<double>(Ak.mat_set_values(), *V, {bc}, 1.0);
    la::MatrixCSR<double> Ad(fem::create_sparsity_pattern(ad));
    fem::assemble_matrix(Ad.mat_add_values(), ad, {bc});
    fem::set_diagonal<double>(Ad.mat_set_values(), *V, {bc}, 1.0);

    la::Vector<double> bk(V->num_nodes()), bd(V->num_nodes());
    bk.set(0.0);
    bd.set(0.0);
    fem::assemble_vector(bk, Lk);
    fem::assemble_vector(bd, Ld);

    // the host loop: tabulate per cell, add through the MatSet functor
    la::MatrixCSR<double> Ah(fem::create_sparsity_pattern(ak));
    {
      const std::vector<double> x = msh->x();
      const std::vector<std::int32_t> cells = msh->cells();
      const std::vector<std::int32_t> dofs = V->dofmap();
      auto mat_add = Ah.mat_add_values();
      std::vector<double> X(12), Ae(16);
      for (std::int32_t c = 0; c < nc; ++c)
      {
        for (int v = 0; v < 4; ++v)
          for (int d = 0; d < 3; ++d)
            X[3 * v + d] = x[3 * std::size_t(cells[4 * std::size_t(c) + v]) + d];
        std::fill(Ae.begin(), Ae.end(), 0.0);
        stiffness_kernel(Ae.data(), &kappa[c], &creac, X.data(), nullptr, nullptr, nullptr);
        std::span<const std::int32_t> dc(dofs.data() + 4 * std::size_t(c), 4);
        mat_add(dc, dc, Ae);
      }
      Ah.scatter_rev();
    }
    la::MatrixCSR<double> An(fem::create_sparsity_pattern(ak));
    fem::assemble_matrix(An.mat_add_values(), ak, {});

    // beta u v ds on the boundary: facet kernel and device kernel
    std::vector<double> beta(nc);
    for (std::int32_t c = 0; c < nc; ++c)
      beta[c] = 2.0 + std::cos(double(orig[c]));
    const std::vector<std::int32_t> pairs = mesh::facet_domain(*msh);
    fem::Form<double> rk(&robin_kernel, 2, V, beta, 1, {}, fem::IntegralType::exterior_facet,
                         pairs);
    fem::Form<double> rd(DFX_K_P1TET_ROBIN_A, V, beta, 1, {}, pairs, true);
    la::MatrixCSR<double> Rk(fem::create_sparsity_pattern(rk));
    fem::assemble_matrix(Rk.mat_add_values(), rk, {});
    la::MatrixCSR<double> Rd(fem::create_sparsity_pattern(rd));
    fem::assemble_matrix(Rd.mat_add_values(), rd, {});

    const std::vector<double> vk = Ak.values(), vd = Ad.values();
    const std::vector<double> vh = Ah.values(), vn = An.values();
    const std::vector<double> rvk = Rk.values(), rvd = Rd.values();
    double dmax = 0, amax = 0;
    for (std::size_t e = 0; e < vk.size(); ++e)
    {
      dmax = std::max(dmax, std::abs(vk[e] - vd[e]));
      amax = std::max(amax, std::abs(vd[e]));
    }
    double hmax = 0;
    for (std::size_t e = 0; e < vh.size(); ++e)
      hmax = std::max(hmax, std::abs(vh[e] - vn[e]));
    std::cout << "{\"nnz\": " << vk.size() << ", \"max_diff_rel\": " << dmax / amax
              << ", \"host_loop_max_diff_rel\": " << hmax / amax << "}" << std::endl;
    std::ofstream out(argv[2], std::ios::binary);
    for (const auto& v : {vk, bk.array(), vd, bd.array(), vh, vn, rvk, rvd})
    {
      const std::int64_t m = static_cast<std::int64_t>(v.size());
      out.write(reinterpret_cast<const char*>(&m), sizeof(m));
      out.write(reinterpret_cast<const char*>(v.data()), sizeof(double) * v.size());
    }
  }
  catch (const std::exception& e)
  {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
