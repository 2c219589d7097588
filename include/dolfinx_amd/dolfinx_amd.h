// C++ drop-in for the DOLFINx assemble-and-solve path on MI355X.
//
// Header-only wrappers over the C ABI of dfx.h that keep the reference's
// call shapes and names, so that a caller of
//   fem::assemble_matrix(A.mat_add_values(), *a, {bc});   (fem/assembler.h:542)
//   fem::set_diagonal(A.mat_set_values(), *V, {bc}, 1.0); (fem/assembler.h:671)
//   fem::assemble_vector(b, *L);                          (fem/assembler.h:247)
//   fem::apply_lifting(b, {a}, {{bc}}, {}, 1.0);          (fem/assembler.h:451)
//   bc->set(b, std::nullopt, 1.0);                        (fem/DirichletBC.h:495)
//   A.mult(x, y);                                         (la/MatrixCSR.h:498)
//   la::inner_product(x, y);                              (la/Vector.h:434)
// switches namespace (dolfinx -> dolfinx_amd) and keeps its code. The data
// lives in HBM: la::Vector and la::MatrixCSR are the reference's
// container-templated types instantiated with a device container
// (la/Vector.h:42-49, la/MatrixCSR.h:67-70). Failures of the C ABI are
// rethrown as std::runtime_error, as the reference throws.
#pragma once

#include "../dfx.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dolfinx_amd
{

inline void check(dfx_status s)
{
  if (s != DFX_OK)
    throw std::runtime_error(dfx_last_error());
}

template <typename T>
constexpr int dtype_code()
{
  static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>,
                "scalar type must be float or double");
  return std::is_same_v<T, double> ? DFX_F64 : DFX_F32;
}

namespace mesh
{
enum class CellType
{
  tetrahedron = DFX_TET,
  hexahedron = DFX_HEX
};

/// mesh::Mesh after create_box (one rank, RCM-reordered cells).
class Mesh
{
public:
  explicit Mesh(dfx_mesh* h, CellType ct) : _h(h, dfx_mesh_destroy), _ct(ct)
  {
    check(dfx_mesh_sizes(h, &_nc, &_nv, &_npc));
  }
  dfx_mesh* handle() const { return _h.get(); }
  CellType cell_type() const { return _ct; }
  std::int32_t num_cells() const { return _nc; }
  std::int32_t num_vertices() const { return _nv; }
  std::int32_t nodes_per_cell() const { return _npc; }
  /// Geometry::x (num_nodes x 3)
  std::vector<double> x() const
  {
    std::vector<double> v(3 * std::size_t(_nv));
    check(dfx_mesh_copy(_h.get(), v.data(), nullptr, nullptr, nullptr));
    return v;
  }
  /// Geometry dofmap = cell-vertex map for P1 geometry
  std::vector<std::int32_t> cells() const
  {
    std::vector<std::int32_t> v(std::size_t(_nc) * _npc);
    check(dfx_mesh_copy(_h.get(), nullptr, v.data(), nullptr, nullptr));
    return v;
  }
  /// Topology::original_cell_index (mesh/Topology.h:218)
  std::vector<std::int64_t> original_cell_index() const
  {
    std::vector<std::int64_t> v(_nc);
    check(dfx_mesh_copy(_h.get(), nullptr, nullptr, v.data(), nullptr));
    return v;
  }
  /// Geometry::input_global_indices
  std::vector<std::int64_t> input_global_indices() const
  {
    std::vector<std::int64_t> v(_nv);
    check(dfx_mesh_copy(_h.get(), nullptr, nullptr, nullptr, v.data()));
    return v;
  }

private:
  std::shared_ptr<dfx_mesh> _h;
  CellType _ct;
  std::int32_t _nc = 0, _nv = 0, _npc = 0;
};

/// mesh::create_box (mesh/generation.h:157-163); comm and partitioner
/// arguments of the reference are not needed on one rank.
inline std::shared_ptr<Mesh> create_box(std::array<std::array<double, 3>, 2> p,
                                        std::array<std::int64_t, 3> n, CellType celltype)
{
  dfx_mesh* h = nullptr;
  check(dfx_mesh_create_box(p[0].data(), p[1].data(), n.data(), static_cast<int>(celltype), &h));
  return std::make_shared<Mesh>(h, celltype);
}

/// fem::compute_integration_domains(IntegralType::exterior_facet, ...)
/// (fem/utils.cpp:154-273) on the exterior facets, or (axis >= 0) on those
/// with every vertex at x[axis] == value: flattened (cell, local facet)
/// pairs in facet-index order.
inline std::vector<std::int32_t> facet_domain(const Mesh& m, int axis = -1, double value = 0.0)
{
  std::int32_t n = 0;
  const int marker = axis >= 0 ? 1 : 0;
  check(dfx_facet_domain(m.handle(), marker, std::max(axis, 0), value, &n, nullptr));
  std::vector<std::int32_t> pairs(2 * std::size_t(n));
  check(dfx_facet_domain(m.handle(), marker, std::max(axis, 0), value, &n, pairs.data()));
  return pairs;
}

/// fem::compute_integration_domains(IntegralType::interior_facet, ...)
/// (fem/utils.cpp:223-253) on every interior facet: flattened (cell0,
/// local facet0, cell1, local facet1) quads in facet-index order.
inline std::vector<std::int32_t> interior_facet_domain(const Mesh& m)
{
  std::int32_t n = 0;
  check(dfx_interior_facet_domain(m.handle(), &n, nullptr));
  std::vector<std::int32_t> quads(4 * std::size_t(n));
  check(dfx_interior_facet_domain(m.handle(), &n, quads.data()));
  return quads;
}

/// mesh::create_mesh(comm, cells, element, x, ...) (mesh/utils.h:1336-1356)
/// on one rank: cells as input vertex indices in DOLFINx vertex order
/// (num_cells x 4 or 8), x as num_vertices x 3.
inline std::shared_ptr<Mesh> create_mesh(CellType celltype, std::span<const std::int64_t> cells,
                                         std::span<const double> x)
{
  const int npc = celltype == CellType::tetrahedron ? 4 : 8;
  if (cells.size() % npc != 0 || x.size() % 3 != 0)
    throw std::runtime_error("cells and x sizes do not match the cell type");
  dfx_mesh* h = nullptr;
  check(dfx_mesh_create(static_cast<int>(celltype), static_cast<std::int32_t>(cells.size() / npc),
                        cells.data(), static_cast<std::int32_t>(x.size() / 3), x.data(), &h));
  return std::make_shared<Mesh>(h, celltype);
}
} // namespace mesh

namespace la
{
template <typename T>
class Vector;

/// la::SparsityPattern after finalize (block pattern over nodes).
class SparsityPattern
{
public:
  explicit SparsityPattern(dfx_sparsity* h) : _h(h, dfx_sparsity_destroy)
  {
    std::int64_t nnz = 0;
    check(dfx_sparsity_sizes(h, &_nrows, &nnz, &_bs));
    _row_ptr.resize(std::size_t(_nrows) + 1);
    _cols.resize(nnz);
    check(dfx_sparsity_copy(h, _row_ptr.data(), _cols.data()));
  }
  std::span<const std::int64_t> graph_offsets() const { return _row_ptr; }
  std::span<const std::int32_t> graph_columns() const { return _cols; }
  int block_size() const { return _bs; }
  std::int32_t num_rows() const { return _nrows; }

private:
  std::shared_ptr<dfx_sparsity> _h;
  std::int32_t _nrows = 0, _bs = 1;
  std::vector<std::int64_t> _row_ptr;
  std::vector<std::int32_t> _cols;
};

/// la::Vector<T, DeviceContainer>: values in HBM.
template <typename T>
class Vector
{
public:
  using value_type = T;
  explicit Vector(std::int32_t n)
  {
    dfx_vec* h = nullptr;
    check(dfx_vec_create(n, dtype_code<T>(), &h));
    _h.reset(h, dfx_vec_destroy);
    _n = n;
  }
  /// Copy-convert constructor (la/Vector.h:168-187)
  template <typename U>
  explicit Vector(const Vector<U>& x) : Vector(x.size())
  {
    check(dfx_vec_copy(_h.get(), x.handle()));
  }
  dfx_vec* handle() const { return _h.get(); }
  std::int32_t size() const { return _n; }
  /// Host copy of the values
  std::vector<T> array() const
  {
    std::vector<T> v(_n);
    check(dfx_vec_download(_h.get(), v.data()));
    return v;
  }
  void set_array(std::span<const T> v)
  {
    if (static_cast<std::int32_t>(v.size()) != _n)
      throw std::runtime_error("size mismatch");
    check(dfx_vec_upload(_h.get(), v.data()));
  }
  void set(T v) { check(dfx_vec_set(_h.get(), static_cast<double>(v))); }
  /// No ghosts on one GPU: the scatters are no-ops (la/Vector.h:371-379)
  void scatter_fwd() {}
  void scatter_rev() {}

private:
  std::shared_ptr<dfx_vec> _h;
  std::int32_t _n = 0;
};

/// Host staging of MatrixCSR::add calls (la/MatrixCSR.h:157-174): blocks
/// of one shape are collected in call order and inserted on the device
/// (dfx_csr_add_blocks, each entry summed in call order) when the shape
/// changes, at scatter_rev(), or before the matrix is next used on the
/// device.
template <typename T>
struct AddStage
{
  dfx_csr* A = nullptr;
  int bs = 1;
  int m = 0, n = 0;
  std::int64_t nblocks = 0;
  std::vector<std::int32_t> rows, cols;
  std::vector<T> vals;
  void add(std::span<const std::int32_t> r, std::span<const std::int32_t> c, std::span<const T> x)
  {
    if (x.size() != r.size() * c.size() * std::size_t(bs * bs))
      throw std::runtime_error("MatrixCSR::add: data size does not match the block");
    if (nblocks > 0 && (int(r.size()) != m || int(c.size()) != n))
      flush();
    m = int(r.size());
    n = int(c.size());
    rows.insert(rows.end(), r.begin(), r.end());
    cols.insert(cols.end(), c.begin(), c.end());
    vals.insert(vals.end(), x.begin(), x.end());
    ++nblocks;
  }
  void flush()
  {
    if (nblocks == 0)
      return;
    const std::int64_t nb = nblocks;
    nblocks = 0;
    const dfx_status st
        = dfx_csr_add_blocks(A, nb, m, n, rows.data(), cols.data(), vals.data());
    rows.clear();
    cols.clear();
    vals.clear();
    check(st);
  }
};

/// mat_add_values(): the reference's MatSet functor (la/utils.h:26-29,
/// la/MatrixCSR.h:224-237) -- mat_add(rows, cols, data) adds the dense
/// (rows x bs) x (cols x bs) block, staged on the host as above -- and the
/// handle through which fem::assemble_matrix routes to the device kernels.
template <typename T>
struct DeviceMatAdd
{
  dfx_csr* A;
  std::shared_ptr<AddStage<T>> stage;
  int operator()(std::span<const std::int32_t> rows, std::span<const std::int32_t> cols,
                 std::span<const T> data) const
  {
    stage->add(rows, cols, data);
    return 0;
  }
  void flush() const
  {
    if (stage)
      stage->flush();
  }
};
template <typename T>
struct DeviceMatSet
{
  dfx_csr* A;
  std::shared_ptr<AddStage<T>> stage;
  void flush() const
  {
    if (stage)
      stage->flush();
  }
};

/// la::MatrixCSR<T, DeviceContainer> in compact block mode.
template <typename T>
class MatrixCSR
{
public:
  using value_type = T;
  explicit MatrixCSR(const SparsityPattern& p)
  {
    dfx_csr* h = nullptr;
    check(dfx_csr_create(p.graph_offsets().data(), p.graph_columns().data(), p.num_rows(),
                         p.num_rows(), p.block_size(), dtype_code<T>(), &h));
    _h.reset(h, dfx_csr_destroy);
    _bs = p.block_size();
    _nrows = p.num_rows();
    _nnz = static_cast<std::int64_t>(p.graph_columns().size());
    _stage = std::make_shared<AddStage<T>>();
    _stage->A = h;
    _stage->bs = _bs;
  }
  /// the device matrix, with any staged MatrixCSR::add calls inserted
  dfx_csr* handle() const
  {
    _stage->flush();
    return _h.get();
  }
  DeviceMatAdd<T> mat_add_values() { return {_h.get(), _stage}; }
  DeviceMatSet<T> mat_set_values() { return {_h.get(), _stage}; }
  /// MatrixCSR::scatter_rev (la/MatrixCSR.h:589-626): one rank, so only the
  /// staged additions are inserted
  void scatter_rev() { _stage->flush(); }
  /// MatrixCSR::set(x): only zero is supported on the device (staged
  /// additions made before it are dropped, as set(0) overwrites them)
  void set(T x)
  {
    if (x != T(0))
      throw std::runtime_error("MatrixCSR::set: only zero is supported");
    _stage->rows.clear();
    _stage->cols.clear();
    _stage->vals.clear();
    _stage->nblocks = 0;
    check(dfx_csr_zero(_h.get()));
  }
  /// y += A x (la/MatrixCSR.h:877-946)
  void mult(Vector<T>& x, Vector<T>& y) const { check(dfx_spmv(handle(), x.handle(), y.handle())); }
  /// Host copy of the values (row-major blocks)
  std::vector<T> values() const
  {
    std::vector<T> v(_nnz * _bs * _bs);
    check(dfx_csr_get_values(handle(), v.data()));
    return v;
  }
  std::array<int, 2> block_size() const { return {_bs, _bs}; }
  std::int32_t num_rows() const { return _nrows; }
  /// Frobenius norm squared (la/MatrixCSR.h:470-490)
  double squared_norm() const
  {
    double s = 0;
    for (T v : values())
      s += double(v) * double(v);
    return s;
  }
  /// Pick the SpMV variant by timing on this matrix
  std::pair<int, float> autotune()
  {
    std::int32_t l = 0;
    float ms = 0;
    check(dfx_csr_autotune(handle(), &l, &ms));
    return {l, ms};
  }

private:
  std::shared_ptr<dfx_csr> _h;
  std::shared_ptr<AddStage<T>> _stage;
  int _bs = 1;
  std::int32_t _nrows = 0;
  std::int64_t _nnz = 0;
};

/// la::inner_product (la/Vector.h:434-460); FP64 accumulation
template <typename T>
double inner_product(const Vector<T>& a, const Vector<T>& b)
{
  double r = 0;
  check(dfx_inner_product(a.handle(), b.handle(), &r));
  return r;
}
/// la::squared_norm (la/Vector.h:462-471)
template <typename T>
double squared_norm(const Vector<T>& a)
{
  double r = 0;
  check(dfx_squared_norm(a.handle(), &r));
  return r;
}
/// la::norm(x, Norm::l2) (la/Vector.h:478-514)
template <typename T>
double norm(const Vector<T>& a)
{
  return std::sqrt(squared_norm(a));
}
} // namespace la

namespace fem
{
/// Lagrange function space with its (block) dofmap.
class FunctionSpace
{
public:
  FunctionSpace(std::shared_ptr<mesh::Mesh> m, int degree, int bs = 1) : _mesh(std::move(m))
  {
    dfx_space* h = nullptr;
    check(dfx_space_create(_mesh->handle(), degree, bs, &h));
    _h.reset(h, dfx_space_destroy);
    check(dfx_space_sizes(h, &_nn, &_k, &_bs));
    _degree = degree;
  }
  dfx_space* handle() const { return _h.get(); }
  std::shared_ptr<mesh::Mesh> mesh() const { return _mesh; }
  int degree() const { return _degree; }
  int bs() const { return _bs; }
  std::int32_t num_nodes() const { return _nn; }
  std::int32_t dofs_per_cell() const { return _k; }
  std::vector<std::int32_t> dofmap() const
  {
    std::vector<std::int32_t> d(std::size_t(_mesh->num_cells()) * _k);
    check(dfx_space_dofmap(_h.get(), d.data()));
    return d;
  }

private:
  std::shared_ptr<mesh::Mesh> _mesh;
  std::shared_ptr<dfx_space> _h;
  std::int32_t _nn = 0, _k = 0, _bs = 1;
  int _degree = 1;
};

/// fem::create_functionspace for ("Lagrange", degree[, shape])
inline std::shared_ptr<FunctionSpace> create_functionspace(std::shared_ptr<mesh::Mesh> m,
                                                           int degree, int bs = 1)
{
  return std::make_shared<FunctionSpace>(std::move(m), degree, bs);
}

/// locate_dofs_topological on the exterior facets (all, or those with every
/// vertex at x[axis] == value), unrolled: returns dof indices.
inline std::vector<std::int32_t> locate_boundary_dofs(const FunctionSpace& V,
                                                      std::optional<int> axis = std::nullopt,
                                                      double value = 0.0)
{
  std::int32_t n = 0;
  const int marker = axis ? 1 : 0;
  check(dfx_locate_boundary_dofs(V.handle(), marker, axis.value_or(0), value, &n, nullptr));
  std::vector<std::int32_t> d(n);
  check(dfx_locate_boundary_dofs(V.handle(), marker, axis.value_or(0), value, &n, d.data()));
  return d;
}

/// fem::DirichletBC on unrolled dofs with values g (empty: g = 0).
template <typename T>
class DirichletBC
{
public:
  DirichletBC(std::shared_ptr<const FunctionSpace> V, std::vector<std::int32_t> dofs,
              std::vector<double> g = {})
      : _V(std::move(V)), _dofs(std::move(dofs))
  {
    const std::int32_t n = _V->num_nodes() * _V->bs();
    if (!g.empty() && static_cast<std::int32_t>(g.size()) != n)
      throw std::runtime_error("g must have one value per dof");
    dfx_bc* h = nullptr;
    check(dfx_bc_create(_dofs.data(), static_cast<std::int32_t>(_dofs.size()), n,
                        g.empty() ? nullptr : g.data(), &h));
    _h.reset(h, dfx_bc_destroy);
  }
  dfx_bc* handle() const { return _h.get(); }
  std::shared_ptr<const FunctionSpace> function_space() const { return _V; }
  /// DirichletBC::set(x, x0, alpha): x[bc] = alpha (g - x0)
  void set(la::Vector<T>& x, std::optional<std::reference_wrapper<const la::Vector<T>>> x0,
           T alpha = 1) const
  {
    check(dfx_bc_set(x.handle(), _h.get(), x0 ? x0->get().handle() : nullptr, alpha));
  }
  std::span<const std::int32_t> dof_indices() const { return _dofs; }

private:
  std::shared_ptr<const FunctionSpace> _V;
  std::vector<std::int32_t> _dofs;
  std::shared_ptr<dfx_bc> _h;
};

/// fem::IntegralType (fem/Form.h:33-39): where an integral is taken
enum class IntegralType : std::int32_t
{
  cell = 0,
  exterior_facet = 1,
  interior_facet = 2
};

/// fem::Form whose single cell integral is one of the hand-written device
/// kernels (DFX_K_*), with packed coefficients (cells x cstride, in the
/// mesh's cell order) and constants.
template <typename T>
class Form
{
public:
  /// With facets the form's integral is a facet integral over them:
  /// flattened (cell, local facet) pairs for the exterior facet kernels
  /// (DFX_K_*_NEUMANN, DFX_K_*_ROBIN_A), (cell0, lf0, cell1, lf1) quads
  /// for the interior facet kernel (DFX_K_P1TET_GRADJUMP_A).
  Form(int kernel, std::shared_ptr<const FunctionSpace> V, std::vector<double> coeffs,
       int cstride, std::vector<double> constants,
       std::span<const std::int32_t> facets = {}, bool facet_integral = false)
      : _V(std::move(V))
  {
    auto m = _V->mesh();
    const std::vector<double> x = m->x();
    const std::vector<std::int32_t> cells = m->cells();
    const std::vector<std::int32_t> dm = _V->dofmap();
    dfx_form_desc d{};
    d.kernel = kernel;
    d.dtype = dtype_code<T>();
    d.num_cells = m->num_cells();
    d.num_nodes = m->num_vertices();
    d.nodes_per_cell = m->nodes_per_cell();
    d.x = x.data();
    d.x_dofmap = cells.data();
    d.dofs_per_cell = _V->dofs_per_cell();
    d.bs = _V->bs();
    d.dofmap = dm.data();
    d.cstride = cstride;
    d.coeffs = coeffs.empty() ? nullptr : coeffs.data();
    d.num_constants = static_cast<std::int32_t>(constants.size());
    d.constants = constants.empty() ? nullptr : constants.data();
    dfx_form* h = nullptr;
    if (kernel == DFX_K_P1TET_GRADJUMP_A)
    {
      _interior.assign(facets.begin(), facets.end());
      check(dfx_form_create_interior_facets(&d, static_cast<std::int32_t>(facets.size() / 4),
                                            facets.data(), &h));
    }
    else if (facet_integral || !facets.empty())
      check(dfx_form_create_facets(&d, static_cast<std::int32_t>(facets.size() / 2),
                                   facets.data(), &h));
    else
      check(dfx_form_create(&d, &h));
    _h.reset(h, dfx_form_destroy);
    switch (kernel)
    {
    case DFX_K_P1TET_L:
    case DFX_K_Q1HEX_L:
    case DFX_K_P2TET_ELAST_L:
    case DFX_K_P3TET_L:
    case DFX_K_ELEMENT_L:
    case DFX_K_P1TET_NEUMANN:
    case DFX_K_Q1HEX_NEUMANN:
      _rank = 1;
      break;
    default:
      _rank = 2;
    }
  }
  /// The element kernel signature of the reference
  /// (fem::impl::kernelptr_t<T>, fem/kernel.h:18-20).
  using kernel_ptr = void (*)(T*, const T*, const T*, const T*, const int*, const std::uint8_t*,
                              void*);
  /// A form whose cell integral is an element kernel with the reference's
  /// ABI -- a compiled tabulate_tensor or a hand-written kernel, as in
  /// integral_data{kernel, cells} (fem/Form.h:52-87,
  /// cpp/demo/custom_kernel/main.cpp:77-81): rank 2 (bilinear, (bs k)^2
  /// row-major tensors) or 1. The kernel runs per cell on the host; its
  /// tensors are inserted on the device. With type exterior_facet or
  /// interior_facet the integral is taken over `entities` instead
  /// (integral_data{kernel, entities}: flattened (cell, local facet) pairs,
  /// or (cell0, lf0, cell1, lf1) quads for a bilinear interior facet
  /// integral), the kernel called per facet with the local facet
  /// index (indices) as assemble_exterior_facets / assemble_interior_facets
  /// call it; coefficients stay packed per cell.
  Form(kernel_ptr kernel, int rank, std::shared_ptr<const FunctionSpace> V,
       std::vector<double> coeffs, int cstride, std::vector<double> constants,
       IntegralType type = IntegralType::cell, std::span<const std::int32_t> entities = {})
      : _V(std::move(V)), _rank(rank)
  {
    auto m = _V->mesh();
    const std::vector<double> x = m->x();
    const std::vector<std::int32_t> cells = m->cells();
    const std::vector<std::int32_t> dm = _V->dofmap();
    dfx_form_desc d{};
    d.kernel = rank == 2 ? DFX_K_ELEMENT_A : DFX_K_ELEMENT_L;
    d.dtype = dtype_code<T>();
    d.num_cells = m->num_cells();
    d.num_nodes = m->num_vertices();
    d.nodes_per_cell = m->nodes_per_cell();
    d.x = x.data();
    d.x_dofmap = cells.data();
    d.dofs_per_cell = _V->dofs_per_cell();
    d.bs = _V->bs();
    d.dofmap = dm.data();
    d.cstride = cstride;
    d.coeffs = coeffs.empty() ? nullptr : coeffs.data();
    d.num_constants = static_cast<std::int32_t>(constants.size());
    d.constants = constants.empty() ? nullptr : constants.data();
    dfx_form* h = nullptr;
    if (type == IntegralType::cell)
      check(dfx_form_create_kernel(&d, reinterpret_cast<void*>(kernel), nullptr, &h));
    else
    {
      const std::size_t per = type == IntegralType::interior_facet ? 4 : 2;
      if (per == 4)
        _interior.assign(entities.begin(), entities.end());
      check(dfx_form_create_kernel_facets(&d, reinterpret_cast<void*>(kernel), nullptr,
                                          static_cast<std::int32_t>(type),
                                          static_cast<std::int32_t>(entities.size() / per),
                                          entities.data(), &h));
    }
    _h.reset(h, dfx_form_destroy);
  }
  /// The (cell0, lf0, cell1, lf1) quads of an interior facet integral.
  const std::vector<std::int32_t>& interior_facets() const { return _interior; }
  dfx_form* handle() const { return _h.get(); }
  int rank() const { return _rank; }
  std::vector<std::shared_ptr<const FunctionSpace>> function_spaces() const
  {
    return std::vector<std::shared_ptr<const FunctionSpace>>(_rank, _V);
  }
  /// fem::pack_coefficients (fem/pack.h:265-388) on the device: the
  /// coefficients from values per input cell (DFX_PACK_CELL, DG0) or per
  /// input vertex (DFX_PACK_VERTEX, P1) of the form's mesh.
  void pack_coefficients(std::span<const double> values, int layout)
  {
    check(dfx_form_pack_coefficients(_h.get(), _V->mesh()->handle(), layout, values.data(),
                                     static_cast<std::int64_t>(values.size())));
  }
  /// Pick the assembly kernel variant for the pattern of A by timing
  std::pair<int, float> autotune(la::MatrixCSR<T>& A, const DirichletBC<T>* bc = nullptr)
  {
    std::int32_t v = 0;
    float ms = 0;
    check(dfx_form_autotune(_h.get(), A.handle(), bc ? bc->handle() : nullptr, &v, &ms));
    return {v, ms};
  }

private:
  std::shared_ptr<const FunctionSpace> _V;
  std::shared_ptr<dfx_form> _h;
  std::vector<std::int32_t> _interior;
  int _rank = 2;
};

/// fem::create_sparsity_pattern (fem/utils.h:197-314): the cell couplings
/// and those of the form's interior facets
template <typename T>
la::SparsityPattern create_sparsity_pattern(const Form<T>& a)
{
  auto V = a.function_spaces().at(0);
  const std::vector<std::int32_t>& q = a.interior_facets();
  dfx_sparsity* h = nullptr;
  check(dfx_sparsity_create_facets(V->handle(), V->handle(),
                                   static_cast<std::int32_t>(q.size() / 4), q.data(), &h));
  return la::SparsityPattern(h);
}

template <typename T>
using BCs = std::vector<std::reference_wrapper<const DirichletBC<T>>>;

/// fem::assemble_matrix(mat_add, a, bcs) (fem/assembler.h:542-602)
template <typename T>
void assemble_matrix(la::DeviceMatAdd<T> mat_add, const Form<T>& a, const BCs<T>& bcs = {})
{
  if (bcs.size() > 1)
    throw std::runtime_error("pass the union of Dirichlet dofs as one DirichletBC");
  mat_add.flush();
  check(dfx_assemble_matrix(mat_add.A, a.handle(), bcs.empty() ? nullptr : bcs[0].get().handle()));
}

/// fem::set_diagonal(set_fn, V, bcs, diagonal) (fem/assembler.h:671-686)
template <typename T>
void set_diagonal(la::DeviceMatSet<T> mat_set, const FunctionSpace&, const BCs<T>& bcs,
                  T diagonal = 1)
{
  mat_set.flush();
  for (auto& bc : bcs)
    check(dfx_set_diagonal(mat_set.A, bc.get().handle(), diagonal));
}

/// fem::assemble_vector(b, L) (fem/assembler.h:247-257): b += L
template <typename T>
void assemble_vector(la::Vector<T>& b, const Form<T>& L)
{
  check(dfx_assemble_vector(b.handle(), L.handle()));
}

/// fem::apply_lifting(b, a, bcs1, x0, alpha) (fem/assembler.h:451-493)
template <typename T>
void apply_lifting(la::Vector<T>& b, const std::vector<std::reference_wrapper<const Form<T>>>& a,
                   const std::vector<BCs<T>>& bcs1,
                   const std::vector<std::reference_wrapper<const la::Vector<T>>>& x0, T alpha)
{
  if (!x0.empty() && x0.size() != a.size())
    throw std::runtime_error("Mismatch in size between x0 and bilinear form in assembler.");
  if (a.size() != bcs1.size())
    throw std::runtime_error("Mismatch in size between a and bcs in assembler.");
  for (std::size_t i = 0; i < a.size(); ++i)
    for (auto& bc : bcs1[i])
      check(dfx_apply_lifting(b.handle(), a[i].get().handle(), bc.get().handle(),
                              x0.empty() ? nullptr : x0[i].get().handle(), alpha));
}
/// fem::assemble_matrix / assemble_vector with a custom kernel
/// (cpp/demo/custom_kernel/main.cpp:73-118): a Form of kernel
/// DFX_K_ELEMENT_A / DFX_K_ELEMENT_L, the kernel's outputs E (cells x k x k)
/// and e (cells x k) in device vectors of the form's scalar type.
template <typename T>
void assemble_matrix_elements(la::DeviceMatAdd<T> mat_add, const Form<T>& a,
                              const la::Vector<T>& E, const BCs<T>& bcs = {})
{
  if (bcs.size() > 1)
    throw std::runtime_error("pass the union of Dirichlet dofs as one DirichletBC");
  mat_add.flush();
  check(dfx_assemble_matrix_elements(mat_add.A, a.handle(),
                                     bcs.empty() ? nullptr : bcs[0].get().handle(), E.handle()));
}

template <typename T>
void assemble_vector_elements(la::Vector<T>& b, const Form<T>& L, const la::Vector<T>& e)
{
  check(dfx_assemble_vector_elements(b.handle(), L.handle(), e.handle()));
}

/// apply_lifting for a Form with caller-supplied element matrices E
template <typename T>
void apply_lifting_elements(la::Vector<T>& b, const Form<T>& a, const la::Vector<T>& E,
                            const BCs<T>& bcs, const la::Vector<T>* x0 = nullptr, T alpha = 1)
{
  for (auto& bc : bcs)
    check(dfx_apply_lifting_elements(b.handle(), a.handle(), bc.get().handle(), E.handle(),
                                     x0 ? x0->handle() : nullptr, alpha));
}

/// The demos' system sequence (cpp/demo/poisson/main.cpp:207-226) in one
/// call: assemble_matrix + set_diagonal, b = 0 + assemble_vector,
/// apply_lifting, DirichletBC::set; the vector assembly overlaps the
/// matrix assembly on a second stream.
template <typename T>
void assemble_system(la::MatrixCSR<T>& A, const Form<T>& a, la::Vector<T>& b, const Form<T>& L,
                     const BCs<T>& bcs = {}, T diagonal = 1)
{
  if (bcs.size() > 1)
    throw std::runtime_error("pass the union of Dirichlet dofs as one DirichletBC");
  check(dfx_assemble_system(A.handle(), a.handle(), bcs.empty() ? nullptr : bcs[0].get().handle(),
                            diagonal, b.handle(), L.handle(), 1));
}
} // namespace fem

namespace linalg
{
/// CG of cpp/demo/poisson_matrix_free/main.cpp:84-132 on an assembled
/// device matrix, optionally Jacobi-preconditioned; returns iterations.
template <typename T>
int cg(la::Vector<T>& x, const la::Vector<T>& b, const la::MatrixCSR<T>& A, int kmax = 50,
       double rtol = 1e-8, bool jacobi = false)
{
  dfx_cg_opts o{kmax, rtol, jacobi ? 1 : 0, 32, 0, 0};
  dfx_cg_result r{};
  check(dfx_cg_solve(A.handle(), b.handle(), x.handle(), &o, &r));
  return r.iterations;
}
} // namespace linalg

/// A sequence of library calls recorded into one HIP graph (no counterpart
/// in the reference): every launch the library makes from this thread
/// inside capture(f) is recorded instead of run, and launch() replays the
/// sequence as one graph launch. Run the sequence once before recording it
/// (calls that need the host part-way -- allocation, plan uploads, waits --
/// throw inside a capture).
class Graph
{
public:
  template <typename F>
  static Graph capture(F&& f)
  {
    check(dfx_graph_begin());
    try
    {
      f();
    }
    catch (...)
    {
      dfx_graph* g = nullptr;
      (void)dfx_graph_end(&g); // closes (and discards) the capture
      dfx_graph_destroy(g);
      throw;
    }
    dfx_graph* g = nullptr;
    check(dfx_graph_end(&g));
    return Graph(g);
  }
  void launch() const { check(dfx_graph_launch(_g.get())); }

private:
  explicit Graph(dfx_graph* g) : _g(g, dfx_graph_destroy) {}
  std::shared_ptr<dfx_graph> _g;
};

} // namespace dolfinx_amd
