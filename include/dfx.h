/*
 * dfx.h - C ABI of the MI355X-native DOLFINx assemble-and-solve path.
 *
 * This is the only boundary between host code and the HIP kernels. It takes
 * plain pointers and sizes, returns integer status codes (0 = success), and
 * keeps the message of the last failure in a per-thread buffer
 * (dfx_last_error). The C++ drop-in headers in include/dolfinx_amd/ wrap it
 * with the reference's fem:: and la:: names and rethrow failures as
 * std::runtime_error, as the reference does.
 *
 * Every entry point names the reference interface it replaces
 * (paths relative to the DOLFINx tree cpp/dolfinx/ unless stated).
 *
 * Ownership: host arrays passed in are only read during the call. Handles
 * (dfx_mesh, dfx_space, dfx_sparsity, dfx_csr, dfx_vec, dfx_form, dfx_bc)
 * own their memory and are released by the matching *_destroy.
 * Threading: one host thread per library instance; device work is ordered
 * on the library's HIP stream (dfx_stream).
 */
#ifndef DFX_H
#define DFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dfx_status;
enum
{
  DFX_OK = 0,
  DFX_ERR_ARG = 1,     /* invalid argument (size mismatch, null handle) */
  DFX_ERR_PATTERN = 2, /* entry not in sparsity (la/matrix_csr_impl.h:93) */
  DFX_ERR_HIP = 3,     /* HIP runtime failure */
  DFX_ERR_NOCONV = 4,  /* iterative solver did not converge */
  DFX_ERR_NODEV = 5,   /* no GPU device available */
  DFX_ERR_CAPTURE = 6  /* the call cannot be recorded into a graph capture */
};

enum { DFX_F32 = 0, DFX_F64 = 1 };
enum { DFX_TET = 0, DFX_HEX = 1 };

/* Hand-written element kernels. Each replaces one FFCx tabulate_tensor
 * (fem/kernel.h:18-60) for a config form, batched over all cells. */
enum
{
  DFX_K_P1TET_A = 0,       /* kappa grad u.grad v + c u v; w=[kappa] c=[c] */
  DFX_K_P1TET_L = 1,       /* f v; w = [f at the 4 vertices] (P1 coeff.) */
  DFX_K_Q1HEX_A = 2,       /* kappa grad u.grad v + c u v; w=[kappa] c=[c] */
  DFX_K_Q1HEX_L = 3,       /* f v; w = [f at the 8 vertices] (Q1 coeff.) */
  DFX_K_P2TET_ELAST_A = 4, /* sigma(u):eps(v), w = [lambda, mu] (DG0) */
  DFX_K_P2TET_ELAST_L = 5, /* f.v, c = [fx, fy, fz] */
  DFX_K_P3TET_A = 6,       /* kappa grad u.grad v + c u v; w=[kappa] c=[c] */
  DFX_K_P3TET_L = 7,       /* f v; c = [f] */
  DFX_K_ELEMENT_A = 8,     /* element matrices supplied by the caller */
  DFX_K_ELEMENT_L = 9,     /* element vectors supplied by the caller */
  DFX_K_P1TET_NEUMANN = 10,/* g v ds on exterior facets; w = [g at the 4 vertices] */
  DFX_K_Q1HEX_NEUMANN = 11,/* g v ds on exterior facets; w = [g at the 8 vertices] */
  DFX_K_P1TET_ROBIN_A = 12,/* beta u v ds on exterior facets; w = [beta] (DG0) */
  DFX_K_Q1HEX_ROBIN_A = 13,/* beta u v ds on exterior facets; w = [beta] (DG0) */
  DFX_K_P1TET_GRADJUMP_A = 14 /* gamma [grad u.n][grad v.n] dS on interior
                                 facets; c = [gamma] */
};

const char* dfx_last_error(void);
/* Library version string. */
const char* dfx_version(void);
/* Initialise the device (select device 0, create the stream). Idempotent. */
dfx_status dfx_init(void);
/* Block until all queued device work has finished. */
dfx_status dfx_synchronize(void);
/* The library's HIP stream as an opaque pointer (hipStream_t). */
void* dfx_stream(void);

/* ------------------------------------------------------------------ */
/* Host setup: mesh, dofmaps, Dirichlet dofs, sparsity (bit-exact).   */
/* ------------------------------------------------------------------ */
typedef struct dfx_mesh dfx_mesh;
typedef struct dfx_space dfx_space;
typedef struct dfx_sparsity dfx_sparsity;

/* mesh::create_box (mesh/generation.h:104-134) on one rank, with the
 * default graph::reorder_rcm cell reordering (graph/ordering.cpp:148). */
dfx_status dfx_mesh_create_box(const double p0[3], const double p1[3],
                               const int64_t n[3], int cell, dfx_mesh** out);
/* mesh::create_mesh (mesh/utils.h:1336-1356 -> :1074-1259) on one rank:
 * cells (num_cells x 4 tetrahedra / 8 hexahedra, input vertex indices in
 * the DOLFINx vertex order) and geometry x (num_vertices x 3, input
 * order); every vertex must belong to a cell. Cells are reordered by RCM
 * on the dual graph and vertices numbered by first appearance, as for
 * create_box. */
dfx_status dfx_mesh_create(int cell, int32_t num_cells, const int64_t* cells,
                           int32_t num_vertices, const double* x, dfx_mesh** out);
dfx_status dfx_mesh_destroy(dfx_mesh* m);
dfx_status dfx_mesh_sizes(const dfx_mesh* m, int32_t* num_cells,
                          int32_t* num_vertices, int32_t* vertices_per_cell);
/* Copies of Geometry::x (nodes x 3), the cell->vertex map, the original
 * (input) cell index (Topology::original_cell_index, mesh/Topology.h:218)
 * and the input vertex index of each vertex. Any pointer may be NULL. */
dfx_status dfx_mesh_copy(const dfx_mesh* m, double* x, int32_t* cells,
                         int64_t* original_cell_index,
                         int64_t* input_vertex_index);

/* fem::compute_integration_domains(IntegralType::exterior_facet, topology,
 * facets) (fem/utils.cpp:154-273) on mesh::exterior_facet_indices
 * (mesh/utils.h:180-194): the (cell, local facet) pair of every exterior
 * facet in facet-index order, or (marker 1) of those whose vertices all
 * lie at x[axis] == value. With pairs NULL only *count is set; otherwise
 * *count is the capacity in pairs on entry. */
dfx_status dfx_facet_domain(const dfx_mesh* m, int marker, int axis, double value,
                            int32_t* count, int32_t* pairs);

/* fem::compute_integration_domains(IntegralType::interior_facet, ...)
 * (fem/utils.cpp:223-253) on every interior facet, in facet-index order:
 * (cell0, local facet0, cell1, local facet1) with cell0 < cell1 (the
 * facet-to-cell links). With quads NULL only *count is set; otherwise
 * *count is the capacity in facets on entry. */
dfx_status dfx_interior_facet_domain(const dfx_mesh* m, int32_t* count, int32_t* quads);

/* fem::create_functionspace -> fem::create_dofmap (fem/utils.cpp:32-76)
 * for Lagrange degree 1-3 (tetrahedra) or 1 (hexahedra), block size bs. */
dfx_status dfx_space_create(dfx_mesh* m, int degree, int bs, dfx_space** out);
dfx_status dfx_space_destroy(dfx_space* V);
dfx_status dfx_space_sizes(const dfx_space* V, int32_t* num_nodes,
                           int32_t* dofs_per_cell, int32_t* bs);
dfx_status dfx_space_dofmap(const dfx_space* V, int32_t* dofmap);
/* fem::locate_dofs_topological on exterior facets (fem/DirichletBC.cpp:
 * 201-310, mesh/utils.cpp:61-89), unrolled to bs*node+comp as a DirichletBC
 * on the whole space does (fem/DirichletBC.h:273-281). marker 0: all
 * exterior facets; marker 1: exterior facets with every vertex at
 * x[axis] == value (mesh::locate_entities_boundary). Call with dofs == NULL
 * to query the count. */
dfx_status dfx_locate_boundary_dofs(const dfx_space* V, int marker, int axis,
                                    double value, int32_t* count,
                                    int32_t* dofs);
/* fem::create_sparsity_pattern + SparsityPattern::finalize (fem/utils.h:
 * 197-314, la/SparsityPattern.cpp:264-491). Block pattern over nodes. */
dfx_status dfx_sparsity_create(const dfx_space* V0, const dfx_space* V1,
                               dfx_sparsity** out);
/* The same pattern with the couplings of interior facet integrals over the
 * nfacets (cell0, lf0, cell1, lf1) quads added: every dof of both cells of
 * a facet couples to every other (sparsitybuild::interior_facets,
 * fem/sparsitybuild.cpp:16-, called from fem/utils.h:288). */
dfx_status dfx_sparsity_create_facets(const dfx_space* V0, const dfx_space* V1,
                                      int32_t nfacets, const int32_t* quads,
                                      dfx_sparsity** out);
dfx_status dfx_sparsity_destroy(dfx_sparsity* sp);
dfx_status dfx_sparsity_sizes(const dfx_sparsity* sp, int32_t* nrows,
                              int64_t* nnz, int32_t* bs);
dfx_status dfx_sparsity_copy(const dfx_sparsity* sp, int64_t* row_ptr,
                             int32_t* cols);

/* ------------------------------------------------------------------ */
/* Device linear algebra: la::MatrixCSR and la::Vector on HBM.         */
/* ------------------------------------------------------------------ */
typedef struct dfx_csr dfx_csr;
typedef struct dfx_vec dfx_vec;

/* la::MatrixCSR(sp, BlockMode::compact) (la/MatrixCSR.h:628-850) with a
 * device container (la/MatrixCSR.h:67-70,210-237): values zeroed. */
dfx_status dfx_csr_create(const int64_t* row_ptr, const int32_t* cols,
                          int32_t nrows, int32_t ncols, int bs, int dtype,
                          dfx_csr** out);
/* A matrix with its own values on the pattern of P (shared, not copied;
 * P must outlive it): many systems on one mesh keep one pattern in HBM. */
dfx_status dfx_csr_create_shared(const dfx_csr* P, int dtype, dfx_csr** out);
dfx_status dfx_csr_destroy(dfx_csr* A);
dfx_status dfx_csr_sizes(const dfx_csr* A, int32_t* nrows, int64_t* nnz,
                         int32_t* bs, int32_t* dtype);
/* MatrixCSR::set(0) */
dfx_status dfx_csr_zero(dfx_csr* A);
/* Copy values (nnz*bs*bs of the matrix dtype) to / from the host. */
dfx_status dfx_csr_get_values(const dfx_csr* A, void* values);
/* MatrixCSR::add / mat_add_values<bs, bs> (la/MatrixCSR.h:157-174,
 * la/matrix_csr_impl.h:67-109) for nblocks dense blocks staged on the host:
 * block q adds its (m bs) x (n bs) row-major values (host, dtype of A) at
 * block rows rows[q m .. q m + m) and block columns cols[q n .. q n + n).
 * Every entry receives its additions in staging order (block, then row,
 * then column), as the reference's call-by-call insertion; an entry
 * outside the pattern returns DFX_ERR_PATTERN ("Entry not in sparsity"). */
dfx_status dfx_csr_add_blocks(dfx_csr* A, int64_t nblocks, int32_t m, int32_t n,
                              const int32_t* rows, const int32_t* cols, const void* values);
dfx_status dfx_csr_set_values(dfx_csr* A, const void* values);
/* Copy the device pattern back (row_ptr as int64 like the reference). */
dfx_status dfx_csr_get_pattern(const dfx_csr* A, int64_t* row_ptr,
                               int32_t* cols);

/* SpMV autotuner (no counterpart in the reference): times the SpMV
 * variants on this matrix and keeps the fastest for every later SpMV/CG on
 * it: 1 the row-block stream kernel, 2..64 lanes per row (scalar CSR), 2
 * one thread per (block row, component) (block CSR). Returns the choice
 * and its time per SpMV. */
dfx_status dfx_csr_autotune(dfx_csr* A, int32_t* lanes, float* ms);
/* Fix the SpMV variant as above (0 restores the built-in heuristic). */
dfx_status dfx_csr_set_lanes(dfx_csr* A, int32_t lanes);

/* la::Vector (la/Vector.h:47-422) with a device container. */
dfx_status dfx_vec_create(int32_t n, int dtype, dfx_vec** out);
dfx_status dfx_vec_destroy(dfx_vec* v);
dfx_status dfx_vec_size(const dfx_vec* v, int32_t* n, int32_t* dtype);
dfx_status dfx_vec_upload(dfx_vec* v, const void* host);
dfx_status dfx_vec_download(const dfx_vec* v, void* host);
dfx_status dfx_vec_set(dfx_vec* v, double value);
/* Copy-convert (la/Vector.h:168-187): dst = (dst dtype) src. */
dfx_status dfx_vec_copy(dfx_vec* dst, const dfx_vec* src);

/* MatrixCSR::mult (la/MatrixCSR.h:877-946 -> impl::spmv,
 * la/matrix_csr_impl.h:259-286): y += A x. x, y share A's dtype. */
dfx_status dfx_spmv(const dfx_csr* A, const dfx_vec* x, dfx_vec* y);
/* la::inner_product (la/Vector.h:434-460); accumulated in FP64. */
dfx_status dfx_inner_product(const dfx_vec* a, const dfx_vec* b,
                             double* out);
/* la::squared_norm (la/Vector.h:462-471). */
dfx_status dfx_squared_norm(const dfx_vec* a, double* out);
/* r = alpha x + y (cpp/demo/poisson_matrix_free/main.cpp:65-69). */
dfx_status dfx_axpy(dfx_vec* r, double alpha, const dfx_vec* x,
                    const dfx_vec* y);

/* ------------------------------------------------------------------ */
/* Forms, Dirichlet BCs and assembly.                                  */
/* ------------------------------------------------------------------ */
typedef struct dfx_form dfx_form;
typedef struct dfx_bc dfx_bc;

/* The data fem::impl::assemble_cells_matrix receives
 * (fem/assemble_matrix_impl.h:92-118): geometry dofmap and x, the
 * (test = trial) dofmap with its block size, packed coefficients
 * (fem/pack.h:265-388, cells x cstride) and constants
 * (fem/pack.h:578-611), plus the element kernel. */
typedef struct
{
  int32_t kernel;         /* DFX_K_* */
  int32_t dtype;          /* DFX_F32 / DFX_F64: scalar type of the form */
  int32_t num_cells;
  int32_t num_nodes;      /* geometry nodes */
  int32_t nodes_per_cell; /* geometry nodes per cell */
  const double* x;        /* num_nodes x 3 */
  const int32_t* x_dofmap;/* num_cells x nodes_per_cell */
  int32_t dofs_per_cell;  /* nodes per cell of the (block) dofmap */
  int32_t bs;             /* block size of the dofmap */
  const int32_t* dofmap;  /* num_cells x dofs_per_cell */
  int32_t cstride;        /* coefficients per cell */
  const double* coeffs;   /* num_cells x cstride (may be NULL if 0) */
  int32_t num_constants;  /* <= 8 */
  const double* constants;
} dfx_form_desc;

/* fem::Form (fem/Form.h:116-) holding device copies of the form data. */
dfx_status dfx_form_create(const dfx_form_desc* desc, dfx_form** out);
/* A form with an exterior facet integral (kernels DFX_K_*_NEUMANN and
 * DFX_K_*_ROBIN_A): its integration entities are the npairs (cell, local
 * facet) pairs (fem/Form.h integral_data entities, e.g. from
 * dfx_facet_domain). assemble_vector adds the facets in this order
 * (fem/assemble_vector_impl.h:118-174); assemble_matrix adds the facet
 * matrices, BC rows and columns zeroed, in this order on top of A's values
 * (assemble_entities, fem/assemble_matrix_impl.h:264-383, run after the
 * cell integrals at :844-880: assemble the cell form first); apply_lifting
 * subtracts them like lift_bc (fem/assemble_vector_impl.h:361-414). */
dfx_status dfx_form_create_facets(const dfx_form_desc* desc, int32_t npairs,
                                  const int32_t* pairs, dfx_form** out);
/* A form with an interior facet integral (kernel DFX_K_P1TET_GRADJUMP_A)
 * over the nfacets (cell0, lf0, cell1, lf1) quads of
 * dfx_interior_facet_domain: assemble_matrix adds each facet's 8 x 8
 * macro-element matrix (dofs of cell0 then cell1), BC rows and columns
 * zeroed, in facet order (assemble_interior_facets,
 * fem/assemble_matrix_impl.h:442-667, run after the cell integrals at
 * :799-842); apply_lifting subtracts them. The matrix needs the pattern of
 * dfx_sparsity_create_facets. */
dfx_status dfx_form_create_interior_facets(const dfx_form_desc* desc, int32_t nfacets,
                                           const int32_t* quads, dfx_form** out);
/* The element kernel ABI of the reference (kernelptr_t, fem/kernel.h:18-20;
 * the tabulate_tensor_float64 / _float32 fields of a ufcx_integral,
 * extracted by fem/kernel.h:34-60): A (zeroed by the caller), w = the
 * cell's packed coefficients, c = the constants, coordinate_dofs = the
 * cell's geometry nodes (3 values each), entity_local_index,
 * quadrature_permutation, custom_data. */
typedef void (*dfx_kernel_f64)(double* A, const double* w, const double* c,
                               const double* coordinate_dofs, const int* entity_local_index,
                               const uint8_t* quadrature_permutation, void* custom_data);
typedef void (*dfx_kernel_f32)(float* A, const float* w, const float* c,
                               const float* coordinate_dofs, const int* entity_local_index,
                               const uint8_t* quadrature_permutation, void* custom_data);
/* fem::Form with a cell integral given as integral_data{kernel, cells}
 * (fem/Form.h:52-87; cpp/demo/custom_kernel/main.cpp:77-81): kernel is a
 * dfx_kernel_f64 (dtype DFX_F64) or dfx_kernel_f32 (DFX_F32, geometry
 * given to it in FP32, as U = scalar_value_t<T>, fem/kernel.h:37-38).
 * desc->kernel is DFX_K_ELEMENT_A (bilinear: (bs k)^2 row-major tensors)
 * or DFX_K_ELEMENT_L (linear: bs k); block size 1 or 3, any k <= 32.
 * dfx_assemble_matrix / dfx_assemble_vector / dfx_apply_lifting on this
 * form call the kernel on the host for every cell in cell order, exactly
 * as assemble_cells_matrix / assemble_cells do
 * (fem/assemble_matrix_impl.h:144-152, fem/assemble_vector_impl.h:99-107,
 * entity_local_index and quadrature_permutation NULL, custom_data as
 * given), copy the tensors to the device once and insert them there as
 * dfx_assemble_matrix_elements / _vector_elements / dfx_apply_lifting_elements
 * do. The kernel must be callable from the calling thread. */
dfx_status dfx_form_create_kernel(const dfx_form_desc* desc, void* kernel, void* custom_data,
                                  dfx_form** out);
/* fem::Form with an exterior (type 1) or interior (type 2) facet integral
 * given as integral_data{kernel, entities} (fem/Form.h:52-87): entities
 * are (cell, local facet) pairs (type 1) or (cell0, local facet0, cell1,
 * local facet1) quads (type 2) in integration order. The kernel has the
 * reference's ABI and is called on the host per entity in that order as
 * assemble_exterior_facets / assemble_interior_facets call it
 * (fem/assemble_matrix_impl.h:264-383, :442-667;
 * fem/assemble_vector_impl.h:118-174): coordinate_dofs of the cell (type 2:
 * both cells, cell0 first), w the cell's packed coefficients (type 2: both
 * cells'), entity_local_index the local facet(s), quadrature_permutation
 * zeros. Matrices ((bs k)^2, type 2: (2 bs k)^2 over the joint dofmap of
 * the two cells) get the rows and columns of Dirichlet dofs zeroed and are
 * added on the device with every entry summed in entity order, as
 * mat_set(dmapjoint0, dmapjoint1, Ae); vectors (type 1 only) likewise.
 * dfx_apply_lifting through such a form is supported only for zero data. */
dfx_status dfx_form_create_kernel_facets(const dfx_form_desc* desc, void* kernel,
                                         void* custom_data, int32_t type, int32_t nentities,
                                         const int32_t* entities, dfx_form** out);
dfx_status dfx_form_destroy(dfx_form* f);
/* Replace packed coefficients / constants (same sizes as at creation). */
dfx_status dfx_form_set_coefficients(dfx_form* f, const double* coeffs,
                                     const double* constants);
/* fem::pack_coefficients (fem/pack.h:265-388) on the device: the form's
 * packed coefficients (cells x cstride, in the mesh's cell order) gathered
 * from values given per input cell of mesh m (DFX_PACK_CELL, cstride values
 * per cell: DG0) or per input vertex (DFX_PACK_VERTEX, one value per
 * vertex, cstride = vertices per cell: P1). nvalues must match. */
#define DFX_PACK_CELL 0
#define DFX_PACK_VERTEX 1
dfx_status dfx_form_pack_coefficients(dfx_form* f, const dfx_mesh* m, int layout,
                                      const double* values, int64_t nvalues);

/* fem::DirichletBC (fem/DirichletBC.h:55-618) on unrolled dof indices with
 * values g (NULL: g = 0). n_total = size of the space (nodes*bs). */
dfx_status dfx_bc_create(const int32_t* dofs, int32_t ndofs, int32_t n_total,
                         const double* g, dfx_bc** out);
dfx_status dfx_bc_destroy(dfx_bc* bc);

/* fem::assemble_matrix(mat_add, a, bcs) (fem/assembler.h:542-602):
 * A += a with rows and columns of the BC dofs zeroed (bc may be NULL). */
dfx_status dfx_assemble_matrix(dfx_csr* A, const dfx_form* a,
                               const dfx_bc* bc);
/* Batched fem::assemble_matrix over n independent systems of one affine
 * P1 or Q1 stiffness kernel (no counterpart in the reference; for many
 * meshes of uneven size): one launch forms the cell factors of all systems
 * and one assembles the rows of all systems. A[s] += a[s] with the rows and
 * columns of bc[s] (may be NULL) zeroed, as dfx_assemble_matrix. */
dfx_status dfx_assemble_matrix_batched(int32_t n, dfx_csr* const* A, const dfx_form* const* a,
                                       const dfx_bc* const* bc);
/* Assembly autotuner (no counterpart in the reference): times the
 * row-kernel variants (rows per workgroup, lanes per row) of form a on
 * matrix A and keeps the fastest for later assemblies of a. Overwrites A's
 * values and leaves them zero. */
dfx_status dfx_form_autotune(dfx_form* a, dfx_csr* A, const dfx_bc* bc,
                             int32_t* variant, float* ms);
/* Fix the row-kernel variant of a form: 0-2 one thread per row with 64,
 * 128, 256 rows per workgroup (four incidences in flight); 3 one incidence
 * per lane; 4-6 as 0-2 with, for Q1 hexahedra, element entries in closed
 * form instead of reference tensors (other elements: eight incidences in
 * flight); 7 element matrices of all cells as one matrix-core product,
 * gathered by rows; 8 one incidence per thread, gathered by rows; 9 Q1 on a
 * structured box of hexahedra: one thread per row accumulating its 27
 * entries in registers (error if the mesh is not such a box); 10 as 4 with
 * the cell factors formed from the geometry by every incidence (no cell
 * pass; affine P1/Q1); 11 as 1 with eight incidences in flight; 12-14 Q1
 * hexahedra whose vertices lie in at most 8 cells: one thread per row with
 * its plan at fixed positions, 64/128/256 rows per workgroup; 15 as 9 with
 * the rows of a workgroup stored through an LDS slice; 16 a team of lanes
 * per row (lane j: column j of each incidence) forming the entries from the
 * cell factors; 17 such teams gathering the matrix-core element matrices
 * of 7. The P2 elasticity form has a single row kernel and ignores the
 * variant. */
dfx_status dfx_form_set_variant(dfx_form* a, int32_t variant);
/* fem::assemble_matrix / fem::assemble_vector with a custom kernel
 * (cpp/demo/custom_kernel/main.cpp:73-118): the form (kernel
 * DFX_K_ELEMENT_A / DFX_K_ELEMENT_L) holds only a dofmap (block size 1 or
 * 3; no geometry or coefficients needed), and the kernel's outputs are
 * given on the device: E = num_cells row-major (bs k) x (bs k) element
 * matrices (dofs interleaved per node, as the reference's blocked
 * kernels order them), e = num_cells x (bs k) element vectors.
 * A[dofs_c, dofs_c] += E_c and b[dofs_c] += e_c as la::MatrixCSR::add
 * (la/matrix_csr_impl.h:67-109; mat_add_values<3, 3> for block size 3,
 * la/MatrixCSR.h:157-174) and assemble_cells
 * (fem/assemble_vector_impl.h:72-116) insert them, each entry summed in
 * cell order; rows and columns of bc's dofs are skipped
 * (fem/assemble_matrix_impl.h:156-180); "Entry not in sparsity" if a pair
 * is missing from A's pattern. */
dfx_status dfx_assemble_matrix_elements(dfx_csr* A, const dfx_form* a, const dfx_bc* bc,
                                        const dfx_vec* E);
dfx_status dfx_assemble_vector_elements(dfx_vec* b, const dfx_form* L, const dfx_vec* e);
/* fem::apply_lifting (fem/assembler.h:336-493 -> impl::lift_bc,
 * fem/assemble_vector_impl.h:361-414) for a DFX_K_ELEMENT_A form, with the
 * element matrices E the kernel would recompute for the cells with a BC
 * column dof: b -= alpha E[:, bc] (g - x0). */
dfx_status dfx_apply_lifting_elements(dfx_vec* b, const dfx_form* a, const dfx_bc* bc,
                                      const dfx_vec* E, const dfx_vec* x0, double alpha);
/* The system sequence of the reference's demos (cpp/demo/poisson/
 * main.cpp:207-226): assemble_matrix(A, a, {bc}) and set_diagonal(A, bc,
 * diag); b = 0 (if zero_b) and assemble_vector(b, L); apply_lifting(b,
 * {a}, {{bc}}); bc.set(b). On meshes of 2^20 cells or more the vector
 * assembly runs on a second stream while the matrix is assembled (it reads
 * L's data and writes b only); lifting and set_bc follow both. Results
 * equal the separate calls. bc may be NULL. */
dfx_status dfx_assemble_system(dfx_csr* A, const dfx_form* a, const dfx_bc* bc, double diag,
                               dfx_vec* b, const dfx_form* L, int32_t zero_b);
/* fem::set_diagonal (fem/assembler.h:644-686): A[d,d] = diag for BC dofs */
dfx_status dfx_set_diagonal(dfx_csr* A, const dfx_bc* bc, double diag);
/* fem::assemble_vector(b, L) (fem/assembler.h:230-257): b += L */
dfx_status dfx_assemble_vector(dfx_vec* b, const dfx_form* L);
/* fem::apply_lifting(b, {a}, {{bc}}, {x0}, alpha) (fem/assembler.h:
 * 336-493 -> impl::lift_bc, fem/assemble_vector_impl.h:361-414):
 * b -= alpha A[:, bc] (g - x0) over cells with a BC column dof. */
dfx_status dfx_apply_lifting(dfx_vec* b, const dfx_form* a, const dfx_bc* bc,
                             const dfx_vec* x0, double alpha);
/* DirichletBC::set (fem/DirichletBC.h:495-578): b[bc] = alpha (g - x0) */
dfx_status dfx_bc_set(dfx_vec* b, const dfx_bc* bc, const dfx_vec* x0,
                      double alpha);

/* ------------------------------------------------------------------ */
/* Krylov solvers.                                                     */
/* ------------------------------------------------------------------ */
typedef struct
{
  int32_t max_it;      /* maximum iterations */
  double rtol;         /* stop when ||r||^2 / ||r0||^2 < rtol^2 */
  int32_t jacobi;      /* 1: Jacobi preconditioner, 0: plain CG */
  int32_t check_every; /* host convergence poll interval (iterations) */
  int32_t graph;       /* 1: capture check_every iterations once as a HIP graph
                          and replay it (same kernels, same results);
                          0: launch every kernel */
  int32_t classic;     /* 0: a scalar system of at most 65,536 rows is solved
                          in one cooperative launch (resident blocks, grid
                          barriers; check_every and graph do not apply);
                          1: always the three-kernel iteration */
} dfx_cg_opts;

typedef struct
{
  int32_t iterations;
  int32_t converged;
  double rnorm0;       /* ||r0||_2 (recursive, FP64) */
  double rnorm;        /* ||r_k||_2 (recursive, FP64) */
  double true_rel_residual; /* ||b - A x||_2 / ||b||_2 evaluated in FP64 */
  double seconds;      /* wall time of the solve */
  int32_t cooperative; /* 1: solved in one cooperative launch */
} dfx_cg_result;

/* CG of the reference (cpp/demo/poisson_matrix_free/main.cpp:84-132,
 * python/test/conftest.py:161-205), optionally Jacobi-preconditioned,
 * device resident. x holds the initial guess and receives the solution.
 * A, b, x share one dtype; reductions are FP64. */
dfx_status dfx_cg_solve(const dfx_csr* A, const dfx_vec* b, dfx_vec* x,
                        const dfx_cg_opts* opts, dfx_cg_result* res);

/* dfx_cg_solve without any device allocation (opts->graph is ignored:
 * kernels are launched one by one). No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The work vectors, solver state and reduction partials live
 * in a fixed device array of the library (one per GPU, held until the
 * process ends): at most 1,048,576 scalar rows (nrows x bs); larger
 * systems return DFX_ERR_ARG. The SpMV is the matrix's CSR kernel (no
 * sliced-ELL copy, which would be allocated). Calls that overlap on one
 * GPU share the array and race. */
dfx_status dfx_cg_solve_noalloc(const dfx_csr* A, const dfx_vec* b, dfx_vec* x,
                                const dfx_cg_opts* opts, dfx_cg_result* res);

typedef struct
{
  int32_t max_outer;   /* refinement steps */
  int32_t max_inner;   /* iterations per inner FP32 solve */
  double rtol;         /* target ||b - A x||_2/||b||_2 (FP64 residual) */
  double inner_rtol;   /* refine when the FP32 residual fell by this */
  int32_t check_every;
  int32_t restart;     /* 1: restart the FP32 CG at every refinement;
                          0: keep its search direction (reliable updates) */
  int32_t graph;       /* 1: inner iterations replayed from a captured HIP graph */
  int32_t fallback;    /* 1: when the refinement stalls, finish with FP64
                          Jacobi-PCG on A64 from the current x */
} dfx_mixed_opts;

typedef struct
{
  int32_t outer;
  int32_t inner_total;
  int32_t converged;
  double true_rel_residual;
  double seconds;
  int32_t fallback_iterations; /* FP64 CG iterations after a stall (0: none) */
} dfx_mixed_result;

/* Mixed-precision iterative refinement (no counterpart in the reference;
 * it wraps the same CG): Jacobi-PCG on A32 in FP32 (FP64 reductions)
 * accumulates a correction d; each time its residual has fallen by
 * inner_rtol (or meets half the outer target, so that the last refinement
 * is not solved further than the outer test needs), x += d in FP64 and the
 * residual is replaced by the FP64 true residual b - A64 x. With restart = 0 the FP32 CG keeps its search
 * direction across replacements (reliable updates, van der Vorst & Ye
 * 2000); with restart = 1 every refinement starts a new inner CG.
 * The refinement stalls when a refinement step reduces the FP64 residual
 * norm by less than half, or when max_outer or max_inner runs out first;
 * with fallback = 1 the solve then continues with FP64 Jacobi-PCG on A64
 * from the current x down to rtol (relative to ||b||), otherwise it
 * returns DFX_ERR_NOCONV. */
dfx_status dfx_mixed_solve(const dfx_csr* A64, const dfx_csr* A32,
                           const dfx_vec* b, dfx_vec* x,
                           const dfx_mixed_opts* opts, dfx_mixed_result* res);

/* ------------------------------------------------------------------ */
/* Batched solves (no counterpart in the reference).                   */
/* ------------------------------------------------------------------ */
typedef struct dfx_batch dfx_batch;

/* Segmented FP64 Jacobi-CG over up to max_systems independent systems
 * with at most max_rows rows in total. Every system is cut into chunks
 * that depend only on itself: its iterates are bitwise the same alone and
 * in any mix. */
dfx_status dfx_batch_create(int32_t max_systems, int64_t max_rows, dfx_batch** out);
dfx_status dfx_batch_destroy(dfx_batch* B);
/* Join system A x = b (x is zeroed; A, b, x must outlive membership);
 * returns its slot. */
dfx_status dfx_batch_add(dfx_batch* B, const dfx_csr* A, const dfx_vec* b, dfx_vec* x,
                         double rtol, int32_t max_it, int32_t* slot);
dfx_status dfx_batch_remove(dfx_batch* B, int32_t slot);
/* Queue `iterations` CG iterations for every active system; use_graph
 * replays them as one captured hipGraph. Asynchronous. */
dfx_status dfx_batch_step(dfx_batch* B, int32_t iterations, int32_t use_graph);
/* Per-slot state (synchronises): done, converged, iterations and the
 * recursive relative residual ||r||/||r0||. Output arrays may be NULL. */
dfx_status dfx_batch_status(dfx_batch* B, int32_t n, const int32_t* slots, int32_t* done,
                            int32_t* converged, int32_t* iterations,
                            double* rel_residual);
/* Restart every system of the batch from x = 0 with the current values of
 * its A and b (Jacobi diagonal, r = b, p = z, reductions and counters as
 * dfx_batch_add; rtol and max_it as given there). Kernels only, driven by
 * the batch's device tables, so it can be recorded into a graph and stays
 * valid while systems join and leave. Asynchronous. */
dfx_status dfx_batch_restart(dfx_batch* B);

/* ------------------------------------------------------------------ */
/* Graph capture of library calls (no counterpart in the reference).   */
/* ------------------------------------------------------------------ */
/* Everything the library launches on its stream between dfx_graph_begin
 * and dfx_graph_end, from the calling thread, is recorded into one HIP
 * graph instead of running; dfx_graph_launch replays it as one launch on
 * the library stream. It is for fixed sequences replayed many times, such
 * as many meshes assembled and solved at once:
 *   zero the matrices, dfx_assemble_matrix_batched, the vectors and BCs,
 *   dfx_batch_restart, dfx_batch_step(K iterations)
 * with new coefficients copied in (dfx_form_set_coefficients) before each
 * replay. A replay reads the device buffers the recorded calls used, so
 * they must stay allocated and keep their addresses. Calls that need the
 * host part-way (allocation, uploads of plans or tables, synchronisation:
 * first-use set-up, solves that poll, status queries) fail with
 * DFX_ERR_CAPTURE inside a capture: run the sequence once before
 * capturing it, so that its plans exist. After a failed call the capture
 * stays open until dfx_graph_end, which then discards it and reports the
 * first failure. A replay repeats the recorded work as recorded: an
 * assembly recorded after MatrixCSR::set(0) stores (A = a), one recorded
 * without it adds (A += a) on every replay. */
typedef struct dfx_graph dfx_graph;
dfx_status dfx_graph_begin(void);
dfx_status dfx_graph_end(dfx_graph** out);
dfx_status dfx_graph_launch(const dfx_graph* g);
dfx_status dfx_graph_destroy(dfx_graph* g);

/* ------------------------------------------------------------------ */
/* Timing on the library stream (hipEvent).                            */
/* ------------------------------------------------------------------ */
/* Record event slot (0..63) on the library stream. */
dfx_status dfx_event_record(int slot);
/* Milliseconds between two recorded slots (synchronises on 'b'). */
dfx_status dfx_event_elapsed(int a, int b, float* ms);

#ifdef __cplusplus
}
#endif
#endif
