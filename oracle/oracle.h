/*
 * CPU oracle for the DOLFINx assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithms, used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py as the checker. The product library
 * (dolfinx_amd/) never links, loads or calls anything in this directory.
 *
 * Every function names the reference file:line it restates (paths are
 * relative to the DOLFINx source tree, cpp/dolfinx/... unless stated).
 *
 * Pinning (what checks each function against something other than itself;
 * tests/test_oracle.py unless stated):
 *   orc_box_*            closed form of generation.h: counts, lexicographic
 *                        coordinates, sum |K| = 1, |det J| = h^3/6.
 *   orc_dual_graph       brute-force facet matching on tiny meshes.
 *   orc_rcm              independent Python restatement of ordering.cpp on
 *                        box dual graphs and random (disconnected) graphs.
 *   orc_mesh_create_box  first-appearance vertex numbering invariant.
 *   orc_mesh_create      the box through it equals orc_mesh_create_box; on
 *                        other inputs it composes the pinned dual graph,
 *                        RCM and numbering (invariants on a bicone fan).
 *   orc_entities         edges = lexicographic ranks of sorted vertex pairs;
 *                        Euler characteristic V - E + F - C = 1.
 *   orc_dofmap           independent Python restatement (P1, P2, P3); P3
 *                        edge dofs agree between the cells of an edge.
 *   orc_boundary_nodes   geometric brute force.
 *   orc_sparsity         brute-force row sets.
 *   orc_tabulate         P1 barycentric closed form; Q1 exact integration of
 *                        trilinear fields on a parallelepiped; P3 exact
 *                        quadratic forms of cubic fields; P2 elasticity rigid
 *                        modes in the kernel and exact energy of linear
 *                        fields; source totals int f (1 and 1/2 for f = 1,
 *                        x_0, python/test/unit/fem/test_assembler.py:67-78).
 *   orc_facet_domain     brute-force facet vertex sets, facet-index order.
 *   orc_assemble_vector_facets  int_Gamma g ds closed forms: surface area,
 *                        one face, linear g (exact in P1/Q1).
 *   orc_assemble_matrix_facets  sum of entries = beta |Gamma|, 1' A x =
 *                        int_Gamma x ds = 3 on the unit cube, a unit facet's
 *                        block = the facet mass matrix; BC rows/columns by
 *                        brute force; orc_lift_bc_facets = dense column
 *                        product of the facet matrix.
 *   orc_interior_facet_domain  brute-force shared facet vertex sets.
 *   orc_assemble_matrix_interior  A u = 0 for u linear (no gradient jump),
 *                        u' A u = gamma |plane| for a kink across a mesh
 *                        plane, symmetric positive semidefinite; BC
 *                        rows/columns and lifting by brute force.
 *   orc_add_element_*    equal to orc_assemble_* fed the tabulated tensors;
 *                        a dense scatter of random tensors on tiny meshes.
 *   orc_assemble_*       P1 mass total = |Omega|, A 1 = 0, "Entry not in
 *                        sparsity", lifting = dense column product,
 *                        manufactured linear solution exact at the nodes;
 *                        frozen fixtures tests/golden/ (test_golden.py).
 *   orc_cg, orc_pcg      scipy.sparse.linalg.spsolve; non-convergence flag.
 * Parity unpinned (no reference run is possible here): the position of the
 * P3 GLL edge nodes (the tests pin the P3 space, not its nodal basis), and
 * whether the reference itself yields these dofmaps at the large sizes.
 */
#ifndef DFX_ORACLE_H
#define DFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TET = 0, ORC_HEX = 1 };

typedef struct orc_mesh
{
  int celltype;                 /* ORC_TET / ORC_HEX */
  int nvpc;                     /* vertices per cell (4 / 8) */
  int32_t num_cells;
  int32_t num_vertices;
  double* x;                    /* (num_vertices x 3), geometry order */
  int32_t* cells;               /* (num_cells x nvpc), local vertex indices */
  int64_t* original_cell_index; /* input (lexicographic) cell of each cell */
  int64_t* input_vertex_index;  /* input (lexicographic) vertex of each vertex */
} orc_mesh;

/* ---- mesh (mesh/generation.h, mesh/utils.h, graphbuild.cpp, ordering.cpp,
 *      Topology.cpp, topologycomputation.cpp) ---- */
int64_t orc_box_num_vertices(const int64_t n[3]);
int64_t orc_box_num_cells(const int64_t n[3], int celltype);
void orc_box_geometry(const double p0[3], const double p1[3],
                      const int64_t n[3], double* x);
void orc_box_cells(const int64_t n[3], int celltype, int64_t* cells);
/* Local dual graph; returns number of edges (links = 2x). Caller frees. */
int orc_dual_graph(int celltype, const int64_t* cells, int32_t num_cells,
                   int32_t** offsets, int32_t** data);
void orc_rcm(const int32_t* offsets, const int32_t* data, int32_t n,
             int32_t* remap);
orc_mesh* orc_mesh_create_box(const double p0[3], const double p1[3],
                              const int64_t n[3], int celltype);
/* mesh::create_mesh on one rank from input cells (nc x nvpc input vertex
 * indices) and geometry (nv x 3). */
orc_mesh* orc_mesh_create(int celltype, const int64_t* cells_in, int32_t nc,
                          const double* x_in, int32_t nv);
void orc_mesh_free(orc_mesh* m);
int orc_cell_num_entities(int celltype, int dim);
/* Entities of dimension dim (1 or 2): c2e (num_cells x ne), e2v (n x nvpe).
 * Returns the number of entities. Caller frees both. */
int32_t orc_entities(const orc_mesh* m, int dim, int32_t** c2e,
                     int32_t** e2v);
/* Per-cell edge reflection bits (bit e set iff edge e is reflected). */
void orc_edge_reflections(const orc_mesh* m, const int32_t* c2e,
                          const int32_t* e2v, uint32_t* bits);
/* Interior facets as (cell0, lf0, cell1, lf1), facet-index order, cell0 <
 * cell1; returns count, caller frees. */
int32_t orc_interior_facet_domain(const orc_mesh* m, int32_t** quads);
/* Exterior facets as (cell, local facet) pairs; returns count. Caller frees */
int32_t orc_exterior_facets(const orc_mesh* m, int32_t** cell_facet);
/* fem::compute_integration_domains(exterior_facet) (fem/utils.cpp:154-273)
 * on mesh::exterior_facet_indices, optionally only the facets with all
 * vertices at x[axis] == value (marker 1): (cell, local facet) pairs in
 * facet-index order. Returns count, caller frees. */
int32_t orc_facet_domain(const orc_mesh* m, int marker, int axis, double value,
                         int32_t** pairs);

/* ---- dofmaps, BCs, sparsity (fem/dofmapbuilder.cpp, fem/utils.cpp,
 *      fem/DirichletBC.cpp, la/SparsityPattern.cpp) ---- */
int orc_element_ndofs(int celltype, int degree);
/* Scalar (node) dofmap for Lagrange degree 1..3 (tet) / 1 (hex).
 * Returns number of nodes; dofmap is (num_cells x ndofs). */
int32_t orc_dofmap(const orc_mesh* m, int degree, int32_t* dofmap);
/* Sorted unique node indices in the closure of the given exterior facets.
 * marker: 0 = all exterior facets, 1 = facets with all vertices at
 * x[axis] == value. Returns count, caller frees. */
int32_t orc_boundary_nodes(const orc_mesh* m, int degree,
                           const int32_t* dofmap, int marker, int axis,
                           double value, int32_t** nodes);
/* Sparsity of the (block) dofmap outer products: row_ptr (nrows+1) int64,
 * cols sorted unique. Returns nnz, caller frees cols. */
int64_t orc_sparsity(const int32_t* dofmap0, int k0, const int32_t* dofmap1,
                     int k1, int32_t num_cells, int32_t nrows,
                     int64_t* row_ptr, int32_t** cols);

/* ---- element kernels (FFCx tabulate_tensor ABI, fem/kernel.h:18-27) ---- */
typedef void (*orc_kernel_fn)(double* A, const double* w, const double* c,
                              const double* coordinate_dofs,
                              const int* entity_local_index,
                              const uint8_t* quadrature_permutation,
                              void* custom_data);
enum
{
  ORC_K_P1TET_A = 0,   /* a = kappa grad u.grad v + c u v; w=[kappa], c=[c] */
  ORC_K_P1TET_L = 1,   /* L = f v; w = [f at 4 vertices] */
  ORC_K_Q1HEX_A = 2,   /* a = kappa grad u.grad v + c u v; w=[kappa], c=[c] */
  ORC_K_Q1HEX_L = 3,   /* L = f v; w = [f at 8 vertices] */
  ORC_K_P2TET_ELAST_A = 4, /* sigma(u):eps(v); w = [lambda, mu] */
  ORC_K_P2TET_ELAST_L = 5, /* f.v; c = [fx, fy, fz] */
  ORC_K_P3TET_A = 6,   /* kappa grad u.grad v + c u v; w=[kappa], c=[c] */
  ORC_K_P3TET_L = 7,   /* f v; c = [f] */
  ORC_K_P1TET_NEUMANN = 10, /* g v ds on an exterior facet; w = [g at 4 vertices] */
  ORC_K_Q1HEX_NEUMANN = 11, /* g v ds on an exterior facet; w = [g at 8 vertices] */
  ORC_K_P1TET_ROBIN_A = 12, /* beta u v ds on an exterior facet; w = [beta] */
  ORC_K_Q1HEX_ROBIN_A = 13, /* beta u v ds on an exterior facet; w = [beta] */
  ORC_K_P1TET_GRADJUMP_A = 14, /* gamma [grad u.n][grad v.n] dS, c = [gamma] */
  ORC_K_COUNT = 15
};
orc_kernel_fn orc_get_kernel(int id);
void orc_tabulate(int id, double* A, const double* w, const double* c,
                  const double* coordinate_dofs);

/* ---- assembly (fem/assemble_matrix_impl.h, assemble_vector_impl.h,
 *      assembler.h, la/matrix_csr_impl.h) ---- */
/* Returns 0, or -1 if an entry is not in the sparsity pattern. */
int orc_assemble_matrix(int kernel, const double* x, const int32_t* x_dofmap,
                        int nodes_per_cell, int32_t num_cells,
                        const int32_t* dofmap, int k, int bs,
                        const double* coeffs, int cstride,
                        const double* constants, const int8_t* bc0,
                        const int8_t* bc1, const int64_t* row_ptr,
                        const int32_t* cols, double* values);
/* Element tensors supplied by the caller (num_cells row-major (bs k)^2
 * matrices / bs k vectors), inserted exactly as orc_assemble_matrix /
 * orc_assemble_vector insert the kernel's outputs. */
int orc_add_element_matrices(int32_t num_cells, const int32_t* dofmap, int k,
                             int bs, const double* E, const int8_t* bc0,
                             const int8_t* bc1, const int64_t* row_ptr,
                             const int32_t* cols, double* values);
void orc_add_element_vectors(int32_t num_cells, const int32_t* dofmap, int k,
                             int bs, const double* e, double* b);
/* Exterior facet matrices over (cell, local facet) pairs in the given
 * order, added into the CSR values with BC rows/columns zeroed; -1 if an
 * entry is not in the sparsity. */
int orc_assemble_matrix_facets(int kernel, const double* x,
                               const int32_t* x_dofmap, int npc,
                               const int32_t* dofmap, int k, int bs,
                               const double* coeffs, int cstride,
                               const double* constants, const int32_t* pairs,
                               int32_t npairs, const int8_t* bc0,
                               const int8_t* bc1, const int64_t* row_ptr,
                               const int32_t* cols, double* values);
/* Interior facet integral over (cell0, lf0, cell1, lf1) quads in order,
 * added into the CSR values with BC rows/columns zeroed; -1 if an entry is
 * not in the sparsity. Lifting over the same integral. */
int orc_assemble_matrix_interior(int kernel, const double* x,
                                 const int32_t* x_dofmap, int npc,
                                 const int32_t* dofmap, int k,
                                 const double* constants, const int32_t* quads,
                                 int32_t nfacets, const int8_t* bc0,
                                 const int8_t* bc1, const int64_t* row_ptr,
                                 const int32_t* cols, double* values);
void orc_lift_bc_interior(int kernel, const double* x, const int32_t* x_dofmap,
                          int npc, const int32_t* dofmap, int k,
                          const double* constants, const int32_t* quads,
                          int32_t nfacets, const double* bc_values,
                          const int8_t* bc_markers, const double* x0,
                          double alpha, double* b);
/* Lifting by exterior facet matrices (lift_bc over the facet integral). */
void orc_lift_bc_facets(int kernel, const double* x, const int32_t* x_dofmap,
                        int npc, const int32_t* dofmap, int k, int bs,
                        const double* coeffs, int cstride,
                        const double* constants, const int32_t* pairs,
                        int32_t npairs, const double* bc_values,
                        const int8_t* bc_markers, const double* x0,
                        double alpha, double* b);
/* Exterior facet loads over (cell, local facet) pairs in the given order. */
void orc_assemble_vector_facets(int kernel, const double* x,
                                const int32_t* x_dofmap, int npc,
                                const int32_t* dofmap, int k, int bs,
                                const double* coeffs, int cstride,
                                const double* constants, const int32_t* pairs,
                                int32_t npairs, double* b);
void orc_assemble_vector(int kernel, const double* x, const int32_t* x_dofmap,
                         int nodes_per_cell, int32_t num_cells,
                         const int32_t* dofmap, int k, int bs,
                         const double* coeffs, int cstride,
                         const double* constants, double* b);
void orc_lift_bc(int kernel, const double* x, const int32_t* x_dofmap,
                 int nodes_per_cell, int32_t num_cells, const int32_t* dofmap,
                 int k, int bs, const double* coeffs, int cstride,
                 const double* constants, const double* bc_values,
                 const int8_t* bc_markers, const double* x0, double alpha,
                 double* b);
void orc_bc_set(double* b, const int32_t* dofs, int32_t ndofs,
                const double* g, const double* x0, double alpha);
int orc_set_diagonal(const int32_t* rows, int32_t nrows, int bs, double diag,
                     const int64_t* row_ptr, const int32_t* cols,
                     double* values);

/* ---- linear algebra (la/matrix_csr_impl.h, la/Vector.h, demo CG) ---- */
void orc_spmv(int32_t nrows, int bs, const int64_t* row_ptr,
              const int32_t* cols, const double* values, const double* x,
              double* y);
double orc_dot(int32_t n, const double* a, const double* b);
/* Reference CG (no preconditioner). Returns iterations, or -1 if kmax hit. */
int orc_cg(int32_t nrows, int bs, const int64_t* row_ptr, const int32_t* cols,
           const double* values, const double* b, double* x, int kmax,
           double rtol);
/* Jacobi-preconditioned CG with the reference's stopping test. */
int orc_pcg_jacobi(int32_t nrows, int bs, const int64_t* row_ptr,
                   const int32_t* cols, const double* values, const double* b,
                   double* x, int kmax, double rtol);

void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
