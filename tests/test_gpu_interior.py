"""Interior facet integrals on P1 tetrahedra (assemble_interior_facets,
fem/assemble_matrix_impl.h:442-667, after the cell integrals at :776-842;
lifting through it, fem/assemble_vector_impl.h:361-414) against the
oracle: the gradient-jump penalty gamma [grad u.n][grad v.n] dS alone and
after a cell integral, BC rows/columns, lifting, the closed forms (A u = 0
for linear u, the energy of a kink), and a pattern without the facet
couplings raising."""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, la, mesh
from dolfinx_amd._lib import DfxError

from conftest import rel_err
from test_mesh_create import box_input
from test_oracle import pattern_with_interior

pytestmark = pytest.mark.gpu


def setup(n=(4, 3, 2), shuffled=False):
    if shuffled:
        cells, x = box_input(n, O.TET, 13)
        om = O.Mesh.from_cells(O.TET, cells, x)
        pm = mesh.create_mesh(mesh.CellType.tetrahedron, cells, x)
    else:
        om = O.Mesh(n, O.TET)
        pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    assert np.array_equal(V.dofmap.list, dm)
    q = fem.interior_facet_domain(pm)
    assert np.array_equal(q, om.interior_facet_domain())
    rp, cols = pattern_with_interior(dm, nd, q)
    return om, pm, V, dm, nd, q, rp, cols


@pytest.mark.parametrize("shuffled", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_gradient_jump_matches_oracle(shuffled, dtype, tol):
    om, pm, V, dm, nd, q, rp, cols = setup(shuffled=shuffled)
    gamma = 0.8
    a = fem.form("p1_gradient_jump", V, None, [gamma], dtype, facets=q)
    bn = om.boundary_nodes(1, dm, marker=1, axis=1, value=0.0)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    bc = fem.dirichletbc(V, bn)
    A = fem.assemble_matrix(a, [bc])
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    ref = O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark, mark, rp,
                                     cols, [gamma])
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data, ref) <= tol
    # again into the same pattern (entries already checked): adds
    fem.assemble_matrix(a, [bc], 1.0, A)
    O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark, mark, rp, cols,
                               [gamma], ref)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data, ref) <= tol
    # stiffness + gradient jump (cells first), kappa per cell
    kap_in = np.random.default_rng(2).uniform(1.0, 2.0, om.num_cells)
    aC = fem.form("p1_stiffness_reaction", V, fem.pack_cellwise(pm, kap_in), [0.1], dtype)
    A2 = fem.assemble_matrix((a, aC), [bc])
    kap = kap_in[om.original_cell_index].reshape(-1, 1)
    ref2 = O.assemble_matrix(O.K_P1TET_A, om.x, om.cells, dm, 1, kap, np.array([0.1]), mark, mark,
                             rp, cols)
    O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark, mark, rp, cols,
                               [gamma], ref2)
    O.set_diagonal(ref2, bn, 1, 1.0, rp, cols)
    assert rel_err(A2.data, ref2) <= tol


def test_gradient_jump_closed_forms_on_gpu():
    om, pm, V, dm, nd, q, rp, cols = setup((4, 2, 3))
    gamma = 2.0
    A = fem.assemble_matrix(fem.form("p1_gradient_jump", V, None, [gamma], facets=q)).to_scipy()
    x = pm.geometry.x
    lin = 1.0 - x[:, 0] + 0.25 * x[:, 1] + 3.0 * x[:, 2]
    assert np.abs(A @ lin).max() <= 1e-12 * abs(A).max()
    kink = np.maximum(x[:, 0] - 0.5, 0.0)
    assert np.isclose(kink @ (A @ kink), gamma, rtol=1e-12)


def test_gradient_jump_lifting_matches_oracle():
    om, pm, V, dm, nd, q, rp, cols = setup((3, 3, 2))
    gamma = 1.2
    a = fem.form("p1_gradient_jump", V, None, [gamma], facets=q)
    bn = om.boundary_nodes(1, dm)
    g = np.zeros(nd)
    g[bn] = 1.0 + np.cos(np.arange(len(bn)))
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    x0 = np.random.default_rng(6).standard_normal(nd)
    b0 = np.random.default_rng(7).standard_normal(nd)
    b = la.Vector.from_array(b0)
    fem.apply_lifting(b, [a], [[fem.dirichletbc(V, bn, g)]], [la.Vector.from_array(x0)],
                      alpha=0.9)
    ref = O.lift_bc_interior(b0.copy(), O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, [gamma], g,
                             mark, x0=x0, alpha=0.9)
    assert rel_err(b.array, ref) <= 1e-12


def test_cell_pattern_raises_for_interior_facets():
    om, pm, V, dm, nd, q, rp, cols = setup((2, 2, 2))
    a = fem.form("p1_gradient_jump", V, None, [1.0], facets=q)
    aC = fem.form("p1_stiffness_reaction", V, np.ones((om.num_cells, 1)), [0.0])
    A = fem.create_matrix(aC)   # cell couplings only
    with pytest.raises(DfxError, match="Entry not in sparsity"):
        fem.assemble_matrix(a, [], 1.0, A)
    with pytest.raises(DfxError, match="out of range"):
        bad = q.copy()
        bad[0, 1] = 7
        fem.form("p1_gradient_jump", V, None, [1.0], facets=bad)
