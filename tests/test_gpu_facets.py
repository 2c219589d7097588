"""Exterior facet integrals on P1 tetrahedra and Q1 hexahedra against the
oracle: loads int_Gamma g v ds (fem::assemble_vector over an exterior_facet
integral, fem/assemble_vector_impl.h:118-174) with g a P1/Q1 field given at
the vertices, on all exterior facets and on one face; facet matrices beta u
v ds (assemble_entities, fem/assemble_matrix_impl.h:264-383) alone and after
a cell integral, with BC rows/columns, and their lifting."""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, la, mesh, solvers
from dolfinx_amd._lib import DfxError

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ct,name,kN", [(O.TET, "p1_neumann", O.K_P1TET_NEUMANN),
                                        (O.HEX, "q1_neumann", O.K_Q1HEX_NEUMANN)])
@pytest.mark.parametrize("axis,value", [(None, 0.0), (0, 1.0)])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_facet_loads_match_oracle(ct, name, kN, axis, value, dtype, tol):
    n = (4, 3, 5)
    om = O.Mesh(n, ct)
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    xin = np.empty_like(om.x)
    xin[om.input_vertex_index] = om.x
    g_in = 1.0 + xin[:, 0] + 2.0 * xin[:, 1] ** 2 - xin[:, 2]   # per input vertex
    pairs = fem.facet_domain(pm, axis, value)
    L = fem.form(name, V, fem.pack_vertexwise(pm, g_in), None, dtype, facets=pairs)
    b = fem.assemble_vector(L)
    g = g_in[om.input_vertex_index][om.cells]
    ref = O.assemble_vector_facets(kN, om.x, om.cells, dm, g, pairs, nd)
    assert rel_err(b.array, ref) <= tol
    # adds onto b
    fem.assemble_vector(L, b)
    assert rel_err(b.array, 2 * ref) <= tol
    if axis is None and dtype == np.float64:
        one = fem.form(name, V, np.ones((om.num_cells, om.nvpc)), None, dtype, facets=pairs)
        assert np.isclose(fem.assemble_vector(one).array.sum(), 6.0, rtol=1e-13)


def test_facet_pairs_out_of_range_raise():
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (2, 2, 2), mesh.CellType.tetrahedron)
    V = fem.functionspace(pm, ("Lagrange", 1))
    nc = pm.topology.cells.shape[0]
    with pytest.raises(Exception, match="out of range"):
        fem.form("p1_neumann", V, np.ones((nc, 4)), facets=np.array([[0, 4]], np.int32))


ROBIN = [(O.TET, "p1_robin", O.K_P1TET_ROBIN_A, "p1_stiffness_reaction", O.K_P1TET_A),
         (O.HEX, "q1_robin", O.K_Q1HEX_ROBIN_A, "q1_stiffness_reaction", O.K_Q1HEX_A)]


def robin_setup(ct, n=(4, 3, 3)):
    om = O.Mesh(n, ct)
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    assert np.array_equal(V.dofmap.list, dm)
    rp, cols = O.sparsity(dm, dm, nd)
    rng = np.random.default_rng(21)
    beta_in = rng.uniform(0.5, 2.0, om.num_cells)
    kap_in = rng.uniform(1.0, 3.0, om.num_cells)
    return om, pm, V, dm, nd, rp, cols, beta_in, kap_in


@pytest.mark.parametrize("ct,rname,kR,aname,kA", ROBIN)
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
def test_robin_matrix_matches_oracle(ct, rname, kR, aname, kA, dtype, tol):
    """beta u v ds on every exterior facet, BC rows/columns of the x = 0
    face zeroed, unit diagonal; then the same facets added after the cell
    integral (a form with two integrals, assembled cells first)."""
    om, pm, V, dm, nd, rp, cols, beta_in, kap_in = robin_setup(ct)
    pairs = fem.facet_domain(pm)
    beta = fem.pack_cellwise(pm, beta_in)
    aR = fem.form(rname, V, beta, None, dtype, facets=pairs)
    bn = om.boundary_nodes(1, dm, marker=1, axis=0, value=0.0)
    bc = fem.dirichletbc(V, bn)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    A = fem.assemble_matrix(aR, [bc])
    co = beta_in[om.original_cell_index].reshape(-1, 1)
    ref = O.assemble_matrix_facets(kR, om.x, om.cells, dm, co, pairs, mark, mark, rp, cols)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    assert rel_err(A.data, ref) <= tol
    # a second assembly into the same pattern (entries already checked) adds
    fem.assemble_matrix(aR, [bc], 1.0, A)
    O.assemble_matrix_facets(kR, om.x, om.cells, dm, co, pairs, mark, mark, rp, cols, ref)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data, ref) <= tol
    # cell integral + facet integral on the face x = 1
    right = fem.facet_domain(pm, 0, 1.0)
    aC = fem.form(aname, V, fem.pack_cellwise(pm, kap_in), [0.25], dtype)
    aR1 = fem.form(rname, V, beta, None, dtype, facets=right)
    A2 = fem.assemble_matrix((aR1, aC), [bc])
    kap = kap_in[om.original_cell_index].reshape(-1, 1)
    ref2 = O.assemble_matrix(kA, om.x, om.cells, dm, 1, kap, np.array([0.25]), mark, mark, rp,
                             cols)
    O.assemble_matrix_facets(kR, om.x, om.cells, dm, co, right, mark, mark, rp, cols, ref2)
    O.set_diagonal(ref2, bn, 1, 1.0, rp, cols)
    assert rel_err(A2.data, ref2) <= tol


@pytest.mark.parametrize("ct,rname,kR,aname,kA", ROBIN)
def test_robin_lifting_matches_oracle(ct, rname, kR, aname, kA):
    """apply_lifting over a form with a cell and a facet integral (nonzero g,
    x0, alpha) equals the oracle's lift_bc over both integrals."""
    om, pm, V, dm, nd, rp, cols, beta_in, kap_in = robin_setup(ct, (3, 4, 3))
    pairs = fem.facet_domain(pm)
    aC = fem.form(aname, V, fem.pack_cellwise(pm, kap_in), [0.25])
    aR = fem.form(rname, V, fem.pack_cellwise(pm, beta_in), facets=pairs)
    bn = om.boundary_nodes(1, dm, marker=1, axis=2, value=0.0)
    g = np.zeros(nd)
    g[bn] = 1.0 + np.sin(np.arange(len(bn)))
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    x0 = np.random.default_rng(4).standard_normal(nd)
    b0 = np.random.default_rng(5).standard_normal(nd)
    b = la.Vector.from_array(b0)
    fem.apply_lifting(b, [(aC, aR)], [[fem.dirichletbc(V, bn, g)]], [la.Vector.from_array(x0)],
                      alpha=0.7)
    kap = kap_in[om.original_cell_index].reshape(-1, 1)
    co = beta_in[om.original_cell_index].reshape(-1, 1)
    ref = O.lift_bc(b0.copy(), kA, om.x, om.cells, dm, 1, kap, np.array([0.25]), g, mark, x0=x0,
                    alpha=0.7)
    O.lift_bc_facets(ref, kR, om.x, om.cells, dm, co, pairs, g, mark, x0=x0, alpha=0.7)
    assert rel_err(b.array, ref) <= 1e-12


@pytest.mark.parametrize("ct,rname,kR,aname,kA", ROBIN)
def test_robin_problem_has_constant_solution(ct, rname, kR, aname, kA):
    """-div(kappa grad u) = 0 with kappa du/dn + beta u = beta on the whole
    boundary has the solution u = 1, which P1 and Q1 hold exactly: CG on
    the assembled system returns 1 at every node."""
    om, pm, V, dm, nd, rp, cols, beta_in, kap_in = robin_setup(ct, (5, 4, 6))
    pairs = fem.facet_domain(pm)
    beta = fem.pack_cellwise(pm, beta_in)
    a = (fem.form(aname, V, fem.pack_cellwise(pm, kap_in), [0.0]),
         fem.form(rname, V, beta, facets=pairs))
    nv = 4 if ct == O.TET else 8
    L = fem.form("p1_neumann" if ct == O.TET else "q1_neumann", V, np.repeat(beta, nv, axis=1),
                 facets=pairs)
    A, b = fem.assemble_system(a, L, [])
    x = la.Vector(nd)
    res = solvers.cg(A, b, x, rtol=1e-12, max_it=5000)
    assert res.converged
    assert np.abs(x.array - 1.0).max() <= 1e-8


def test_robin_forms_reject_misuse():
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (2, 2, 2), mesh.CellType.hexahedron)
    V = fem.functionspace(pm, ("Lagrange", 1))
    nc = pm.topology.cells.shape[0]
    aR = fem.form("q1_robin", V, np.ones((nc, 1)), facets=fem.facet_domain(pm))
    with pytest.raises(DfxError, match="linear form"):
        fem.assemble_vector(aR)
    diag = la.MatrixCSR(np.arange(V.num_dofs + 1, dtype=np.int64),
                        np.arange(V.num_dofs, dtype=np.int32))
    with pytest.raises(DfxError, match="Entry not in sparsity"):
        fem.assemble_matrix(aR, [], 1.0, diag)


def _padded(rp, cols, nd, width):
    """The pattern with every row's columns united with 0..width-1 (rows
    wider than the element couplings: the slot search paths for rows of
    more than 64 and more than 256 entries)."""
    rows = []
    for r in range(nd):
        rows.append(np.union1d(cols[rp[r]:rp[r + 1]], np.arange(min(width, nd))))
    prp = np.zeros(nd + 1, np.int64)
    prp[1:] = np.cumsum([len(x) for x in rows])
    return prp, np.concatenate(rows).astype(np.int32)


@pytest.mark.parametrize("width", [100, 300])
def test_facet_forms_twice_into_a_superset_pattern(width):
    """Robin and gradient-jump matrices assembled twice into a pattern whose
    rows hold more than 64 (width 100) or more than 256 (width 300) entries:
    the first assembly checks and caches the slots, the second reuses them;
    both against the oracle."""
    from test_gpu_interior import setup as interior_setup

    n = (8, 8, 8)
    om, pm, V, dm, nd, q, rp0, cols0 = interior_setup(n)
    rp, cols = _padded(rp0, cols0, nd, width)
    assert np.diff(rp).max() > (256 if width == 300 else 64)
    bn = om.boundary_nodes(1, dm, marker=1, axis=0, value=0.0)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    bc = fem.dirichletbc(V, bn)
    beta_in = np.random.default_rng(8).uniform(0.5, 2.0, om.num_cells)
    co = beta_in[om.original_cell_index].reshape(-1, 1)
    pairs = fem.facet_domain(pm)
    aR = fem.form("p1_robin", V, fem.pack_cellwise(pm, beta_in), facets=pairs)
    aJ = fem.form("p1_gradient_jump", V, None, [0.6], facets=q)
    for form, kind in ((aR, "robin"), (aJ, "jump")):
        A = la.MatrixCSR(rp, cols)
        ref = np.zeros(len(cols))
        for _ in range(2):
            fem.assemble_matrix(form, [bc], 1.0, A)
            if kind == "robin":
                O.assemble_matrix_facets(O.K_P1TET_ROBIN_A, om.x, om.cells, dm, co, pairs, mark,
                                         mark, rp, cols, ref)
            else:
                O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark,
                                           mark, rp, cols, [0.6], ref)
            O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
            assert rel_err(A.data, ref) <= 1e-13, kind
