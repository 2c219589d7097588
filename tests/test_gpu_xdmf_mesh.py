"""Assembly and solve on the reference's own test meshes, read from XDMF
(python/test/data/create_unit_cube_{tetra,hexahedron}.xdmf, the meshes of
python/test/unit/fem/test_fem_pipeline.py:457-470): an unstructured
tetrahedral mesh and a mesh of non-affine hexahedra. A and the lifted b
equal the oracle's on the same mesh; the Dirichlet problem with a linear
exact solution (the degree-1 case of run_scalar_test,
test_fem_pipeline.py:55-108: P1 and Q1 hold linear functions, and 2x2x2
Gauss passes the patch test on trilinear hexahedra) is solved to the
nodal values, and the solution written to XDMF reads back."""

import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, io, la, solvers

from conftest import rel_err
from test_xdmf import read_fixture

pytestmark = pytest.mark.gpu


def exact(X):
    return 1.0 + X[0] + 2.0 * X[1] + 3.0 * X[2]


@pytest.mark.parametrize("ct,name,kA", [(O.TET, "p1_stiffness_reaction", O.K_P1TET_A),
                                        (O.HEX, "q1_stiffness_reaction", O.K_Q1HEX_A)])
def test_fixture_mesh_assembly_and_linear_solution(ct, name, kA, tmp_path):
    pm, cells, x = read_fixture(ct)
    om = O.Mesh.from_cells(ct, cells, x)
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    assert np.array_equal(V.dofmap.list, dm)
    rp, cols = O.sparsity(dm, dm, nd)
    kap = 1.7
    a = fem.form(name, V, np.full((om.num_cells, 1), kap), [0.0])
    bn = fem.locate_boundary_dofs(V)
    g = np.zeros(nd)
    g[bn] = exact(pm.geometry.x[bn].T)
    bc = fem.dirichletbc(V, bn, g)
    A = fem.assemble_matrix(a, [bc])
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    Ar = O.assemble_matrix(kA, om.x, om.cells, dm, 1, np.full((om.num_cells, 1), kap),
                           np.array([0.0]), mark, mark, rp, cols)
    O.set_diagonal(Ar, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data, Ar) <= 1e-12
    b = la.Vector(nd)
    fem.apply_lifting(b, [a], [[bc]])
    fem.set_bc(b, [bc])
    br = O.lift_bc(np.zeros(nd), kA, om.x, om.cells, dm, 1, np.full((om.num_cells, 1), kap),
                   np.array([0.0]), g, mark)
    O.bc_set(br, bn, g)
    assert rel_err(b.array, br) <= 1e-12
    u = fem.Function(V, name="u")
    res = solvers.cg(A, b, u.x, rtol=1e-13, max_it=5000)
    assert res.converged
    ue = exact(pm.geometry.x.T)
    assert np.abs(u.x.array - ue).max() <= 1e-10
    fn = tmp_path / "u.xdmf"
    with io.XDMFFile(fn, "w") as f:
        f.write_mesh(pm)
        f.write_function(u, 1.0)
    att = ET.parse(fn).getroot().find("Domain/Grid[@Name='u']/Grid/Attribute/DataItem")
    assert np.allclose(np.array(att.text.split(), float), ue, rtol=0, atol=1e-10)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_general_hexahedra_are_assembled_deterministically(dtype, tol):
    """Q1 on hexahedra with perturbed vertices (non-affine): the element
    matrices by 2x2x2 Gauss are added by rows in cell order, so repeated
    assemblies and liftings are bitwise equal, and they equal the oracle."""
    from dolfinx_amd import mesh
    from test_mesh_create import box_input
    n = (10, 9, 8)
    cells, x = box_input(n, O.HEX, 5)
    rng = np.random.default_rng(11)
    inner = np.all((x > 1e-12) & (x < 1 - 1e-12), axis=1)
    x = x.copy()
    x[inner] += rng.uniform(-0.2, 0.2, (inner.sum(), 3)) / np.array(n)
    pm = mesh.create_mesh(mesh.CellType.hexahedron, cells, x)
    om = O.Mesh.from_cells(O.HEX, cells, x)
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    assert np.array_equal(V.dofmap.list, dm)
    rp, cols = O.sparsity(dm, dm, nd)
    kap_in = rng.uniform(1.0, 3.0, om.num_cells)
    a = fem.form("q1_stiffness_reaction", V, fem.pack_cellwise(pm, kap_in), [0.4], dtype)
    bn = fem.locate_boundary_dofs(V)
    g = np.zeros(nd)
    g[bn] = exact(pm.geometry.x[bn].T)
    bc = fem.dirichletbc(V, bn, g)
    A1 = fem.assemble_matrix(a, [bc])
    A2 = fem.assemble_matrix(a, [bc])
    assert np.array_equal(A1.data, A2.data)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    kap = kap_in[om.original_cell_index].reshape(-1, 1)
    Ar = O.assemble_matrix(O.K_Q1HEX_A, om.x, om.cells, dm, 1, kap, np.array([0.4]), mark, mark,
                           rp, cols)
    O.set_diagonal(Ar, bn, 1, 1.0, rp, cols)
    assert rel_err(A1.data.astype(np.float64), Ar) <= tol
    bs = []
    for _ in range(2):
        b = la.Vector(nd, dtype)
        fem.apply_lifting(b, [a], [[bc]])
        bs.append(b.array.copy())
    assert np.array_equal(bs[0], bs[1])
    br = O.lift_bc(np.zeros(nd), O.K_Q1HEX_A, om.x, om.cells, dm, 1, kap, np.array([0.4]), g,
                   mark)
    assert rel_err(bs[0].astype(np.float64), br) <= tol
