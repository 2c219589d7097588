"""GPU assembly on meshes from mesh::create_mesh against the oracle: P2
elasticity on a fan of 40 tetrahedra around one vertex (block rows with 40
incidences), P1/P3 Poisson and Q1 on boxes given in shuffled input order.
FP64 within 1e-12 of the oracle."""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, mesh, problems

from conftest import rel_err
from test_mesh_create import bicone, box_input

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m", [20, 24])
def test_elasticity_on_fan_matches_oracle(m):
    cells, x = bicone(m)
    om = O.Mesh.from_cells(O.TET, cells, x)
    pm = mesh.create_mesh(mesh.CellType.tetrahedron, cells, x)
    V = fem.functionspace(pm, ("Lagrange", 2, (3,)))
    dm, nn = om.dofmap(2)
    assert np.array_equal(V.dofmap.list, dm)
    nc = om.num_cells
    E = np.random.default_rng(m).uniform(0.5, 2.0, nc)
    lam, mu = problems.lame_from_E(E)
    a = fem.form("p2_elasticity", V, fem.pack_cellwise(pm, np.stack([lam, mu], 1)))
    nodes = om.boundary_nodes(2, dm)
    bdofs = (3 * nodes[:, None] + np.arange(3)).ravel()
    bc = fem.dirichletbc(V, bdofs)
    A = fem.assemble_matrix(a, [bc])
    rp, cols = O.sparsity(dm, dm, nn)
    mark = np.zeros(3 * nn, np.int8)
    mark[bdofs] = 1
    co = np.stack([lam, mu], 1)[om.original_cell_index]
    Ar = O.assemble_matrix(O.K_P2TET_ELAST_A, om.x, om.cells, dm, 3, co, None, mark, mark, rp,
                           cols)
    O.set_diagonal(Ar, bdofs, 3, 1.0, rp, cols)
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    assert rel_err(A.data, Ar) <= 1e-12
    # the centre's block row gathers 2 m incidences
    centre = int(np.flatnonzero(np.all(om.x == 0.0, axis=1))[0])
    assert np.count_nonzero(dm == centre) == 2 * m


@pytest.mark.parametrize("ct,deg,kA", [(O.TET, 1, O.K_P1TET_A), (O.TET, 3, O.K_P3TET_A),
                                       (O.HEX, 1, O.K_Q1HEX_A)])
def test_poisson_on_shuffled_box_matches_oracle(ct, deg, kA):
    cells, x = box_input((3, 2, 3), ct, 7)
    om = O.Mesh.from_cells(ct, cells, x)
    pm = mesh.create_mesh(mesh.CellType(ct), cells, x)
    V = fem.functionspace(pm, ("Lagrange", deg))
    dm, nd = om.dofmap(deg)
    kap_in = np.random.default_rng(3).uniform(1.0, 3.0, om.num_cells)
    name = {O.K_P1TET_A: "p1_stiffness_reaction", O.K_P3TET_A: "p3_stiffness_reaction",
            O.K_Q1HEX_A: "q1_stiffness_reaction"}[kA]
    a = fem.form(name, V, fem.pack_cellwise(pm, kap_in), [0.5])
    bn = om.boundary_nodes(deg, dm)
    bc = fem.dirichletbc(V, bn)
    A = fem.assemble_matrix(a, [bc])
    rp, cols = O.sparsity(dm, dm, nd)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    kap = kap_in[om.original_cell_index].reshape(-1, 1)
    Ar = O.assemble_matrix(kA, om.x, om.cells, dm, 1, kap, np.array([0.5]), mark, mark, rp, cols)
    O.set_diagonal(Ar, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data, Ar) <= 1e-12
