"""Host setup of libdfx (mesh, reordering, dofmaps, boundary dofs, sparsity)
against the oracle, bit for bit. The setup entry points are host code, so
this runs without a GPU; it calls the C ABI directly where the Python
mirror would need a device form."""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import _lib, fem, mesh

CASES = [(O.TET, (3, 4, 2)), (O.HEX, (4, 3, 5)), (O.TET, (7, 6, 5)), (O.HEX, (9, 8, 7)),
         (O.TET, (1, 1, 1)), (O.HEX, (1, 1, 1)), (O.TET, (16, 1, 3))]


def product_sparsity(V):
    h = C.c_void_p()
    _lib.call("dfx_sparsity_create", V.handle, V.handle, C.byref(h))
    from dolfinx_amd.la import SparsityPattern

    sp = SparsityPattern(h)
    return sp.row_ptr, sp.cols


@pytest.mark.parametrize("ct,n", CASES)
def test_mesh_matches_oracle(ct, n):
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    om = O.Mesh(n, ct)
    assert np.array_equal(m.geometry.x, om.x)
    assert np.array_equal(m.topology.cells, om.cells)
    assert np.array_equal(m.topology.original_cell_index, om.original_cell_index)
    assert np.array_equal(m.geometry.input_global_indices, om.input_vertex_index)


def test_mesh_on_a_shifted_box_matches_oracle():
    p0, p1 = (-1.0, 0.5, 2.0), (0.25, 3.0, 2.5)
    m = mesh.create_box([p0, p1], (5, 3, 4), mesh.CellType.tetrahedron)
    om = O.Mesh((5, 3, 4), O.TET, p0, p1)
    assert np.array_equal(m.geometry.x, om.x)
    assert np.array_equal(m.topology.cells, om.cells)


@pytest.mark.parametrize("ct,n", CASES)
def test_dofmaps_boundary_and_sparsity_match_oracle(ct, n):
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    om = O.Mesh(n, ct)
    degrees = [1, 2, 3] if ct == O.TET else [1]
    for deg in degrees:
        V = fem.functionspace(m, ("Lagrange", deg))
        dm, nd = om.dofmap(deg)
        assert V.dofmap.index_map_size == nd
        assert np.array_equal(V.dofmap.list, dm), deg
        assert np.array_equal(fem.locate_boundary_dofs(V), om.boundary_nodes(deg, dm))
        rp, cols = O.sparsity(dm, dm, nd)
        prp, pcols = product_sparsity(V)
        assert np.array_equal(prp, rp) and np.array_equal(pcols, cols), deg


@pytest.mark.parametrize("n", [(2, 2, 2), (5, 4, 3)])
def test_blocked_p2_space_matches_oracle(n):
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    om = O.Mesh(n, O.TET)
    V = fem.functionspace(m, ("Lagrange", 2, (3,)))
    dm, nn = om.dofmap(2)
    assert V.dofmap.bs == 3 and V.num_dofs == 3 * nn
    assert np.array_equal(V.dofmap.list, dm)
    nodes = om.boundary_nodes(2, dm, 1, 0, 0.0)
    expect = (3 * nodes[:, None] + np.arange(3)).ravel()
    assert np.array_equal(fem.locate_boundary_dofs(V, axis=0, value=0.0), expect)
    rp, cols = O.sparsity(dm, dm, nn)
    prp, pcols = product_sparsity(V)
    assert np.array_equal(prp, rp) and np.array_equal(pcols, cols)


@pytest.mark.parametrize("axis,value", [(0, 1.0), (1, 0.0), (2, 1.0)])
def test_boundary_dofs_on_one_face_match_oracle(axis, value):
    n = (4, 5, 3)
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.hexahedron)
    om = O.Mesh(n, O.HEX)
    V = fem.functionspace(m, ("Lagrange", 1))
    dm, _ = om.dofmap(1)
    got = fem.locate_boundary_dofs(V, axis=axis, value=value)
    assert np.array_equal(got, om.boundary_nodes(1, dm, 1, axis, value))
    # every located dof sits on the face, and the face holds the whole vertex grid
    dof_x = np.zeros((V.dofmap.index_map_size, 3))
    dof_x[V.dofmap.list.ravel()] = m.geometry.x[m.geometry.dofmap.ravel()]
    assert np.all(dof_x[got, axis] == value)
    others = [n[i] + 1 for i in range(3) if i != axis]
    assert len(got) == others[0] * others[1]


def test_config_counts():
    """Closed-form sizes of the configurations (nnz = (3n+1)^3 for Q1 hex,
    (n+1)^3 + 2E(n) for P1 tets; boundary (n+1)^3 - (n-1)^3)."""
    for n in (6, 10):
        m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (n, n, n), mesh.CellType.hexahedron)
        V = fem.functionspace(m, ("Lagrange", 1))
        rp, _ = product_sparsity(V)
        assert rp[-1] == (3 * n + 1) ** 3
        assert len(fem.locate_boundary_dofs(V)) == (n + 1) ** 3 - (n - 1) ** 3
        m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (n, n, n),
                            mesh.CellType.tetrahedron)
        V = fem.functionspace(m, ("Lagrange", 1))
        rp, _ = product_sparsity(V)
        E = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
        assert rp[-1] == (n + 1) ** 3 + 2 * E
        assert len(fem.locate_boundary_dofs(V)) == (n + 1) ** 3 - (n - 1) ** 3
