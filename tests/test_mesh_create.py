"""mesh::create_mesh from caller cells and geometry (mesh/utils.h:1336-1356):
the product's host setup equals the oracle's bit for bit on inputs in
arbitrary order and on a mesh unlike a box (a fan of 40 tetrahedra around
one vertex), and a box fed through create_mesh equals create_box. CPU only.
"""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import _lib, fem, mesh
from dolfinx_amd._lib import DfxError
from dolfinx_amd.la import SparsityPattern


def bicone(m):
    """Tetrahedra (C, P_k, P_k+1, T) and (C, P_k, P_k+1, B) around the centre
    C = 0 of a ring of m points: vertex C belongs to 2 m cells."""
    t = 2 * np.pi * np.arange(m) / m
    ring = np.stack([np.cos(t), np.sin(t), np.zeros(m)], 1)
    x = np.vstack([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0], ring])
    cells = []
    for k in range(m):
        a, b = 3 + k, 3 + (k + 1) % m
        cells.append([0, a, b, 1])
        cells.append([0, a, b, 2])
    return np.array(cells, np.int64), x


def box_input(n, ct, seed):
    """The cells and geometry of a box in a shuffled input order."""
    om = O.Mesh(n, ct)
    cells_in = np.empty_like(om.cells, dtype=np.int64)
    cells_in[om.original_cell_index] = om.input_vertex_index[om.cells]
    x_in = np.empty_like(om.x)
    x_in[om.input_vertex_index] = om.x
    rng = np.random.default_rng(seed)
    cp = rng.permutation(len(cells_in))
    vp = rng.permutation(len(x_in))  # new label of each input vertex
    x_new = np.empty_like(x_in)
    x_new[vp] = x_in
    return vp[cells_in[cp]], x_new


def compare(ct, cells, x, degrees):
    om = O.Mesh.from_cells(ct, cells, x)
    pm = mesh.create_mesh(mesh.CellType(ct), cells, x)
    assert np.array_equal(pm.geometry.x, om.x)
    assert np.array_equal(pm.topology.cells, om.cells)
    assert np.array_equal(pm.topology.original_cell_index, om.original_cell_index)
    assert np.array_equal(pm.geometry.input_global_indices, om.input_vertex_index)
    for deg in degrees:
        V = fem.functionspace(pm, ("Lagrange", deg))
        dm, nd = om.dofmap(deg)
        assert V.num_dofs == nd and np.array_equal(V.dofmap.list, dm)
        assert np.array_equal(fem.locate_boundary_dofs(V), om.boundary_nodes(deg, dm))
        h = C.c_void_p()
        _lib.call("dfx_sparsity_create", V.handle, V.handle, C.byref(h))
        sp = SparsityPattern(h)
        rp, cols = O.sparsity(dm, dm, nd)
        assert np.array_equal(sp.row_ptr, rp) and np.array_equal(sp.cols, cols)
    return om, pm


@pytest.mark.parametrize("ct,n", [(O.TET, (3, 2, 4)), (O.HEX, (4, 3, 2))])
def test_box_through_create_mesh_equals_create_box(ct, n):
    om = O.Mesh(n, ct)
    cells_in = np.empty_like(om.cells, dtype=np.int64)
    cells_in[om.original_cell_index] = om.input_vertex_index[om.cells]
    x_in = np.empty_like(om.x)
    x_in[om.input_vertex_index] = om.x
    a = mesh.create_mesh(mesh.CellType(ct), cells_in, x_in)
    b = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    assert np.array_equal(a.geometry.x, b.geometry.x)
    assert np.array_equal(a.topology.cells, b.topology.cells)
    assert np.array_equal(a.topology.original_cell_index, b.topology.original_cell_index)
    om2 = O.Mesh.from_cells(ct, cells_in, x_in)
    assert np.array_equal(om2.cells, om.cells) and np.array_equal(om2.x, om.x)


@pytest.mark.parametrize("ct,n,degrees", [(O.TET, (3, 3, 2), (1, 2, 3)), (O.HEX, (3, 4, 2), (1,))])
@pytest.mark.parametrize("seed", [0, 1])
def test_shuffled_input_matches_oracle(ct, n, degrees, seed):
    cells, x = box_input(n, ct, seed)
    compare(ct, cells, x, degrees)


def test_bicone_fan_matches_oracle():
    cells, x = bicone(20)
    om, pm = compare(O.TET, cells, x, (1, 2, 3))
    # the centre vertex is in all 40 cells
    centre = int(np.flatnonzero(np.all(om.x == 0.0, axis=1))[0])
    assert np.count_nonzero(om.cells == centre) == 40


def test_create_mesh_rejects_bad_input():
    cells, x = bicone(6)
    bad = cells.copy()
    bad[0, 0] = len(x)
    with pytest.raises(DfxError, match="out of range"):
        mesh.create_mesh(mesh.CellType.tetrahedron, bad, x)
    bad = cells.copy()
    bad[1, 2] = bad[1, 1]
    with pytest.raises(DfxError, match="repeats a vertex"):
        mesh.create_mesh(mesh.CellType.tetrahedron, bad, x)
    with pytest.raises(DfxError, match="belong to a cell"):
        mesh.create_mesh(mesh.CellType.tetrahedron, cells, np.vstack([x, [[5.0, 5.0, 5.0]]]))


@pytest.mark.parametrize("ct,n", [(O.TET, (3, 2, 4)), (O.HEX, (2, 3, 3))])
@pytest.mark.parametrize("axis,value", [(None, 0.0), (0, 1.0), (2, 0.0)])
def test_facet_domain_matches_oracle(ct, n, axis, value):
    """Exterior facet integration entities (fem::compute_integration_domains)
    equal the oracle's pairs, in facet-index order, on a box and on the same
    box given in shuffled input order."""
    for cells, x in [(None, None), box_input(n, ct, 5)]:
        if cells is None:
            om = O.Mesh(n, ct)
            pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
        else:
            om = O.Mesh.from_cells(ct, cells, x)
            pm = mesh.create_mesh(mesh.CellType(ct), cells, x)
        ref = om.facet_domain() if axis is None else om.facet_domain(axis, value)
        assert np.array_equal(fem.facet_domain(pm, axis, value), ref)


@pytest.mark.parametrize("shuffled", [False, True])
def test_interior_facet_domain_and_pattern_match_oracle(shuffled):
    """Interior facet quads (fem::compute_integration_domains(interior_facet))
    equal the oracle's, and the pattern with their couplings
    (sparsitybuild::interior_facets) equals the brute-force rows."""
    from test_oracle import pattern_with_interior
    n = (3, 2, 3)
    if shuffled:
        cells, x = box_input(n, O.TET, 9)
        om = O.Mesh.from_cells(O.TET, cells, x)
        pm = mesh.create_mesh(mesh.CellType.tetrahedron, cells, x)
    else:
        om = O.Mesh(n, O.TET)
        pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    q = fem.interior_facet_domain(pm)
    assert np.array_equal(q, om.interior_facet_domain())
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    h = C.c_void_p()
    _lib.call("dfx_sparsity_create_facets", V.handle, V.handle, q.shape[0], _lib.ptr(q),
              C.byref(h))
    sp = SparsityPattern(h)
    rp, cols = pattern_with_interior(dm, nd, q)
    assert np.array_equal(sp.row_ptr, rp) and np.array_equal(sp.cols, cols)
    bad = q.copy()
    bad[0, 3] = 4
    with pytest.raises(DfxError, match="out of range"):
        _lib.call("dfx_sparsity_create_facets", V.handle, V.handle, 1, _lib.ptr(bad),
                  C.byref(h))
