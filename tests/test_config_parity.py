"""Host setup at the full sizes of the five workloads, product against the
oracle, bit for bit: mesh (coordinates, reordered cells, original cell index,
input vertex index), dofmaps, Dirichlet dofs and the CSR pattern.

These sizes run the product setup's threaded paths (counting sorts and row
unions split over threads above 200,000 items), which the small boxes of
test_setup_parity.py do not reach. Sizes (cells / dofs / nnz):
  P1 tetrahedra 32^3   196,608 / 35,937 / 513,313
  P3 tetrahedra 24^3   82,944 / 389,017 / 18,388,729
  P2 vector tetrahedra 48^3  663,552 / 912,673 nodes (bs 3) / 25,755,265 blocks
  Q1 hexahedra 128^3   2,097,152 / 2,146,689 / 57,066,625
"""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import _lib, fem, mesh
from dolfinx_amd.la import SparsityPattern

# (cell type, n, degree, block size, Dirichlet marker: None = whole boundary,
#  (axis, value) = one face), expected (cells, nodes, nnz)
CONFIGS = {
    "p1_tet_32": (O.TET, 32, 1, 1, None, (196_608, 35_937, 513_313)),
    "p3_tet_24": (O.TET, 24, 3, 1, None, (82_944, 389_017, 18_388_729)),
    "p2v_tet_48": (O.TET, 48, 2, 3, (0, 0.0), (663_552, 912_673, 25_755_265)),
    "q1_hex_128": (O.HEX, 128, 1, 1, None, (2_097_152, 2_146_689, 57_066_625)),
}


def _sparsity(V):
    h = C.c_void_p()
    _lib.call("dfx_sparsity_create", V.handle, V.handle, C.byref(h))
    sp = SparsityPattern(h)
    return sp.row_ptr, sp.cols


@pytest.mark.parametrize("name", list(CONFIGS))
def test_setup_matches_oracle_at_config_size(name):
    ct, n, deg, bs, face, (ncells, nnodes, nnz) = CONFIGS[name]
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (n, n, n), mesh.CellType(ct))
    om = O.Mesh((n, n, n), ct)
    assert m.topology.cells.shape[0] == ncells
    assert np.array_equal(m.geometry.x, om.x)
    assert np.array_equal(m.topology.cells, om.cells)
    assert np.array_equal(m.topology.original_cell_index, om.original_cell_index)
    assert np.array_equal(m.geometry.input_global_indices, om.input_vertex_index)
    del om.x

    V = fem.functionspace(m, ("Lagrange", deg) if bs == 1 else ("Lagrange", deg, (bs,)))
    dm, nd = om.dofmap(deg)
    assert nd == nnodes and V.dofmap.index_map_size == nd
    assert np.array_equal(V.dofmap.list, dm)

    if face is None:
        got = fem.locate_boundary_dofs(V)
        want = om.boundary_nodes(deg, dm)
        assert len(want) == (n + 1) ** 3 - (n - 1) ** 3 if deg == 1 else len(want) > 0
    else:
        got = fem.locate_boundary_dofs(V, axis=face[0], value=face[1])
        nodes = om.boundary_nodes(deg, dm, 1, face[0], face[1])
        want = (bs * nodes[:, None] + np.arange(bs)).ravel()
    assert np.array_equal(got, want)

    prp, pcols = _sparsity(V)
    rp, cols = O.sparsity(dm, dm, nd)
    assert rp[-1] == nnz
    assert np.array_equal(prp, rp)
    assert np.array_equal(pcols, cols)
