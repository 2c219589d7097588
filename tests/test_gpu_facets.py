"""Exterior facet loads int_Gamma g v ds (fem::assemble_vector over an
exterior_facet integral, fem/assemble_vector_impl.h:118-174) on P1
tetrahedra and Q1 hexahedra against the oracle, on all exterior facets and
on one face; g is a P1/Q1 field given at the vertices."""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, la, mesh

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
