"""GPU parity of assembly from caller-supplied element tensors (the
reference's custom-kernel path, cpp/demo/custom_kernel/main.cpp:73-118)
against the oracle's cell-by-cell insertion: FP64 matrices and vectors bit
for bit (every entry is summed in cell order on both sides), FP32 within
5e-6; a pair missing from the pattern raises."""

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from dolfinx_amd import fem, la, mesh
from dolfinx_amd._lib import DfxError

from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = {
    # name: (cell type, degree, oracle matrix kernel or None for random tensors)
    "p1_tet": (O.TET, 1, O.K_P1TET_A),
    "q1_hex": (O.HEX, 1, O.K_Q1HEX_A),
    "p2_tet": (O.TET, 2, None),
    "p3_tet": (O.TET, 3, O.K_P3TET_A),
}


def setup(name, n, seed):
    ct, deg, kern = CASES[name]
    om = O.Mesh(n, ct)
    dm, nd = om.dofmap(deg)
    msh = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    V = fem.functionspace(msh, ("Lagrange", deg))
    assert np.array_equal(V.dofmap.list, dm)
    k = dm.shape[1]
    rng = np.random.default_rng(seed)
    if kern is None:
        E = rng.standard_normal((om.num_cells, k, k))
    else:
        kap = rng.uniform(0.5, 2.0, om.num_cells)
        E = np.stack([O.tabulate(kern, om.x[om.cells[c]], w=[kap[c]], c=[0.3], ndim=k * k)
                      for c in range(om.num_cells)]).reshape(-1, k, k)
    return om, dm, nd, V, E


@pytest.mark.parametrize("name", list(CASES))
def test_element_matrices_bitwise(name):
    om, dm, nd, V, E = setup(name, (3, 2, 2), 1)
    rp, cols = O.sparsity(dm, dm, nd)
    bn = om.boundary_nodes(CASES[name][1], dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    ref = O.add_element_matrices(dm, 1, E, mark, mark, rp, cols)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    a = fem.element_form(V, 2)
    bc = fem.dirichletbc(V, bn)
    A = fem.assemble_matrix_elements(a, E, [bc])
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    assert np.array_equal(A.data, ref)
    # device-resident tensors, store mode after zeroing, then add mode
    Ed = la.Vector.from_array(E.ravel())
    A.set_value(0.0)
    fem.assemble_matrix_elements(a, Ed, [bc], 1.0, A)
    assert np.array_equal(A.data, ref)
    B = fem.assemble_matrix_elements(a, Ed)
    plain = B.data.copy()
    assert np.array_equal(plain, O.add_element_matrices(dm, 1, E, None, None, rp, cols))
    fem.assemble_matrix_elements(a, Ed, [], 1.0, B)
    assert np.array_equal(B.data, plain + plain)


@pytest.mark.parametrize("name", ["p1_tet", "p3_tet"])
def test_element_matrices_fp32(name):
    om, dm, nd, V, E = setup(name, (2, 3, 2), 2)
    rp, cols = O.sparsity(dm, dm, nd)
    ref = O.add_element_matrices(dm, 1, E, None, None, rp, cols)
    a = fem.element_form(V, 2, np.float32)
    A = fem.assemble_matrix_elements(a, E.astype(np.float32))
    assert rel_err(A.data, ref) <= 5e-6


@pytest.mark.parametrize("name", ["p1_tet", "q1_hex", "p3_tet"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_element_vectors(name, dtype):
    om, dm, nd, V, _ = setup(name, (3, 3, 2), 3)
    e = np.random.default_rng(5).standard_normal((om.num_cells, dm.shape[1]))
    ref = O.add_element_vectors(dm, 1, e, nd)
    L = fem.element_form(V, 1, dtype)
    b = fem.assemble_vector_elements(L, e.astype(dtype))
    if dtype == np.float64:
        assert np.array_equal(b.array, ref)
    else:
        assert rel_err(b.array, ref) <= 5e-6
    # adds onto b (assemble_vector semantics)
    fem.assemble_vector_elements(L, e.astype(dtype), b)
    if dtype == np.float64:
        assert np.array_equal(b.array, O.add_element_vectors(dm, 1, e, nd, ref.copy()))


def test_element_lifting_p1_matches_oracle():
    """apply_lifting with the supplied element matrices equals the oracle's
    lift_bc, which recomputes them with the kernel (nonzero g, x0, alpha)."""
    n = (3, 3, 2)
    om = O.Mesh(n, O.TET)
    dm, nd = om.dofmap(1)
    msh = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    V = fem.functionspace(msh, ("Lagrange", 1))
    kap = np.random.default_rng(7).uniform(0.5, 2.0, om.num_cells)
    E = np.stack([O.tabulate(O.K_P1TET_A, om.x[om.cells[c]], w=[kap[c]], c=[0.4], ndim=16)
                  for c in range(om.num_cells)])
    bn = om.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    g = np.zeros(nd)
    g[bn] = 1.0 + np.cos(np.arange(len(bn)))
    x0 = np.random.default_rng(8).standard_normal(nd)
    b0 = np.random.default_rng(9).standard_normal(nd)
    ref = O.lift_bc(b0.copy(), O.K_P1TET_A, om.x, om.cells, dm, 1, kap.reshape(-1, 1),
                    np.array([0.4]), g, mark, x0=x0, alpha=0.7)
    a = fem.element_form(V, 2)
    b = la.Vector.from_array(b0)
    fem.apply_lifting_elements(b, a, E, [fem.dirichletbc(V, bn, g)],
                               la.Vector.from_array(x0), alpha=0.7)
    assert rel_err(b.array, ref) <= 1e-12


def test_element_lifting_dense_column_product():
    """Random P2 element matrices: lifting = -alpha A_full[:, bc] g with
    A_full the assembled matrix without BCs (brute force)."""
    om, dm, nd, V, E = setup("p2_tet", (2, 2, 2), 10)
    rp, cols = O.sparsity(dm, dm, nd)
    Afull = O.add_element_matrices(dm, 1, E, None, None, rp, cols)
    bn = om.boundary_nodes(2, dm)
    g = np.zeros(nd)
    g[bn] = np.sin(np.arange(len(bn)) + 1.0)
    D = sp.csr_matrix((Afull, cols, rp), shape=(nd, nd))
    b = la.Vector(nd)
    fem.apply_lifting_elements(b, fem.element_form(V, 2), E, [fem.dirichletbc(V, bn, g)],
                               alpha=1.5)
    assert rel_err(b.array, -1.5 * (D @ g)) <= 1e-12


def test_element_matrix_outside_pattern_raises():
    """A pair missing from the pattern raises like insert_csr
    (la/matrix_csr_impl.h:93-94)."""
    om, dm, nd, V, E = setup("p1_tet", (2, 2, 2), 4)
    diag = la.MatrixCSR(np.arange(nd + 1, dtype=np.int64), np.arange(nd, dtype=np.int32))
    with pytest.raises(DfxError, match="Entry not in sparsity"):
        fem.assemble_matrix_elements(fem.element_form(V, 2), E, [], 1.0, diag)


def test_element_forms_reject_other_entry_points():
    om, dm, nd, V, E = setup("q1_hex", (2, 2, 2), 6)
    a = fem.element_form(V, 2)
    with pytest.raises(DfxError, match="dfx_assemble_matrix_elements"):
        fem.assemble_matrix(a, [])
    with pytest.raises(ValueError, match="values expected"):
        fem.assemble_matrix_elements(a, E[:-1])


@pytest.mark.parametrize("bs", [1, 3])
def test_matrix_add_blocks_matches_cell_by_cell_insertion(bs):
    """MatrixCSR::add of host blocks (la.MatrixCSR.add, dfx_csr_add_blocks):
    square blocks on a dofmap equal the oracle's cell-by-cell insertion bit
    for bit (entries summed in block order); rectangular blocks with a
    repeated row add in call order, checked against a dense sum; a column
    outside the pattern raises."""
    n = (3, 3, 2)
    om = O.Mesh(n, O.TET)
    dm, nd = om.dofmap(2)
    rp, cols = O.sparsity(dm, dm, nd)
    k = dm.shape[1]
    rng = np.random.default_rng(11 + bs)
    E = rng.standard_normal((om.num_cells, k * bs, k * bs))
    A = la.MatrixCSR(rp, cols, bs=bs)
    A.add(dm, dm, E)
    ref = O.add_element_matrices(dm, bs, E, None, None, rp, cols)
    assert np.array_equal(A.data, ref)
    # rectangular blocks (2 rows x 3 columns of one cell), rows repeated
    B = la.MatrixCSR(rp, cols, bs=bs)
    rows = np.stack([dm[:, [0, 1]], dm[:, [1, 1]]], axis=1).reshape(-1, 2)
    cl = np.repeat(dm[:, [0, 2, 3]], 2, axis=0)
    V = rng.standard_normal((rows.shape[0], 2 * bs, 3 * bs))
    B.add(rows, cl, V)
    dense = np.zeros((nd * bs, nd * bs))
    for r, c, v in zip(rows, cl, V):
        for i in range(2):
            for j in range(3):
                dense[r[i] * bs:(r[i] + 1) * bs, c[j] * bs:(c[j] + 1) * bs] += \
                    v[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs]
    Bd = np.zeros_like(dense)
    vals = B.data.reshape(-1, bs, bs)
    for r in range(nd):
        for p in range(rp[r], rp[r + 1]):
            Bd[r * bs:(r + 1) * bs, cols[p] * bs:(cols[p] + 1) * bs] = vals[p]
    assert np.allclose(Bd, dense, rtol=1e-13, atol=1e-13)
    # a pair outside the pattern: the first and the last node never share a cell
    far = np.array([[0]]), np.array([[nd - 1]])
    assert not np.any(cols[rp[0]:rp[1]] == nd - 1)
    with pytest.raises(DfxError, match="Entry not in sparsity"):
        A.add(far[0], far[1], np.ones((1, bs, bs)))
