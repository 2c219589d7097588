"""GPU parity of assembly (P1 tetrahedra, Q1 hexahedra) against the oracle.

Tolerances: FP64 values within 1e-12 relative (norm-wise and entry-wise
against max |A|); FP32 within 2e-6 relative (a few FP32 ulps of the
element computation plus atomics reassociation).
"""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import _lib, fem, la, mesh, problems
from dolfinx_amd._lib import DfxError

from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [(O.TET, (3, 4, 2)), (O.HEX, (4, 3, 5)), (O.TET, (12, 12, 12)), (O.HEX, (20, 20, 20))]
TOL = {np.float64: 1e-12, np.float32: 2e-6}


def oracle_system(p, n, with_bc=True, g=None):
    """Assemble the same problem with the oracle (on the oracle's own mesh)."""
    om = O.Mesh(n, int(p.mesh.topology.cell_type))
    dm, nd = om.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    kap = p.kappa[om.original_cell_index].reshape(-1, 1)
    fv = p.f[om.input_vertex_index][om.cells]
    kA = O.K_P1TET_A if om.celltype == O.TET else O.K_Q1HEX_A
    kL = O.K_P1TET_L if om.celltype == O.TET else O.K_Q1HEX_L
    bnodes = om.boundary_nodes(1, dm) if with_bc else np.zeros(0, np.int32)
    mark = np.zeros(nd, np.int8)
    mark[bnodes] = 1
    A = O.assemble_matrix(kA, om.x, om.cells, dm, 1, kap, np.array([p.c]), mark, mark, rp, cols)
    O.set_diagonal(A, bnodes, 1, 1.0, rp, cols)
    b = O.assemble_vector(kL, om.x, om.cells, dm, 1, fv, None, nd)
    gg = np.zeros(nd) if g is None else g
    O.lift_bc(b, kA, om.x, om.cells, dm, 1, kap, np.array([p.c]), gg, mark)
    O.bc_set(b, bnodes, gg)
    return dict(mesh=om, rp=rp, cols=cols, A=A, b=b, bnodes=bnodes, nd=nd)


def build(cell, n, dtype, kappa="exp_normal", c=1.5, f="normal"):
    return problems.scalar_problem(n, cell, kappa=kappa, c=c, f=f, seed=7, dtype=dtype)


@pytest.mark.parametrize("cell,n", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_assemble_system_matches_oracle(cell, n, dtype):
    p = build(cell, n, dtype)
    A, b = problems.assemble_system(p)
    ref = oracle_system(p, n)
    rp, ci = A.pattern()
    assert np.array_equal(rp, ref["rp"]) and np.array_equal(ci, ref["cols"])
    tol = TOL[dtype]
    Av = A.data.astype(np.float64)
    assert rel_err(Av, ref["A"]) <= tol
    assert np.max(np.abs(Av - ref["A"])) <= tol * np.max(np.abs(ref["A"]))
    assert rel_err(b.array, ref["b"]) <= tol


@pytest.mark.parametrize("cell,n", CASES)
@pytest.mark.parametrize("variant", list(range(18)))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_every_row_kernel_variant_matches_oracle(cell, n, variant, dtype):
    """Variants 0-3 use the reference tensors, 4-6 closed-form entries for
    Q1 hexahedra (eight incidences in flight for tetrahedra), 7 element
    matrices on the matrix cores gathered by rows, 8 incidence-parallel
    gather, 9 structured-box Q1 rows in registers (hexahedra only), 10 rows
    forming the cell factors from the geometry (no cell pass), 12-14
    Q1 rows with their plan at fixed positions, 15 structured rows in
    registers through an LDS slice (hexahedra only), 16 a team of lanes per
    row forming the entries from the cell records, 17 such teams gathering
    the element matrices of the matrix-core product."""
    p = build(cell, n, dtype)
    p.a.set_variant(variant)
    if variant in (9, 12, 13, 14, 15) and cell == O.TET:
        with pytest.raises(DfxError, match="not available"):
            problems.assemble_system(p)
        return
    A, b = problems.assemble_system(p)
    ref = oracle_system(p, n)
    tol = TOL[dtype]
    Av = A.data.astype(np.float64)
    assert rel_err(Av, ref["A"]) <= tol
    assert np.max(np.abs(Av - ref["A"])) <= tol * np.max(np.abs(ref["A"]))


def test_autotune_keeps_results():
    p = build(O.HEX, (20, 20, 20), np.float64)
    A = fem.create_matrix(p.a)
    variant, ms = p.a.autotune(A, p.bc)
    assert 0 <= variant <= 17 and ms > 0
    assert np.all(A.data == 0)
    fem.assemble_matrix(p.a, [p.bc], 1.0, A)
    ref = oracle_system(p, (20, 20, 20))
    assert rel_err(A.data, ref["A"]) <= 1e-12


@pytest.mark.parametrize("cell,n", CASES[:2])
def test_lifting_nonzero_g_matches_oracle(cell, n):
    """Manufactured nonzero Dirichlet data pins apply_lifting + set_bc."""
    p = build(cell, n, np.float64)
    x = p.mesh.geometry.x
    g = 1.0 + x[:, 0] ** 2 + 0.5 * x[:, 1] * x[:, 2]
    bcg = fem.dirichletbc(p.V, p.bc.dofs, g)
    b = fem.assemble_vector(p.L)
    fem.apply_lifting(b, [p.a], [[bcg]])
    fem.set_bc(b, [bcg])
    ref = oracle_system(p, n, g=g)
    assert rel_err(b.array, ref["b"]) <= 1e-12


@pytest.mark.parametrize("cell,n", CASES + [(O.HEX, (104, 104, 100))])
def test_fused_system_equals_separate_calls(cell, n):
    """assemble_system gives the bits of the separate calls, with nonzero
    Dirichlet data, repeated into the same A and b (the stiffness and source
    forms share their mesh, so one cell pass serves both); the
    1,081,600-cell box takes the path with the vector assembly on a second
    stream."""
    p = build(cell, n, np.float64)
    x = p.mesh.geometry.x
    g = 1.0 + x[:, 0] ** 2 + 0.5 * x[:, 1] * x[:, 2]
    bcg = fem.dirichletbc(p.V, p.bc.dofs, g)
    A1 = fem.assemble_matrix(p.a, [bcg])
    b1 = fem.assemble_vector(p.L)
    fem.apply_lifting(b1, [p.a], [[bcg]])
    fem.set_bc(b1, [bcg])
    A2, b2 = fem.assemble_system(p.a, p.L, [bcg])
    assert np.array_equal(A2.data, A1.data) and np.array_equal(b2.array, b1.array)
    for _ in range(2):
        fem.assemble_system(p.a, p.L, [bcg], 1.0, A2, b2)
        assert np.array_equal(A2.data, A1.data) and np.array_equal(b2.array, b1.array)


@pytest.mark.parametrize("cell", [O.TET, O.HEX])
def test_system_with_forms_on_different_geometry(cell):
    """A source form on another mesh of the same size (same cells, moved
    vertices) does not share the cell pass: assemble_system still gives the
    separate calls' bits, before and after a system call with matching
    forms."""
    n = (6, 5, 4)
    p = build(cell, n, np.float64)
    ct = mesh.CellType.tetrahedron if cell == O.TET else mesh.CellType.hexahedron
    m2 = mesh.create_box([[0.0, 0.0, 0.0], [2.0, 1.5, 1.0]], n, ct)
    V2 = fem.functionspace(m2, ("Lagrange", 1))
    nc = m2.topology.cells.shape[0]
    nv = 4 if cell == O.TET else 8
    f2 = np.random.default_rng(3).standard_normal((nc, nv))
    L2 = fem.form("p1_source" if cell == O.TET else "q1_source", V2, f2, None)
    A1 = fem.assemble_matrix(p.a, [p.bc])
    b1 = fem.assemble_vector(L2)
    fem.apply_lifting(b1, [p.a], [[p.bc]])
    fem.set_bc(b1, [p.bc])
    A2, b2 = fem.assemble_system(p.a, L2, [p.bc])
    assert np.array_equal(A2.data, A1.data) and np.array_equal(b2.array, b1.array)
    fem.assemble_system(p.a, p.L, [p.bc], 1.0, A2, b2)
    fem.assemble_system(p.a, L2, [p.bc], 1.0, A2, b2)
    assert np.array_equal(A2.data, A1.data) and np.array_equal(b2.array, b1.array)


def test_entry_not_in_sparsity_raises():
    """A pattern missing entries must raise, like insert_csr
    (la/matrix_csr_impl.h:93-94)."""
    p = build(O.TET, (2, 2, 2), np.float64)
    sp = fem.create_sparsity_pattern(p.a)
    n = sp.num_rows
    diag = la.MatrixCSR(np.arange(n + 1, dtype=np.int64), np.arange(n, dtype=np.int32))
    with pytest.raises(DfxError, match="Entry not in sparsity"):
        fem.assemble_matrix(p.a, [], 1.0, diag)


def test_reassembly_is_linear_in_kappa():
    """Assembling with 2*kappa doubles the operator (test_assembler.py
    re-assembly linearity)."""
    p1 = build(O.HEX, (6, 5, 4), np.float64, c=0.0)
    p2 = problems.scalar_problem((6, 5, 4), O.HEX, kappa=2 * p1.kappa, c=0.0, f=p1.f,
                                 msh=p1.mesh, V=p1.V)
    A1 = fem.assemble_matrix(p1.a, [])
    A2 = fem.assemble_matrix(p2.a, [])
    assert rel_err(A2.data, 2 * A1.data) <= 1e-14


def test_stiffness_annihilates_constants():
    """A 1 = 0 for the pure stiffness without BCs (cpp/test/matrix.cpp:116-119)."""
    for cell, n in CASES[:2]:
        p = build(cell, n, np.float64, c=0.0)
        A = fem.assemble_matrix(p.a, [])
        y = la.Vector(p.V.num_dofs)
        A.mult(la.Vector.from_array(np.ones(p.V.num_dofs)), y)
        assert np.max(np.abs(y.array)) <= 1e-13 * np.max(np.abs(A.data))


@pytest.mark.parametrize("cell,n", CASES[:2])
def test_functionals_of_the_source_form(cell, n):
    """sum_i b_i = int f dx: 1 for f = 1 and 1/2 for f = x_0 on the unit cube
    (python/test/unit/fem/test_assembler.py:67-78)."""
    p = build(cell, n, np.float64, f="one")
    b = fem.assemble_vector(p.L)
    assert abs(b.array.sum() - 1.0) <= 1e-13
    x0 = np.zeros(p.mesh.geometry.x.shape[0])
    x0[p.mesh.geometry.input_global_indices] = p.mesh.geometry.x[:, 0]
    p2 = problems.scalar_problem(n, cell, f=x0, msh=p.mesh, V=p.V)
    b2 = fem.assemble_vector(p2.L)
    assert abs(b2.array.sum() - 0.5) <= 1e-13


def test_zeroing_and_accumulation_semantics():
    """MatrixCSR::set(0) then assemble_matrix adds into zeros; a second
    assembly without zeroing accumulates; zeroed values read back as zeros
    through every reader (values, SpMV, set_diagonal)."""
    p = build(O.HEX, (5, 4, 3), np.float64)
    A = fem.assemble_matrix(p.a, [])
    ref = A.data.copy()
    fem.assemble_matrix(p.a, [], 1.0, A)            # add mode: 2 A
    assert rel_err(A.data, 2 * ref) <= 1e-15
    A.set_value(0.0)
    assert not np.any(A.data)
    A.set_value(0.0)
    fem.assemble_matrix(p.a, [], 1.0, A)            # store mode after zeroing
    assert np.array_equal(A.data, ref)
    A.set_value(0.0)
    y = la.Vector.from_array(np.ones(p.V.num_dofs))
    A.mult(la.Vector.from_array(np.ones(p.V.num_dofs)), y)
    assert np.all(y.array == 1.0)
    A.set_value(0.0)
    _lib.call("dfx_set_diagonal", A.handle, p.bc.handle, 2.0)
    d = A.to_scipy().diagonal()
    expect = np.zeros(p.V.num_dofs)
    expect[p.bc.dofs] = 2.0
    assert np.array_equal(d, expect) and A.to_scipy().nnz == A.nnz
    S = la.MatrixCSR.shared(A, np.float64)
    assert not np.any(S.data)


@pytest.mark.parametrize("cell,n", CASES[:2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_packing_equals_host_packing(cell, n, dtype):
    """pack_coefficients on the device (values per input cell / vertex)
    gives the same assembled A and b as packing on the host."""
    p = build(cell, n, dtype)
    A1, b1 = problems.assemble_system(p)
    rng = np.random.default_rng(5)
    kap = np.exp(rng.standard_normal(p.kappa.shape))
    f = rng.standard_normal(p.f.shape)
    p.a.pack(p.mesh, kap, "cell")
    p.L.pack(p.mesh, f, "vertex")
    A2, b2 = problems.assemble_system(p)
    q = problems.scalar_problem(n, cell, kappa=kap, c=p.c, f=f, msh=p.mesh, V=p.V, dtype=dtype)
    A3, b3 = problems.assemble_system(q)
    assert np.array_equal(A2.data, A3.data) and np.array_equal(b2.array, b3.array)
    assert not np.array_equal(A1.data, A2.data)
    with pytest.raises(DfxError, match="wrong number of values"):
        p.a.pack(p.mesh, kap[:-1], "cell")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batched_assembly_equals_single(dtype):
    """One batched call over meshes of uneven size (tetrahedra; and
    hexahedra in a second batch) gives the matrices of the single calls."""
    for cell, sizes in [(O.TET, [(3, 4, 2), (6, 5, 4), (2, 2, 2), (8, 7, 6)]),
                        (O.HEX, [(4, 3, 5), (7, 6, 5), (2, 3, 2)])]:
        probs = [problems.scalar_problem(n, cell, kappa="exp_normal", c=1.5, f="normal",
                                         seed=i, dtype=dtype) for i, n in enumerate(sizes)]
        if cell == O.HEX:
            for p in probs:   # the batched kernel forms Q1 entries in closed form
                p.a.set_variant(4)
        single = [fem.assemble_matrix(p.a, [p.bc]).data for p in probs]
        mats = [fem.create_matrix(p.a) for p in probs]
        bcs = [p.bc if i % 2 == 0 else None for i, p in enumerate(probs)]
        fem.assemble_matrix_batched([p.a for p in probs], bcs, mats)
        for i, (A, ref, p) in enumerate(zip(mats, single, probs)):
            if bcs[i] is None:
                ref = fem.assemble_matrix(p.a, []).data
            assert np.array_equal(A.data, ref), (cell, i)
        # a second call accumulates
        fem.assemble_matrix_batched([p.a for p in probs], [None] * len(probs), mats, 1.0)
        for i, (A, p) in enumerate(zip(mats, probs)):
            base = fem.assemble_matrix(p.a, [p.bc] if bcs[i] is not None else []).data
            add = fem.assemble_matrix(p.a, []).data
            assert np.allclose(A.data, base + add, rtol=1e-6 if dtype == np.float32 else 1e-14)


@pytest.mark.parametrize("cell", [O.TET, O.HEX])
def test_single_cell_box_and_zero_rhs(cell):
    """1x1x1 boxes (every dof on the boundary, fewer rows than a workgroup):
    the system matches the oracle, A is the identity and CG returns x = 0 at
    once for b = 0."""
    from dolfinx_amd import solvers

    p = build(cell, (1, 1, 1), np.float64, f="one")
    A, b = problems.assemble_system(p)
    ref = oracle_system(p, (1, 1, 1))
    assert np.array_equal(A.data, ref["A"]) or rel_err(A.data, ref["A"]) <= 1e-12
    assert np.allclose(A.to_dense(), np.eye(p.V.num_dofs))
    assert not np.any(b.array)
    x = la.Vector(p.V.num_dofs)
    res = solvers.cg(A, b, x, rtol=1e-10)
    assert res.converged and res.iterations == 0 and not np.any(x.array)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("variant", [9, 12, 15])
def test_structured_q1_variants_on_shuffled_and_l_shaped_input(dtype, variant):
    """The structured Q1 row variants (corner cells / fixed-width plans) on
    a box given in a shuffled input order (other cell and vertex numbers)
    and on an L-shaped block of hexahedra (rows with absent corners),
    against the oracle."""
    from dolfinx_amd import mesh as dmesh
    from test_mesh_create import box_input

    om2 = O.Mesh((2, 2, 1), O.HEX)
    cin = np.empty_like(om2.cells, dtype=np.int64)
    cin[om2.original_cell_index] = om2.input_vertex_index[om2.cells]
    xin = np.empty_like(om2.x)
    xin[om2.input_vertex_index] = om2.x
    cl = cin[:3]
    used = np.unique(cl)
    relabel = np.full(len(xin), -1)
    relabel[used] = np.arange(len(used))
    for cells, x in (box_input((6, 5, 7), O.HEX, 3), (relabel[cl], xin[used])):
        m = dmesh.create_mesh(dmesh.CellType.hexahedron, cells, x)
        nc = m.topology.cells.shape[0]
        kappa = np.exp(np.random.default_rng(9).standard_normal(nc))
        p = problems.scalar_problem(None, O.HEX, kappa=kappa, c=0.4, f="normal", seed=9,
                                    dtype=dtype, msh=m)
        p.a.set_variant(variant)
        A, b = problems.assemble_system(p)
        om = O.Mesh.from_cells(O.HEX, cells, x)
        dm, nd = om.dofmap(1)
        rp, cols = O.sparsity(dm, dm, nd)
        bn = om.boundary_nodes(1, dm)
        mark = np.zeros(nd, np.int8)
        mark[bn] = 1
        kap = kappa[om.original_cell_index].reshape(-1, 1)
        Ar = O.assemble_matrix(O.K_Q1HEX_A, om.x, om.cells, dm, 1, kap, np.array([0.4]), mark,
                               mark, rp, cols)
        O.set_diagonal(Ar, bn, 1, 1.0, rp, cols)
        assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
        assert rel_err(A.data.astype(np.float64), Ar) <= TOL[dtype]
