"""GPU parity of the P2 vector elasticity (3x3-block CSR) and P3 forms
against the oracle: block pattern bit-exact, FP64 values within 1e-12
relative, FP32 within 5e-6, CG solutions within 1e-8."""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, la, problems, solvers

from conftest import rel_err

pytestmark = pytest.mark.gpu


def oracle_elasticity(p, n):
    m = O.Mesh(n, O.TET)
    dm, nn = m.dofmap(2)
    rp, cols = O.sparsity(dm, dm, nn)
    lam, mu = problems.lame_from_E(p.E)
    co = np.stack([lam, mu], 1)[m.original_cell_index]
    nodes = m.boundary_nodes(2, dm, 1, 0, 0.0)
    bdofs = (3 * nodes[:, None] + np.arange(3)).ravel()
    mark = np.zeros(3 * nn, np.int8)
    mark[bdofs] = 1
    A = O.assemble_matrix(O.K_P2TET_ELAST_A, m.x, m.cells, dm, 3, co, None, mark, mark, rp, cols)
    O.set_diagonal(A, bdofs, 3, 1.0, rp, cols)
    b = O.assemble_vector(O.K_P2TET_ELAST_L, m.x, m.cells, dm, 3, None, np.array([0, 0, -1.0]),
                          3 * nn)
    O.bc_set(b, bdofs, np.zeros(3 * nn))
    return rp, cols, A, b, bdofs


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 5e-6)])
def test_elasticity_matches_oracle(dtype, tol):
    n = (3, 2, 3)
    p = problems.elasticity_problem(n, seed=5, dtype=dtype)
    A, b = problems.assemble_system(p)
    rp, cols, Ar, br, bdofs = oracle_elasticity(p, n)
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    assert np.array_equal(np.sort(p.bc.dofs), bdofs)
    assert rel_err(A.data, Ar) <= tol
    assert np.max(np.abs(A.data - Ar)) <= tol * np.max(np.abs(Ar))
    assert rel_err(b.array, br) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_elasticity_sums_are_deterministic(dtype):
    """Every entry is summed in cell order by one lane: repeated assemblies
    are bitwise equal, a second assembly without zeroing doubles the
    values exactly (x + x), and zeroing then assembling stores the same
    bits again."""
    n = (6, 5, 4)
    p = problems.elasticity_problem(n, seed=11, dtype=dtype)
    A = fem.assemble_matrix(p.a, [p.bc])
    ref = A.data.copy()
    A2 = fem.assemble_matrix(p.a, [p.bc])
    assert np.array_equal(A2.data, ref)
    A.set_value(0.0)
    fem.assemble_matrix(p.a, [p.bc], 1.0, A)
    assert np.array_equal(A.data, ref)
    B = fem.assemble_matrix(p.a, [])
    plain = B.data.copy()
    fem.assemble_matrix(p.a, [], 1.0, B)          # add mode
    assert np.array_equal(B.data, plain + plain)


def test_elasticity_solve_matches_oracle():
    n = (4, 3, 3)
    p = problems.elasticity_problem(n, seed=6)
    A, b = problems.assemble_system(p)
    x = la.Vector(p.V.num_dofs)
    res = solvers.cg(A, b, x, rtol=1e-12, max_it=20000)
    assert res.converged
    rp, cols, Ar, br, _ = oracle_elasticity(p, n)
    xr, it = O.cg(rp, cols, Ar, 3, br, kmax=20000, rtol=1e-12)
    assert it > 0
    assert rel_err(x.array, xr) <= 1e-8
    # both block SpMV variants (1 row-block stream, 2 one thread per block
    # row component) against the oracle
    v = np.random.default_rng(0).standard_normal(p.V.num_dofs)
    yr = O.spmv(rp, cols, Ar, 3, v)
    for lanes in (1, 2):
        A.set_lanes(lanes)
        y = la.Vector(p.V.num_dofs)
        A.mult(la.Vector.from_array(v), y)
        assert rel_err(y.array, yr) <= 1e-13, lanes
        x2 = la.Vector(p.V.num_dofs)
        assert solvers.cg(A, b, x2, rtol=1e-12, max_it=20000).converged
        assert rel_err(x2.array, xr) <= 1e-8, lanes


def oracle_p3(p, n, c):
    m = O.Mesh(n, O.TET)
    dm, nd = m.dofmap(3)
    rp, cols = O.sparsity(dm, dm, nd)
    kap = p.kappa[m.original_cell_index].reshape(-1, 1)
    A = O.assemble_matrix(O.K_P3TET_A, m.x, m.cells, dm, 1, kap, np.array([c]), None, None, rp,
                          cols)
    b = O.assemble_vector(O.K_P3TET_L, m.x, m.cells, dm, 1, None, np.array([1.0]), nd)
    return rp, cols, A, b


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 5e-6)])
@pytest.mark.parametrize("variant", [3, 7, 8, 16, 17])
def test_p3_matches_oracle(dtype, tol, variant):
    """P3 variants: 3 one incidence per lane, 7 element matrices as one
    batched product on the matrix cores then gathered by rows, 8
    incidence-parallel gather from the cell factors, 16 a team of 20 lanes
    per row forming the entries from the cell factors, 17 such teams
    gathering the element matrices of variant 7."""
    n = (3, 2, 2)
    p = problems.p3_problem(n, seed=8, dtype=dtype)
    p.a.set_variant(variant)
    A, b = problems.assemble_system(p)
    rp, cols, Ar, br = oracle_p3(p, n, 1.0)
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    assert rel_err(A.data, Ar) <= tol
    assert np.max(np.abs(A.data - Ar)) <= tol * np.max(np.abs(Ar))
    assert rel_err(b.array, br) <= tol


@pytest.mark.parametrize("variant", [3, 7, 8, 16, 17])
def test_p3_solve_matches_oracle(variant):
    n = (4, 4, 3)
    p = problems.p3_problem(n, seed=9)
    p.a.set_variant(variant)
    A, b = problems.assemble_system(p)
    x = la.Vector(p.V.num_dofs)
    res = solvers.cg(A, b, x, rtol=1e-12, max_it=20000)
    assert res.converged
    rp, cols, Ar, br = oracle_p3(p, n, 1.0)
    xr, _ = O.cg(rp, cols, Ar, 1, br, kmax=20000, rtol=1e-12)
    assert rel_err(x.array, xr) <= 1e-8


def test_p3_lifting_nonzero_g():
    """Lifting with nonzero boundary data on the P3 space vs the oracle."""
    n = (2, 2, 2)
    p = problems.p3_problem(n, seed=10)
    dofs = fem.locate_boundary_dofs(p.V)
    g = np.zeros(p.V.num_dofs)
    g[dofs] = np.cos(np.arange(len(dofs)))
    bcg = fem.dirichletbc(p.V, dofs, g)
    b = fem.assemble_vector(p.L)
    fem.apply_lifting(b, [p.a], [[bcg]])
    fem.set_bc(b, [bcg])
    m = O.Mesh(n, O.TET)
    dm, nd = m.dofmap(3)
    kap = p.kappa[m.original_cell_index].reshape(-1, 1)
    br = O.assemble_vector(O.K_P3TET_L, m.x, m.cells, dm, 1, None, np.array([1.0]), nd)
    mark = np.zeros(nd, np.int8)
    mark[dofs] = 1
    O.lift_bc(br, O.K_P3TET_A, m.x, m.cells, dm, 1, kap, np.array([1.0]), g, mark)
    O.bc_set(br, dofs, g)
    assert rel_err(b.array, br) <= 1e-12


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 5e-6)])
def test_elasticity_lifting_nonzero_g_matches_oracle(dtype, tol):
    """apply_lifting of the P2 vector form (block size 3) with nonzero
    Dirichlet values on the clamped face and a nonzero x0, alpha != 1,
    then bc.set, against the oracle's lift_bc (assemble_vector_impl.h
    lift_bc over the cells with a Dirichlet column dof)."""
    n = (3, 3, 2)
    p = problems.elasticity_problem(n, seed=12, dtype=dtype)
    N = p.V.num_dofs
    dofs = np.asarray(p.bc.dofs)
    g = np.zeros(N)
    g[dofs] = 0.1 + np.sin(np.arange(len(dofs)))
    x0 = np.random.default_rng(13).standard_normal(N)
    alpha = -0.7
    bcg = fem.dirichletbc(p.V, dofs, g)
    b = fem.assemble_vector(p.L)
    x0v = la.Vector.from_array(x0.astype(dtype))
    fem.apply_lifting(b, [p.a], [[bcg]], x0=[x0v], alpha=alpha)
    fem.set_bc(b, [bcg], x0v, alpha)
    m = O.Mesh(n, O.TET)
    dm, nn = m.dofmap(2)
    lam, mu = problems.lame_from_E(p.E)
    co = np.stack([lam, mu], 1)[m.original_cell_index]
    mark = np.zeros(3 * nn, np.int8)
    mark[dofs] = 1
    br = O.assemble_vector(O.K_P2TET_ELAST_L, m.x, m.cells, dm, 3, None, np.array([0, 0, -1.0]),
                           3 * nn)
    x0r = x0.astype(dtype).astype(np.float64)
    O.lift_bc(br, O.K_P2TET_ELAST_A, m.x, m.cells, dm, 3, co, None, g, mark, x0r, alpha)
    O.bc_set(br, dofs, g, x0r, alpha)
    assert rel_err(b.array, br) <= tol
    # the lifting changed the rows next to the face (not only the BC rows)
    b0 = fem.assemble_vector(p.L)
    fem.set_bc(b0, [bcg], x0v, alpha)
    touched = np.flatnonzero(np.abs(b.array - b0.array) > 0)
    assert np.setdiff1d(touched, dofs).size > 0
