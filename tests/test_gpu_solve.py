"""GPU parity of SpMV, reductions and the CG solvers against the oracle.

Tolerances: SpMV FP64 within 1e-14 relative (summation order differs from
the row-serial reference); FP32 SpMV within 1e-6; solutions within 1e-8
relative L2 of the oracle's FP64 Jacobi-PCG solved to 1e-12.
"""

import numpy as np
import pytest
import scipy.sparse as sps

import oracle as O
from dolfinx_amd import la, problems, solvers
from dolfinx_amd._lib import DfxError

from conftest import rel_err

pytestmark = pytest.mark.gpu


def system(cell, n, dtype=np.float64, **kw):
    p = problems.scalar_problem(n, cell, dtype=dtype, seed=11, **kw)
    A, b = problems.assemble_system(p)
    return p, A, b


@pytest.mark.parametrize("cell,n", [(O.TET, (10, 9, 8)), (O.HEX, (24, 20, 16))])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
def test_spmv_matches_oracle(cell, n, dtype, tol, rng):
    p, A, b = system(cell, n, dtype, kappa="exp_normal")
    x = rng.standard_normal(p.V.num_dofs)
    y0 = rng.standard_normal(p.V.num_dofs)
    xv = la.Vector.from_array(x.astype(dtype))
    yv = la.Vector.from_array(y0.astype(dtype))
    A.mult(xv, yv)  # y += A x
    Av = A.data.astype(np.float64)
    ref = O.spmv(A.indptr, A.indices, Av, 1, x.astype(dtype).astype(np.float64),
                 y0.astype(dtype).astype(np.float64))
    assert rel_err(yv.array, ref) <= tol


def test_inner_product_and_axpy(rng):
    n = 1_000_003
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    va, vb = la.Vector.from_array(a), la.Vector.from_array(b)
    assert abs(la.inner_product(va, vb) - O.lib().orc_dot(n, O._p(a), O._p(b))) <= 1e-12 * n ** 0.5
    assert abs(la.squared_norm(va) - float(a @ a)) <= 1e-12 * float(a @ a)
    r = la.Vector(n)
    la.axpy(r, -2.5, va, vb)
    assert rel_err(r.array, -2.5 * a + b) <= 1e-15


@pytest.mark.parametrize("cell,n", [(O.TET, (12, 12, 12)), (O.HEX, (20, 20, 20))])
@pytest.mark.parametrize("jacobi", [True, False])
def test_cg_matches_oracle(cell, n, jacobi):
    p, A, b = system(cell, n, kappa="exp_normal", c=0.0, f="one")
    x = la.Vector(p.V.num_dofs)
    res = solvers.cg(A, b, x, rtol=1e-12, max_it=20000, jacobi=jacobi, check_every=16)
    assert res.converged and res.true_rel_residual <= 1e-11
    xr, it = O.cg(A.indptr, A.indices, A.data, 1, b.array, kmax=20000, rtol=1e-12, jacobi=jacobi)
    assert it > 0
    assert rel_err(x.array, xr) <= 1e-8
    # same recurrence: iteration counts agree up to round-off effects
    assert abs(res.iterations - it) <= max(2, it // 50)


def test_cg_raises_when_not_converged():
    """conftest.py cg_solver raises after maxit (python/test/conftest.py:200-202)."""
    p, A, b = system(O.TET, (8, 8, 8))
    x = la.Vector(p.V.num_dofs)
    with pytest.raises(DfxError, match="exceeded max iterations"):
        solvers.cg(A, b, x, rtol=1e-14, max_it=3)


@pytest.mark.parametrize("cell,n", [(O.TET, (12, 12, 12)), (O.HEX, (24, 24, 24))])
@pytest.mark.parametrize("restart", [False, True])
def test_mixed_refinement_matches_fp64(cell, n, restart):
    kw = dict(kappa="exp_normal", c=0.0, f="one")
    p64 = problems.scalar_problem(n, cell, dtype=np.float64, seed=3, **kw)
    p32 = problems.scalar_problem(n, cell, dtype=np.float32, seed=3, msh=p64.mesh, V=p64.V, **kw)
    A64, b = problems.assemble_system(p64)
    A32, _ = problems.assemble_system(p32)
    x = la.Vector(p64.V.num_dofs)
    res = solvers.mixed_refinement(A64, A32, b, x, rtol=1e-10, inner_rtol=1e-4, restart=restart)
    assert res.converged and res.true_rel_residual <= 1e-10
    xr, _ = O.cg(A64.indptr, A64.indices, A64.data, 1, b.array, kmax=50000, rtol=1e-13)
    assert rel_err(x.array, xr) <= 1e-8


def test_mixed_refinement_falls_back_to_fp64_when_it_stalls():
    """A refinement step that does not halve the FP64 residual (inner_rtol
    0.9 here) is a stall: with fallback the solve finishes in FP64
    Jacobi-PCG from the current x and meets rtol against the FP64 matrix;
    without it the solve raises."""
    kw = dict(kappa="exp_normal", c=0.0, f="one")
    n = (16, 14, 12)
    p64 = problems.scalar_problem(n, O.HEX, dtype=np.float64, seed=7, **kw)
    p32 = problems.scalar_problem(n, O.HEX, dtype=np.float32, seed=7, msh=p64.mesh, V=p64.V, **kw)
    A64, b = problems.assemble_system(p64)
    A32, _ = problems.assemble_system(p32)
    x = la.Vector(p64.V.num_dofs)
    res = solvers.mixed_refinement(A64, A32, b, x, rtol=1e-10, inner_rtol=0.9)
    assert res.converged and res.true_rel_residual <= 1e-10
    assert res.outer == 1 and res.fallback_iterations > 0
    xr, _ = O.cg(A64.indptr, A64.indices, A64.data, 1, b.array, kmax=50000, rtol=1e-13)
    assert rel_err(x.array, xr) <= 1e-8
    # the FP64 residual of the returned x, recomputed on the host
    Ah = sps.csr_matrix((A64.data, A64.indices, A64.indptr), shape=A64.shape)
    r = b.array - Ah @ x.array
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(b.array)
    # the budget running out is a stall too
    x2 = la.Vector(p64.V.num_dofs)
    res2 = solvers.mixed_refinement(A64, A32, b, x2, rtol=1e-10, inner_rtol=1e-4, max_outer=1)
    assert res2.converged and res2.outer == 1 and res2.fallback_iterations > 0
    x3 = la.Vector(p64.V.num_dofs)
    with pytest.raises(DfxError, match="did not converge"):
        solvers.mixed_refinement(A64, A32, b, x3, rtol=1e-10, inner_rtol=0.9, fallback=False)


def test_poisson_tet_32_end_to_end():
    """The CPU-runnable config: P1 Poisson on create_box 32^3 tets, f = 1,
    u = 0 on the boundary, Jacobi-CG to 1e-10."""
    p, A, b = system(O.TET, (32, 32, 32), kappa="one", c=0.0, f="one")
    assert A.nnz == 513_313 and len(p.bc.dofs) == 6_146 and p.V.num_dofs == 35_937
    x = la.Vector(p.V.num_dofs)
    res = solvers.cg(A, b, x, rtol=1e-10)
    assert res.converged
    xr, _ = O.cg(A.indptr, A.indices, A.data, 1, b.array, rtol=1e-12)
    assert rel_err(x.array, xr) <= 1e-8
    # maximum of the solution of -lap u = 1 on the unit cube (P1, h = 1/32)
    assert 0.0555 < x.array.max() < 0.0565


@pytest.mark.parametrize("cell,n,jacobi", [(O.TET, (12, 10, 9), True), (O.HEX, (16, 12, 10), False)])
def test_cg_graph_replay_equals_eager(cell, n, jacobi):
    """Batches of iterations replayed from a captured HIP graph run the same
    kernels in the same order as the eager launches: the iterates are
    bitwise equal, over several replays per solve (host polls and copies in
    between) and over repeated solves."""
    p, A, b = system(cell, n, kappa="exp_normal", c=0.3, f="normal")
    xe = la.Vector(p.V.num_dofs)
    re = solvers.cg(A, b, xe, rtol=1e-11, max_it=20000, jacobi=jacobi, check_every=8, graph=False,
                    classic=True)
    for _ in range(2):
        xg = la.Vector(p.V.num_dofs)
        rg = solvers.cg(A, b, xg, rtol=1e-11, max_it=20000, jacobi=jacobi, check_every=8,
                        graph=True, classic=True)
        assert rg.iterations == re.iterations and rg.iterations > 3 * 8
        assert np.array_equal(xg.array, xe.array)
        assert rg.true_rel_residual == re.true_rel_residual
    # max_it inside a replayed batch stops at max_it exactly
    xm = la.Vector(p.V.num_dofs)
    with pytest.raises(DfxError, match="exceeded max iterations"):
        solvers.cg(A, b, xm, rtol=1e-14, max_it=13, check_every=8, graph=True, classic=True)


@pytest.mark.parametrize("cell,n,jacobi", [(O.TET, (16, 14, 12), True), (O.HEX, (24, 20, 18), True),
                                           (O.TET, (10, 9, 8), False)])
@pytest.mark.parametrize("dtype,rtol,tol", [(np.float64, 1e-11, 1e-9), (np.float32, 1e-5, 2e-4)])
def test_cooperative_cg_matches_classic_and_oracle(cell, n, jacobi, dtype, rtol, tol):
    """The one-launch cooperative CG of small scalar systems (grid barriers
    between the phases) against the three-kernel iteration and the
    oracle's PCG, from x = 0 and from a nonzero initial guess; repeated
    solves are bitwise equal."""
    p = problems.scalar_problem(n, cell, kappa="exp_normal", c=0.3, f="normal", seed=9,
                                dtype=dtype)
    A, b = problems.assemble_system(p)
    N = p.V.num_dofs
    assert N <= 65536
    xc = la.Vector(N, dtype)
    rc = solvers.cg(A, b, xc, rtol=rtol, max_it=20000, jacobi=jacobi, classic=True)
    xs = []
    for _ in range(2):
        x = la.Vector(N, dtype)
        r = solvers.cg(A, b, x, rtol=rtol, max_it=20000, jacobi=jacobi)
        assert r.cooperative == 1 and rc.cooperative == 0
        assert r.converged and abs(r.iterations - rc.iterations) <= 2
        xs.append(x.array.copy())
    assert np.array_equal(xs[0], xs[1])
    assert rel_err(xs[0], xc.array) <= tol
    if dtype == np.float64:
        xr, it = O.cg(A.indptr, A.indices, A.data, 1, b.array, kmax=20000, rtol=1e-12,
                      jacobi=jacobi)
        assert rel_err(xs[0], xr) <= 1e-8
    # a nonzero initial guess
    x0 = np.random.default_rng(3).standard_normal(N).astype(dtype)
    xa, xb = la.Vector.from_array(x0.copy()), la.Vector.from_array(x0.copy())
    ra = solvers.cg(A, b, xa, rtol=rtol, max_it=20000, jacobi=jacobi)
    rb = solvers.cg(A, b, xb, rtol=rtol, max_it=20000, jacobi=jacobi, classic=True)
    assert abs(ra.iterations - rb.iterations) <= 2
    # from a random guess the initial residual is large: FP32 iterates
    # stop at rtol 1e-5 of it, so the two solutions differ more
    assert rel_err(xa.array, xb.array) <= (tol if dtype == np.float64 else 10 * tol)
    assert ra.rnorm0 == pytest.approx(rb.rnorm0, rel=1e-6)
    # max_it reached inside the launch
    with pytest.raises(DfxError, match="exceeded max iterations"):
        solvers.cg(A, b, la.Vector(N, dtype), rtol=1e-30, max_it=7, jacobi=jacobi)


@pytest.mark.parametrize("restart", [True, False])
def test_mixed_graph_replay_equals_eager(restart):
    kw = dict(kappa="exp_normal", c=0.0, f="one")
    n = (14, 12, 10)
    p64 = problems.scalar_problem(n, O.HEX, dtype=np.float64, seed=5, **kw)
    p32 = problems.scalar_problem(n, O.HEX, dtype=np.float32, seed=5, msh=p64.mesh, V=p64.V, **kw)
    A64, b = problems.assemble_system(p64)
    A32, _ = problems.assemble_system(p32)
    out = []
    for graph in (False, True):
        x = la.Vector(p64.V.num_dofs)
        r = solvers.mixed_refinement(A64, A32, b, x, rtol=1e-10, inner_rtol=1e-4, check_every=8,
                                     restart=restart, graph=graph)
        assert r.converged
        out.append((x.array, r.outer, r.inner_total))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1:] == out[1][1:]


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
def test_sliced_ell_snapshot_spmv(dtype, tol, rng):
    """lanes = -1: the SpMV through the solvers' sliced-ELL copy, taken when
    the mode is set, against the oracle."""
    p, A, b = system(O.HEX, (20, 18, 16), dtype, kappa="exp_normal")
    A.set_lanes(-1)
    x = rng.standard_normal(p.V.num_dofs)
    xv = la.Vector.from_array(x.astype(dtype))
    yv = la.Vector(p.V.num_dofs, dtype)
    A.mult(xv, yv)
    ref = O.spmv(A.indptr, A.indices, A.data.astype(np.float64), 1,
                 x.astype(dtype).astype(np.float64))
    assert rel_err(yv.array, ref) <= tol
    # the snapshot is what the solvers use: CG gives the same iterates
    xa = la.Vector(p.V.num_dofs, dtype)
    solvers.cg(A, b, xa, rtol=1e-5 if dtype == np.float32 else 1e-10)
    A.set_lanes(1)
    xb = la.Vector(p.V.num_dofs, dtype)
    solvers.cg(A, b, xb, rtol=1e-5 if dtype == np.float32 else 1e-10)
    assert np.array_equal(xa.array, xb.array)
