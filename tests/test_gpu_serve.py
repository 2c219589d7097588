"""Batched solves and the continuous-batching service: results match the
single-system CG and the oracle, and each request's solution is bitwise the
same alone and in the mix."""

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import la, problems, serve, solvers

from conftest import rel_err

pytestmark = pytest.mark.gpu


def test_batch_equals_single_and_alone():
    sizes = [(6, 6, 6), (9, 7, 8), (12, 12, 12)]
    systems = []
    for i, n in enumerate(sizes):
        p = problems.scalar_problem(n, O.TET, kappa="lognormal_0.5", c=3.0 * i, f="normal",
                                    seed=20 + i)
        A, b = problems.assemble_system(p)
        systems.append((p, A, b))
    # the batch of all three
    B = serve.BatchSolver(4, sum(p.V.num_dofs for p, _, _ in systems))
    xs = [la.Vector(p.V.num_dofs) for p, _, _ in systems]
    slots = [B.add(A, b, x, 1e-10) for (p, A, b), x in zip(systems, xs)]
    for _ in range(400):
        B.step(8, True)
        done, conv, its, rel = B.status(slots)
        if done.all():
            break
    assert conv.all()
    mix = [x.array for x in xs]
    for (p, A, b), xm, it in zip(systems, mix, its):
        # single-system CG (same recurrence)
        xv = la.Vector(p.V.num_dofs)
        res = solvers.cg(A, b, xv, rtol=1e-10)
        assert rel_err(xm, xv.array) <= 1e-9
        assert abs(res.iterations - it) <= 2
        # alone in its own batch: bitwise identical
        B1 = serve.BatchSolver(1, p.V.num_dofs)
        x1 = la.Vector(p.V.num_dofs)
        s1 = B1.add(A, b, x1, 1e-10)
        while not B1.status([s1])[0][0]:
            B1.step(8, False)
        assert np.array_equal(x1.array, xm)
        # oracle
        xr, _ = O.cg(A.indptr, A.indices, A.data, 1, b.array, rtol=1e-12)
        assert rel_err(xm, xr) <= 1e-8


def test_service_continuous_batching_small():
    svc = serve.SolveService(hbm_budget_bytes=64 << 20, max_slots=8, rtol=1e-8,
                             iters_per_step=8, sizes=range(6, 11))
    reqs = [serve.make_request(i, seed=7, n_range=(6, 10)) for i in range(24)]
    keep = [0, 5, 11]
    stats = svc.run(reqs, rate=2000.0, keep_solutions=keep)
    assert stats.completed == len(reqs)
    assert all(r.converged for r in reqs)
    assert stats.max_resident > 1  # requests overlapped on the device
    for r in reqs:
        if r.rid in keep:
            assert np.array_equal(svc.solve_alone(r), r.x)
    # one request against the oracle
    r = reqs[keep[0]]
    m = O.Mesh((r.n, r.n, r.n), O.TET)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    bn = m.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    kap = r.kappa[m.original_cell_index].reshape(-1, 1)
    A = O.assemble_matrix(O.K_P1TET_A, m.x, m.cells, dm, 1, kap, np.array([r.c]), mark, mark,
                          rp, cols)
    O.set_diagonal(A, bn, 1, 1.0, rp, cols)
    fv = r.f[m.input_vertex_index][m.cells]
    b = O.assemble_vector(O.K_P1TET_L, m.x, m.cells, dm, 1, fv, None, nd)
    O.bc_set(b, bn, np.zeros(nd))
    xr, _ = O.cg(rp, cols, A, 1, b, rtol=1e-12)
    assert rel_err(r.x, xr) <= 1e-7


def test_service_hbm_budget_binds():
    """A budget that holds about one request of the largest size: requests
    wait for memory (their latency includes the wait), resident bytes never
    exceed the budget, and results are bitwise those of the request alone."""
    sizes = range(6, 11)
    probe = serve.SolveService(hbm_budget_bytes=1 << 30, max_slots=8, sizes=sizes)
    biggest = max(t.footprint() for t in probe.templates.values())
    smallest = min(t.footprint() for t in probe.templates.values())
    del probe
    budget = biggest + smallest // 2
    svc = serve.SolveService(hbm_budget_bytes=budget, max_slots=8, rtol=1e-8, iters_per_step=8,
                             sizes=sizes)
    reqs = [serve.make_request(i, seed=31, n_range=(6, 10)) for i in range(20)]
    keep = [2, 9, 17]
    stats = svc.run(reqs, rate=5000.0, keep_solutions=keep)
    assert stats.completed == len(reqs) and all(r.converged for r in reqs)
    assert stats.max_bytes <= budget
    waits = [r.admitted - r.arrival for r in reqs]
    assert all(r.arrival <= r.admitted <= r.completed for r in reqs)
    assert max(waits) > 1e-3            # some requests waited for memory
    lat = np.array(stats.latencies)
    assert np.all(lat >= 0) and lat.max() >= max(waits)
    for r in reqs:
        if r.rid in keep:
            assert np.array_equal(svc.solve_alone(r), r.x)


def test_service_same_size_requests_share_the_batch():
    """Requests of one mesh size arriving together are all admitted (in
    waves of one batched assembly each) instead of one per step."""
    svc = serve.SolveService(hbm_budget_bytes=256 << 20, max_slots=8, rtol=1e-8,
                             iters_per_step=8, sizes=[8])
    reqs = [serve.make_request(i, seed=41, n_range=(8, 8)) for i in range(8)]
    stats = svc.run(reqs, rate=1e6, keep_solutions=[0, 7])
    assert stats.completed == 8 and all(r.converged for r in reqs)
    assert stats.max_resident >= 4
    for r in reqs:
        if r.rid in (0, 7):
            assert np.array_equal(svc.solve_alone(r), r.x)
