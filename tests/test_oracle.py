"""Pins the CPU oracle (oracle/) against checks that do not restate its code:
closed forms of create_box, brute-force facet/edge matching and an
independent Python restatement of the RCM rules on tiny meshes, exact
polynomial integrals for the element kernels, rigid-body modes for
elasticity, and the reference tests' known answers (A 1 = 0, total mass,
manufactured solutions). CPU only.
"""

from itertools import combinations
from math import factorial

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle as O

TET_EDGES = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]
TET_FACETS = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
HEX_FACETS = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6), (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)]


# ---------------------------------------------------------------- mesh
@pytest.mark.parametrize("ct,n", [(O.TET, (2, 3, 4)), (O.HEX, (3, 2, 5)), (O.TET, (5, 5, 5))])
def test_box_closed_form(ct, n):
    m = O.Mesh(n, ct)
    nv = np.prod(np.array(n) + 1)
    assert m.num_vertices == nv
    assert m.num_cells == (6 if ct == O.TET else 1) * np.prod(n)
    # geometry is the lexicographic grid, permuted by the vertex numbering
    v = m.input_vertex_index
    idx = np.stack([v % (n[0] + 1), (v // (n[0] + 1)) % (n[1] + 1), v // ((n[0] + 1) * (n[1] + 1))], 1)
    assert np.array_equal(m.x, idx * (1.0 / np.array(n, float)))
    assert np.array_equal(np.sort(v), np.arange(nv))
    assert np.array_equal(np.sort(m.original_cell_index), np.arange(m.num_cells))


def test_tet_volumes_and_orientation():
    n = (4, 3, 5)
    m = O.Mesh(n, O.TET)
    X = m.x[m.cells]
    det = np.linalg.det(X[:, 1:, :] - X[:, :1, :])
    h3 = 1.0 / np.prod(n)
    assert np.allclose(np.abs(det), h3, rtol=1e-13)
    assert np.isclose(np.abs(det).sum() / 6, 1.0, rtol=1e-13)


def test_vertex_numbering_is_first_appearance():
    for ct in (O.TET, O.HEX):
        m = O.Mesh((4, 4, 3), ct)
        seen, nxt = set(), 0
        for v in m.cells.ravel():
            if v not in seen:
                assert v == nxt
                seen.add(v)
                nxt += 1


def brute_dual_graph(ct, cells):
    """Cells sharing a facet; each cell's links ordered by the shared
    facet's sorted vertex key."""
    facets = TET_FACETS if ct == O.TET else HEX_FACETS
    owner = {}
    for c, cv in enumerate(cells):
        for f in facets:
            owner.setdefault(tuple(sorted(cv[list(f)])), []).append(c)
    links = [[] for _ in cells]
    for key in sorted(owner):
        cs = owner[key]
        for a, b in combinations(cs, 2):
            links[a].append((key, b))
            links[b].append((key, a))
    return [[b for _, b in sorted(lk)] for lk in links]


@pytest.mark.parametrize("ct,n", [(O.TET, (2, 2, 2)), (O.HEX, (3, 3, 2)), (O.TET, (3, 2, 1))])
def test_dual_graph_against_brute_force(ct, n):
    cells = np.zeros((int(np.prod(n)) * (6 if ct == O.TET else 1), 4 if ct == O.TET else 8), np.int64)
    O.lib().orc_box_cells(O._p(np.array(n, np.int64)), ct, O._p(cells))
    off, data = O.dual_graph(ct, cells)
    ref = brute_dual_graph(ct, cells)
    for c in range(len(cells)):
        assert list(data[off[c]:off[c + 1]]) == ref[c]


def py_rcm(off, adj):
    """Reverse Cuthill-McKee as graph/ordering.cpp describes it: root = first
    node of minimum degree, George-Liu pseudo-peripheral search (first
    minimum-degree node of the last level while the depth grows), each
    node's unvisited neighbours appended stably by degree, then reversed."""
    n = len(off) - 1
    deg = np.diff(off)
    nb = [list(adj[off[i]:off[i + 1]]) for i in range(n)]

    def levels(s):
        lev, seen, cur = [], {s}, [s]
        while cur:
            lev.append(cur)
            nxt = []
            for u in cur:
                for w in nb[u]:
                    if w not in seen:
                        seen.add(w)
                        nxt.append(w)
            cur = nxt
        return lev

    remap = -np.ones(n, int)
    count = 0
    while count < n:
        free = [i for i in range(n) if remap[i] == -1]
        v = min(free, key=lambda i: (deg[i], i))
        lv = levels(v)
        while True:
            s = min(lv[-1], key=lambda i: deg[i])  # first minimum (stable)
            l2 = levels(s)
            if len(l2) > len(lv):
                v, lv = s, l2
            else:
                break
        order, seen = [v], {v}
        i = 0
        while i < len(order):
            new = [w for w in nb[order[i]] if w not in seen]
            seen.update(new)
            order += sorted(new, key=lambda w: deg[w])  # Python sort is stable
            i += 1
        for q in reversed(order):
            remap[q] = count
            count += 1
    return remap


@pytest.mark.parametrize("ct,n", [(O.TET, (2, 2, 2)), (O.HEX, (4, 3, 3)), (O.TET, (3, 3, 2))])
def test_rcm_against_python_restatement(ct, n):
    cells = np.zeros((int(np.prod(n)) * (6 if ct == O.TET else 1), 4 if ct == O.TET else 8), np.int64)
    O.lib().orc_box_cells(O._p(np.array(n, np.int64)), ct, O._p(cells))
    off, data = O.dual_graph(ct, cells)
    r = O.rcm(off, data)
    assert np.array_equal(np.sort(r), np.arange(len(off) - 1))
    assert np.array_equal(r, py_rcm(off, data))


def test_rcm_disconnected_random_graphs():
    rng = np.random.default_rng(4)
    for trial in range(20):
        n = int(rng.integers(3, 40))
        edges = {tuple(sorted(rng.choice(n, 2, replace=False))) for _ in range(n)}
        nb = [[] for _ in range(n)]
        for a, b in sorted(edges):
            nb[a].append(b)
            nb[b].append(a)
        off = np.zeros(n + 1, np.int32)
        off[1:] = np.cumsum([len(x) for x in nb])
        adj = np.array([w for x in nb for w in x], np.int32)
        assert np.array_equal(O.rcm(off, adj), py_rcm(off, adj)), trial


# ---------------------------------------------------------------- entities, dofmaps
def test_edges_are_lexicographic_ranks():
    m = O.Mesh((3, 2, 2), O.TET)
    c2e, e2v = m.entities(1)
    keys = sorted({tuple(sorted(m.cells[c][list(e)])) for c in range(m.num_cells) for e in TET_EDGES})
    assert [tuple(k) for k in e2v] == keys
    for c in range(m.num_cells):
        for e, (a, b) in enumerate(TET_EDGES):
            assert tuple(e2v[c2e[c, e]]) == tuple(sorted((m.cells[c, a], m.cells[c, b])))
    # Euler characteristic of a ball: V - E + F - C = 1
    _, f2v = m.entities(2)
    assert m.num_vertices - len(e2v) + len(f2v) - m.num_cells == 1


def py_dofmap(m, degree):
    """Independent restatement: entity numbering, first-traversal renumbering,
    P3 reflected-edge swap."""
    nv = m.num_vertices
    c2e, e2v = m.entities(1)
    c2f, f2v = m.entities(2)
    ede = degree - 1
    basic = []
    for c in range(m.num_cells):
        d = list(m.cells[c])
        for e in range(6):
            d += [nv + ede * c2e[c, e] + j for j in range(ede)]
        if degree == 3:
            d += [nv + ede * len(e2v) + c2f[c, f] for f in range(4)]
        basic.append(d)
    o2n = {}
    for d in basic:
        for x in d:
            o2n.setdefault(x, len(o2n))
    dm = np.array([[o2n[x] for x in d] for d in basic])
    if degree == 3:
        for c in range(m.num_cells):
            for e, (a, b) in enumerate(TET_EDGES):
                if m.cells[c, a] > m.cells[c, b]:
                    dm[c, [4 + 2 * e, 5 + 2 * e]] = dm[c, [5 + 2 * e, 4 + 2 * e]]
    return dm


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_dofmap_against_restatement(degree):
    m = O.Mesh((2, 3, 2), O.TET)
    dm, n = m.dofmap(degree)
    assert np.array_equal(dm, py_dofmap(m, degree))
    _, e2v = m.entities(1)
    _, f2v = m.entities(2)
    assert n == m.num_vertices + (degree - 1) * len(e2v) + (len(f2v) if degree == 3 else 0)
    if degree == 1:
        assert np.array_equal(dm, m.cells)


def test_p3_edge_dofs_agree_between_cells():
    """After the reflection swap, two cells sharing an edge list its two
    dofs in the order of the global edge orientation (low -> high vertex)."""
    m = O.Mesh((2, 2, 2), O.TET)
    dm, _ = m.dofmap(3)
    seen = {}
    for c in range(m.num_cells):
        for e, (a, b) in enumerate(TET_EDGES):
            va, vb = m.cells[c, a], m.cells[c, b]
            d = (dm[c, 4 + 2 * e], dm[c, 5 + 2 * e])
            key = (min(va, vb), max(va, vb))
            # position 4+2e is the point nearer local vertex a
            ordered = d if va < vb else d[::-1]  # (nearer lower, nearer higher vertex)
            seen.setdefault(key, ordered)
            assert seen[key] == ordered


def test_boundary_dofs_brute_force():
    m = O.Mesh((3, 3, 2), O.TET)
    for degree in (1, 2):
        dm, _ = m.dofmap(degree)
        got = m.boundary_nodes(degree, dm)
        x = m.x
        # P1/P2 nodes lie on the boundary iff their position does
        pts = {}
        for c in range(m.num_cells):
            X = x[m.cells[c]]
            nodes = list(X) + [(X[a] + X[b]) / 2 for a, b in TET_EDGES][: 6 * (degree - 1)]
            for k, p in enumerate(nodes):
                pts[dm[c, k]] = p
        on = sorted(d for d, p in pts.items() if np.any(np.isclose(p, 0) | np.isclose(p, 1)))
        assert list(got) == on


def test_sparsity_brute_force():
    m = O.Mesh((2, 2, 3), O.TET)
    dm, n = m.dofmap(2)
    rp, cols = O.sparsity(dm, dm, n)
    rows = [set() for _ in range(n)]
    for d in dm:
        for i in d:
            rows[i].update(d)
    assert np.array_equal(np.diff(rp), [len(r) for r in rows])
    assert np.array_equal(cols, np.concatenate([sorted(r) for r in rows]))


# ---------------------------------------------------------------- kernels
def random_tet(rng):
    while True:
        X = rng.standard_normal((4, 3))
        if abs(np.linalg.det(X[1:] - X[0])) > 0.3:
            return X


def monomial_integral_tet(X, f_coeffs):
    """Exact integral over the tet X of a polynomial given as {(a,b,c): coeff}
    in reference coordinates: int_ref xi^a eta^b zeta^c = a!b!c!/(a+b+c+3)!."""
    det = abs(np.linalg.det(X[1:] - X[0]))
    return det * sum(v * factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                     for (a, b, c), v in f_coeffs.items())


def test_p1_kernel_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(5):
        X = random_tet(rng)
        kap, c = 2.5, 0.75
        A = O.tabulate(O.K_P1TET_A, X, w=[kap], c=[c], ndim=16).reshape(4, 4)
        B = np.linalg.inv(np.hstack([np.ones((4, 1)), X]))  # barycentric coefficients
        G = B[1:, :].T  # gradients of barycentric coordinates
        vol = abs(np.linalg.det(X[1:] - X[0])) / 6
        M = vol / 20 * (np.ones((4, 4)) + np.eye(4))
        assert np.allclose(A, kap * vol * G @ G.T + c * M, rtol=1e-13, atol=1e-14)
        b = O.tabulate(O.K_P1TET_L, X, w=np.ones(4), ndim=4)
        assert np.allclose(b, vol / 4, rtol=1e-14)


def q1_nodal_values(X, f):
    return np.array([f(*X[i]) for i in range(8)])


def test_q1_kernel_exact_on_parallelepipeds():
    """On a parallelepiped every trilinear u is exactly representable:
    u^T A u = kappa int |grad u|^2 + c int u^2, integrated by numpy with a
    4x4x4 Gauss rule (exact for these polynomials)."""
    rng = np.random.default_rng(2)
    e = rng.standard_normal((3, 3)) + 2 * np.eye(3)
    x0 = rng.standard_normal(3)
    X = np.array([x0 + ((i & 1) * e[0] + ((i >> 1) & 1) * e[1] + ((i >> 2) & 1) * e[2])
                  for i in range(8)])
    kap, c = 1.7, 0.4
    A = O.tabulate(O.K_Q1HEX_A, X, w=[kap], c=[c], ndim=64).reshape(8, 8)
    coef = rng.standard_normal(8)  # u = sum coef_i phi_i
    g, w = np.polynomial.legendre.leggauss(4)
    g, w = (g + 1) / 2, w / 2
    J = np.stack([e[0], e[1], e[2]], 1)
    detJ = abs(np.linalg.det(J))
    val = 0.0
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for cc, wc in zip(g, w):
                xi = np.array([a, b, cc])
                phi = np.array([np.prod([xi[d] if (i >> d) & 1 else 1 - xi[d] for d in range(3)])
                                for i in range(8)])
                dphi = np.array([[np.prod([(1 if (i >> d) & 1 else -1) if d == k else
                                           (xi[d] if (i >> d) & 1 else 1 - xi[d]) for d in range(3)])
                                  for k in range(3)] for i in range(8)])
                gu = np.linalg.solve(J.T, dphi.T @ coef)
                val += wa * wb * wc * detJ * (kap * gu @ gu + c * (phi @ coef) ** 2)
    assert np.isclose(coef @ A @ coef, val, rtol=1e-12)
    assert np.allclose(A, A.T, rtol=1e-14, atol=1e-14)
    Ak = O.tabulate(O.K_Q1HEX_A, X, w=[kap], c=[0.0], ndim=64).reshape(8, 8)
    assert np.allclose(Ak @ np.ones(8), 0, atol=1e-12)


P3_TERMS = {(3, 0, 0): 1.0, (1, 2, 0): -0.5, (0, 1, 1): 2.0, (0, 0, 0): 0.3}


def poly_eval(terms, x):
    return sum(v * x[0] ** a * x[1] ** b * x[2] ** c for (a, b, c), v in terms.items())


def poly_mul(p, q):
    out = {}
    for (a, b, c), v in p.items():
        for (d, e, f), w in q.items():
            k = (a + d, b + e, c + f)
            out[k] = out.get(k, 0.0) + v * w
    return out


def poly_grad(p):
    gs = [{}, {}, {}]
    for (a, b, c), v in p.items():
        for k, (e, newk) in enumerate([(a, (a - 1, b, c)), (b, (a, b - 1, c)), (c, (a, b, c - 1))]):
            if e:
                gs[k][newk] = gs[k].get(newk, 0.0) + v * e
    return gs


def p3_nodes_reference():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    t = [0.5 - 0.5 / np.sqrt(5), 0.5 + 0.5 / np.sqrt(5)]
    pts = list(V)
    for a, b in TET_EDGES:
        pts += [V[a] + s * (V[b] - V[a]) for s in t]
    pts += [V[list(f)].mean(0) for f in TET_FACETS]
    return np.array(pts)


def test_p3_kernel_exact_quadratic_forms():
    """On the reference tet (identity map) a cubic u is interpolated exactly,
    so u^T A u equals kappa int |grad u|^2 + c int u^2 exactly."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    kap, c = 3.0, 2.0
    A = O.tabulate(O.K_P3TET_A, X, w=[kap], c=[c], ndim=400).reshape(20, 20)
    u = np.array([poly_eval(P3_TERMS, p) for p in p3_nodes_reference()])
    g = poly_grad(P3_TERMS)
    grad2 = {}
    for gk in g:
        for k, v in poly_mul(gk, gk).items():
            grad2[k] = grad2.get(k, 0.0) + v
    exact = kap * monomial_integral_tet(X, grad2) + c * monomial_integral_tet(X, poly_mul(P3_TERMS, P3_TERMS))
    assert np.isclose(u @ A @ u, exact, rtol=1e-12)
    A0 = O.tabulate(O.K_P3TET_A, X, w=[1.0], c=[0.0], ndim=400).reshape(20, 20)
    assert np.allclose(A0 @ np.ones(20), 0, atol=1e-12)
    M = O.tabulate(O.K_P3TET_A, X, w=[0.0], c=[1.0], ndim=400).reshape(20, 20)
    assert np.isclose(M.sum(), 1 / 6, rtol=1e-13)


def test_p2_elasticity_rigid_modes_and_symmetry():
    rng = np.random.default_rng(3)
    X = random_tet(rng)
    A = O.tabulate(O.K_P2TET_ELAST_A, X, w=[1.3, 0.7], ndim=900).reshape(30, 30)
    assert np.allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    nodes = list(X) + [(X[a] + X[b]) / 2 for a, b in TET_EDGES]
    modes = []
    for d in range(3):
        t = np.zeros(3)
        t[d] = 1
        modes.append(np.concatenate([t for _ in nodes]))
    for w in np.eye(3):
        modes.append(np.concatenate([np.cross(w, p) for p in nodes]))
    for mvec in modes:
        assert np.allclose(A @ mvec, 0, atol=1e-11 * np.abs(A).max())
    # a linear displacement with strain eps: u^T A u = vol (2 mu eps:eps + lam tr(eps)^2)
    Gm = rng.standard_normal((3, 3))
    u = np.concatenate([Gm @ p for p in nodes])
    eps = 0.5 * (Gm + Gm.T)
    vol = abs(np.linalg.det(X[1:] - X[0])) / 6
    assert np.isclose(u @ A @ u, vol * (2 * 0.7 * np.sum(eps * eps) + 1.3 * np.trace(eps) ** 2),
                      rtol=1e-12)
    b = O.tabulate(O.K_P2TET_ELAST_L, X, c=[0.0, 0.0, -1.0], ndim=30)
    assert np.isclose(b[2::3].sum(), -vol, rtol=1e-13)
    assert np.allclose(b[0::3], 0) and np.allclose(b[1::3], 0)


# ---------------------------------------------------------------- assembly, solve
def assemble_p1(n, kappa=1.0, c=0.0, with_bc=True):
    m = O.Mesh(n, O.TET)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    bn = m.boundary_nodes(1, dm) if with_bc else np.zeros(0, np.int32)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    A = O.assemble_matrix(O.K_P1TET_A, m.x, m.cells, dm, 1, np.full((m.num_cells, 1), kappa),
                          np.array([c]), mark if with_bc else None, mark if with_bc else None,
                          rp, cols)
    return m, dm, nd, rp, cols, A, bn, mark


def test_mass_total_and_stiffness_nullspace():
    """P1 mass total = |Omega| (test_assembler.py:1506-1535); A 1 = 0
    (cpp/test/matrix.cpp:116-119)."""
    m, dm, nd, rp, cols, A, _, _ = assemble_p1((3, 4, 2), kappa=0.0, c=1.0, with_bc=False)
    assert np.isclose(A.sum(), 1.0, rtol=1e-13)
    m, dm, nd, rp, cols, K, _, _ = assemble_p1((3, 4, 2), kappa=1.0, c=0.0, with_bc=False)
    assert np.allclose(O.spmv(rp, cols, K, 1, np.ones(nd)), 0, atol=1e-13)


@pytest.mark.parametrize("ct,kL", [(O.TET, O.K_P1TET_L), (O.HEX, O.K_Q1HEX_L)])
def test_source_functionals(ct, kL):
    """sum b = int f dx = 1 for f = 1 and 1/2 for f = x_0
    (python/test/unit/fem/test_assembler.py:67-78); P3 source total = |Omega| c."""
    m = O.Mesh((3, 4, 2), ct)
    dm, nd = m.dofmap(1)
    one = np.ones((m.num_cells, m.nvpc))
    assert np.isclose(O.assemble_vector(kL, m.x, m.cells, dm, 1, one, None, nd).sum(), 1.0,
                      rtol=1e-14)
    fx = m.x[:, 0][m.cells]
    assert np.isclose(O.assemble_vector(kL, m.x, m.cells, dm, 1, fx, None, nd).sum(), 0.5,
                      rtol=1e-14)
    if ct == O.TET:
        dm3, nd3 = m.dofmap(3)
        b3 = O.assemble_vector(O.K_P3TET_L, m.x, m.cells, dm3, 1, None, np.array([2.5]), nd3)
        assert np.isclose(b3.sum(), 2.5, rtol=1e-13)


def test_insert_outside_pattern_raises():
    m, dm, nd, rp, cols, A, _, _ = assemble_p1((2, 2, 2), with_bc=False)
    diag_rp = np.arange(nd + 1, dtype=np.int64)
    with pytest.raises(RuntimeError, match="Entry not in sparsity"):
        O.assemble_matrix(O.K_P1TET_A, m.x, m.cells, dm, 1, np.ones((m.num_cells, 1)),
                          np.zeros(1), None, None, diag_rp, np.arange(nd, dtype=np.int32))


def test_manufactured_linear_solution_is_exact():
    """u = 1 + x + 2y + 3z lies in P1: the discrete solution of -lap u = 0
    with u = g on the boundary (lifting + bc.set + CG) is exact at nodes."""
    m, dm, nd, rp, cols, A, bn, mark = assemble_p1((6, 5, 4))
    O.set_diagonal(A, bn, 1, 1.0, rp, cols)
    g = 1 + m.x[:, 0] + 2 * m.x[:, 1] + 3 * m.x[:, 2]
    b = np.zeros(nd)
    O.lift_bc(b, O.K_P1TET_A, m.x, m.cells, dm, 1, np.ones((m.num_cells, 1)), np.zeros(1), g, mark)
    O.bc_set(b, bn, g)
    x, it = O.cg(rp, cols, A, 1, b, rtol=1e-13)
    assert it > 0
    assert np.allclose(x, g, atol=1e-10)


def test_lifting_equals_dense_column_product():
    """lift_bc subtracts A_full[:, bc] g over every row (brute force)."""
    m, dm, nd, rp, cols, _, bn, mark = assemble_p1((3, 3, 3))
    Afull = O.assemble_matrix(O.K_P1TET_A, m.x, m.cells, dm, 1, np.ones((m.num_cells, 1)),
                              np.array([0.5]), None, None, rp, cols)
    g = np.zeros(nd)
    g[bn] = np.sin(np.arange(len(bn)))
    b = O.lift_bc(np.zeros(nd), O.K_P1TET_A, m.x, m.cells, dm, 1, np.ones((m.num_cells, 1)),
                  np.array([0.5]), g, mark)
    D = sp.csr_matrix((Afull, cols, rp), shape=(nd, nd))
    assert np.allclose(b, -(D @ g), atol=1e-14)


def tabulate_cells(kernel, m, coeffs, constants, ndim):
    """Element tensors of every cell of m from the oracle's kernel."""
    return np.stack([O.tabulate(kernel, m.x[m.cells[c]], w=coeffs[c], c=constants, ndim=ndim)
                     for c in range(m.num_cells)])


@pytest.mark.parametrize("ct,kA,kL", [(O.TET, O.K_P1TET_A, O.K_P1TET_L),
                                      (O.HEX, O.K_Q1HEX_A, O.K_Q1HEX_L)])
def test_element_insertion_equals_kernel_assembly(ct, kA, kL):
    """Inserting the tabulated element tensors reproduces the assembly that
    calls the kernel inside the cell loop, bit for bit (with BCs)."""
    m = O.Mesh((3, 2, 2), ct)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    rng = np.random.default_rng(3)
    kap = rng.uniform(0.5, 2.0, (m.num_cells, 1))
    c = np.array([0.7])
    bn = m.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    k = dm.shape[1]
    E = tabulate_cells(kA, m, kap, c, k * k)
    A = O.assemble_matrix(kA, m.x, m.cells, dm, 1, kap, c, mark, mark, rp, cols)
    assert np.array_equal(O.add_element_matrices(dm, 1, E, mark, mark, rp, cols), A)
    f = rng.standard_normal((m.num_cells, k))
    e = tabulate_cells(kL, m, f, None, k)
    assert np.array_equal(O.add_element_vectors(dm, 1, e, nd),
                          O.assemble_vector(kL, m.x, m.cells, dm, 1, f, None, nd))


def test_element_insertion_dense_scatter():
    """Random element tensors against a dense scatter D[dofs, dofs] += E_c
    with BC rows and columns removed (brute force), and a missing pattern
    entry raises."""
    m = O.Mesh((2, 1, 2), O.HEX)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    rng = np.random.default_rng(4)
    E = rng.standard_normal((m.num_cells, 8, 8))
    mark = (rng.uniform(size=nd) < 0.3).astype(np.int8)
    D = np.zeros((nd, nd))
    for cc in range(m.num_cells):
        d = dm[cc]
        Ec = E[cc].copy()
        Ec[mark[d] == 1, :] = 0
        Ec[:, mark[d] == 1] = 0
        D[np.ix_(d, d)] += Ec
    vals = O.add_element_matrices(dm, 1, E, mark, mark, rp, cols)
    assert np.allclose(sp.csr_matrix((vals, cols, rp), shape=(nd, nd)).toarray(), D,
                       rtol=1e-15, atol=1e-15)
    e = rng.standard_normal((m.num_cells, 8))
    b = np.zeros(nd)
    np.add.at(b, dm.ravel(), e.ravel())
    assert np.allclose(O.add_element_vectors(dm, 1, e, nd), b, rtol=1e-15, atol=1e-15)
    with pytest.raises(RuntimeError, match="Entry not in sparsity"):
        O.add_element_matrices(dm, 1, E, None, None, np.arange(nd + 1, dtype=np.int64),
                               np.arange(nd, dtype=np.int32))


@pytest.mark.parametrize("ct,n", [(O.TET, (3, 2, 4)), (O.HEX, (2, 3, 3))])
def test_facet_domain_brute_force(ct, n):
    """(cell, local facet) of every exterior facet, in facet-index order:
    facets found by matching sorted vertex tuples; marked facets have all
    vertices on the plane."""
    m = O.Mesh(n, ct)
    table = TET_FACETS if ct == O.TET else HEX_FACETS
    seen = {}
    for c in range(m.num_cells):
        for lf, fv in enumerate(table):
            key = tuple(sorted(m.cells[c][list(fv)]))
            seen.setdefault(key, []).append((c, lf))
    ext = {v[0] for v in seen.values() if len(v) == 1}
    pairs = m.facet_domain()
    assert {tuple(p) for p in pairs} == ext and len(pairs) == len(ext)
    c2f, _ = m.entities(2)
    idx = [c2f[c, lf] for c, lf in pairs]
    assert all(a < b for a, b in zip(idx, idx[1:]))
    on = m.facet_domain(axis=2, value=1.0)
    assert {tuple(p) for p in on} == {
        (c, lf) for (c, lf) in ext if np.all(m.x[m.cells[c][list(table[lf])], 2] == 1.0)}


@pytest.mark.parametrize("ct,kN,nv", [(O.TET, O.K_P1TET_NEUMANN, 4), (O.HEX, O.K_Q1HEX_NEUMANN, 8)])
def test_facet_loads_closed_forms(ct, kN, nv):
    """int_Gamma g v ds summed over v: the unit cube's surface area 6 for
    g = 1; int over x = 1 of 1 + 2y + 3z = 3.5 and over y = 0 of 1 = 1
    (g linear is exact in P1 and Q1)."""
    m = O.Mesh((3, 2, 4), ct)
    dm, nd = m.dofmap(1)
    one = np.ones((m.num_cells, nv))
    b = O.assemble_vector_facets(kN, m.x, m.cells, dm, one, m.facet_domain(), nd)
    assert np.isclose(b.sum(), 6.0, rtol=1e-14)
    g = (m.x[:, 0] + 2 * m.x[:, 1] + 3 * m.x[:, 2])[m.cells]
    b1 = O.assemble_vector_facets(kN, m.x, m.cells, dm, g, m.facet_domain(0, 1.0), nd)
    assert np.isclose(b1.sum(), 3.5, rtol=1e-14)
    b2 = O.assemble_vector_facets(kN, m.x, m.cells, dm, one, m.facet_domain(1, 0.0), nd)
    assert np.isclose(b2.sum(), 1.0, rtol=1e-14)
    # only nodes on the face receive load
    assert np.all(b2[m.x[:, 1] != 0.0] == 0.0)


@pytest.mark.parametrize("jacobi", [True, False])
def test_cg_against_scipy(jacobi):
    m, dm, nd, rp, cols, A, bn, mark = assemble_p1((8, 8, 8), kappa=1.0)
    O.set_diagonal(A, bn, 1, 1.0, rp, cols)
    b = O.assemble_vector(O.K_P1TET_L, m.x, m.cells, dm, 1, np.ones((m.num_cells, 4)), None, nd)
    O.bc_set(b, bn, np.zeros(nd))
    x, it = O.cg(rp, cols, A, 1, b, rtol=1e-12, jacobi=jacobi)
    xs = sla.spsolve(sp.csr_matrix((A, cols, rp), shape=(nd, nd)).tocsc(), b)
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
    # the reference CG raises after maxit without convergence
    _, it = O.cg(rp, cols, A, 1, b, kmax=2, rtol=1e-14, jacobi=jacobi)
    assert it == -1


@pytest.mark.parametrize("ct,kR", [(O.TET, O.K_P1TET_ROBIN_A), (O.HEX, O.K_Q1HEX_ROBIN_A)])
def test_facet_matrices_closed_forms(ct, kR):
    """beta u v ds over the unit cube's surface: 1' A 1 = beta |Gamma| = 6
    beta; 1' A x = beta int_Gamma x ds = 3 beta (x is in P1 and Q1);
    1' A y = 3 beta on one face y = 1 only for its area 1; the entries of a
    single unit facet form the facet mass matrix (|F| (1 + d_ab) / 12 on a
    triangle, the tensor product of [2 1; 1 2] / 6 on a square)."""
    n = (3, 2, 4)
    m = O.Mesh(n, ct)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    beta = 2.5
    co = np.full((m.num_cells, 1), beta)
    A = O.assemble_matrix_facets(kR, m.x, m.cells, dm, co, m.facet_domain(), None, None, rp, cols)
    D = sp.csr_matrix((A, cols, rp), shape=(nd, nd))
    one = np.ones(nd)
    assert np.isclose(one @ (D @ one), 6.0 * beta, rtol=1e-14)
    assert np.isclose(one @ (D @ m.x[:, 0]), 3.0 * beta, rtol=1e-14)
    assert abs(D - D.T).max() <= 1e-15
    # the face y = 1: 1' A 1 = beta (its area), 1' A z = beta / 2
    A1 = O.assemble_matrix_facets(kR, m.x, m.cells, dm, co, m.facet_domain(1, 1.0), None, None,
                                  rp, cols)
    D1 = sp.csr_matrix((A1, cols, rp), shape=(nd, nd))
    assert np.isclose(one @ (D1 @ one), beta, rtol=1e-14)
    assert np.isclose(one @ (D1 @ m.x[:, 2]), beta / 2, rtol=1e-14)
    off = m.x[:, 1] != 1.0
    assert np.all(D1[off].data == 0.0)
    # one facet of a single cell with unit edges
    s = O.Mesh((1, 1, 1), ct)
    sdm, snd = s.dofmap(1)
    srp, scols = O.sparsity(sdm, sdm, snd)
    pairs = s.facet_domain(2, 0.0)[:1]
    As = O.assemble_matrix_facets(kR, s.x, s.cells, sdm, np.ones((s.num_cells, 1)), pairs, None,
                                  None, srp, scols)
    Ds = sp.csr_matrix((As, scols, srp), shape=(snd, snd)).toarray()
    table = TET_FACETS if ct == O.TET else HEX_FACETS
    fv = sdm[pairs[0][0]][list(table[pairs[0][1]])]
    F = Ds[np.ix_(fv, fv)]
    if ct == O.TET:
        area = 0.5 * np.linalg.norm(np.cross(s.x[fv[1]] - s.x[fv[0]], s.x[fv[2]] - s.x[fv[0]]))
        assert np.allclose(F, area * (np.ones((3, 3)) + np.eye(3)) / 12, rtol=1e-14, atol=0)
    else:
        M1 = np.array([[2.0, 1.0], [1.0, 2.0]]) / 6
        # facet vertices in tensor order: (0,0), (1,0), (0,1), (1,1)
        assert np.allclose(F, np.kron(M1, M1), rtol=1e-14, atol=0)
    mask = np.ones(snd, bool)
    mask[fv] = False
    assert np.all(Ds[mask] == 0.0) and np.all(Ds[:, mask] == 0.0)


@pytest.mark.parametrize("ct,kR", [(O.TET, O.K_P1TET_ROBIN_A), (O.HEX, O.K_Q1HEX_ROBIN_A)])
def test_facet_matrices_bcs_and_lifting(ct, kR):
    """With BC markers the facet matrices lose exactly the marked rows and
    columns (entry by entry equal to the unmarked assembly elsewhere);
    lifting = -alpha A[:, bc] (g - x0) with A assembled without BCs; an
    entry outside the pattern raises."""
    m = O.Mesh((3, 3, 2), ct)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    rng = np.random.default_rng(11)
    co = rng.uniform(0.5, 2.0, (m.num_cells, 1))
    pairs = m.facet_domain()
    A0 = O.assemble_matrix_facets(kR, m.x, m.cells, dm, co, pairs, None, None, rp, cols)
    bn = m.boundary_nodes(1, dm, marker=1, axis=0, value=0.0)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    Ab = O.assemble_matrix_facets(kR, m.x, m.cells, dm, co, pairs, mark, mark, rp, cols)
    rows = np.repeat(np.arange(nd), np.diff(rp))
    keep = (mark[rows] == 0) & (mark[cols] == 0)
    assert np.array_equal(Ab[keep], A0[keep]) and np.all(Ab[~keep] == 0.0)
    g = np.zeros(nd)
    g[bn] = np.cos(np.arange(len(bn)))
    x0 = rng.standard_normal(nd)
    b = O.lift_bc_facets(np.zeros(nd), kR, m.x, m.cells, dm, co, pairs, g, mark, x0=x0, alpha=0.6)
    D = sp.csr_matrix((A0, cols, rp), shape=(nd, nd))
    w = np.where(mark == 1, g - x0, 0.0)
    assert np.allclose(b, -0.6 * (D @ w), rtol=0, atol=1e-14 * np.abs(D @ w).max())
    with pytest.raises(RuntimeError, match="Entry not in sparsity"):
        O.assemble_matrix_facets(kR, m.x, m.cells, dm, co, pairs, None, None,
                                 np.arange(nd + 1, dtype=np.int64), np.arange(nd, dtype=np.int32))


def pattern_with_interior(dm, nd, quads):
    """Brute-force rows: cell couplings plus all dofs of both cells of each
    interior facet."""
    rows = [set() for _ in range(nd)]
    for d in dm:
        for r in d:
            rows[r].update(d.tolist())
    for c0, _, c1, _ in quads:
        joint = np.concatenate([dm[c0], dm[c1]])
        for r in joint:
            rows[r].update(joint.tolist())
    rp = np.zeros(nd + 1, np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    cols = np.array([c for r in rows for c in sorted(r)], np.int32)
    return rp, cols


def test_interior_facet_domain_brute_force():
    """Every facet shared by two cells once, as (cell0, lf0, cell1, lf1)
    with cell0 < cell1 and both local facets naming the same vertex set, in
    facet-index order."""
    m = O.Mesh((3, 2, 2), O.TET)
    seen = {}
    for c in range(m.num_cells):
        for lf, fv in enumerate(TET_FACETS):
            seen.setdefault(tuple(sorted(m.cells[c][list(fv)])), []).append((c, lf))
    shared = {tuple(v[0] + v[1]) for v in seen.values() if len(v) == 2}
    q = m.interior_facet_domain()
    assert {tuple(r) for r in q} == shared and len(q) == len(shared)
    assert np.all(q[:, 0] < q[:, 2])
    c2f, _ = m.entities(2)
    idx = [c2f[c, lf] for c, lf, _, _ in q]
    assert all(a < b for a, b in zip(idx, idx[1:]))
    assert all(c2f[c0, l0] == c2f[c1, l1] for c0, l0, c1, l1 in q)


def test_interior_gradient_jump_closed_forms():
    """gamma [grad u.n][grad v.n] dS on P1: A u = 0 for every linear u;
    u = max(x - 1/2, 0) (piecewise linear on a box with a mesh plane at
    x = 1/2) has u' A u = gamma * 1 (a unit jump over the unit square);
    A symmetric positive semidefinite; the pattern needs the facet
    couplings (the cell pattern raises)."""
    m = O.Mesh((4, 3, 2), O.TET)
    dm, nd = m.dofmap(1)
    q = m.interior_facet_domain()
    rp, cols = pattern_with_interior(dm, nd, q)
    gamma = 1.75
    A = O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, m.x, m.cells, dm, q, None, None, rp,
                                   cols, [gamma])
    D = sp.csr_matrix((A, cols, rp), shape=(nd, nd))
    lin = 0.3 + m.x[:, 0] - 2.0 * m.x[:, 1] + 0.5 * m.x[:, 2]
    assert np.abs(D @ lin).max() <= 1e-12 * np.abs(D).max()
    kink = np.maximum(m.x[:, 0] - 0.5, 0.0)
    assert np.isclose(kink @ (D @ kink), gamma, rtol=1e-13)
    assert abs(D - D.T).max() <= 1e-14 * np.abs(D).max()
    assert np.linalg.eigvalsh(D.toarray()).min() >= -1e-12 * np.abs(D).max()
    crp, ccols = O.sparsity(dm, dm, nd)
    with pytest.raises(RuntimeError, match="Entry not in sparsity"):
        O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, m.x, m.cells, dm, q, None, None, crp,
                                   ccols, [gamma])
    # BC rows and columns vanish, the rest is unchanged; lifting is the
    # dense column product
    bn = m.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    Ab = O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, m.x, m.cells, dm, q, mark, mark, rp,
                                    cols, [gamma])
    rows = np.repeat(np.arange(nd), np.diff(rp))
    keep = (mark[rows] == 0) & (mark[cols] == 0)
    assert np.array_equal(Ab[keep], A[keep]) and np.all(Ab[~keep] == 0.0)
    g = np.zeros(nd)
    g[bn] = np.sin(np.arange(len(bn)) + 0.5)
    b = O.lift_bc_interior(np.zeros(nd), O.K_P1TET_GRADJUMP_A, m.x, m.cells, dm, q, [gamma], g,
                           mark, alpha=1.3)
    assert np.allclose(b, -1.3 * (D @ g), rtol=0, atol=1e-13 * np.abs(D @ g).max())
