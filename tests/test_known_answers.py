"""Known answers derived by hand from the reference's rules, written here as
literals. They pin the oracle (and, for the meshes, the product setup too)
independently of any restatement of the same code.

RCM (graph/ordering.cpp:65-144; reorder_rcm :150-170): the root is the FIRST
vertex of minimum degree among the unlabelled ones (strict <, :79-87); the
pseudo-peripheral search moves to the FIRST minimum-degree vertex of the last
level (std::ranges::min_element, :95) only while that gives MORE levels
(strict >, :97); Cuthill-McKee sorts each processed vertex's own newly found
neighbours by degree (:117-136, not each BFS level); the order is reversed;
disconnected parts are numbered one after the other. Each graph below is
chosen so that one plausible misreading gives a different permutation; the
misreading's answer is noted beside the test.

Meshes: create_box (mesh/generation.h:404-471) -> dual graph of shared facets
in the order of the stable facet sort (mesh/graphbuild.cpp:600-896) -> RCM
-> reorder_list (mesh/utils.h:282-306) -> vertices numbered by first
appearance (mesh/Topology.cpp:1222-1258) -> entity dofs (vertices, then
edges by lexicographic rank of their sorted vertex pair,
mesh/topologycomputation.cpp:822-892) renumbered in first-traversal order
over the cells (fem/dofmapbuilder.cpp:445-459; no reorder_fn, :476).

P3 nodes: DOLFINx builds Lagrange elements through basix.ufl, whose default
variant for the P family is gll_warped (Basix >= 0.5; fenics-basix 0.12 is
the reference's pin, python/pyproject.toml:30-32). At degree 3 that places the
two interior nodes of every edge at the Gauss-Lobatto-Legendre points, the
roots of P_3' mapped to [0, 1], ordered from the edge's first vertex to its
second, and one node at each face centroid.
"""

from itertools import product as iproduct

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, mesh


def _csr(adj):
    off = np.zeros(len(adj) + 1, np.int32)
    off[1:] = np.cumsum([len(a) for a in adj])
    return off, np.array([w for a in adj for w in a], np.int32)


def _adj(n, edges):
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    return [sorted(a) for a in adj]


# A path 0-1-2-3-4 ending in a fork 4-{5,6}; 5-7 where 7 carries three leaves
# (9, 10, 11) and 6-8-12. From root 0 the pseudo-peripheral search stays at 0
# (the far leaf 9 gives 8 levels, not more). Per-vertex sorting appends 7
# (found from 5) before 8 (found from 6): CM = 0..12, so r[q] = 12 - q.
# Sorting each BFS level by degree would put 8 (degree 2) before 7 (degree 4):
# r = [12,11,10,9,8,7,6,4,5,2,1,0,3]. Taking the LAST minimum-degree root (12)
# gives r = [0,4,6,8,9,7,10,5,11,3,2,1,12].
TREE13 = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (4, 6), (5, 7), (6, 8), (7, 9), (7, 10),
          (7, 11), (8, 12)]

# The 5-cycle 0-3-2-1-4-0 (all degree 2). Root 0, levels {0},{3,4},{2,1};
# the last level's first vertex 2 gives 3 levels too, so the root stays
# (moving on ties, >=, would wander). CM: 0, 3, 4, 2, 1; reversed.
CYCLE5 = [(0, 3), (0, 4), (1, 2), (1, 4), (2, 3)]

# A star centre 4 with arms 0, 1-3, 2-5. From root 0 the last level is
# {3, 5}; its FIRST vertex 3 gives 5 levels (> 4), so the root moves to 3,
# and 5 then gives no more. CM from 3: 3, 1, 4, {0, 2} by degree, 5.
# Taking the last vertex of the level (5) would give r = [2,1,4,0,3,5].
STAR6 = [(0, 4), (1, 3), (1, 4), (2, 4), (2, 5)]


@pytest.mark.parametrize("n,edges,expect", [
    (13, TREE13, list(range(12, -1, -1))),
    (5, CYCLE5, [4, 0, 1, 3, 2]),
    (6, STAR6, [2, 4, 1, 5, 3, 0]),
])
def test_rcm_known_answers(n, edges, expect):
    off, data = _csr(_adj(n, edges))
    assert O.rcm(off, data).tolist() == expect


def test_rcm_disconnected_known_answer():
    """The 5-cycle on vertices 0-4 and the star on 5-10. The first root is
    the first vertex of minimum degree over the whole graph, 5 (degree 1), so
    the star is numbered first (0..5, as STAR6 shifted by 5), then the
    cycle's root is the first unlabelled vertex of minimum degree, 0, and it
    takes 6..10 in its own order."""
    edges = CYCLE5 + [(a + 5, b + 5) for a, b in STAR6]
    off, data = _csr(_adj(11, edges))
    assert O.rcm(off, data).tolist() == [10, 6, 7, 9, 8, 2, 4, 1, 5, 3, 0]


# ---- the unit cube as one box of six tetrahedra -------------------------
# Input vertices v = ix + 2 iy + 4 iz; cells (generation.h:418-420)
#   c0 [0,1,3,7] c1 [0,1,7,5] c2 [0,5,7,4] c3 [0,3,2,7] c4 [0,6,4,7] c5 [0,2,6,7].
# Shared facets in sorted key order: (0,1,7) c0-c1, (0,2,7) c3-c5,
# (0,3,7) c0-c3, (0,4,7) c2-c4, (0,5,7) c1-c2, (0,6,7) c4-c5, which fill the
# adjacency lists in that order. RCM: root 0; levels {0},{1,3},{2,5},{4};
# from 4 also 4 levels, so root 0; CM 0,1,3,2,5,4; reversed -> remap
# [5,4,2,3,0,1]: the new cells are c4, c5, c2, c3, c1, c0.
TET1_DUAL = [[1, 3], [0, 2], [4, 1], [5, 0], [2, 5], [3, 4]]
TET1_ORIGINAL = [4, 5, 2, 3, 1, 0]
# first appearance over c4, c5, c2, c3, c1: input 0,6,4,7,2,5,3,1 -> 0..7
TET1_INPUT_VERTEX = [0, 6, 4, 7, 2, 5, 3, 1]
TET1_CELLS = [[0, 1, 2, 3], [0, 4, 1, 3], [0, 5, 3, 2], [0, 6, 4, 3], [0, 7, 3, 5],
              [0, 7, 6, 3]]
# P2: vertex dofs 0..7, edge e -> 8 + rank of its sorted vertex pair among the
# 19 edges; cell dofs = 4 vertices then edges [2,3],[1,3],[1,2],[0,3],[0,2],
# [0,1]; renumbered by first traversal.
TET1_P2 = [[0, 1, 2, 3, 4, 5, 6, 7, 8, 9],
           [0, 10, 1, 3, 5, 11, 12, 7, 9, 13],
           [0, 14, 3, 2, 4, 15, 16, 8, 7, 17],
           [0, 18, 10, 3, 11, 19, 20, 7, 13, 21],
           [0, 22, 3, 14, 16, 23, 24, 17, 7, 25],
           [0, 22, 18, 3, 19, 24, 26, 7, 21, 25]]

# ---- two hexahedra along x ----------------------------------------------
# Input vertices v = ix + 3 iy + 6 iz; c0 [0,1,3,4,6,7,9,10],
# c1 [1,2,4,5,7,8,10,11]. Dual graph 0-1; RCM root 0, CM 0,1, reversed:
# the new cells are c1, c0; first appearance numbers c1's vertices 0..7.
HEX2_ORIGINAL = [1, 0]
HEX2_INPUT_VERTEX = [1, 2, 4, 5, 7, 8, 10, 11, 0, 3, 6, 9]
HEX2_CELLS = [[0, 1, 2, 3, 4, 5, 6, 7], [8, 0, 9, 2, 10, 4, 11, 6]]


def _grid_x(n, iv):
    nx, ny, _ = n
    iv = np.asarray(iv)
    ix, iy, iz = iv % (nx + 1), (iv // (nx + 1)) % (ny + 1), iv // ((nx + 1) * (ny + 1))
    return np.stack([ix / n[0], iy / n[1], iz / n[2]], 1)


def test_tet_box_1x1x1_oracle():
    om = O.Mesh((1, 1, 1), O.TET)
    cells0 = np.array([[0, 1, 3, 7], [0, 1, 7, 5], [0, 5, 7, 4], [0, 3, 2, 7], [0, 6, 4, 7],
                       [0, 2, 6, 7]])
    off, data = O.dual_graph(O.TET, cells0)
    assert [data[off[i]:off[i + 1]].tolist() for i in range(6)] == TET1_DUAL
    assert om.original_cell_index.tolist() == TET1_ORIGINAL
    assert om.input_vertex_index.tolist() == TET1_INPUT_VERTEX
    assert om.cells.tolist() == TET1_CELLS
    assert np.array_equal(om.x, _grid_x((1, 1, 1), TET1_INPUT_VERTEX))
    dm1, n1 = om.dofmap(1)
    assert n1 == 8 and dm1.tolist() == TET1_CELLS
    dm2, n2 = om.dofmap(2)
    assert n2 == 27 and dm2.tolist() == TET1_P2


def test_tet_box_1x1x1_product():
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (1, 1, 1), mesh.CellType.tetrahedron)
    assert m.topology.original_cell_index.tolist() == TET1_ORIGINAL
    assert m.geometry.input_global_indices.tolist() == TET1_INPUT_VERTEX
    assert m.topology.cells.tolist() == TET1_CELLS
    assert np.array_equal(m.geometry.x, _grid_x((1, 1, 1), TET1_INPUT_VERTEX))
    assert fem.functionspace(m, ("Lagrange", 1)).dofmap.list.tolist() == TET1_CELLS
    V2 = fem.functionspace(m, ("Lagrange", 2))
    assert V2.dofmap.index_map_size == 27 and V2.dofmap.list.tolist() == TET1_P2


def test_hex_box_2x1x1_oracle_and_product():
    om = O.Mesh((2, 1, 1), O.HEX)
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (2, 1, 1), mesh.CellType.hexahedron)
    for orig, inp, cells, x in [
            (om.original_cell_index, om.input_vertex_index, om.cells, om.x),
            (m.topology.original_cell_index, m.geometry.input_global_indices, m.topology.cells,
             m.geometry.x)]:
        assert orig.tolist() == HEX2_ORIGINAL
        assert inp.tolist() == HEX2_INPUT_VERTEX
        assert cells.tolist() == HEX2_CELLS
        assert np.array_equal(x, _grid_x((2, 1, 1), HEX2_INPUT_VERTEX))
    dm, n = om.dofmap(1)
    assert n == 12 and dm.tolist() == HEX2_CELLS
    assert fem.functionspace(m, ("Lagrange", 1)).dofmap.list.tolist() == HEX2_CELLS


# ---- P3 element: nodes from the Legendre polynomial ----------------------
TET_V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
BASIX_TET_EDGES = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]
BASIX_TET_FACES = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
CUBIC = [(a, b, c) for a, b, c in iproduct(range(4), repeat=3) if a + b + c <= 3]


def _p3_nodes():
    # interior Gauss-Lobatto-Legendre points of degree 3: roots of P_3'
    r = np.sort(np.polynomial.legendre.Legendre.basis(3).deriv().roots())
    t = (r + 1) / 2
    pts = list(TET_V)
    for a, b in BASIX_TET_EDGES:
        pts += [TET_V[a] + s * (TET_V[b] - TET_V[a]) for s in t]
    pts += [TET_V[list(f)].mean(0) for f in BASIX_TET_FACES]
    return np.array(pts)


def _tet_rule(order=6):
    """Collapsed (Duffy) Gauss-Legendre rule on the reference tetrahedron,
    exact for polynomials of degree <= 2 order - 3."""
    g, w = np.polynomial.legendre.leggauss(order)
    g, w = (g + 1) / 2, w / 2
    pts, wts = [], []
    for (u, wu), (v, wv), (s, ws) in iproduct(zip(g, w), zip(g, w), zip(g, w)):
        x = u
        y = v * (1 - u)
        z = s * (1 - u) * (1 - v)
        pts.append((x, y, z))
        wts.append(wu * wv * ws * (1 - u) ** 2 * (1 - v))
    return np.array(pts), np.array(wts)


def _p3_tensors(X):
    """Mass and stiffness matrices of the nodal P3 basis on the physical tet
    X (affine map from the reference tet), exact for cubic polynomials."""
    nodes = _p3_nodes()
    V = np.array([[p[0] ** a * p[1] ** b * p[2] ** c for a, b, c in CUBIC] for p in nodes])
    C = np.linalg.inv(V)  # column i = monomial coefficients of basis i
    q, w = _tet_rule()
    J = (X[1:] - X[0]).T
    det = abs(np.linalg.det(J))
    Jinv = np.linalg.inv(J)
    mono = np.array([[x ** a * y ** b * z ** c for a, b, c in CUBIC] for x, y, z in q])

    def dmono(k):
        out = []
        for x, y, z in q:
            row = []
            for e in CUBIC:
                p = [x, y, z]
                if e[k] == 0:
                    row.append(0.0)
                    continue
                ee = list(e)
                ee[k] -= 1
                row.append(e[k] * p[0] ** ee[0] * p[1] ** ee[1] * p[2] ** ee[2])
            out.append(row)
        return np.array(out)

    phi = mono @ C
    dphi = np.stack([dmono(k) @ C for k in range(3)], 2)  # (q, i, reference k)
    gphys = dphi @ Jinv  # (q, i, physical)
    M = det * np.einsum("q,qi,qj->ij", w, phi, phi)
    K = det * np.einsum("q,qid,qjd->ij", w, gphys, gphys)
    return M, K


def test_p3_nodes_are_gauss_lobatto_points():
    t = (np.sort(np.polynomial.legendre.Legendre.basis(3).deriv().roots()) + 1) / 2
    assert np.allclose(t, [0.5 - 0.5 / np.sqrt(5), 0.5 + 0.5 / np.sqrt(5)], rtol=0, atol=1e-15)
    rng = np.random.default_rng(7)
    for X in [TET_V, TET_V @ np.diag([2.0, 0.5, 1.5]) + 0.3,
              rng.standard_normal((4, 3)) + np.vstack([np.zeros(3), 2 * np.eye(3)])]:
        M, K = _p3_tensors(X)
        Mo = O.tabulate(O.K_P3TET_A, X, w=[0.0], c=[1.0], ndim=400).reshape(20, 20)
        Ko = O.tabulate(O.K_P3TET_A, X, w=[1.0], c=[0.0], ndim=400).reshape(20, 20)
        s = np.abs(M).max()
        assert np.abs(Mo - M).max() <= 1e-12 * s
        assert np.abs(Ko - K).max() <= 1e-12 * np.abs(K).max()


def test_p3_equispaced_nodes_would_fail():
    """The check above separates the node families: with equispaced edge
    nodes (1/3, 2/3) the mass matrix differs at the 1e-3 level."""
    nodes = _p3_nodes()
    eq = nodes.copy()
    for e, (a, b) in enumerate(BASIX_TET_EDGES):
        for k, s in enumerate([1 / 3, 2 / 3]):
            eq[4 + 2 * e + k] = TET_V[a] + s * (TET_V[b] - TET_V[a])
    V = np.array([[p[0] ** a * p[1] ** b * p[2] ** c for a, b, c in CUBIC] for p in eq])
    C = np.linalg.inv(V)
    q, w = _tet_rule()
    mono = np.array([[x ** a * y ** b * z ** c for a, b, c in CUBIC] for x, y, z in q])
    phi = mono @ C
    Meq = np.einsum("q,qi,qj->ij", w, phi, phi)
    Mo = O.tabulate(O.K_P3TET_A, TET_V, w=[0.0], c=[1.0], ndim=400).reshape(20, 20)
    assert np.abs(Mo - Meq).max() > 1e-4
