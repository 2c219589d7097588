"""Facet integrals with a host kernel in the reference's kernel ABI
(integral_data{kernel, entities} of an exterior_facet / interior_facet
integral, fem/Form.h:52-87): the kernel runs per facet on the host as
assemble_exterior_facets and assemble_interior_facets call it
(fem/assemble_matrix_impl.h:264-383, :442-667;
fem/assemble_vector_impl.h:118-174), its tensors are added on the device
in facet order.

Kernels passed: the oracle's facet kernels (Robin and Neumann on an
exterior facet, the gradient-jump penalty on an interior facet) through
their C function pointers, and, for FP32, a ctypes kernel that calls the
same oracle kernel in double. Checked against the oracle's serial facet
assembly (FP64 within 1e-12, FP32 within 2e-6, patterns bit-exact) and
against the device forms of the same integrals; an interior facet kernel
on Q1 hexahedra (written here) against the same kernel assembled on the
host in facet order."""

import ctypes as C

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import fem, la, mesh
from dolfinx_amd._lib import DfxError

from conftest import rel_err
from test_gpu_kernel_abi import _oracle_kernel
from test_oracle import pattern_with_interior

pytestmark = pytest.mark.gpu

_ORC_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double),
                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                      C.POINTER(C.c_uint8), C.c_void_p)


def _f32_kernel(kid, nA, nw, nc, nx):
    """A KERNEL_F32 function computing in double through the oracle's
    kernel kid: nA tensor entries, nw coefficients, nc constants, nx
    coordinate values."""
    fn = _ORC_FN(_oracle_kernel(kid))

    def k(A, w, c, xc, ent, perm, data):
        Ad = np.zeros(nA)
        wd = np.ctypeslib.as_array(w, (max(nw, 1),)).astype(np.float64)
        cd = np.ctypeslib.as_array(c, (max(nc, 1),)).astype(np.float64)
        xd = np.ctypeslib.as_array(xc, (nx,)).astype(np.float64)
        dp = C.POINTER(C.c_double)
        fn(Ad.ctypes.data_as(dp), wd.ctypes.data_as(dp), cd.ctypes.data_as(dp),
           xd.ctypes.data_as(dp), ent, perm, None)
        out = np.ctypeslib.as_array(A, (nA,))
        out[:] += Ad.astype(np.float32)
    return fem.KERNEL_F32(k)


def _kernel(kid, dtype, nA, nw, nc, nx):
    if dtype == np.float64:
        return _oracle_kernel(kid)
    return _f32_kernel(kid, nA, nw, nc, nx)


def _tet(n=(4, 3, 2)):
    om = O.Mesh(n, O.TET)
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    assert np.array_equal(V.dofmap.list, dm)
    return om, pm, V, dm, nd


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_interior_kernel_matches_oracle_and_device_form(dtype, tol):
    om, pm, V, dm, nd = _tet()
    q = fem.interior_facet_domain(pm)
    rp, cols = pattern_with_interior(dm, nd, q)
    gamma = 0.8
    aK = fem.kernel_facet_form(V, 2, _kernel(O.K_P1TET_GRADJUMP_A, dtype, 64, 0, 1, 24),
                               "interior_facet", q, constants=[gamma], dtype=dtype)
    bn = om.boundary_nodes(1, dm, marker=1, axis=1, value=0.0)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    bc = fem.dirichletbc(V, bn)
    A = fem.assemble_matrix(aK, [bc])
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    ref = O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark, mark, rp,
                                     cols, [gamma])
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data.astype(np.float64), ref) <= tol
    # the device form of the same integral
    Ad = fem.assemble_matrix(fem.form("p1_gradient_jump", V, None, [gamma], dtype, facets=q),
                             [bc])
    assert rel_err(A.data.astype(np.float64), Ad.data.astype(np.float64)) <= tol
    # adds into the same pattern
    fem.assemble_matrix(aK, [bc], 1.0, A)
    O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark, mark, rp, cols,
                               [gamma], ref)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data.astype(np.float64), ref) <= tol


def test_cell_kernel_plus_interior_kernel():
    """A form with a cell integral and an interior facet integral, both
    host kernels: cells first, then the interior facets."""
    om, pm, V, dm, nd = _tet((3, 3, 2))
    q = fem.interior_facet_domain(pm)
    rp, cols = pattern_with_interior(dm, nd, q)
    kap_in = np.random.default_rng(5).uniform(1.0, 2.0, om.num_cells)
    aC = fem.kernel_form(V, 2, _oracle_kernel(O.K_P1TET_A),
                         coeffs=fem.pack_cellwise(pm, kap_in), constants=[0.1])
    aF = fem.kernel_facet_form(V, 2, _oracle_kernel(O.K_P1TET_GRADJUMP_A), "interior_facet", q,
                               constants=[1.5])
    bn = om.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    A = fem.assemble_matrix((aF, aC), [fem.dirichletbc(V, bn)])
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    kap = kap_in[om.original_cell_index].reshape(-1, 1)
    ref = O.assemble_matrix(O.K_P1TET_A, om.x, om.cells, dm, 1, kap, np.array([0.1]), mark, mark,
                            rp, cols)
    O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, om.x, om.cells, dm, q, mark, mark, rp, cols,
                               [1.5], ref)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data, ref) <= 1e-12


ROBIN = [(O.TET, "p1_robin", O.K_P1TET_ROBIN_A, 4), (O.HEX, "q1_robin", O.K_Q1HEX_ROBIN_A, 8)]


@pytest.mark.parametrize("ct,rname,kR,npc", ROBIN)
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_exterior_kernel_matrix_matches_oracle(ct, rname, kR, npc, dtype, tol):
    n = (4, 3, 3)
    om = O.Mesh(n, ct)
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType(ct))
    V = fem.functionspace(pm, ("Lagrange", 1))
    dm, nd = om.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    beta_in = np.random.default_rng(21).uniform(0.5, 2.0, om.num_cells)
    beta = fem.pack_cellwise(pm, beta_in)
    pairs = fem.facet_domain(pm)
    aK = fem.kernel_facet_form(V, 2, _kernel(kR, dtype, npc * npc, 1, 0, 3 * npc),
                               "exterior_facet", pairs, coeffs=beta, dtype=dtype)
    bn = om.boundary_nodes(1, dm, marker=1, axis=0, value=0.0)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    bc = fem.dirichletbc(V, bn)
    A = fem.assemble_matrix(aK, [bc])
    assert np.array_equal(A.indptr, rp) and np.array_equal(A.indices, cols)
    co = beta_in[om.original_cell_index].reshape(-1, 1)
    ref = O.assemble_matrix_facets(kR, om.x, om.cells, dm, co, pairs, mark, mark, rp, cols)
    O.set_diagonal(ref, bn, 1, 1.0, rp, cols)
    assert rel_err(A.data.astype(np.float64), ref) <= tol
    Ad = fem.assemble_matrix(fem.form(rname, V, beta, None, dtype, facets=pairs), [bc])
    assert rel_err(A.data.astype(np.float64), Ad.data.astype(np.float64)) <= tol


@pytest.mark.parametrize("axis,value", [(None, 0.0), (0, 1.0)])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_exterior_kernel_vector_matches_oracle(axis, value, dtype, tol):
    om, pm, V, dm, nd = _tet((4, 3, 5))
    xin = np.empty_like(om.x)
    xin[om.input_vertex_index] = om.x
    g_in = 1.0 + xin[:, 0] + 2.0 * xin[:, 1] ** 2 - xin[:, 2]
    pairs = fem.facet_domain(pm, axis, value)
    gv = fem.pack_vertexwise(pm, g_in)
    LK = fem.kernel_facet_form(V, 1, _kernel(O.K_P1TET_NEUMANN, dtype, 4, 4, 0, 12),
                               "exterior_facet", pairs, coeffs=gv, dtype=dtype)
    b = fem.assemble_vector(LK)
    g = g_in[om.input_vertex_index][om.cells]
    ref = O.assemble_vector_facets(O.K_P1TET_NEUMANN, om.x, om.cells, dm, g, pairs, nd)
    assert rel_err(b.array.astype(np.float64), ref) <= tol
    bd = fem.assemble_vector(fem.form("p1_neumann", V, gv, None, dtype, facets=pairs))
    assert rel_err(b.array.astype(np.float64), bd.array.astype(np.float64)) <= tol
    fem.assemble_vector(LK, b)
    assert rel_err(b.array.astype(np.float64), 2 * ref) <= tol


def test_kernel_facet_forms_reject_misuse():
    om, pm, V, dm, nd = _tet((2, 2, 2))
    q = fem.interior_facet_domain(pm)
    k = _oracle_kernel(O.K_P1TET_GRADJUMP_A)
    with pytest.raises(DfxError, match="bilinear"):
        fem.kernel_facet_form(V, 1, k, "interior_facet", q, constants=[1.0])
    with pytest.raises(DfxError, match="out of range"):
        fem.kernel_facet_form(V, 2, k, "exterior_facet", np.array([[0, 4]], np.int32))
    with pytest.raises(ValueError):
        fem.kernel_facet_form(V, 2, k, "cell", q)
    aF = fem.kernel_facet_form(V, 2, k, "interior_facet", q, constants=[1.0])
    # a pattern of the cell couplings only has no room for the facet couplings
    aC = fem.kernel_form(V, 2, _oracle_kernel(O.K_P1TET_A), coeffs=np.ones((om.num_cells, 1)),
                         constants=[0.0])
    A = fem.create_matrix(aC)
    with pytest.raises(DfxError, match="Entry not in sparsity"):
        fem.assemble_matrix(aF, [], 1.0, A)
    # lifting with nonzero Dirichlet data through such a form
    bn = om.boundary_nodes(1, dm)
    g = np.ones(nd)
    b = la.Vector(nd)
    with pytest.raises(DfxError, match="host kernel"):
        fem.apply_lifting(b, [aF], [[fem.dirichletbc(V, bn, g)]])


# DOLFINx's hexahedron facets as local vertex numbers (facet f of the
# reference cell), the first three spanning the face
_HEX_FACETS = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6), (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)]


def _hex_face_tensor(X, lf0, lf1, gamma):
    """A 16 x 16 interior facet tensor on the joint dofs (cell 0's 8, then
    cell 1's 8): gamma |F| times a pattern coupling the face nodes of both
    sides, different per local facet pair, so every index of the joint
    layout is exercised."""
    f0, f1 = _HEX_FACETS[lf0], _HEX_FACETS[lf1]
    p = X[f0[0]]
    area = np.linalg.norm(np.cross(X[f0[1]] - p, X[f0[2]] - p))
    M = np.zeros((16, 16))
    for u, a in enumerate(f0):
        for v, b in enumerate(f1):
            w = gamma * area * (1.0 + 0.1 * u + 0.01 * v + 0.001 * lf0 + 0.0001 * lf1)
            M[a, 8 + b] += w
            M[8 + b, a] += w
            M[a, a] += 0.5 * w
            M[8 + b, 8 + b] += 0.25 * w
    return M


def test_interior_kernel_on_hexahedra_matches_host_assembly():
    """An interior facet integral on Q1 hexahedra given as a ctypes kernel:
    the device path (both cells' geometry, local facets, joint dofmap, BC
    rows and columns, insertion in facet order) against the same kernel
    assembled on the host in facet order."""
    n = (3, 3, 2)
    pm = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.hexahedron)
    V = fem.functionspace(pm, ("Lagrange", 1))
    q = fem.interior_facet_domain(pm)
    assert len(q) == 2 * 3 * 2 + 3 * 2 * 2 + 3 * 3 * 1   # faces between neighbours
    gamma = 1.3

    def k(A, w, c, xc, ent, perm, data):
        X = np.ctypeslib.as_array(xc, (48,)).reshape(16, 3)
        out = np.ctypeslib.as_array(A, (256,))
        out[:] += _hex_face_tensor(X, ent[0], ent[1], c[0]).ravel()

    kern = fem.KERNEL_F64(k)
    aF = fem.kernel_facet_form(V, 2, kern, "interior_facet", q, constants=[gamma])
    bn = fem.locate_boundary_dofs(V, 0, 0.0)
    A = fem.assemble_matrix(aF, [fem.dirichletbc(V, bn)])
    x = pm.geometry.x
    xd = pm.geometry.dofmap
    dm = V.dofmap.list
    N = V.num_dofs
    ref = np.zeros((N, N))
    mark = np.zeros(N, bool)
    mark[bn] = True
    for c0, l0, c1, l1 in q:
        X = np.concatenate([x[xd[c0]], x[xd[c1]]])
        M = _hex_face_tensor(X, l0, l1, gamma)
        joint = np.concatenate([dm[c0], dm[c1]])
        M[mark[joint], :] = 0.0
        M[:, mark[joint]] = 0.0
        for i in range(16):
            for jj in range(16):
                ref[joint[i], joint[jj]] += M[i, jj]
    ref[bn, bn] = 1.0
    D = A.to_scipy().toarray()
    assert np.abs(D - ref).max() <= 1e-14 * np.abs(ref).max()
