"""CPU oracle for the DOLFINx assemble-and-solve path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, always as the checker. The product package
``dolfinx_amd`` never imports it.

The C sources (mesh.c, fem.c, la.c) restate the reference algorithms; each
function cites the reference file:line it follows.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TET, HEX = 0, 1
K_P1TET_A, K_P1TET_L, K_Q1HEX_A, K_Q1HEX_L = 0, 1, 2, 3
K_P2TET_ELAST_A, K_P2TET_ELAST_L, K_P3TET_A, K_P3TET_L = 4, 5, 6, 7
K_P1TET_NEUMANN, K_Q1HEX_NEUMANN = 10, 11
K_P1TET_ROBIN_A, K_Q1HEX_ROBIN_A = 12, 13
K_P1TET_GRADJUMP_A = 14

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def use_native() -> str:
    """Rebuild the oracle with -O3 -march=native on this machine (for the
    CPU baseline) and use that copy from now on; returns the -march used
    ("native", or "x86-64-v3" when the native build is not possible). Must
    run before any other call of this module."""
    global _lib
    if _lib is not None:
        raise RuntimeError("oracle.use_native() must run before the oracle is first used")
    try:
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "native"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        _lib = C.CDLL(os.path.join(_HERE, "_native", "liboracle.so"))
        _declare(_lib)
        return "native"
    except (OSError, subprocess.CalledProcessError):
        lib()
        return "x86-64-v3"


P = C.c_void_p
I32, I64, D = C.c_int32, C.c_int64, C.c_double


class _Mesh(C.Structure):
    _fields_ = [
        ("celltype", C.c_int),
        ("nvpc", C.c_int),
        ("num_cells", I32),
        ("num_vertices", I32),
        ("x", C.POINTER(D)),
        ("cells", C.POINTER(I32)),
        ("original_cell_index", C.POINTER(I64)),
        ("input_vertex_index", C.POINTER(I64)),
    ]


def _declare(L):
    L.orc_mesh_create_box.restype = C.POINTER(_Mesh)
    L.orc_mesh_create_box.argtypes = [P, P, P, C.c_int]
    L.orc_mesh_free.argtypes = [C.POINTER(_Mesh)]
    L.orc_mesh_create.restype = C.POINTER(_Mesh)
    L.orc_mesh_create.argtypes = [C.c_int, P, I32, P, I32]
    L.orc_dual_graph.argtypes = [C.c_int, P, I32, C.POINTER(P), C.POINTER(P)]
    L.orc_rcm.argtypes = [P, P, I32, P]
    L.orc_entities.restype = I32
    L.orc_entities.argtypes = [C.POINTER(_Mesh), C.c_int, C.POINTER(P), C.POINTER(P)]
    L.orc_edge_reflections.argtypes = [C.POINTER(_Mesh), P, P, P]
    L.orc_exterior_facets.restype = I32
    L.orc_exterior_facets.argtypes = [C.POINTER(_Mesh), C.POINTER(P)]
    L.orc_element_ndofs.argtypes = [C.c_int, C.c_int]
    L.orc_dofmap.restype = I32
    L.orc_dofmap.argtypes = [C.POINTER(_Mesh), C.c_int, P]
    L.orc_boundary_nodes.restype = I32
    L.orc_boundary_nodes.argtypes = [C.POINTER(_Mesh), C.c_int, P, C.c_int, C.c_int, D, C.POINTER(P)]
    L.orc_sparsity.restype = I64
    L.orc_sparsity.argtypes = [P, C.c_int, P, C.c_int, I32, I32, P, C.POINTER(P)]
    L.orc_tabulate.argtypes = [C.c_int, P, P, P, P]
    L.orc_assemble_matrix.restype = C.c_int
    L.orc_assemble_matrix.argtypes = [C.c_int, P, P, C.c_int, I32, P, C.c_int, C.c_int, P,
                                      C.c_int, P, P, P, P, P, P]
    L.orc_assemble_vector.argtypes = [C.c_int, P, P, C.c_int, I32, P, C.c_int, C.c_int, P,
                                      C.c_int, P, P]
    L.orc_facet_domain.restype = I32
    L.orc_facet_domain.argtypes = [C.POINTER(_Mesh), C.c_int, C.c_int, D, C.POINTER(P)]
    L.orc_assemble_vector_facets.argtypes = [C.c_int, P, P, C.c_int, P, C.c_int, C.c_int, P,
                                             C.c_int, P, P, I32, P]
    L.orc_interior_facet_domain.restype = I32
    L.orc_interior_facet_domain.argtypes = [C.POINTER(_Mesh), C.POINTER(P)]
    L.orc_assemble_matrix_interior.restype = C.c_int
    L.orc_assemble_matrix_interior.argtypes = [C.c_int, P, P, C.c_int, P, C.c_int, P, P, I32, P,
                                               P, P, P, P]
    L.orc_lift_bc_interior.argtypes = [C.c_int, P, P, C.c_int, P, C.c_int, P, P, I32, P, P, P, D,
                                       P]
    L.orc_assemble_matrix_facets.restype = C.c_int
    L.orc_assemble_matrix_facets.argtypes = [C.c_int, P, P, C.c_int, P, C.c_int, C.c_int, P,
                                             C.c_int, P, P, I32, P, P, P, P, P]
    L.orc_lift_bc_facets.argtypes = [C.c_int, P, P, C.c_int, P, C.c_int, C.c_int, P, C.c_int,
                                     P, P, I32, P, P, P, D, P]
    L.orc_add_element_matrices.restype = C.c_int
    L.orc_add_element_matrices.argtypes = [I32, P, C.c_int, C.c_int, P, P, P, P, P, P]
    L.orc_add_element_vectors.argtypes = [I32, P, C.c_int, C.c_int, P, P]
    L.orc_lift_bc.argtypes = [C.c_int, P, P, C.c_int, I32, P, C.c_int, C.c_int, P, C.c_int,
                              P, P, P, P, D, P]
    L.orc_bc_set.argtypes = [P, P, I32, P, P, D]
    L.orc_set_diagonal.restype = C.c_int
    L.orc_set_diagonal.argtypes = [P, I32, C.c_int, D, P, P, P]
    L.orc_spmv.argtypes = [I32, C.c_int, P, P, P, P, P]
    L.orc_dot.restype = D
    L.orc_dot.argtypes = [I32, P, P]
    L.orc_cg.restype = C.c_int
    L.orc_cg.argtypes = [I32, C.c_int, P, P, P, P, P, C.c_int, D]
    L.orc_pcg_jacobi.restype = C.c_int
    L.orc_pcg_jacobi.argtypes = [I32, C.c_int, P, P, P, P, P, C.c_int, D]
    L.orc_free.argtypes = [P]


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _take(ptr, n, dtype):
    """Copy a malloc'ed C array into numpy and free it."""
    ct = {np.int32: I32, np.int64: I64, np.float64: D}[dtype]
    if n == 0:
        out = np.zeros(0, dtype)
    else:
        out = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()
    lib().orc_free(ptr)
    return out


class Mesh:
    """Oracle mesh: create_box -> dual graph -> RCM -> topology."""

    def __init__(self, n, celltype, p0=(0.0, 0.0, 0.0), p1=(1.0, 1.0, 1.0), _ptr=None):
        L = lib()
        if _ptr is None:
            p0 = np.asarray(p0, np.float64)
            p1 = np.asarray(p1, np.float64)
            nn = np.asarray(n, np.int64)
            mp = L.orc_mesh_create_box(_p(p0), _p(p1), _p(nn), celltype)
        else:
            mp = _ptr
        m = mp.contents
        self.celltype = celltype
        self.nvpc = m.nvpc
        self.num_cells = m.num_cells
        self.num_vertices = m.num_vertices
        self.x = np.ctypeslib.as_array(m.x, shape=(m.num_vertices, 3)).copy()
        self.cells = np.ctypeslib.as_array(m.cells, shape=(m.num_cells, m.nvpc)).copy()
        self.original_cell_index = np.ctypeslib.as_array(
            m.original_cell_index, shape=(m.num_cells,)).copy()
        self.input_vertex_index = np.ctypeslib.as_array(
            m.input_vertex_index, shape=(m.num_vertices,)).copy()
        self._ptr = mp

    @classmethod
    def from_cells(cls, celltype, cells, x):
        """mesh::create_mesh on one rank from input cells (input vertex
        indices, DOLFINx vertex order) and input geometry."""
        cells = np.ascontiguousarray(cells, np.int64)
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        mp = lib().orc_mesh_create(celltype, _p(cells), cells.shape[0], _p(x), x.shape[0])
        return cls(None, celltype, _ptr=mp)

    def __del__(self):
        try:
            lib().orc_mesh_free(self._ptr)
        except Exception:
            pass

    def entities(self, dim):
        c2e, e2v = P(), P()
        n = lib().orc_entities(self._ptr, dim, C.byref(c2e), C.byref(e2v))
        ne = lib().orc_element_ndofs  # placeholder to keep lib alive
        del ne
        nce = {(TET, 1): 6, (TET, 2): 4, (HEX, 1): 12, (HEX, 2): 6}[(self.celltype, dim)]
        nve = 2 if dim == 1 else (3 if self.celltype == TET else 4)
        a = _take(c2e.value, self.num_cells * nce, np.int32).reshape(self.num_cells, nce)
        b = _take(e2v.value, n * nve, np.int32).reshape(n, nve)
        return a, b

    def edge_reflections(self):
        c2e, e2v = self.entities(1)
        bits = np.zeros(self.num_cells, np.uint32)
        lib().orc_edge_reflections(self._ptr, _p(np.ascontiguousarray(c2e)),
                                   _p(np.ascontiguousarray(e2v)), _p(bits))
        return bits

    def exterior_facets(self):
        out = P()
        n = lib().orc_exterior_facets(self._ptr, C.byref(out))
        return _take(out.value, 2 * n, np.int32).reshape(n, 2)

    def interior_facet_domain(self):
        """(cell0, lf0, cell1, lf1) of every interior facet, facet-index
        order."""
        out = P()
        n = lib().orc_interior_facet_domain(self._ptr, C.byref(out))
        return _take(out.value, 4 * n, np.int32).reshape(n, 4)

    def facet_domain(self, axis=None, value=0.0):
        """(cell, local facet) of the exterior facets in facet-index order,
        or of those with every vertex at x[axis] == value."""
        out = P()
        n = lib().orc_facet_domain(self._ptr, 0 if axis is None else 1,
                                   0 if axis is None else int(axis), float(value), C.byref(out))
        return _take(out.value, 2 * n, np.int32).reshape(n, 2)

    def dofmap(self, degree):
        k = lib().orc_element_ndofs(self.celltype, degree)
        dm = np.zeros((self.num_cells, k), np.int32)
        n = lib().orc_dofmap(self._ptr, degree, _p(dm))
        return dm, n

    def boundary_nodes(self, degree, dofmap, marker=0, axis=0, value=0.0):
        out = P()
        n = lib().orc_boundary_nodes(self._ptr, degree, _p(np.ascontiguousarray(dofmap)),
                                     marker, axis, value, C.byref(out))
        return _take(out.value, n, np.int32)


def dual_graph(celltype, cells):
    cells = np.ascontiguousarray(cells, np.int64)
    nc = cells.shape[0]
    off, data = P(), P()
    ne = lib().orc_dual_graph(celltype, _p(cells), nc, C.byref(off), C.byref(data))
    return _take(off.value, nc + 1, np.int32), _take(data.value, 2 * ne, np.int32)


def rcm(offsets, data):
    n = len(offsets) - 1
    remap = np.zeros(n, np.int32)
    lib().orc_rcm(_p(np.ascontiguousarray(offsets, np.int32)),
                  _p(np.ascontiguousarray(data, np.int32)), n, _p(remap))
    return remap


def sparsity(dofmap0, dofmap1, nrows):
    d0 = np.ascontiguousarray(dofmap0, np.int32)
    d1 = np.ascontiguousarray(dofmap1, np.int32)
    row_ptr = np.zeros(nrows + 1, np.int64)
    cols = P()
    nnz = lib().orc_sparsity(_p(d0), d0.shape[1], _p(d1), d1.shape[1], d0.shape[0], nrows,
                             _p(row_ptr), C.byref(cols))
    return row_ptr, _take(cols.value, nnz, np.int32)


def tabulate(kernel, coords, w=None, c=None, ndim=None):
    coords = np.ascontiguousarray(coords, np.float64)
    w = np.ascontiguousarray(np.zeros(1) if w is None else w, np.float64)
    c = np.ascontiguousarray(np.zeros(3) if c is None else c, np.float64)
    A = np.zeros(ndim, np.float64)
    lib().orc_tabulate(kernel, _p(A), _p(w), _p(c), _p(coords))
    return A


def assemble_matrix(kernel, x, x_dofmap, dofmap, bs, coeffs, constants, bc0, bc1,
                    row_ptr, cols):
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    co = None if coeffs is None else np.ascontiguousarray(coeffs, np.float64)
    cs = 0 if co is None else (co.shape[1] if co.ndim == 2 else 1)
    kc = None if constants is None else np.ascontiguousarray(constants, np.float64)
    b0 = None if bc0 is None else np.ascontiguousarray(bc0, np.int8)
    b1 = None if bc1 is None else np.ascontiguousarray(bc1, np.int8)
    vals = np.zeros(len(cols) * bs * bs, np.float64)
    rc = lib().orc_assemble_matrix(kernel, _p(x), _p(xd), xd.shape[1], xd.shape[0], _p(dm),
                                   dm.shape[1], bs, _p(co), cs, _p(kc), _p(b0), _p(b1),
                                   _p(np.ascontiguousarray(row_ptr, np.int64)),
                                   _p(np.ascontiguousarray(cols, np.int32)), _p(vals))
    if rc != 0:
        raise RuntimeError("Entry not in sparsity")
    return vals


def assemble_vector(kernel, x, x_dofmap, dofmap, bs, coeffs, constants, n):
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    co = None if coeffs is None else np.ascontiguousarray(coeffs, np.float64)
    cs = 0 if co is None else (co.shape[1] if co.ndim == 2 else 1)
    kc = None if constants is None else np.ascontiguousarray(constants, np.float64)
    b = np.zeros(n, np.float64)
    lib().orc_assemble_vector(kernel, _p(x), _p(xd), xd.shape[1], xd.shape[0], _p(dm),
                              dm.shape[1], bs, _p(co), cs, _p(kc), _p(b))
    return b


def assemble_vector_facets(kernel, x, x_dofmap, dofmap, coeffs, pairs, n, constants=None):
    """b from an exterior-facet kernel over (cell, local facet) pairs."""
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    co = np.ascontiguousarray(coeffs, np.float64)
    kc = None if constants is None else np.ascontiguousarray(constants, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32)
    b = np.zeros(n, np.float64)
    lib().orc_assemble_vector_facets(kernel, _p(x), _p(xd), xd.shape[1], _p(dm), dm.shape[1], 1,
                                     _p(co), co.shape[1], _p(kc), _p(pr), pr.shape[0], _p(b))
    return b


def assemble_matrix_facets(kernel, x, x_dofmap, dofmap, coeffs, pairs, bc0, bc1, row_ptr, cols,
                           values=None, constants=None):
    """values += the exterior facet matrices over (cell, local facet) pairs,
    BC rows (bc0) and columns (bc1) zeroed (scalar spaces)."""
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    co = np.ascontiguousarray(coeffs, np.float64).reshape(xd.shape[0], -1)
    kc = None if constants is None else np.ascontiguousarray(constants, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32)
    b0 = None if bc0 is None else np.ascontiguousarray(bc0, np.int8)
    b1 = None if bc1 is None else np.ascontiguousarray(bc1, np.int8)
    if values is None:
        values = np.zeros(len(cols), np.float64)
    rc = lib().orc_assemble_matrix_facets(
        kernel, _p(x), _p(xd), xd.shape[1], _p(dm), dm.shape[1], 1, _p(co), co.shape[1], _p(kc),
        _p(pr), pr.shape[0], _p(b0), _p(b1), _p(np.ascontiguousarray(row_ptr, np.int64)),
        _p(np.ascontiguousarray(cols, np.int32)), _p(values))
    if rc != 0:
        raise RuntimeError("Entry not in sparsity")
    return values


def assemble_matrix_interior(kernel, x, x_dofmap, dofmap, quads, bc0, bc1, row_ptr, cols,
                             constants, values=None):
    """values += an interior facet integral over (cell0, lf0, cell1, lf1)
    quads, BC rows (bc0) and columns (bc1) zeroed (scalar spaces)."""
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    q = np.ascontiguousarray(quads, np.int32).reshape(-1, 4)
    kc = np.ascontiguousarray(np.atleast_1d(constants), np.float64)
    b0 = None if bc0 is None else np.ascontiguousarray(bc0, np.int8)
    b1 = None if bc1 is None else np.ascontiguousarray(bc1, np.int8)
    if values is None:
        values = np.zeros(len(cols), np.float64)
    rc = lib().orc_assemble_matrix_interior(
        kernel, _p(x), _p(xd), xd.shape[1], _p(dm), dm.shape[1], _p(kc), _p(q), q.shape[0],
        _p(b0), _p(b1), _p(np.ascontiguousarray(row_ptr, np.int64)),
        _p(np.ascontiguousarray(cols, np.int32)), _p(values))
    if rc != 0:
        raise RuntimeError("Entry not in sparsity")
    return values


def lift_bc_interior(b, kernel, x, x_dofmap, dofmap, quads, constants, bc_values, bc_markers,
                     x0=None, alpha=1.0):
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    q = np.ascontiguousarray(quads, np.int32).reshape(-1, 4)
    kc = np.ascontiguousarray(np.atleast_1d(constants), np.float64)
    lib().orc_lift_bc_interior(kernel, _p(x), _p(xd), xd.shape[1], _p(dm), dm.shape[1], _p(kc),
                               _p(q), q.shape[0],
                               _p(np.ascontiguousarray(bc_values, np.float64)),
                               _p(np.ascontiguousarray(bc_markers, np.int8)),
                               None if x0 is None else _p(np.ascontiguousarray(x0, np.float64)),
                               alpha, _p(b))
    return b


def lift_bc_facets(b, kernel, x, x_dofmap, dofmap, coeffs, pairs, bc_values, bc_markers,
                   x0=None, alpha=1.0):
    """b -= alpha F (g - x0) over the BC columns of each facet matrix F."""
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    co = np.ascontiguousarray(coeffs, np.float64).reshape(xd.shape[0], -1)
    pr = np.ascontiguousarray(pairs, np.int32)
    lib().orc_lift_bc_facets(kernel, _p(x), _p(xd), xd.shape[1], _p(dm), dm.shape[1], 1, _p(co),
                             co.shape[1], None, _p(pr), pr.shape[0],
                             _p(np.ascontiguousarray(bc_values, np.float64)),
                             _p(np.ascontiguousarray(bc_markers, np.int8)),
                             None if x0 is None else _p(np.ascontiguousarray(x0, np.float64)),
                             alpha, _p(b))
    return b


def add_element_matrices(dofmap, bs, E, bc0, bc1, row_ptr, cols, values=None):
    """values += the element matrices E (cells x (bs k) x (bs k)) inserted
    cell by cell, BC rows (bc0) and columns (bc1) zeroed first."""
    dm = np.ascontiguousarray(dofmap, np.int32)
    E = np.ascontiguousarray(E, np.float64)
    b0 = None if bc0 is None else np.ascontiguousarray(bc0, np.int8)
    b1 = None if bc1 is None else np.ascontiguousarray(bc1, np.int8)
    if values is None:
        values = np.zeros(len(cols) * bs * bs, np.float64)
    rc = lib().orc_add_element_matrices(dm.shape[0], _p(dm), dm.shape[1], bs, _p(E), _p(b0),
                                        _p(b1), _p(np.ascontiguousarray(row_ptr, np.int64)),
                                        _p(np.ascontiguousarray(cols, np.int32)), _p(values))
    if rc != 0:
        raise RuntimeError("Entry not in sparsity")
    return values


def add_element_vectors(dofmap, bs, e, n, b=None):
    """b += the element vectors e (cells x bs k) scattered cell by cell."""
    dm = np.ascontiguousarray(dofmap, np.int32)
    e = np.ascontiguousarray(e, np.float64)
    if b is None:
        b = np.zeros(n, np.float64)
    lib().orc_add_element_vectors(dm.shape[0], _p(dm), dm.shape[1], bs, _p(e), _p(b))
    return b


def lift_bc(b, kernel, x, x_dofmap, dofmap, bs, coeffs, constants, bc_values, bc_markers,
            x0=None, alpha=1.0):
    x = np.ascontiguousarray(x, np.float64)
    xd = np.ascontiguousarray(x_dofmap, np.int32)
    dm = np.ascontiguousarray(dofmap, np.int32)
    co = None if coeffs is None else np.ascontiguousarray(coeffs, np.float64)
    cs = 0 if co is None else (co.shape[1] if co.ndim == 2 else 1)
    kc = None if constants is None else np.ascontiguousarray(constants, np.float64)
    lib().orc_lift_bc(kernel, _p(x), _p(xd), xd.shape[1], xd.shape[0], _p(dm), dm.shape[1], bs,
                      _p(co), cs, _p(kc), _p(np.ascontiguousarray(bc_values, np.float64)),
                      _p(np.ascontiguousarray(bc_markers, np.int8)),
                      None if x0 is None else _p(np.ascontiguousarray(x0, np.float64)),
                      alpha, _p(b))
    return b


def bc_set(b, dofs, g, x0=None, alpha=1.0):
    lib().orc_bc_set(_p(b), _p(np.ascontiguousarray(dofs, np.int32)), len(dofs),
                     _p(np.ascontiguousarray(g, np.float64)),
                     None if x0 is None else _p(np.ascontiguousarray(x0, np.float64)), alpha)
    return b


def set_diagonal(values, rows, bs, diag, row_ptr, cols):
    rc = lib().orc_set_diagonal(_p(np.ascontiguousarray(rows, np.int32)), len(rows), bs, diag,
                                _p(np.ascontiguousarray(row_ptr, np.int64)),
                                _p(np.ascontiguousarray(cols, np.int32)), _p(values))
    if rc != 0:
        raise RuntimeError("Entry not in sparsity")
    return values


def spmv(row_ptr, cols, values, bs, x, y=None):
    nrows = len(row_ptr) - 1
    y = np.zeros(nrows * bs) if y is None else y
    lib().orc_spmv(nrows, bs, _p(np.ascontiguousarray(row_ptr, np.int64)),
                   _p(np.ascontiguousarray(cols, np.int32)),
                   _p(np.ascontiguousarray(values, np.float64)),
                   _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def cg(row_ptr, cols, values, bs, b, x=None, kmax=10000, rtol=1e-10, jacobi=True):
    nrows = len(row_ptr) - 1
    x = np.zeros(nrows * bs) if x is None else np.array(x, np.float64)
    fn = lib().orc_pcg_jacobi if jacobi else lib().orc_cg
    it = fn(nrows, bs, _p(np.ascontiguousarray(row_ptr, np.int64)),
            _p(np.ascontiguousarray(cols, np.int32)),
            _p(np.ascontiguousarray(values, np.float64)),
            _p(np.ascontiguousarray(b, np.float64)), _p(x), kmax, rtol)
    return x, it
