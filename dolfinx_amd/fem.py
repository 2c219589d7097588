"""Finite element assembly (mirror of dolfinx.fem for this path).

The UFL/FFCx front end is replaced by a fixed set of hand-written device
kernels (one per config form); ``form`` selects one by name and packs its
coefficients like fem::pack_coefficients (cpp/dolfinx/fem/pack.h:265-388).
Assembly, lifting and BC application run on the GPU through include/dfx.h.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .la import MatrixCSR, SparsityPattern, Vector
from .mesh import CellType, Mesh


class DofMap:
    def __init__(self, dofmap, bs, num_nodes):
        self.list = dofmap            # (num_cells, dofs_per_cell) node indices
        self.bs = bs
        self.index_map_size = num_nodes

    def cell_dofs(self, c):
        return self.list[c]


class FunctionSpace:
    """Lagrange space (fem::create_functionspace, fem/utils.h:925-950)."""

    def __init__(self, mesh: Mesh, degree: int, bs: int = 1):
        self.mesh = mesh
        self.degree = degree
        self._h = C.c_void_p()
        _lib.call("dfx_space_create", mesh.handle, degree, bs, C.byref(self._h))
        nn, k, b = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.call("dfx_space_sizes", self._h, C.byref(nn), C.byref(k), C.byref(b))
        dm = np.zeros((mesh.topology.cells.shape[0], k.value), np.int32)
        _lib.call("dfx_space_dofmap", self._h, _lib.ptr(dm))
        self.dofmap = DofMap(dm, b.value, nn.value)

    @property
    def handle(self):
        return self._h

    @property
    def num_dofs(self):
        """Unrolled size (nodes * block size)."""
        return self.dofmap.index_map_size * self.dofmap.bs

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_space_destroy(h)
            except Exception:
                pass


class Function:
    """fem::Function (fem/Function.h): a space and its dof values, held on
    the GPU in a la.Vector (x.array copies them to the host)."""

    def __init__(self, V: FunctionSpace, x: Vector | None = None, name="f",
                 dtype=np.float64):
        self.function_space = V
        self.x = Vector(V.num_dofs, dtype) if x is None else x
        if self.x.size != V.num_dofs:
            raise ValueError(f"{V.num_dofs} values expected, got {self.x.size}")
        self.name = name

    def interpolate(self, f) -> None:
        """Nodal interpolation of f(x) (x: (3, n) points; f returns (n,) or
        (bs, n)) on a degree-1 space, whose dofs sit at the mesh nodes."""
        V = self.function_space
        if V.degree != 1:
            raise NotImplementedError("interpolation is implemented for degree-1 spaces")
        X = np.zeros((V.dofmap.index_map_size, 3))
        X[V.dofmap.list.ravel()] = V.mesh.geometry.x[V.mesh.geometry.dofmap.ravel()]
        v = np.asarray(f(X.T), np.float64).reshape(V.dofmap.bs, -1)
        self.x.array = np.ascontiguousarray(v.T).ravel()


def functionspace(mesh: Mesh, element) -> FunctionSpace:
    """functionspace(mesh, ("Lagrange", degree[, (3,)]))."""
    family, degree = element[0], element[1]
    if family not in ("Lagrange", "P", "Q"):
        raise ValueError(f"unsupported family {family}")
    bs = 1
    if len(element) > 2 and element[2]:
        bs = int(np.prod(element[2]))
    return FunctionSpace(mesh, int(degree), bs)


def locate_boundary_dofs(V: FunctionSpace, axis=None, value=0.0) -> np.ndarray:
    """Unrolled dofs on exterior facets (all of them, or those with every
    vertex at x[axis] == value): locate_dofs_topological on the facets of
    exterior_facet_indices / locate_entities_boundary."""
    marker = 0 if axis is None else 1
    n = C.c_int32()
    _lib.call("dfx_locate_boundary_dofs", V.handle, marker, 0 if axis is None else axis,
              float(value), C.byref(n), None)
    out = np.zeros(n.value, np.int32)
    _lib.call("dfx_locate_boundary_dofs", V.handle, marker, 0 if axis is None else axis,
              float(value), C.byref(n), _lib.ptr(out))
    return out


class DirichletBC:
    """fem::DirichletBC on unrolled dofs with values g (fem/DirichletBC.h)."""

    def __init__(self, V: FunctionSpace, dofs, g=None):
        self.V = V
        self.dofs = np.ascontiguousarray(dofs, np.int32)
        n = V.num_dofs
        self.g = None if g is None else np.ascontiguousarray(g, np.float64)
        if self.g is not None and self.g.size != n:
            raise ValueError("g must have one value per dof of V")
        self._h = C.c_void_p()
        _lib.call("dfx_bc_create", _lib.ptr(self.dofs), len(self.dofs), n, _lib.ptr(self.g),
                  C.byref(self._h))

    @property
    def handle(self):
        return self._h

    def markers(self):
        m = np.zeros(self.V.num_dofs, np.int8)
        m[self.dofs] = 1
        return m

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_bc_destroy(h)
            except Exception:
                pass


def dirichletbc(V: FunctionSpace, dofs, g=None) -> DirichletBC:
    return DirichletBC(V, dofs, g)


# Kernels of exterior facet integrals (integration entities: (cell, local
# facet) pairs)
FACET_KERNELS = {"p1_neumann", "q1_neumann", "p1_robin", "q1_robin"}
# Kernels of interior facet integrals (integration entities: (cell0, local
# facet0, cell1, local facet1) quads)
INTERIOR_KERNELS = {"p1_gradient_jump"}

# name -> (kernel id, cell type, degree, block size, rank)
KERNELS = {
    "p1_stiffness_reaction": (_lib.K_P1TET_A, CellType.tetrahedron, 1, 1, 2),
    "p1_source": (_lib.K_P1TET_L, CellType.tetrahedron, 1, 1, 1),
    "q1_stiffness_reaction": (_lib.K_Q1HEX_A, CellType.hexahedron, 1, 1, 2),
    "q1_source": (_lib.K_Q1HEX_L, CellType.hexahedron, 1, 1, 1),
    "p2_elasticity": (_lib.K_P2TET_ELAST_A, CellType.tetrahedron, 2, 3, 2),
    "p2_body_force": (_lib.K_P2TET_ELAST_L, CellType.tetrahedron, 2, 3, 1),
    "p3_stiffness_reaction": (_lib.K_P3TET_A, CellType.tetrahedron, 3, 1, 2),
    "p1_neumann": (_lib.K_P1TET_NEUMANN, CellType.tetrahedron, 1, 1, 1),
    "q1_neumann": (_lib.K_Q1HEX_NEUMANN, CellType.hexahedron, 1, 1, 1),
    "p1_robin": (_lib.K_P1TET_ROBIN_A, CellType.tetrahedron, 1, 1, 2),
    "q1_robin": (_lib.K_Q1HEX_ROBIN_A, CellType.hexahedron, 1, 1, 2),
    "p1_gradient_jump": (_lib.K_P1TET_GRADJUMP_A, CellType.tetrahedron, 1, 1, 2),
    "p3_source": (_lib.K_P3TET_L, CellType.tetrahedron, 3, 1, 1),
}


def pack_cellwise(mesh: Mesh, values_by_input_cell) -> np.ndarray:
    """DG0 coefficient(s) given per input (lexicographic) cell, packed in the
    mesh's cell order: (num_cells, ncomp)."""
    v = np.asarray(values_by_input_cell, np.float64)
    v = v.reshape(v.shape[0], -1)
    return np.ascontiguousarray(v[mesh.topology.original_cell_index])


def pack_vertexwise(mesh: Mesh, values_by_input_vertex) -> np.ndarray:
    """P1/Q1 coefficient given per input vertex, packed per cell vertex:
    (num_cells, vertices_per_cell)."""
    f = np.asarray(values_by_input_vertex, np.float64)
    fv = f[mesh.geometry.input_global_indices]  # by local vertex
    return np.ascontiguousarray(fv[mesh.topology.cells])


class Form:
    """fem::Form holding device copies of geometry, dofmap and packed
    coefficients (fem/Form.h:116-)."""

    def __init__(self, name, V: FunctionSpace, coeffs=None, constants=None, dtype=np.float64,
                 facets=None):
        if name in ELEMENT_FORMS:
            self._init_elements(name, V, dtype)
            return
        if name not in KERNELS:
            raise ValueError(f"unknown form {name}")
        kid, ct, deg, bs, rank = KERNELS[name]
        mesh = V.mesh
        if mesh.topology.cell_type != ct or V.degree != deg or V.dofmap.bs != bs:
            raise ValueError(f"form {name} does not match the space")
        self.name, self.V, self.rank, self.dtype = name, V, rank, np.dtype(dtype)
        self.integral = ("exterior_facet" if name in FACET_KERNELS else
                         "interior_facet" if name in INTERIOR_KERNELS else "cell")
        self.facets = None
        self.coeffs = None if coeffs is None else np.ascontiguousarray(coeffs, np.float64)
        nc = mesh.topology.cells.shape[0]
        if self.coeffs is not None:
            self.coeffs = self.coeffs.reshape(nc, -1)
        self.constants = np.ascontiguousarray(
            np.zeros(1) if constants is None else np.atleast_1d(constants), np.float64)
        x = np.ascontiguousarray(mesh.geometry.x)
        xd = np.ascontiguousarray(mesh.geometry.dofmap)
        dm = np.ascontiguousarray(V.dofmap.list)
        d = _lib.FormDesc()
        d.kernel, d.dtype = kid, _lib.code_of(dtype)
        d.num_cells, d.num_nodes, d.nodes_per_cell = nc, x.shape[0], xd.shape[1]
        d.x, d.x_dofmap = _lib.ptr(x), _lib.ptr(xd)
        d.dofs_per_cell, d.bs, d.dofmap = dm.shape[1], bs, _lib.ptr(dm)
        d.cstride = 0 if self.coeffs is None else self.coeffs.shape[1]
        d.coeffs = _lib.ptr(self.coeffs)
        d.num_constants = len(self.constants)
        d.constants = _lib.ptr(self.constants)
        self._h = C.c_void_p()
        if self.integral == "interior_facet":
            q = np.ascontiguousarray(np.zeros((0, 4)) if facets is None else facets,
                                     np.int32).reshape(-1, 4)
            self.facets = q
            _lib.call("dfx_form_create_interior_facets", C.byref(d), q.shape[0], _lib.ptr(q),
                      C.byref(self._h))
        elif facets is not None:
            pr = np.ascontiguousarray(facets, np.int32).reshape(-1, 2)
            self.facets = pr
            _lib.call("dfx_form_create_facets", C.byref(d), pr.shape[0], _lib.ptr(pr),
                      C.byref(self._h))
        else:
            _lib.call("dfx_form_create", C.byref(d), C.byref(self._h))

    def _init_elements(self, name, V, dtype):
        kid, rank = ELEMENT_FORMS[name]
        self.integral = "cell"
        if V.dofmap.bs != 1:
            raise ValueError("element tensors need a scalar space")
        self.name, self.V, self.rank, self.dtype = name, V, rank, np.dtype(dtype)
        self.coeffs, self.constants = None, np.zeros(1)
        dm = np.ascontiguousarray(V.dofmap.list)
        d = _lib.FormDesc()
        d.kernel, d.dtype = kid, _lib.code_of(dtype)
        d.num_cells, d.dofs_per_cell, d.bs, d.dofmap = dm.shape[0], dm.shape[1], 1, _lib.ptr(dm)
        self._dm = dm
        self._h = C.c_void_p()
        _lib.call("dfx_form_create", C.byref(d), C.byref(self._h))

    @property
    def handle(self):
        return self._h

    def autotune(self, A: MatrixCSR, bc: DirichletBC | None = None):
        """Pick the fastest assembly kernel variant for this form and the
        pattern of A by timing (A's values are overwritten and left zero).
        Returns (variant, ms per assembly)."""
        v, ms = C.c_int32(), C.c_float()
        _lib.call("dfx_form_autotune", self._h, A.handle, None if bc is None else bc.handle,
                  C.byref(v), C.byref(ms))
        return v.value, ms.value

    def pack(self, mesh: Mesh, values, layout="cell"):
        """fem::pack_coefficients (fem/pack.h:265-388) on the device: the
        form's coefficients from values per input cell ("cell": DG0, cstride
        values per cell) or per input vertex ("vertex": P1)."""
        v = np.ascontiguousarray(values, np.float64).ravel()
        code = {"cell": 0, "vertex": 1}[layout]
        _lib.call("dfx_form_pack_coefficients", self._h, mesh.handle, code, _lib.ptr(v), v.size)

    def set_variant(self, variant: int):
        """Fix the assembly kernel variant (0-2 thread per row with 64/128/256
        rows per workgroup, 3 one incidence per lane, 4-6 as 0-2 with
        closed-form Q1 entries, 7 element matrices on the matrix cores then
        gathered by rows, 8 one incidence per thread, 9 structured-box Q1
        rows in registers, 10 factors formed from the geometry per
        incidence, 11 as 1 with eight incidences in flight, 12-14 Q1 rows
        with their plan at fixed positions, 15 structured Q1 rows in
        registers stored through an LDS slice, 16 a team of lanes per row
        forming the entries, 17 such teams gathering the element matrices
        of 7)."""
        _lib.call("dfx_form_set_variant", self._h, int(variant))

    @property
    def function_spaces(self):
        return [self.V] * self.rank

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_form_destroy(h)
            except Exception:
                pass


def form(name, V, coeffs=None, constants=None, dtype=np.float64, facets=None) -> Form:
    """A form of a named kernel; exterior facet kernels ("p1_neumann",
    "q1_neumann": g v ds with g packed per cell vertex; "p1_robin",
    "q1_robin": beta u v ds with DG0 beta) take their integration entities as
    (cell, local facet) pairs (facet_domain); the interior facet kernel
    "p1_gradient_jump" (gamma [grad u.n][grad v.n] dS, constants [gamma])
    takes (cell0, local facet0, cell1, local facet1) quads
    (interior_facet_domain). A form with several integrals is passed to the
    assemblers as a tuple of such forms."""
    return Form(name, V, coeffs, constants, dtype, facets)


def interior_facet_domain(mesh: Mesh) -> np.ndarray:
    """fem::compute_integration_domains(interior_facet) on every interior
    facet: (cell0, local facet0, cell1, local facet1), cell0 < cell1, in
    facet-index order (fem/utils.cpp:223-253)."""
    n = C.c_int32()
    _lib.call("dfx_interior_facet_domain", mesh.handle, C.byref(n), None)
    out = np.zeros((n.value, 4), np.int32)
    _lib.call("dfx_interior_facet_domain", mesh.handle, C.byref(n), _lib.ptr(out))
    return out


_ORDER = {"cell": 0, "interior_facet": 1, "exterior_facet": 2}


def _integrals(a):
    """The forms of a (one form or a tuple of forms of one form's
    integrals) in the reference's order: cell integrals, interior facets,
    then exterior facets (fem/assemble_matrix_impl.h:776-880,
    fem/assemble_vector_impl.h:540-615)."""
    forms = list(a) if isinstance(a, (list, tuple)) else [a]
    if not forms:
        raise ValueError("no forms")
    return sorted(forms, key=lambda f: _ORDER[f.integral])


def facet_domain(mesh: Mesh, axis=None, value=0.0) -> np.ndarray:
    """fem::compute_integration_domains(exterior_facet) on the exterior
    facets, or on those with every vertex at x[axis] == value: (cell, local
    facet) pairs in facet-index order (fem/utils.cpp:154-273)."""
    marker, ax = (0, 0) if axis is None else (1, int(axis))
    n = C.c_int32()
    _lib.call("dfx_facet_domain", mesh.handle, marker, ax, float(value), C.byref(n), None)
    out = np.zeros((n.value, 2), np.int32)
    _lib.call("dfx_facet_domain", mesh.handle, marker, ax, float(value), C.byref(n),
              _lib.ptr(out))
    return out


# Forms whose element tensors the caller supplies (the reference's custom
# kernels, cpp/demo/custom_kernel/main.cpp:73-118): name -> (kernel, rank)
ELEMENT_FORMS = {
    "element_matrix": (_lib.K_ELEMENT_A, 2),
    "element_vector": (_lib.K_ELEMENT_L, 1),
}


def element_form(V: FunctionSpace, rank: int, dtype=np.float64) -> Form:
    """A form on V whose cell tensors come from the caller: rank 2 takes
    element matrices (cells x k x k, k in {4, 8, 10, 20}), rank 1
    element vectors (cells x k), in V's cell order and local dof order."""
    return Form("element_matrix" if rank == 2 else "element_vector", V, dtype=dtype)


# The reference's element kernel signature (kernelptr_t, fem/kernel.h:18-20)
# for FP64 and FP32 forms: A, w, c, coordinate_dofs, entity_local_index,
# quadrature_permutation, custom_data.
KERNEL_F64 = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double),
                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int),
                         C.POINTER(C.c_uint8), C.c_void_p)
KERNEL_F32 = C.CFUNCTYPE(None, C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int),
                         C.POINTER(C.c_uint8), C.c_void_p)


class KernelForm(Form):
    """fem::Form with a cell integral given as integral_data{kernel, cells}
    (fem/Form.h:52-87; the shape of cpp/demo/custom_kernel/main.cpp:77-81):
    `kernel` is a function with the reference's element kernel ABI -- the
    address of a compiled tabulate_tensor (an int) or a KERNEL_F64 /
    KERNEL_F32 ctypes function -- called on the host for every cell in cell
    order; its tensors are inserted on the device. rank 2: (bs k)^2
    row-major element matrices; rank 1: bs k element vectors."""

    def __init__(self, V: FunctionSpace, rank: int, kernel, coeffs=None, constants=None,
                 dtype=np.float64, custom_data=None):
        self.name, self.V, self.rank, self.dtype = "kernel", V, rank, np.dtype(dtype)
        self.integral = "cell"
        self.facets = None
        mesh = V.mesh
        nc = mesh.topology.cells.shape[0]
        self.coeffs = (None if coeffs is None
                       else np.ascontiguousarray(coeffs, np.float64).reshape(nc, -1))
        self.constants = (np.zeros(0) if constants is None
                          else np.ascontiguousarray(np.atleast_1d(constants), np.float64))
        self._kernel = kernel  # keeps a ctypes callback alive
        addr = kernel if isinstance(kernel, int) else C.cast(kernel, C.c_void_p).value
        x = np.ascontiguousarray(mesh.geometry.x)
        xd = np.ascontiguousarray(mesh.geometry.dofmap)
        dm = np.ascontiguousarray(V.dofmap.list)
        self._dm = dm
        d = _lib.FormDesc()
        d.kernel = _lib.K_ELEMENT_A if rank == 2 else _lib.K_ELEMENT_L
        d.dtype = _lib.code_of(dtype)
        d.num_cells, d.num_nodes, d.nodes_per_cell = nc, x.shape[0], xd.shape[1]
        d.x, d.x_dofmap = _lib.ptr(x), _lib.ptr(xd)
        d.dofs_per_cell, d.bs, d.dofmap = dm.shape[1], V.dofmap.bs, _lib.ptr(dm)
        d.cstride = 0 if self.coeffs is None else self.coeffs.shape[1]
        d.coeffs = _lib.ptr(self.coeffs)
        d.num_constants = len(self.constants)
        d.constants = _lib.ptr(self.constants) if len(self.constants) else None
        self._h = C.c_void_p()
        _lib.call("dfx_form_create_kernel", C.byref(d), C.c_void_p(addr),
                  None if custom_data is None else C.c_void_p(custom_data), C.byref(self._h))


class KernelFacetForm(Form):
    """fem::Form with an exterior or interior facet integral given as
    integral_data{kernel, entities} (fem/Form.h:52-87): `entities` are
    (cell, local facet) pairs ("exterior_facet") or (cell0, local facet0,
    cell1, local facet1) quads ("interior_facet", e.g. from
    interior_facet_domain); the kernel has the reference's element kernel
    ABI (an address or a KERNEL_F64 / KERNEL_F32 function) and is called on
    the host per entity, as assemble_exterior_facets /
    assemble_interior_facets do; the tensors are added on the device in
    entity order. coeffs are packed per cell (cells x cstride); an
    interior kernel gets both cells' coefficients."""

    def __init__(self, V: FunctionSpace, rank: int, kernel, integral: str, entities,
                 coeffs=None, constants=None, dtype=np.float64, custom_data=None):
        if integral not in ("exterior_facet", "interior_facet"):
            raise ValueError(f"integral {integral!r}")
        self.name, self.V, self.rank, self.dtype = "kernel", V, rank, np.dtype(dtype)
        self.integral = integral
        per = 4 if integral == "interior_facet" else 2
        ent = np.ascontiguousarray(entities, np.int32).reshape(-1, per)
        self.facets = ent
        mesh = V.mesh
        nc = mesh.topology.cells.shape[0]
        self.coeffs = (None if coeffs is None
                       else np.ascontiguousarray(coeffs, np.float64).reshape(nc, -1))
        self.constants = (np.zeros(0) if constants is None
                          else np.ascontiguousarray(np.atleast_1d(constants), np.float64))
        self._kernel = kernel
        addr = kernel if isinstance(kernel, int) else C.cast(kernel, C.c_void_p).value
        x = np.ascontiguousarray(mesh.geometry.x)
        xd = np.ascontiguousarray(mesh.geometry.dofmap)
        dm = np.ascontiguousarray(V.dofmap.list)
        self._keep = (x, xd, dm)
        d = _lib.FormDesc()
        d.kernel = _lib.K_ELEMENT_A if rank == 2 else _lib.K_ELEMENT_L
        d.dtype = _lib.code_of(dtype)
        d.num_cells, d.num_nodes, d.nodes_per_cell = nc, x.shape[0], xd.shape[1]
        d.x, d.x_dofmap = _lib.ptr(x), _lib.ptr(xd)
        d.dofs_per_cell, d.bs, d.dofmap = dm.shape[1], V.dofmap.bs, _lib.ptr(dm)
        d.cstride = 0 if self.coeffs is None else self.coeffs.shape[1]
        d.coeffs = _lib.ptr(self.coeffs)
        d.num_constants = len(self.constants)
        d.constants = _lib.ptr(self.constants) if len(self.constants) else None
        self._h = C.c_void_p()
        _lib.call("dfx_form_create_kernel_facets", C.byref(d), C.c_void_p(addr),
                  None if custom_data is None else C.c_void_p(custom_data),
                  1 if per == 2 else 2, ent.shape[0], _lib.ptr(ent), C.byref(self._h))


def kernel_facet_form(V: FunctionSpace, rank: int, kernel, integral: str, entities, coeffs=None,
                      constants=None, dtype=np.float64, custom_data=None) -> KernelFacetForm:
    """A facet integral whose tensors come from a kernel with the reference's
    ABI (see KernelFacetForm)."""
    return KernelFacetForm(V, rank, kernel, integral, entities, coeffs, constants, dtype,
                           custom_data)


def kernel_form(V: FunctionSpace, rank: int, kernel, coeffs=None, constants=None,
                dtype=np.float64, custom_data=None) -> KernelForm:
    """A form whose cell tensors come from an element kernel with the
    reference's ABI (see KernelForm)."""
    return KernelForm(V, rank, kernel, coeffs, constants, dtype, custom_data)


def _device_tensors(x, n, dtype):
    if isinstance(x, Vector):
        if x.size != n:
            raise ValueError(f"{n} values expected, got {x.size}")
        return x
    a = np.ascontiguousarray(x, dtype).ravel()
    if a.size != n:
        raise ValueError(f"{n} values expected, got {a.size}")
    return Vector.from_array(a, dtype)


def create_sparsity_pattern(a) -> SparsityPattern:
    """fem::create_sparsity_pattern (fem/utils.h:197-314): the cell
    couplings, plus those of the interior facets of any interior facet
    integral of a (a Form or a tuple of one form's integrals)."""
    forms = _integrals(a)
    q = [f.facets for f in forms if f.integral == "interior_facet"]
    q = np.ascontiguousarray(np.concatenate(q) if q else np.zeros((0, 4)), np.int32)
    V = forms[0].V
    h = C.c_void_p()
    _lib.call("dfx_sparsity_create_facets", V.handle, V.handle, q.shape[0], _lib.ptr(q),
              C.byref(h))
    return SparsityPattern(h)


def create_matrix(a, dtype=None) -> MatrixCSR:
    """python/dolfinx/fem/assemble.py:127-143 (block_mode compact)."""
    sp = create_sparsity_pattern(a)
    return MatrixCSR.from_sparsity(sp, _integrals(a)[0].dtype if dtype is None else dtype)


def assemble_matrix(a, bcs=(), diag=1.0, A: MatrixCSR | None = None) -> MatrixCSR:
    """fem.assemble_matrix (python/dolfinx/fem/assemble.py:283-376): assemble
    with BC rows/columns zeroed, then insert ``diag`` on BC rows. ``a`` is a
    Form or a tuple of the forms of one form's integrals."""
    forms = _integrals(a)
    if A is None:
        A = create_matrix(forms)
    if len(bcs) > 1:
        raise ValueError("pass the union of Dirichlet dofs as one DirichletBC")
    bc = bcs[0] if bcs else None
    for f in forms:
        _lib.call("dfx_assemble_matrix", A.handle, f.handle, None if bc is None else bc.handle)
    for bc in bcs:
        _lib.call("dfx_set_diagonal", A.handle, bc.handle, float(diag))
    return A


def assemble_matrix_elements(a: Form, E, bcs=(), diag=1.0, A: MatrixCSR | None = None) -> MatrixCSR:
    """fem.assemble_matrix with a custom kernel whose outputs are E (cells x
    k x k, host array or device Vector): A[dofs_c, dofs_c] += E_c in cell
    order with BC rows/columns skipped, then ``diag`` on BC rows."""
    if A is None:
        A = create_matrix(a)
    k = a._dm.shape[1]
    Ed = _device_tensors(E, a._dm.shape[0] * k * k, a.dtype)
    if len(bcs) > 1:
        raise ValueError("pass the union of Dirichlet dofs as one DirichletBC")
    bc = bcs[0] if bcs else None
    _lib.call("dfx_assemble_matrix_elements", A.handle, a.handle,
              None if bc is None else bc.handle, Ed.handle)
    if bc is not None:
        _lib.call("dfx_set_diagonal", A.handle, bc.handle, float(diag))
    return A


def assemble_vector_elements(L: Form, e, b: Vector | None = None) -> Vector:
    """fem.assemble_vector with a custom kernel whose outputs are e (cells x
    k): b[dofs_c] += e_c in cell order."""
    if b is None:
        b = Vector(L.V.num_dofs, L.dtype)
    ed = _device_tensors(e, L._dm.shape[0] * L._dm.shape[1], L.dtype)
    _lib.call("dfx_assemble_vector_elements", b.handle, L.handle, ed.handle)
    return b


def apply_lifting_elements(b: Vector, a: Form, E, bcs, x0: Vector | None = None,
                           alpha=1.0) -> None:
    """fem.apply_lifting for a form with caller-supplied element matrices
    E (cells x k x k): b -= alpha E[:, bc] (g - x0) over the BC columns."""
    k = a._dm.shape[1]
    Ed = _device_tensors(E, a._dm.shape[0] * k * k, a.dtype)
    for bc in bcs:
        _lib.call("dfx_apply_lifting_elements", b.handle, a.handle, bc.handle, Ed.handle,
                  None if x0 is None else x0.handle, float(alpha))


def assemble_system(a: Form, L: Form, bcs=(), diag=1.0, A: MatrixCSR | None = None,
                    b: Vector | None = None):
    """The demos' system sequence (cpp/demo/poisson/main.cpp:207-226):
    A = assemble_matrix(a, bcs) with ``diag`` on BC rows (A zeroed first),
    b = assemble_vector(L) from zero, apply_lifting(b, [a], [bcs]),
    set_bc(b, bcs); the vector assembly overlaps the matrix assembly on a
    second stream. Returns (A, b). Forms with several integrals (tuples)
    are assembled by the separate calls in the same sequence."""
    if isinstance(a, (list, tuple)) or isinstance(L, (list, tuple)):
        if A is not None:
            A.set_value(0.0)
        A = assemble_matrix(a, bcs, diag, A)
        if b is None:
            b = Vector(_integrals(L)[0].V.num_dofs, _integrals(L)[0].dtype)
        else:
            b.set(0.0)
        assemble_vector(L, b)
        apply_lifting(b, [a], [bcs])
        set_bc(b, bcs)
        return A, b
    if A is None:
        A = create_matrix(a)
    else:
        A.set_value(0.0)
    if b is None:
        b = Vector(L.V.num_dofs, L.dtype)
    if len(bcs) > 1:
        raise ValueError("pass the union of Dirichlet dofs as one DirichletBC")
    bc = bcs[0] if bcs else None
    _lib.call("dfx_assemble_system", A.handle, a.handle, None if bc is None else bc.handle,
              float(diag), b.handle, L.handle, 1)
    return A, b


def assemble_matrix_batched(forms, bcs, matrices, diag=1.0) -> None:
    """assemble_matrix for many independent systems of one P1 or Q1
    stiffness kernel in two launches (cell factors of all systems, rows of
    all systems); bcs[i] may be None. Then ``diag`` on each BC diagonal."""
    n = len(forms)
    if not (len(bcs) == len(matrices) == n):
        raise ValueError("forms, bcs and matrices differ in length")
    if n == 0:
        return

    def arr(objs):
        return (C.c_void_p * n)(*[None if o is None else o.handle.value for o in objs])

    _lib.call("dfx_assemble_matrix_batched", n, arr(matrices), arr(forms), arr(bcs))
    for A, bc in zip(matrices, bcs):
        if bc is not None:
            _lib.call("dfx_set_diagonal", A.handle, bc.handle, float(diag))


def assemble_vector(L, b: Vector | None = None) -> Vector:
    """fem.assemble_vector (fem/assembler.h:230-257): b += L (a Form or a
    tuple of the forms of one form's integrals)."""
    forms = _integrals(L)
    if b is None:
        b = Vector(forms[0].V.num_dofs, forms[0].dtype)
    for f in forms:
        _lib.call("dfx_assemble_vector", b.handle, f.handle)
    return b


def apply_lifting(b: Vector, a_list, bcs_list, x0=None, alpha=1.0) -> None:
    """fem.apply_lifting (python/dolfinx/fem/assemble.py:378-493); each
    entry of a_list is a Form or a tuple of the forms of its integrals."""
    for i, (a, bcs) in enumerate(zip(a_list, bcs_list)):
        xx = None if not x0 else x0[i]
        for f in _integrals(a):
            for bc in bcs:
                _lib.call("dfx_apply_lifting", b.handle, f.handle, bc.handle,
                          None if xx is None else xx.handle, float(alpha))


def set_bc(b: Vector, bcs, x0: Vector | None = None, alpha=1.0) -> None:
    """DirichletBC::set (fem/DirichletBC.h:495-578): b[bc] = alpha (g - x0)."""
    for bc in bcs:
        _lib.call("dfx_bc_set", b.handle, bc.handle, None if x0 is None else x0.handle,
                  float(alpha))
