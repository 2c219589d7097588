"""Mesh creation (mirror of dolfinx.mesh for the box meshes of this path).

create_box follows python/dolfinx/mesh.py:1245-1291 on one rank: no
partitioner, cells reordered by reverse Cuthill-McKee on the dual graph, all
work in the C++ host library (include/dfx.h).
"""

from __future__ import annotations

import ctypes as C
from enum import IntEnum

import numpy as np

from . import _lib


class CellType(IntEnum):
    tetrahedron = _lib.TET
    hexahedron = _lib.HEX


class Geometry:
    def __init__(self, x, dofmap, input_global_indices):
        self.x = x                      # (num_nodes, 3)
        self.dofmap = dofmap            # (num_cells, nodes_per_cell)
        self.input_global_indices = input_global_indices


class Topology:
    def __init__(self, cell_type, cells, original_cell_index, num_vertices):
        self.cell_type = cell_type
        self.dim = 3
        self.cells = cells
        self.original_cell_index = original_cell_index
        self.num_vertices = num_vertices


class Mesh:
    """Owns a dfx_mesh handle plus host copies of its arrays."""

    def __init__(self, handle, cell_type):
        self._h = handle
        nc, nv, npc = I32(), I32(), I32()
        _lib.call("dfx_mesh_sizes", handle, C.byref(nc), C.byref(nv), C.byref(npc))
        x = np.zeros((nv.value, 3), np.float64)
        cells = np.zeros((nc.value, npc.value), np.int32)
        orig = np.zeros(nc.value, np.int64)
        inp = np.zeros(nv.value, np.int64)
        _lib.call("dfx_mesh_copy", handle, _lib.ptr(x), _lib.ptr(cells), _lib.ptr(orig),
                  _lib.ptr(inp))
        self.topology = Topology(cell_type, cells, orig, nv.value)
        # P1 geometry: the geometry dofmap is the cell-vertex map
        self.geometry = Geometry(x, cells, inp)
        self.name = "mesh"

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_mesh_destroy(h)
            except Exception:
                pass


I32 = C.c_int32


def create_box(points, n, cell_type=CellType.tetrahedron) -> Mesh:
    """mesh::create_box (cpp/dolfinx/mesh/generation.h:104-134)."""
    p0 = np.ascontiguousarray(points[0], np.float64)
    p1 = np.ascontiguousarray(points[1], np.float64)
    nn = np.ascontiguousarray(n, np.int64)
    h = C.c_void_p()
    _lib.call("dfx_mesh_create_box", _lib.ptr(p0), _lib.ptr(p1), _lib.ptr(nn), int(cell_type),
              C.byref(h))
    return Mesh(h, CellType(int(cell_type)))


def create_mesh(cell_type, cells, x) -> Mesh:
    """mesh::create_mesh (cpp/dolfinx/mesh/utils.h:1336-1356) on one rank:
    cells (num_cells x vertices, input vertex indices, DOLFINx vertex order)
    and x (num_vertices x 3); cells reordered by RCM, vertices numbered by
    first appearance."""
    cells = np.ascontiguousarray(cells, np.int64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    h = C.c_void_p()
    _lib.call("dfx_mesh_create", int(cell_type), cells.shape[0], _lib.ptr(cells), x.shape[0],
              _lib.ptr(x), C.byref(h))
    return Mesh(h, CellType(int(cell_type)))


def create_unit_cube(nx, ny, nz, cell_type=CellType.tetrahedron) -> Mesh:
    return create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], [nx, ny, nz], cell_type)
