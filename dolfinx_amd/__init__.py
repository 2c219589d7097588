"""MI355X-native assemble-and-solve path of DOLFINx.

Host setup (mesh, dofmaps, sparsity) is C++; assembly, SpMV and CG run as
hand-written HIP kernels for gfx950. Everything crosses the C ABI of
include/dfx.h through libdfx.so, built in-tree by dolfinx_amd.build.
"""

from . import _lib  # noqa: F401
from . import fem, graph, la, mesh, solvers  # noqa: F401

__all__ = ["fem", "graph", "la", "mesh", "solvers"]
