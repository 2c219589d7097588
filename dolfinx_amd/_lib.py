"""ctypes binding of libdfx.so (the C ABI declared in include/dfx.h).

The library is loaded from this package directory. There is no fallback: if
libdfx.so is missing or does not load, every call raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdfx.so")

(DFX_OK, DFX_ERR_ARG, DFX_ERR_PATTERN, DFX_ERR_HIP, DFX_ERR_NOCONV, DFX_ERR_NODEV,
 DFX_ERR_CAPTURE) = range(7)
F32, F64 = 0, 1
TET, HEX = 0, 1
(K_P1TET_A, K_P1TET_L, K_Q1HEX_A, K_Q1HEX_L, K_P2TET_ELAST_A, K_P2TET_ELAST_L,
 K_P3TET_A, K_P3TET_L, K_ELEMENT_A, K_ELEMENT_L, K_P1TET_NEUMANN, K_Q1HEX_NEUMANN,
 K_P1TET_ROBIN_A, K_Q1HEX_ROBIN_A, K_P1TET_GRADJUMP_A) = range(15)

P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I32, I64, D, F = C.c_int32, C.c_int64, C.c_double, C.c_float


class FormDesc(C.Structure):
    _fields_ = [
        ("kernel", I32), ("dtype", I32), ("num_cells", I32), ("num_nodes", I32),
        ("nodes_per_cell", I32), ("x", P), ("x_dofmap", P), ("dofs_per_cell", I32),
        ("bs", I32), ("dofmap", P), ("cstride", I32), ("coeffs", P),
        ("num_constants", I32), ("constants", P),
    ]


class CGOpts(C.Structure):
    _fields_ = [("max_it", I32), ("rtol", D), ("jacobi", I32), ("check_every", I32),
                ("graph", I32), ("classic", I32)]


class CGResult(C.Structure):
    _fields_ = [("iterations", I32), ("converged", I32), ("rnorm0", D), ("rnorm", D),
                ("true_rel_residual", D), ("seconds", D), ("cooperative", I32)]


class MixedOpts(C.Structure):
    _fields_ = [("max_outer", I32), ("max_inner", I32), ("rtol", D), ("inner_rtol", D),
                ("check_every", I32), ("restart", I32), ("graph", I32), ("fallback", I32)]


class MixedResult(C.Structure):
    _fields_ = [("outer", I32), ("inner_total", I32), ("converged", I32),
                ("true_rel_residual", D), ("seconds", D), ("fallback_iterations", I32)]


# name -> (restype, argtypes); every function returns dfx_status unless noted
SIGNATURES = {
    "dfx_last_error": (C.c_char_p, []),
    "dfx_version": (C.c_char_p, []),
    "dfx_init": (C.c_int, []),
    "dfx_synchronize": (C.c_int, []),
    "dfx_stream": (P, []),
    "dfx_mesh_create_box": (C.c_int, [P, P, P, C.c_int, PP]),
    "dfx_mesh_create": (C.c_int, [C.c_int, I32, P, I32, P, PP]),
    "dfx_facet_domain": (C.c_int, [P, C.c_int, C.c_int, C.c_double, P, P]),
    "dfx_form_create_facets": (C.c_int, [P, I32, P, PP]),
    "dfx_form_create_kernel": (C.c_int, [P, P, P, PP]),
    "dfx_form_create_interior_facets": (C.c_int, [P, I32, P, PP]),
    "dfx_interior_facet_domain": (C.c_int, [P, P, P]),
    "dfx_sparsity_create_facets": (C.c_int, [P, P, I32, P, PP]),
    "dfx_mesh_destroy": (C.c_int, [P]),
    "dfx_mesh_sizes": (C.c_int, [P, P, P, P]),
    "dfx_mesh_copy": (C.c_int, [P, P, P, P, P]),
    "dfx_space_create": (C.c_int, [P, C.c_int, C.c_int, PP]),
    "dfx_space_destroy": (C.c_int, [P]),
    "dfx_space_sizes": (C.c_int, [P, P, P, P]),
    "dfx_space_dofmap": (C.c_int, [P, P]),
    "dfx_locate_boundary_dofs": (C.c_int, [P, C.c_int, C.c_int, D, P, P]),
    "dfx_sparsity_create": (C.c_int, [P, P, PP]),
    "dfx_sparsity_destroy": (C.c_int, [P]),
    "dfx_sparsity_sizes": (C.c_int, [P, P, P, P]),
    "dfx_sparsity_copy": (C.c_int, [P, P, P]),
    "dfx_csr_create": (C.c_int, [P, P, I32, I32, C.c_int, C.c_int, PP]),
    "dfx_csr_create_shared": (C.c_int, [P, C.c_int, PP]),
    "dfx_csr_destroy": (C.c_int, [P]),
    "dfx_csr_sizes": (C.c_int, [P, P, P, P, P]),
    "dfx_csr_zero": (C.c_int, [P]),
    "dfx_csr_get_values": (C.c_int, [P, P]),
    "dfx_csr_set_values": (C.c_int, [P, P]),
    "dfx_csr_get_pattern": (C.c_int, [P, P, P]),
    "dfx_csr_autotune": (C.c_int, [P, P, P]),
    "dfx_csr_set_lanes": (C.c_int, [P, I32]),
    "dfx_vec_create": (C.c_int, [I32, C.c_int, PP]),
    "dfx_vec_destroy": (C.c_int, [P]),
    "dfx_vec_size": (C.c_int, [P, P, P]),
    "dfx_vec_upload": (C.c_int, [P, P]),
    "dfx_vec_download": (C.c_int, [P, P]),
    "dfx_vec_set": (C.c_int, [P, D]),
    "dfx_vec_copy": (C.c_int, [P, P]),
    "dfx_spmv": (C.c_int, [P, P, P]),
    "dfx_inner_product": (C.c_int, [P, P, P]),
    "dfx_squared_norm": (C.c_int, [P, P]),
    "dfx_axpy": (C.c_int, [P, D, P, P]),
    "dfx_form_create": (C.c_int, [C.POINTER(FormDesc), PP]),
    "dfx_form_destroy": (C.c_int, [P]),
    "dfx_form_set_coefficients": (C.c_int, [P, P, P]),
    "dfx_form_pack_coefficients": (C.c_int, [P, P, C.c_int, P, I64]),
    "dfx_assemble_matrix_batched": (C.c_int, [I32, P, P, P]),
    "dfx_bc_create": (C.c_int, [P, I32, I32, P, PP]),
    "dfx_bc_destroy": (C.c_int, [P]),
    "dfx_assemble_matrix": (C.c_int, [P, P, P]),
    "dfx_assemble_system": (C.c_int, [P, P, P, C.c_double, P, P, I32]),
    "dfx_assemble_matrix_elements": (C.c_int, [P, P, P, P]),
    "dfx_assemble_vector_elements": (C.c_int, [P, P, P]),
    "dfx_apply_lifting_elements": (C.c_int, [P, P, P, P, P, C.c_double]),
    "dfx_set_diagonal": (C.c_int, [P, P, D]),
    "dfx_form_autotune": (C.c_int, [P, P, P, P, P]),
    "dfx_form_set_variant": (C.c_int, [P, I32]),
    "dfx_assemble_vector": (C.c_int, [P, P]),
    "dfx_apply_lifting": (C.c_int, [P, P, P, P, D]),
    "dfx_bc_set": (C.c_int, [P, P, P, D]),
    "dfx_cg_solve": (C.c_int, [P, P, P, C.POINTER(CGOpts), C.POINTER(CGResult)]),
    "dfx_cg_solve_noalloc": (C.c_int, [P, P, P, C.POINTER(CGOpts), C.POINTER(CGResult)]),
    "dfx_mixed_solve": (C.c_int, [P, P, P, P, C.POINTER(MixedOpts), C.POINTER(MixedResult)]),
    "dfx_batch_create": (C.c_int, [I32, I64, PP]),
    "dfx_batch_destroy": (C.c_int, [P]),
    "dfx_batch_add": (C.c_int, [P, P, P, P, D, I32, C.POINTER(I32)]),
    "dfx_batch_remove": (C.c_int, [P, I32]),
    "dfx_batch_step": (C.c_int, [P, I32, I32]),
    "dfx_batch_status": (C.c_int, [P, I32, P, P, P, P, P]),
    "dfx_batch_restart": (C.c_int, [P]),
    "dfx_csr_add_blocks": (C.c_int, [P, I64, I32, I32, P, P, P]),
    "dfx_form_create_kernel_facets": (C.c_int, [P, P, P, I32, I32, P, PP]),
    "dfx_graph_begin": (C.c_int, []),
    "dfx_graph_end": (C.c_int, [PP]),
    "dfx_graph_launch": (C.c_int, [P]),
    "dfx_graph_destroy": (C.c_int, [P]),
    "dfx_event_record": (C.c_int, [C.c_int]),
    "dfx_event_elapsed": (C.c_int, [C.c_int, C.c_int, C.POINTER(F)]),
}

_lib = None


class DfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libdfx.so (raises if it is absent: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; run dolfinx_amd.build.build() first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call a dfx_* entry point; raise DfxError with dfx_last_error on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != DFX_OK:
        msg = L.dfx_last_error().decode()
        raise DfxError(rc, f"{name}: {msg}")
    return rc


def ptr(a):
    """Data pointer of a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(P)


def dtype_of(code):
    return np.float64 if code == F64 else np.float32


def code_of(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise ValueError(f"unsupported dtype {dtype}")
