"""The C ABI library loads and exports every symbol include/dfx.h declares
(no device calls: this runs without a GPU)."""

import ctypes
import os
import re

from dolfinx_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dfx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ["dfx_assemble_matrix", "dfx_assemble_vector", "dfx_apply_lifting",
              "dfx_set_diagonal", "dfx_bc_set", "dfx_spmv", "dfx_inner_product",
              "dfx_cg_solve", "dfx_mixed_solve", "dfx_csr_create", "dfx_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_host_setup_errors_are_reported():
    import numpy as np
    import pytest
    from dolfinx_amd import mesh

    with pytest.raises(_lib.DfxError, match="At least one cell"):
        mesh.create_box([[0, 0, 0], [1, 1, 1]], [0, 2, 2])
    with pytest.raises(_lib.DfxError, match="p\\[0\\] < p\\[1\\]"):
        mesh.create_box([[0, 0, 0], [1, -1, 1]], [2, 2, 2])
    assert np.isfinite(1.0)


def test_csr_pattern_rows_must_be_sorted():
    """dfx_csr_create checks each row's columns are strictly increasing (the
    device kernels binary-search rows); the check runs before any device
    work, so it is reported without a GPU."""
    import numpy as np
    import pytest
    from dolfinx_amd import la

    rp = np.array([0, 2, 4], np.int64)
    with pytest.raises(_lib.DfxError, match="not strictly increasing") as e:
        la.MatrixCSR(rp, np.array([1, 0, 0, 1], np.int32))
    assert e.value.code == _lib.DFX_ERR_ARG
    with pytest.raises(_lib.DfxError, match="not strictly increasing"):
        la.MatrixCSR(rp, np.array([0, 0, 0, 1], np.int32))   # duplicate column
    with pytest.raises(_lib.DfxError, match="column index out of range"):
        la.MatrixCSR(rp, np.array([0, 2, 0, 1], np.int32), ncols=2)


def test_graph_calls_outside_a_capture_are_refused():
    """dfx_graph_end without dfx_graph_begin and a launch of no graph fail
    before touching a device; destroying no graph is a no-op."""
    import ctypes as C
    import pytest

    h = C.c_void_p()
    with pytest.raises(_lib.DfxError, match="no graph capture is open"):
        _lib.call("dfx_graph_end", C.byref(h))
    with pytest.raises(_lib.DfxError, match="null graph"):
        _lib.call("dfx_graph_launch", None)
    assert _lib.lib().dfx_graph_destroy(None) == _lib.DFX_OK
