"""The C++ drop-in (include/dolfinx_amd/dolfinx_amd.h) runs the reference's
Poisson pipeline (cpp/demo/poisson/main.cpp shape) on the GPU; its A, b and
x match the oracle (FP64 values within 1e-12, solution within 1e-8)."""

import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "dolfinx_amd", "_build", "poisson_demo")


def read_arrays(path):
    out = []
    with open(path, "rb") as fh:
        for _ in range(3):
            n = int(np.frombuffer(fh.read(8), np.int64)[0])
            out.append(np.frombuffer(fh.read(8 * n), np.float64))
    return out


@pytest.mark.parametrize("cell,name,n", [(O.TET, "tet", 10), (O.HEX, "hex", 16)])
def test_cpp_poisson_pipeline(tmp_path, cell, name, n):
    assert os.path.exists(EXE), "build the examples first (dolfinx_amd.build)"
    out = tmp_path / "out.bin"
    r = subprocess.run([EXE, str(n), name, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    A, b, x = read_arrays(out)
    m = O.Mesh((n, n, n), cell)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    kap = (1.0 + 0.5 * np.sin(m.original_cell_index.astype(float))).reshape(-1, 1)
    bn = m.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    kA = O.K_P1TET_A if cell == O.TET else O.K_Q1HEX_A
    kL = O.K_P1TET_L if cell == O.TET else O.K_Q1HEX_L
    Ar = O.assemble_matrix(kA, m.x, m.cells, dm, 1, kap, np.zeros(1), mark, mark, rp, cols)
    O.set_diagonal(Ar, bn, 1, 1.0, rp, cols)
    br = O.assemble_vector(kL, m.x, m.cells, dm, 1, np.ones((m.num_cells, m.nvpc)), None, nd)
    O.bc_set(br, bn, np.zeros(nd))
    xr, _ = O.cg(rp, cols, Ar, 1, br, rtol=1e-12)
    assert rel_err(A, Ar) <= 1e-12
    assert rel_err(b, br) <= 1e-12
    assert rel_err(x, xr) <= 1e-8
    assert info["iterations"] > 0


def test_cpp_custom_kernel_forms(tmp_path):
    """Forms built from hand-written C++ element kernels with the
    reference's kernel ABI (fem::Form(kernel_ptr, rank, ...)), assembled
    through the drop-in, equal the library's device kernels for the same
    forms and the oracle; the same kernel in a host loop over the cells,
    its element matrices added through A.mat_add_values() (the MatSet
    functor) and scatter_rev, gives the bits of the form without BCs, and
    the oracle's cell-by-cell insertion. A hand-written exterior facet
    kernel (beta u v ds, fem::Form(kernel_ptr, ..., IntegralType::
    exterior_facet, pairs)) equals the device Robin kernel and the
    oracle's facet assembly."""
    exe = os.path.join(ROOT, "dolfinx_amd", "_build", "custom_kernel_demo")
    assert os.path.exists(exe), "build the examples first (dolfinx_amd.build)"
    n = 6
    out = tmp_path / "out.bin"
    r = subprocess.run([exe, str(n), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["max_diff_rel"] <= 1e-13
    assert info["host_loop_max_diff_rel"] == 0.0
    with open(out, "rb") as fh:
        arrs = []
        for _ in range(8):
            k = int(np.frombuffer(fh.read(8), np.int64)[0])
            arrs.append(np.frombuffer(fh.read(8 * k), np.float64))
    Ak, bk, Ad, bd, Ah, An, Rk, Rd = arrs
    assert np.array_equal(Ah, An)
    assert rel_err(bk, bd) <= 1e-13
    m = O.Mesh((n, n, n), O.TET)
    dm, nd = m.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    kap = (1.0 + 0.5 * np.sin(m.original_cell_index.astype(float))).reshape(-1, 1)
    bn = m.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    Ar = O.assemble_matrix(O.K_P1TET_A, m.x, m.cells, dm, 1, kap, np.array([0.75]), mark, mark,
                           rp, cols)
    O.set_diagonal(Ar, bn, 1, 1.0, rp, cols)
    assert rel_err(Ak, Ar) <= 1e-12
    none = np.zeros(nd, np.int8)
    Arn = O.assemble_matrix(O.K_P1TET_A, m.x, m.cells, dm, 1, kap, np.array([0.75]), none, none,
                            rp, cols)
    assert rel_err(Ah, Arn) <= 1e-12
    assert rel_err(Rk, Rd) <= 1e-13
    beta = (2.0 + np.cos(m.original_cell_index.astype(float))).reshape(-1, 1)
    Rr = O.assemble_matrix_facets(O.K_P1TET_ROBIN_A, m.x, m.cells, dm, beta,
                                  m.facet_domain(), None, None, rp, cols)
    assert rel_err(Rk, Rr) <= 1e-12
