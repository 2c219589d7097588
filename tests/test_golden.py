"""Frozen golden systems (tests/golden/system_*.npz, written by
tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every fixture, and the product's host setup
(mesh, dofmap, boundary dofs, pattern) equals it bit for bit.
GPU: the product's assembled A and b equal the fixture's within 1e-12.
"""

import ctypes as C
import glob
import os

import numpy as np
import pytest

import oracle as O
from dolfinx_amd import _lib, fem, mesh, problems

from conftest import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "system_*.npz")))
IDS = [os.path.basename(f)[7:-4] for f in FILES]


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FILES) == 4, FILES


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_oracle_reproduces_fixture(path):
    g = load(path)
    m = O.Mesh(tuple(g["n"]), int(g["celltype"]))
    assert np.array_equal(m.x, g["x"]) and np.array_equal(m.cells, g["cells"])
    dm, nd = m.dofmap(int(g["degree"]))
    assert nd == int(g["num_nodes"]) and np.array_equal(dm, g["dofmap"])
    rp, cols = O.sparsity(dm, dm, nd)
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(cols, g["cols"])
    bs = int(g["bs"])
    if bs == 1:
        kap = g["kappa"][m.original_cell_index].reshape(-1, 1)
        c = np.array([float(g["c"])])
        mark = np.zeros(nd, np.int8)
        mark[g["bdofs"]] = 1
        if int(g["degree"]) == 1:
            tet = int(g["celltype"]) == O.TET
            kA = O.K_P1TET_A if tet else O.K_Q1HEX_A
            kL = O.K_P1TET_L if tet else O.K_Q1HEX_L
            b = O.assemble_vector(kL, m.x, m.cells, dm, 1,
                                  g["f"][m.input_vertex_index][m.cells], None, nd)
        else:
            kA, kL = O.K_P3TET_A, O.K_P3TET_L
            b = O.assemble_vector(kL, m.x, m.cells, dm, 1, None, g["fconst"], nd)
        A = O.assemble_matrix(kA, m.x, m.cells, dm, 1, kap, c, mark, mark, rp, cols)
        O.set_diagonal(A, g["bdofs"], 1, 1.0, rp, cols)
        O.lift_bc(b, kA, m.x, m.cells, dm, 1, kap, c, g["g"], mark)
        O.bc_set(b, g["bdofs"], g["g"])
    else:
        lam, mu = problems.lame_from_E(g["E"])
        co = np.stack([lam, mu], 1)[m.original_cell_index]
        mark = np.zeros(3 * nd, np.int8)
        mark[g["bdofs"]] = 1
        A = O.assemble_matrix(O.K_P2TET_ELAST_A, m.x, m.cells, dm, 3, co, None, mark, mark, rp,
                              cols)
        O.set_diagonal(A, g["bdofs"], 3, 1.0, rp, cols)
        b = O.assemble_vector(O.K_P2TET_ELAST_L, m.x, m.cells, dm, 3, None, g["fconst"], 3 * nd)
        O.bc_set(b, g["bdofs"], np.zeros(3 * nd))
    assert np.allclose(A, g["A"], rtol=1e-14, atol=1e-15 * np.abs(g["A"]).max())
    assert np.allclose(b, g["b"], rtol=1e-14, atol=1e-15 * max(np.abs(g["b"]).max(), 1.0))


def product_space(g):
    ct = mesh.CellType(int(g["celltype"]))
    m = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], tuple(g["n"]), ct)
    deg, bs = int(g["degree"]), int(g["bs"])
    V = fem.functionspace(m, ("Lagrange", deg, (3,)) if bs == 3 else ("Lagrange", deg))
    return m, V


@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_product_setup_matches_fixture(path):
    g = load(path)
    m, V = product_space(g)
    assert np.array_equal(m.geometry.x, g["x"])
    assert np.array_equal(m.topology.cells, g["cells"])
    assert np.array_equal(m.topology.original_cell_index, g["original_cell_index"])
    assert np.array_equal(V.dofmap.list, g["dofmap"])
    if int(g["bs"]) == 1:
        bd = fem.locate_boundary_dofs(V)
    else:
        bd = fem.locate_boundary_dofs(V, axis=0, value=0.0)
    assert np.array_equal(bd, g["bdofs"])
    h = C.c_void_p()
    _lib.call("dfx_sparsity_create", V.handle, V.handle, C.byref(h))
    from dolfinx_amd.la import SparsityPattern

    sp = SparsityPattern(h)
    assert np.array_equal(sp.row_ptr, g["row_ptr"]) and np.array_equal(sp.cols, g["cols"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=IDS)
def test_product_assembly_matches_fixture(path):
    g = load(path)
    m, V = product_space(g)
    if int(g["bs"]) == 1:
        kap = fem.pack_cellwise(m, g["kappa"])
        if int(g["degree"]) == 1:
            tet = int(g["celltype"]) == O.TET
            a = fem.form("p1_stiffness_reaction" if tet else "q1_stiffness_reaction", V, kap,
                         [float(g["c"])])
            L = fem.form("p1_source" if tet else "q1_source", V, fem.pack_vertexwise(m, g["f"]))
        else:
            a = fem.form("p3_stiffness_reaction", V, kap, [float(g["c"])])
            L = fem.form("p3_source", V, None, list(g["fconst"]))
        bc = fem.dirichletbc(V, g["bdofs"], g["g"])
    else:
        lam, mu = problems.lame_from_E(g["E"])
        a = fem.form("p2_elasticity", V, fem.pack_cellwise(m, np.stack([lam, mu], 1)))
        L = fem.form("p2_body_force", V, None, list(g["fconst"]))
        bc = fem.dirichletbc(V, g["bdofs"])
    A = fem.assemble_matrix(a, [bc])
    rp, cols = A.pattern()
    assert np.array_equal(rp, g["row_ptr"]) and np.array_equal(cols, g["cols"])
    assert rel_err(A.data, g["A"]) <= 1e-12
    assert np.max(np.abs(A.data - g["A"])) <= 1e-12 * np.max(np.abs(g["A"]))
    b = fem.assemble_vector(L)
    fem.apply_lifting(b, [a], [[bc]])
    fem.set_bc(b, [bc])
    assert rel_err(b.array, g["b"]) <= 1e-12
