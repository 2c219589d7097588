"""Write the frozen golden systems under tests/golden/ from the oracle.

Each fixture holds the inputs of one small problem (coefficients per input
cell / vertex, constants, Dirichlet data per dof) and the oracle's outputs:
mesh after reordering, dofmap, boundary dofs, CSR pattern, assembled A
(BC rows/columns zeroed, unit diagonal) and b (lifting + bc.set).

Regenerate with  python tests/golden/make_golden.py  (only when the oracle
is deliberately changed; the tests check the oracle still reproduces them).
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

NU = 0.3


def lame(E):
    return E * NU / ((1 + NU) * (1 - 2 * NU)), E / (2 * (1 + NU))


def scalar_case(name, ct, n, degree, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    m = O.Mesh(n, ct)
    dm, nd = m.dofmap(degree)
    rp, cols = O.sparsity(dm, dm, nd)
    kappa = np.exp(0.5 * rng.standard_normal(m.num_cells))      # per input cell
    c = float(rng.uniform(0.5, 2.0))
    bdofs = m.boundary_nodes(degree, dm)
    g = np.zeros(nd)
    g[bdofs] = np.cos(1.0 + np.arange(len(bdofs)))
    mark = np.zeros(nd, np.int8)
    mark[bdofs] = 1
    kap = kappa[m.original_cell_index].reshape(-1, 1)
    if degree == 1:
        f = rng.standard_normal(m.num_vertices)                   # per input vertex
        kA = O.K_P1TET_A if ct == O.TET else O.K_Q1HEX_A
        kL = O.K_P1TET_L if ct == O.TET else O.K_Q1HEX_L
        b = O.assemble_vector(kL, m.x, m.cells, dm, 1, f[m.input_vertex_index][m.cells], None, nd)
        fconst = np.zeros(0)
    else:
        f = np.zeros(0)
        fconst = np.array([float(rng.uniform(-1.0, 1.0))])
        kA, kL = O.K_P3TET_A, O.K_P3TET_L
        b = O.assemble_vector(kL, m.x, m.cells, dm, 1, None, fconst, nd)
    A = O.assemble_matrix(kA, m.x, m.cells, dm, 1, kap, np.array([c]), mark, mark, rp, cols)
    O.set_diagonal(A, bdofs, 1, 1.0, rp, cols)
    O.lift_bc(b, kA, m.x, m.cells, dm, 1, kap, np.array([c]), g, mark)
    O.bc_set(b, bdofs, g)
    np.savez_compressed(
        os.path.join(HERE, f"system_{name}.npz"),
        celltype=ct, n=np.array(n), degree=degree, bs=1,
        kappa=kappa, c=c, f=f, fconst=fconst, g=g,
        x=m.x, cells=m.cells, original_cell_index=m.original_cell_index,
        input_vertex_index=m.input_vertex_index, dofmap=dm, num_nodes=nd, bdofs=bdofs,
        row_ptr=rp, cols=cols, A=A, b=b)


def elasticity_case(name, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    m = O.Mesh(n, O.TET)
    dm, nn = m.dofmap(2)
    rp, cols = O.sparsity(dm, dm, nn)
    E = rng.uniform(0.5, 2.0, m.num_cells)                        # per input cell
    lam, mu = lame(E)
    co = np.stack([lam, mu], 1)[m.original_cell_index]
    nodes = m.boundary_nodes(2, dm, 1, 0, 0.0)
    bdofs = (3 * nodes[:, None] + np.arange(3)).ravel().astype(np.int32)
    mark = np.zeros(3 * nn, np.int8)
    mark[bdofs] = 1
    A = O.assemble_matrix(O.K_P2TET_ELAST_A, m.x, m.cells, dm, 3, co, None, mark, mark, rp, cols)
    O.set_diagonal(A, bdofs, 3, 1.0, rp, cols)
    fconst = np.array([0.0, 0.0, -1.0])
    b = O.assemble_vector(O.K_P2TET_ELAST_L, m.x, m.cells, dm, 3, None, fconst, 3 * nn)
    O.bc_set(b, bdofs, np.zeros(3 * nn))
    np.savez_compressed(
        os.path.join(HERE, f"system_{name}.npz"),
        celltype=O.TET, n=np.array(n), degree=2, bs=3, E=E, fconst=fconst,
        x=m.x, cells=m.cells, original_cell_index=m.original_cell_index,
        input_vertex_index=m.input_vertex_index, dofmap=dm, num_nodes=nn, bdofs=bdofs,
        row_ptr=rp, cols=cols, A=A, b=b)


CASES = {
    "p1_tet_3x3x3": lambda: scalar_case("p1_tet_3x3x3", O.TET, (3, 3, 3), 1, 101),
    "q1_hex_3x2x4": lambda: scalar_case("q1_hex_3x2x4", O.HEX, (3, 2, 4), 1, 102),
    "p3_tet_2x2x2": lambda: scalar_case("p3_tet_2x2x2", O.TET, (2, 2, 2), 3, 103),
    "p2v_tet_2x2x2": lambda: elasticity_case("p2v_tet_2x2x2", (2, 2, 2), 104),
}


if __name__ == "__main__":
    for name, make in CASES.items():
        make()
        print("wrote", name)
