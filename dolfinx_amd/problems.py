"""Builders for the benchmark problems of this path (seeded, synthetic).

Random fields use numpy's PCG64 generator with a fixed seed per problem and
are indexed by the *input* (lexicographic) cell or vertex of create_box, so
they do not depend on the cell reordering.

  poisson_tet   -div(kappa grad u) + c u = f on create_box tetrahedra, P1
  poisson_hex   the same on hexahedra, Q1
  u = 0 on the whole boundary (DirichletBC on all exterior facet dofs)
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import fem, la, mesh


@dataclass
class ScalarProblem:
    mesh: mesh.Mesh
    V: fem.FunctionSpace
    a: fem.Form
    L: fem.Form
    bc: fem.DirichletBC
    kappa: np.ndarray       # per input cell
    c: float
    f: np.ndarray           # per input vertex
    dtype: np.dtype


def kappa_field(kind, num_cells, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    if kind == "one":
        return np.ones(num_cells)
    if kind == "exp_normal":          # kappa = exp(G), G ~ N(0, 1)
        return np.exp(rng.standard_normal(num_cells))
    if kind == "lognormal_0.5":       # log-normal(0, 0.5)
        return np.exp(0.5 * rng.standard_normal(num_cells))
    if kind == "uniform_1_10":
        return rng.uniform(1.0, 10.0, num_cells)
    raise ValueError(kind)


def scalar_problem(n, cell, kappa="one", c=0.0, f="one", seed=0, dtype=np.float64,
                   bc=True, msh=None, V=None) -> ScalarProblem:
    """Build mesh, space, forms and the homogeneous Dirichlet BC."""
    ct = mesh.CellType(cell)
    if msh is None:
        msh = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, ct)
    if V is None:
        V = fem.functionspace(msh, ("Lagrange", 1))
    nc = msh.topology.cells.shape[0]
    nv = msh.geometry.x.shape[0]
    kap = kappa_field(kappa, nc, seed) if isinstance(kappa, str) else np.asarray(kappa)
    if isinstance(f, str):
        rng = np.random.Generator(np.random.PCG64(seed + 1))
        fv = np.ones(nv) if f == "one" else rng.standard_normal(nv)
    else:
        fv = np.asarray(f)
    tet = ct == mesh.CellType.tetrahedron
    a = fem.form("p1_stiffness_reaction" if tet else "q1_stiffness_reaction", V,
                 fem.pack_cellwise(msh, kap), [c], dtype)
    L = fem.form("p1_source" if tet else "q1_source", V, fem.pack_vertexwise(msh, fv),
                 None, dtype)
    dofs = fem.locate_boundary_dofs(V) if bc else np.zeros(0, np.int32)
    bcs = fem.dirichletbc(V, dofs)
    return ScalarProblem(msh, V, a, L, bcs, kap, c, fv, np.dtype(dtype))


def assemble_system(p: ScalarProblem, A: la.MatrixCSR | None = None, b: la.Vector | None = None):
    """The reference's Poisson pipeline: A = assemble_matrix(a, [bc]) (BC
    rows/cols zeroed, diagonal 1); b = assemble_vector(L); apply_lifting;
    set_bc."""
    return fem.assemble_system(p.a, p.L, [p.bc], 1.0, A, b)


@dataclass
class VectorProblem:
    mesh: mesh.Mesh
    V: fem.FunctionSpace
    a: fem.Form
    L: fem.Form
    bc: fem.DirichletBC
    E: np.ndarray           # Young's modulus per input cell
    dtype: np.dtype


def lame_from_E(E, nu=0.3):
    """Lame parameters scaled by E (fixed Poisson ratio nu)."""
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return lam, mu


def elasticity_problem(n, seed=0, dtype=np.float64, msh=None, V=None) -> VectorProblem:
    """P2 vector elasticity on create_box tetrahedra: sigma(u):eps(v) with
    DG0 lambda, mu from E ~ U[0.5, 2] per cell (nu = 0.3), body force
    (0, 0, -1), the x = 0 face clamped."""
    if msh is None:
        msh = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    if V is None:
        V = fem.functionspace(msh, ("Lagrange", 2, (3,)))
    nc = msh.topology.cells.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    E = rng.uniform(0.5, 2.0, nc)
    lam, mu = lame_from_E(E)
    a = fem.form("p2_elasticity", V, fem.pack_cellwise(msh, np.stack([lam, mu], 1)), None, dtype)
    L = fem.form("p2_body_force", V, None, [0.0, 0.0, -1.0], dtype)
    bc = fem.dirichletbc(V, fem.locate_boundary_dofs(V, axis=0, value=0.0))
    return VectorProblem(msh, V, a, L, bc, E, np.dtype(dtype))


def p3_problem(n, seed=0, dtype=np.float64, c=1.0, msh=None, V=None) -> ScalarProblem:
    """P3 reaction-diffusion on create_box tetrahedra: kappa grad u.grad v
    + c u v with kappa ~ U[1, 10] per cell, f = 1, no Dirichlet condition."""
    if msh is None:
        msh = mesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], n, mesh.CellType.tetrahedron)
    if V is None:
        V = fem.functionspace(msh, ("Lagrange", 3))
    nc = msh.topology.cells.shape[0]
    kap = kappa_field("uniform_1_10", nc, seed)
    a = fem.form("p3_stiffness_reaction", V, fem.pack_cellwise(msh, kap), [c], dtype)
    L = fem.form("p3_source", V, None, [1.0], dtype)
    bcs = fem.dirichletbc(V, np.zeros(0, np.int32))
    return ScalarProblem(msh, V, a, L, bcs, kap, c, np.ones(1), np.dtype(dtype))
