"""Benchmark of the assemble-and-solve path on one MI355X.

Workload (the precision configuration): P1 Poisson on 128^3 unit-cube
hexahedra (Q1, 2,097,152 cells, 2,146,689 dofs, 57,066,625 nonzeros),
kappa = exp(G), G ~ N(0,1) i.i.d. per cell, f = 1, u = 0 on the boundary.

A step is one full value assembly of the system on the GPU: A (FP64) with
BC rows/columns zeroed and unit diagonal, b with lifting and bc.set. The
sparsity pattern and the slot map are built once beforehand.
`value` = scalar nonzeros assembled per second (GNNZ/s) over the timed
steps. Alongside: CG time-to-1e-8 (||b - A x||/||b|| in FP64) for plain FP64
Jacobi-PCG and for FP32 assembly + FP32 inner PCG with FP64 refinement, the
roofline of the assembly kernel and of the FP64 SpMV, and the CPU oracle
timed on the host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes of one FP64 dfx_assemble_matrix call (cell_factors_k + the row
# kernel, store mode) on the 128^3 hexahedra, per row-kernel variant, from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this bench with --variant V,
# scaled by the counters' reading for a 1 GiB device copy (profiles/
# q1hex128_assembly_spmv_cg.md gives the commands and the per-kernel figures;
# variant 15 with the 4-value records of an axis-aligned box).
ASSEMBLY_TRAFFIC_BYTES = {1: 1335891297, 4: 1331244534, 15: 928498144}


def ensure_built():
    if not os.path.exists(os.path.join(ROOT, "dolfinx_amd", "libdfx.so")):
        from dolfinx_amd import build

        build.build()


def event_ms(a, b):
    import ctypes as C

    from dolfinx_amd import _lib

    ms = C.c_float()
    _lib.call("dfx_event_elapsed", a, b, C.byref(ms))
    return ms.value


def record(slot):
    from dolfinx_amd import _lib

    _lib.call("dfx_event_record", slot)


def sync():
    from dolfinx_amd import _lib

    _lib.call("dfx_synchronize")


def host_info():
    """nproc, usable cores, CPU model and clock of this host (/proc/cpuinfo)."""
    model, mhz = None, []
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("cpu MHz"):
                    mhz.append(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cores": len(os.sched_getaffinity(0)),
            "cpu_model": model, "cpu_mhz_median": round(float(np.median(mhz)), 1) if mhz else None}


def cpu_baseline(n, kappa_input, solutions, march):
    """Oracle (CPU restatement of the reference loops, rebuilt here with -O3
    -march=native) on the host, one thread: the FP64 Q1 matrix assembly of
    the same mesh with the binary-search insert, one warm-up then the
    median of 5 (steady clock), and its FP64 Jacobi-PCG to 1e-8, whose
    solution is the reference for the GPU solutions' errors."""
    import oracle as O

    om = O.Mesh((n, n, n), O.HEX)
    dm, nd = om.dofmap(1)
    rp, cols = O.sparsity(dm, dm, nd)
    kap = kappa_input[om.original_cell_index].reshape(-1, 1)
    bn = om.boundary_nodes(1, dm)
    mark = np.zeros(nd, np.int8)
    mark[bn] = 1
    times = []
    for rep in range(6):
        t0 = time.perf_counter()
        A = O.assemble_matrix(O.K_Q1HEX_A, om.x, om.cells, dm, 1, kap, np.zeros(1), mark, mark,
                              rp, cols)
        O.set_diagonal(A, bn, 1, 1.0, rp, cols)
        t1 = time.perf_counter()
        if rep > 0:
            times.append(t1 - t0)
    t_asm = float(np.median(times))
    b = O.assemble_vector(O.K_Q1HEX_L, om.x, om.cells, dm, 1, np.ones((om.num_cells, 8)),
                          None, nd)
    O.lift_bc(b, O.K_Q1HEX_A, om.x, om.cells, dm, 1, kap, np.zeros(1), np.zeros(nd), mark)
    O.bc_set(b, bn, np.zeros(nd))
    t2 = time.perf_counter()
    xr, its = O.cg(rp, cols, A, 1, b, kmax=20000, rtol=1e-8, jacobi=True)
    t3 = time.perf_counter()
    out = {
        "value": round(len(cols) / t_asm / 1e9, 4),
        "unit": "GNNZ/s",
        "cores": 1,
        "kind": "port",
        "sample": f"FP64 Q1 matrix assembly of the {n}^3 hex mesh ({om.num_cells} cells, "
                  f"binary-search CSR insert, BC rows/cols zeroed, diagonal set): median of 5 "
                  f"after 1 warm-up = {t_asm:.3f} s (runs {', '.join(f'{t:.3f}' for t in times)}); "
                  f"Jacobi-PCG to 1e-8: {its} iterations in {t3 - t2:.1f} s",
        "march": march,
        "cpu_pcg_ms_per_iter": round((t3 - t2) / max(its, 1) * 1e3, 2),
        "cpu_pcg_iterations": int(its),
        **host_info(),
    }
    nr = np.linalg.norm(xr)
    for name, x in solutions.items():
        out[f"err_{name}_vs_oracle"] = float(np.linalg.norm(x - xr) / nr)
    return out


def sell_bytes(indptr, indices, s):
    """Bytes one SpMV through the sliced-ELL copy moves: every slot of its
    64-row slices padded to the slice's longest row (value + a 2-byte column
    offset, 4 bytes where the slice's columns span 2^16 or more), the row
    pointers, x, and y read and written."""
    n = len(indptr) - 1
    ns = (n + 63) // 64
    lens = np.diff(indptr)
    padded = np.zeros(ns * 64, np.int64)
    padded[:n] = lens
    width = padded.reshape(ns, 64).max(axis=1)
    first = indptr[np.minimum(np.arange(ns) * 64, n)]
    last = indptr[np.minimum(np.arange(ns) * 64 + 64, n)]
    cb = np.full(ns, 2)
    for sl in range(ns):
        if last[sl] > first[sl]:
            c = indices[first[sl]:last[sl]]
            if int(c.max()) - int(c.min()) >= 65536:
                cb[sl] = 4
    entries = int(np.sum(width * 64))
    return int(np.sum(width * 64 * (s + cb))) + (n + 1) * 4 + 3 * n * s, entries


def sell_spmv(A, x, y, N, nnz, s, reps):
    """Event-timed SpMV through the sliced-ELL copy (lanes -1), with its
    own bytes and the CSR-equivalent bytes nnz (s + 4) + (rows + 1) 4 +
    3 N s over the same time."""
    A.set_lanes(-1)
    for _ in range(3):
        A.mult(x, y)
    record(10)
    for _ in range(reps):
        A.mult(x, y)
    record(11)
    ms = event_ms(10, 11) / reps
    own, entries = sell_bytes(A.indptr, A.indices, s)
    csr = nnz * (s + 4) + (N + 1) * 4 + 3 * N * s
    return {"ms": round(ms, 4), "padded_entries": entries,
            "own_bytes": own, "own_gbs": round(own / ms / 1e6, 1),
            "own_frac": round(own / ms / 1e6 / HBM_PEAK_GBS, 4),
            "csr_equivalent_gbs": round(csr / ms / 1e6, 1),
            "csr_equivalent_frac": round(csr / ms / 1e6 / HBM_PEAK_GBS, 4)}


def timed_assembly(assemble, steps):
    """Seconds per call of `assemble` (device-synchronised, after one warm call)."""
    assemble()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        assemble()
    sync()
    return (time.perf_counter() - t0) / steps


def other_configs(steps):
    """The remaining workloads of the path, each assembled and solved on the
    GPU: P1 Poisson on 32^3 tetrahedra (CG+Jacobi to 1e-10), P2 vector
    elasticity on 48^3 tetrahedra (CG+Jacobi to 1e-8) and P3
    reaction-diffusion on 24^3 tetrahedra (FP64 and FP32 assembly,
    CG+Jacobi to 1e-8). Assembly rates count scalar nonzeros."""
    from dolfinx_amd import la, problems, solvers
    from dolfinx_amd.mesh import CellType

    out = {}

    def run(name, p, rtol, desc, fp32=None):
        t0 = time.perf_counter()
        A, b = problems.assemble_system(p)
        sync()
        setup = time.perf_counter() - t0
        variant = p.a.autotune(A, p.bc)
        s = timed_assembly(lambda: problems.assemble_system(p, A, b), steps)
        scalar_nnz = A.nnz * A.bs * A.bs
        # the first solve in the process also loads the solver kernels and
        # builds the sliced-ELL layout; the figure is the median of three
        # solves after it (each from x = 0)
        x = la.Vector(p.V.num_dofs)
        r = solvers.cg(A, b, x, rtol=rtol, max_it=50000, jacobi=True, check_every=64)
        first_cg_s = r.seconds
        runs = []
        for _ in range(3):
            x = la.Vector(p.V.num_dofs)
            runs.append(solvers.cg(A, b, x, rtol=rtol, max_it=50000, jacobi=True,
                                   check_every=64))
        r = sorted(runs, key=lambda q: q.seconds)[1]
        rec = {
            "workload": desc,
            "cells": int(p.mesh.topology.cells.shape[0]),
            "dofs": int(p.V.num_dofs),
            "scalar_nnz": int(scalar_nnz),
            "assembly_ms": round(s * 1e3, 4),
            "assembly_gnnz_s": round(scalar_nnz / s / 1e9, 3),
            "assembly_variant": variant[0],
            "cg_rtol": rtol,
            "cg_s": round(r.seconds, 4),
            "cg_us_per_iteration": round(r.seconds / max(r.iterations, 1) * 1e6, 1),
            "cg_iterations": r.iterations,
            "cg_rel_residual": r.true_rel_residual,
            "cg_first_solve_s": round(first_cg_s, 4),
            "first_assembly_s": round(setup, 3),
        }
        if fp32 is not None:
            A32, b32 = problems.assemble_system(fp32)
            fp32.a.autotune(A32, fp32.bc)
            s32 = timed_assembly(lambda: problems.assemble_system(fp32, A32, b32), steps)
            rec["assembly_f32_gnnz_s"] = round(scalar_nnz / s32 / 1e9, 3)
        out[name] = rec
        print(f"# {name}: {rec}", file=sys.stderr, flush=True)

    p = problems.scalar_problem((32, 32, 32), CellType.tetrahedron, kappa="one", c=0.0, f="one")
    run("p1_tet_32", p, 1e-10, "P1 Poisson -lap u = 1 on 32^3 unit-cube tetrahedra, u=0 on the "
        "boundary")
    del p
    p = problems.elasticity_problem((48, 48, 48), seed=2026)
    run("p2_elasticity_tet_48", p, 1e-8, "P2 vector elasticity on 48^3 tetrahedra, DG0 lambda, "
        "mu from E~U[0.5,2] (nu=0.3), x=0 clamped, body force (0,0,-1)")
    del p
    p = problems.p3_problem((24, 24, 24), seed=2026)
    p32 = problems.p3_problem((24, 24, 24), seed=2026, msh=p.mesh, V=p.V, dtype=np.float32)
    run("p3_tet_24", p, 1e-8, "P3 reaction-diffusion kappa grad u.grad v + u v on 24^3 "
        "tetrahedra, kappa~U[1,10] per cell, f=1", fp32=p32)
    return out


def facet_and_element_paths(steps, V_hex, A_hex, do_cpu):
    """Facet integrals on the GPU, each one assembly call of a facet form
    timed with events around the call, with the oracle's serial loop on the
    same facets beside it for the matrices: beta u v ds over the 98,304
    boundary facets of the 128^3 hexahedra (into the headline matrix, add
    mode), g v ds on the same facets (vector; per facet its pair, 4 values
    of g, 4 vertices and 4 entries read and written: 200 B), caller-supplied
    element matrices on 64^3 hexahedra (the custom-kernel path), and
    gamma [grad u.n][grad v.n] dS over the 387,072 interior facets of the
    32^3 tetrahedra. Algorithmic bytes per facet: the entity (8 / 16 B), its
    coefficient (8 B), the vertex coordinates the facet matrix uses (4 x 24
    / 8 x 24 B) and a read and a write of every entry of its matrix (16 x 16
    / 64 x 16 B)."""
    from dolfinx_amd import _lib, fem, mesh as dmesh
    from dolfinx_amd.mesh import CellType

    out = {}

    def kernel_ms(call):
        call()
        sync()
        ts = []
        for _ in range(max(steps, 5)):
            record(0)
            call()
            record(1)
            ts.append(event_ms(0, 1))
        return float(np.median(ts))

    pairs = fem.facet_domain(V_hex.mesh)
    nc = V_hex.mesh.topology.cells.shape[0]
    beta = np.full((nc, 1), 2.0)
    aR = fem.form("q1_robin", V_hex, beta, facets=pairs)
    ms = kernel_ms(lambda: _lib.call("dfx_assemble_matrix", A_hex.handle, aR.handle, None))
    nbytes = len(pairs) * (8 + 8 + 4 * 24 + 16 * 16)
    rec = {"workload": "beta u v ds on the exterior facets of 128^3 hexahedra (Q1)",
           "facets": int(len(pairs)), "kernel_ms": round(ms, 4),
           "facets_per_s": round(len(pairs) / ms * 1e3, 1),
           "roofline": {"bound": "hbm", "achieved": round(nbytes / ms / 1e6, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}}
    if do_cpu:
        import oracle as O

        xd = np.ascontiguousarray(V_hex.mesh.geometry.dofmap)
        vals = np.zeros(A_hex.nnz)
        t0 = time.perf_counter()
        O.assemble_matrix_facets(O.K_Q1HEX_ROBIN_A, V_hex.mesh.geometry.x, xd, V_hex.dofmap.list,
                                 beta, pairs, None, None, A_hex.indptr, A_hex.indices, vals)
        rec["cpu_baseline_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    out["exterior_robin_hex_128"] = rec
    del aR

    # g v ds on the same facets (g = 1 at every vertex), into a vector
    from dolfinx_amd import la as dla

    aN = fem.form("q1_neumann", V_hex, np.ones((nc, 8)), facets=pairs)
    bN = dla.Vector(V_hex.num_dofs)
    ms = kernel_ms(lambda: _lib.call("dfx_assemble_vector", bN.handle, aN.handle))
    nbytes = len(pairs) * (8 + 4 * 8 + 4 * 24 + 4 * 16)
    out["exterior_neumann_hex_128"] = {
        "workload": "g v ds on the exterior facets of 128^3 hexahedra (Q1, g at the vertices)",
        "facets": int(len(pairs)), "kernel_ms": round(ms, 4),
        "facets_per_s": round(len(pairs) / ms * 1e3, 1),
        "roofline": {"bound": "hbm", "achieved": round(nbytes / ms / 1e6, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}}
    del aN, bN

    m = dmesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (32, 32, 32), CellType.tetrahedron)
    V = fem.functionspace(m, ("Lagrange", 1))
    q = fem.interior_facet_domain(m)
    aJ = fem.form("p1_gradient_jump", V, None, [0.5], facets=q)
    A = fem.create_matrix(aJ)
    ms = kernel_ms(lambda: _lib.call("dfx_assemble_matrix", A.handle, aJ.handle, None))
    nbytes = len(q) * (16 + 8 * 24 + 64 * 16)
    rec = {"workload": "gamma [grad u.n][grad v.n] dS on the interior facets of 32^3 "
                       "tetrahedra (P1)",
           "facets": int(len(q)), "kernel_ms": round(ms, 4),
           "facets_per_s": round(len(q) / ms * 1e3, 1),
           "roofline": {"bound": "hbm", "achieved": round(nbytes / ms / 1e6, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}}
    if do_cpu:
        import oracle as O

        vals = np.zeros(A.nnz)
        t0 = time.perf_counter()
        O.assemble_matrix_interior(O.K_P1TET_GRADJUMP_A, m.geometry.x, m.geometry.dofmap,
                                   V.dofmap.list, q, None, None, A.indptr, A.indices, [0.5], vals)
        rec["cpu_baseline_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    out["interior_gradjump_tet_32"] = rec
    del aJ, A

    # caller-supplied element matrices (the custom-kernel path): 64^3 Q1
    # hexahedra, 8 x 8 tensors resident on the device; per cell its 64
    # entries, 8 dofs, and the CSR values once (stored)
    mh = dmesh.create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (64, 64, 64), CellType.hexahedron)
    Vh = fem.functionspace(mh, ("Lagrange", 1))
    ncell = mh.topology.cells.shape[0]
    E = np.random.default_rng(1).standard_normal(ncell * 64)
    from dolfinx_amd import la as dla

    Ed = dla.Vector.from_array(E)
    aE = fem.element_form(Vh, 2)
    Ah = fem.create_matrix(aE)
    # store mode each time (zeroing only marks the values; the assembly
    # then overwrites every entry)
    ms = kernel_ms(lambda: (Ah.set_value(0.0),
                            _lib.call("dfx_assemble_matrix_elements", Ah.handle, aE.handle, None,
                                      Ed.handle)))
    nbytes = ncell * (64 * 8 + 8 * 4) + Ah.nnz * 8
    rec = {"workload": "caller-supplied 8x8 element matrices on 64^3 hexahedra (device-resident)",
           "cells": int(ncell), "scalar_nnz": int(Ah.nnz), "kernel_ms": round(ms, 4),
           "gnnz_s": round(Ah.nnz / ms / 1e6, 2),
           "roofline": {"bound": "hbm", "achieved": round(nbytes / ms / 1e6, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}}
    if do_cpu:
        import oracle as O

        t0 = time.perf_counter()
        O.add_element_matrices(Vh.dofmap.list, 1, E, None, None, Ah.indptr, Ah.indices)
        rec["cpu_baseline_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    out["element_matrices_hex_64"] = rec
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--only-configs", action="store_true",
                    help="run only the other workloads (P1 32^3, P2 48^3, P3 24^3) and print "
                         "their figures (no benchmark line; for profiling them)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="fix the FP64/FP32 assembly variant of the timed workload "
                         "(default: autotuned)")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: the assembly path is measured on a single MI355X; --gpus must be 1"}))
        sys.exit(2)

    ensure_built()
    from dolfinx_amd import _lib, fem, la, problems, solvers
    from dolfinx_amd.mesh import CellType

    march = None
    if not args.no_cpu:
        # the oracle timed as the CPU baseline is compiled for this host
        import oracle as O

        march = O.use_native()

    if args.only_configs:
        print(json.dumps({"other_configs": other_configs(args.steps)}))
        return

    n = args.n
    t_setup = time.perf_counter()
    p64 = problems.scalar_problem((n, n, n), CellType.hexahedron, kappa="exp_normal", c=0.0,
                                  f="one", seed=args.seed, dtype=np.float64)
    p32 = problems.scalar_problem((n, n, n), CellType.hexahedron, kappa=p64.kappa, c=0.0,
                                  f=p64.f, msh=p64.mesh, V=p64.V, dtype=np.float32)
    sp = fem.create_sparsity_pattern(p64.a)
    A64 = la.MatrixCSR.from_sparsity(sp, np.float64)
    A32 = la.MatrixCSR.from_sparsity(sp, np.float32)
    b64 = la.Vector(p64.V.num_dofs, np.float64)
    b32 = la.Vector(p64.V.num_dofs, np.float32)
    setup_s = time.perf_counter() - t_setup
    nnz = A64.nnz
    N = p64.V.num_dofs
    C = p64.mesh.topology.cells.shape[0]
    Nv = p64.mesh.geometry.x.shape[0]
    print(f"# setup {setup_s:.1f} s: cells {C} dofs {N} nnz {nnz}", file=sys.stderr, flush=True)

    if args.variant < 0:
        variant64 = p64.a.autotune(A64, p64.bc)
        variant32 = p32.a.autotune(A32, p32.bc)
    else:
        p64.a.set_variant(args.variant)
        p32.a.set_variant(args.variant)
        variant64 = variant32 = (args.variant, float("nan"))
    print(f"# assembly variants: f64 {variant64} f32 {variant32}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        problems.assemble_system(p64, A64, b64)
    sync()

    # ---- timed region: K full FP64 assemblies ----
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        problems.assemble_system(p64, A64, b64)
    sync()
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    gnnz = nnz / step_s / 1e9

    # ---- assembly kernel duration (events bracket the single launch) ----
    reps = 5
    kms = []
    for _ in range(reps):
        A64.set_value(0.0)
        record(0)
        _lib.call("dfx_assemble_matrix", A64.handle, p64.a.handle, p64.bc.handle)
        record(1)
        kms.append(event_ms(0, 1))
    asm_ms = float(np.median(kms))
    # algorithmic bytes of the matrix kernel: dofmap + coords + kappa + values
    asm_bytes = C * 8 * 4 + Nv * 3 * 8 + C * 8 + nnz * 8
    asm_gbs = asm_bytes / (asm_ms * 1e-3) / 1e9

    # ---- apply_lifting with nonzero Dirichlet data g = 1 + x on the
    # boundary (the timed step's lifting has g = 0 and returns at once) ----
    gvals = np.zeros(N)
    bdofs = np.asarray(p64.bc.dofs)
    gvals[bdofs] = 1.0 + p64.mesh.geometry.x[bdofs, 0]
    bcg = fem.dirichletbc(p64.V, bdofs, gvals)
    bl = la.Vector(N)
    _lib.call("dfx_apply_lifting", bl.handle, p64.a.handle, bcg.handle, None, 1.0)
    sync()
    lts = []
    for _ in range(5):
        record(8)
        _lib.call("dfx_apply_lifting", bl.handle, p64.a.handle, bcg.handle, None, 1.0)
        record(9)
        lts.append(event_ms(8, 9))
    lift_ms = float(np.median(lts))
    del bl, bcg

    # ---- FP32 assembly ----
    for _ in range(2):
        problems.assemble_system(p32, A32, b32)
    sync()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        problems.assemble_system(p32, A32, b32)
    sync()
    asm32_s = (time.perf_counter() - t2) / args.steps

    # ---- SpMV lanes per row picked by the autotuner ----
    lanes64, _ = A64.autotune()
    lanes32, _ = A32.autotune()

    # ---- SpMV roofline (FP64) ----
    x = la.Vector.from_array(np.random.default_rng(1).standard_normal(N))
    y = la.Vector(N)
    for _ in range(3):
        A64.mult(x, y)
    record(2)
    nsp = 20
    for _ in range(nsp):
        A64.mult(x, y)
    record(3)
    spmv_ms = event_ms(2, 3) / nsp
    spmv_bytes = nnz * (8 + 4) + (N + 1) * 4 + 2 * N * 8 + N * 8  # y is read and written
    spmv_gbs = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    x32 = la.Vector.from_array(np.random.default_rng(1).standard_normal(N).astype(np.float32))
    y32 = la.Vector(N, np.float32)
    for _ in range(3):
        A32.mult(x32, y32)
    record(4)
    for _ in range(nsp):
        A32.mult(x32, y32)
    record(5)
    spmv32_ms = event_ms(4, 5) / nsp
    spmv32_gbs = (nnz * 8 + (N + 1) * 4 + 3 * N * 4) / (spmv32_ms * 1e-3) / 1e9

    # ---- SpMV through the solvers' sliced-ELL copy (the kernel the CG
    # iterations run), FP64 and FP32 ----
    sell = sell_spmv(A64, x, y, N, nnz, 8, nsp)
    sell32 = sell_spmv(A32, x32, y32, N, nnz, 4, nsp)
    A64.set_lanes(lanes64)
    A32.set_lanes(lanes32)

    # ---- practical streaming bandwidth: device copy of 2^27 doubles ----
    ncopy = 1 << 27
    src = la.Vector(ncopy)
    dst = la.Vector(ncopy)
    src.set(1.0)
    dst.copy_from(src)
    record(6)
    for _ in range(5):
        dst.copy_from(src)
    record(7)
    copy_ms = event_ms(6, 7) / 5
    copy_gbs = 2 * 8 * ncopy / (copy_ms * 1e-3) / 1e9
    del src, dst

    out_solve = {}
    solutions = {}
    if not args.no_solve:
        problems.assemble_system(p64, A64, b64)
        problems.assemble_system(p32, A32, b32)
        sync()
        # warm-up: two iterations of each solver load their kernels (the
        # first launch of a kernel in a process loads it)
        from dolfinx_amd._lib import DfxError

        for warm in (lambda x: solvers.cg(A64, b64, x, rtol=1e-30, max_it=2, check_every=1),
                     lambda x: solvers.mixed_refinement(A64, A32, b64, x, rtol=1e-30,
                                                        max_outer=1, max_inner=2, check_every=1,
                                                        fallback=False)):
            try:
                warm(la.Vector(N))
            except DfxError:
                pass
        xs = la.Vector(N)
        r64 = solvers.cg(A64, b64, xs, rtol=1e-8, max_it=20000, jacobi=True, check_every=64)
        x64 = xs.array
        xm = la.Vector(N)
        rm = solvers.mixed_refinement(A64, A32, b64, xm, rtol=1e-8, inner_rtol=1e-4,
                                      check_every=64, restart=True)
        xmix = xm.array
        xr_ = la.Vector(N)
        rr_ = solvers.mixed_refinement(A64, A32, b64, xr_, rtol=1e-8, inner_rtol=1e-4,
                                       check_every=64, restart=False)
        # reference-accuracy solution for the error figures
        xt = la.Vector(N)
        solvers.cg(A64, b64, xt, rtol=1e-11, max_it=40000, jacobi=True, check_every=128)
        xref = xt.array
        nref = np.linalg.norm(xref)
        solutions = {"fp64": x64, "mixed": xmix}
        out_solve = {
            "cg_fp64_s": round(r64.seconds, 4),
            "cg_fp64_iterations": r64.iterations,
            "cg_fp64_rel_residual": r64.true_rel_residual,
            "mixed_s": round(rm.seconds, 4),
            "mixed_outer": rm.outer,
            "mixed_inner_iterations": rm.inner_total,
            "mixed_fp64_fallback_iterations": rm.fallback_iterations,
            "mixed_rel_residual": rm.true_rel_residual,
            "mixed_speedup_vs_fp64": round(r64.seconds / rm.seconds, 3),
            "mixed_reliable_update_s": round(rr_.seconds, 4),
            "mixed_reliable_update_inner_iterations": rr_.inner_total,
            "mixed_reliable_update_fp64_fallback_iterations": rr_.fallback_iterations,
            "err_fp64_vs_tight": float(np.linalg.norm(x64 - xref) / nref),
            "err_mixed_vs_tight": float(np.linalg.norm(xmix - xref) / nref),
            "err_mixed_vs_fp64": float(np.linalg.norm(xmix - x64) / np.linalg.norm(x64)),
        }

    serve_out = None
    if not args.no_serve:
        from dolfinx_amd import serve

        t_s = time.perf_counter()
        svc = serve.SolveService(hbm_budget_bytes=16 << 30, max_slots=64, rtol=1e-8,
                                 iters_per_step=16, sizes=range(8, 49))
        setup_serve = time.perf_counter() - t_s
        # warm the graph and every template once
        svc.run([serve.make_request(10_000 + i, seed=99) for i in range(8)], rate=1000.0)
        reqs = [serve.make_request(i, seed=2026) for i in range(args.serve_requests)]
        keep = [0, 1, 2, 3, 4]
        st = svc.run(reqs, rate=args.serve_rate, keep_solutions=keep)
        diffs = [float(np.max(np.abs(svc.solve_alone(r) - r.x))) for r in reqs if r.rid in keep]
        serve_out = st.summary(args.serve_rate)
        serve_out.update({
            "workload": "P1 tets n^3, n~U{8..48}, kappa~logN(0,0.5)/cell, c~U[0,10], "
                        "f~N(0,1)/vertex, u=0 on the boundary, CG+Jacobi to 1e-8",
            "all_converged": all(r.converged for r in reqs),
            "mean_iterations": round(float(np.mean([r.iterations for r in reqs])), 1),
            "alone_vs_mix_max_abs_diff": max(diffs),
            "template_setup_s": round(setup_serve, 2),
        })
        del svc

    facets_out = None
    if not args.no_configs:
        facets_out = facet_and_element_paths(args.steps, p64.V, A64, not args.no_cpu)
        print(f"# facet and element paths: {facets_out}", file=sys.stderr, flush=True)

    configs_out = None
    if not args.no_configs:
        del A32, b32
        configs_out = other_configs(args.steps)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, p64.kappa, solutions, march)

    line = {
        "metric": "assembly GNNZ/s + CG time-to-1e-8, mixed vs FP64 on GPU; serve p99 latency",
        "value": round(gnnz, 4),
        "unit": "GNNZ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"Q1 Poisson on {n}^3 unit-cube hexahedra, kappa=exp(N(0,1)) per cell, "
                        "f=1, u=0 on the boundary (precision configuration)",
            "cells": C, "dofs": N, "nnz": nnz,
            "step": "FP64 assemble_matrix (BC rows/cols zeroed, diag 1) + assemble_vector "
                    "+ apply_lifting (g = 0, x0 = 0: returns at once; the reference's lift_bc "
                    "visits the boundary cells and adds exactly zero) + set_bc",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "cell_factors_k + row kernel of the kept variant (FP64 Q1 matrix assembly, one dfx_assemble_matrix call)",
            "achieved": round(asm_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(asm_gbs / HBM_PEAK_GBS, 4),
            "traffic": ASSEMBLY_TRAFFIC_BYTES.get(variant64[0]) if n == 128 else None,
            "algorithmic_bytes": asm_bytes,
            "kernel_ms": round(asm_ms, 4),
        },
        "apply_lifting_nonzero_g_ms": round(lift_ms, 4),
        "copy_gbs": round(copy_gbs, 1),
        "assembly_f32_gnnz_s": round(nnz / asm32_s / 1e9, 4),
        "assembly_variant": {"f64": variant64[0], "f32": variant32[0]},
        "spmv": {
            "lanes_per_row": {"f64": lanes64, "f32": lanes32},
            "f64_ms": round(spmv_ms, 4), "f64_gbs": round(spmv_gbs, 1),
            "f64_frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
            "f32_ms": round(spmv32_ms, 4), "f32_gbs": round(spmv32_gbs, 1),
            "f32_frac": round(spmv32_gbs / HBM_PEAK_GBS, 4),
            "sliced_ell_f64": sell,
            "sliced_ell_f32": sell32,
        },
        **out_solve,
        "serve": serve_out,
        "other_configs": configs_out,
        "facet_and_element_paths": facets_out,
        "cpu_baseline": cpu,
        "setup_s": round(setup_s, 2),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
