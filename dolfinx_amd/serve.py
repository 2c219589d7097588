"""Continuous-batching solve service for parametric P1 Poisson /
reaction-diffusion requests on one GPU.

A request asks for  -div(kappa grad u) + c u = f  on create_box(n^3)
tetrahedra with u = 0 on the boundary: kappa per cell, c a constant, f per
vertex. The service keeps, per mesh size n, the mesh, dofmap, Dirichlet
dofs, sparsity pattern (in HBM, shared by every request on that mesh) and
the device forms; a request only brings its coefficients and owns its
matrix values and vectors while it is resident.

Requests are admitted in arrival order while a batch slot is free and
their HBM footprint fits the budget; the requests admitted together are
assembled by one batched launch pair (dfx_assemble_matrix_batched) and
joined to the running segmented CG batch (dfx_batch_*), so new requests
start iterating while earlier ones are still on the GPU. Converged requests leave the batch and free
their memory. Latency is measured from arrival to the completed solution.
"""

from __future__ import annotations

import ctypes as C
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib, fem, la
from .mesh import CellType, create_box


class BatchSolver:
    """dfx_batch: segmented FP64 Jacobi-CG over many systems."""

    def __init__(self, max_systems: int, max_rows: int):
        self._h = C.c_void_p()
        _lib.call("dfx_batch_create", int(max_systems), int(max_rows), C.byref(self._h))
        self.max_systems = max_systems

    def add(self, A: la.MatrixCSR, b: la.Vector, x: la.Vector, rtol=1e-8, max_it=20000) -> int:
        slot = C.c_int32()
        _lib.call("dfx_batch_add", self._h, A.handle, b.handle, x.handle, float(rtol),
                  int(max_it), C.byref(slot))
        return slot.value

    def remove(self, slot: int) -> None:
        _lib.call("dfx_batch_remove", self._h, int(slot))

    def step(self, iterations: int, graph: bool = True) -> None:
        _lib.call("dfx_batch_step", self._h, int(iterations), 1 if graph else 0)

    def restart(self) -> None:
        """Every system from x = 0 with the current values of its A and b
        (kernels only: can be recorded into a graph.Graph)."""
        _lib.call("dfx_batch_restart", self._h)

    def status(self, slots):
        s = np.ascontiguousarray(slots, np.int32)
        n = len(s)
        done = np.zeros(n, np.int32)
        conv = np.zeros(n, np.int32)
        its = np.zeros(n, np.int32)
        rel = np.zeros(n, np.float64)
        _lib.call("dfx_batch_status", self._h, n, _lib.ptr(s), _lib.ptr(done), _lib.ptr(conv),
                  _lib.ptr(its), _lib.ptr(rel))
        return done, conv, its, rel

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_batch_destroy(h)
            except Exception:
                pass


@dataclass
class Request:
    rid: int
    n: int
    kappa: np.ndarray      # per input cell
    c: float
    f: np.ndarray          # per input vertex
    arrival: float = 0.0   # seconds from the start of the run
    admitted: float = -1.0
    completed: float = -1.0
    iterations: int = 0
    converged: bool = False
    x: np.ndarray | None = None


def make_request(rid: int, seed: int, n_range=(8, 48)) -> Request:
    """n ~ U{8..48}; kappa ~ lognormal(0, 0.5) per cell; c ~ U[0, 10];
    f ~ N(0, 1) per vertex (PCG64 seeded per request)."""
    rng = np.random.Generator(np.random.PCG64([seed, rid]))
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    nc, nv = 6 * n ** 3, (n + 1) ** 3
    return Request(rid, n, np.exp(0.5 * rng.standard_normal(nc)), float(rng.uniform(0.0, 10.0)),
                   rng.standard_normal(nv))


class MeshTemplate:
    """Per-size state kept resident: mesh, space, BC, pattern, forms."""

    def __init__(self, n: int):
        self.n = n
        self.mesh = create_box([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], (n, n, n), CellType.tetrahedron)
        self.V = fem.functionspace(self.mesh, ("Lagrange", 1))
        nc = self.mesh.topology.cells.shape[0]
        self.a = fem.form("p1_stiffness_reaction", self.V, np.ones((nc, 1)), [0.0])
        self.L = fem.form("p1_source", self.V, np.ones((nc, 4)), None)
        self.bc = fem.dirichletbc(self.V, fem.locate_boundary_dofs(self.V))
        self.pattern = fem.create_matrix(self.a)
        # deterministic assembly variant (one thread per row)
        self.a.set_variant(1)
        self.N = self.V.num_dofs
        self.nnz = self.pattern.nnz

    def footprint(self) -> int:
        """HBM bytes a resident request needs: matrix values, b, x and the
        CG work vectors r, p, q, diag^-1."""
        return 8 * self.nnz + 6 * 8 * self.N

    def matrix(self) -> la.MatrixCSR:
        return la.MatrixCSR.shared(self.pattern, np.float64)

    def load(self, req: Request) -> None:
        """The request's coefficients into the template's forms, packed on
        the device from the raw per-cell and per-vertex arrays."""
        self.a.pack(self.mesh, req.kappa, "cell")
        _lib.call("dfx_form_set_coefficients", self.a.handle, None,
                  _lib.ptr(np.array([req.c], np.float64)))
        self.L.pack(self.mesh, req.f, "vertex")

    def rhs(self) -> la.Vector:
        """b of the loaded request: assemble_vector, apply_lifting, set_bc."""
        b = la.Vector(self.N)
        fem.assemble_vector(self.L, b)
        fem.apply_lifting(b, [self.a], [[self.bc]])
        fem.set_bc(b, [self.bc])
        return b

    def assemble(self, req: Request):
        """A, b of the request on the device (reference pipeline: BC rows and
        columns zeroed with unit diagonal; b with lifting and bc.set)."""
        self.load(req)
        A = self.matrix()
        fem.assemble_matrix(self.a, [self.bc], 1.0, A)
        return A, self.rhs()


@dataclass
class ServeStats:
    latencies: list = field(default_factory=list)
    wall: float = 0.0
    completed: int = 0
    max_resident: int = 0
    max_bytes: int = 0

    def summary(self, rate):
        lat = np.array(self.latencies)
        return {
            "requests": self.completed,
            "arrival_rate_per_s": rate,
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3),
            "solves_per_s": round(self.completed / self.wall, 2),
            "max_resident": self.max_resident,
            "max_resident_bytes": self.max_bytes,
        }


class SolveService:
    def __init__(self, hbm_budget_bytes=8 << 30, max_slots=64, rtol=1e-8, iters_per_step=16,
                 graph=True, sizes=range(8, 49)):
        self.budget = int(hbm_budget_bytes)
        self.rtol = rtol
        self.iters = iters_per_step
        self.graph = graph
        self.templates = {n: MeshTemplate(n) for n in sizes}
        max_rows = max_slots * max(t.N for t in self.templates.values())
        self.batch = BatchSolver(max_slots, max_rows)
        self.max_slots = max_slots

    def solve_alone(self, req: Request) -> np.ndarray:
        """The same request on its own (a one-system batch)."""
        t = self.templates[req.n]
        A, b = t.assemble(req)
        x = la.Vector(t.N)
        B = BatchSolver(1, t.N)
        slot = B.add(A, b, x, self.rtol)
        while True:
            B.step(self.iters, self.graph)
            done, conv, its, _ = B.status([slot])
            if done[0]:
                break
        return x.array

    def run(self, requests, rate: float, keep_solutions=()) -> ServeStats:
        """Serve `requests` arriving as a Poisson process of the given rate
        (their arrival times are set here, seeded)."""
        rng = np.random.Generator(np.random.PCG64(12345))
        t_arr = np.cumsum(rng.exponential(1.0 / rate, len(requests)))
        for r, t in zip(requests, t_arr):
            r.arrival = float(t)
        pending = deque(requests)
        queue = deque()
        active = {}   # slot -> (req, A, b, x, bytes)
        used = 0
        stats = ServeStats()
        keep = set(keep_solutions)
        t0 = time.perf_counter()
        while pending or queue or active:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # admission: in arrival order while a slot is free and the
            # footprint fits the HBM budget (a request that does not fit
            # waits, and so do the requests behind it, so memory is not
            # handed past it); the admitted requests are assembled in waves
            # of one request per mesh size (a template's forms hold one
            # request's coefficients), each wave one batched launch pair
            admitted = []
            while queue and len(active) + len(admitted) < self.max_slots:
                req = queue[0]
                tpl = self.templates[req.n]
                fp = tpl.footprint()
                if used + fp > self.budget and (active or admitted):
                    break
                queue.popleft()
                admitted.append((req, tpl, fp))
                used += fp
            while admitted:
                group, rest, seen = [], [], set()
                for item in admitted:
                    if id(item[1]) in seen:
                        rest.append(item)
                    else:
                        seen.add(id(item[1]))
                        group.append(item)
                admitted = rest
                mats = []
                for req, tpl, _ in group:
                    tpl.load(req)
                    mats.append(tpl.matrix())
                fem.assemble_matrix_batched([g[1].a for g in group], [g[1].bc for g in group],
                                            mats)
                for (req, tpl, fp), A in zip(group, mats):
                    b = tpl.rhs()
                    x = la.Vector(tpl.N)
                    slot = self.batch.add(A, b, x, self.rtol)
                    req.admitted = time.perf_counter() - t0
                    active[slot] = (req, A, b, x, fp)
            stats.max_resident = max(stats.max_resident, len(active))
            stats.max_bytes = max(stats.max_bytes, used)
            if not active:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            self.batch.step(self.iters, self.graph)
            slots = list(active)
            done, conv, its, _ = self.batch.status(slots)
            tdone = time.perf_counter() - t0
            for s, d, cv, it in zip(slots, done, conv, its):
                if not d:
                    continue
                req, A, b, x, fp = active.pop(s)
                if req.rid in keep:
                    req.x = x.array
                self.batch.remove(s)
                used -= fp
                req.completed = tdone
                req.iterations = int(it)
                req.converged = bool(cv)
                stats.latencies.append(req.completed - req.arrival)
                stats.completed += 1
        stats.wall = time.perf_counter() - t0
        return stats
