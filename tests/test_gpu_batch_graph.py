"""Many meshes assembled and solved by one recorded graph: the zeroing, the
batched matrix assembly, the vectors with lifting and BCs, the restart of
the segmented CG and its iterations are recorded once (graph.Graph) and
replayed as one launch per coefficient set. Every replay equals the same
sequence run call by call, bitwise, with the program's own eager work and
synchronisations between the replays; the solutions match the
single-system CG. Calls that need the host inside a capture raise
DFX_ERR_CAPTURE and leave the library usable."""

import numpy as np
import pytest

from dolfinx_amd import _lib, fem, graph, la, serve, solvers
from dolfinx_amd._lib import DfxError

from conftest import rel_err

pytestmark = pytest.mark.gpu

SIZES = (6, 9, 12)
ITERS = 300  # more than any of these systems needs at rtol 1e-10


class Batch:
    def __init__(self):
        self.t = [serve.MeshTemplate(n) for n in SIZES]
        for t in self.t:
            _lib.call("dfx_form_set_coefficients", t.a.handle, None,
                      _lib.ptr(np.array([2.0], np.float64)))
        self.A = [t.matrix() for t in self.t]
        self.b = [la.Vector(t.N) for t in self.t]
        self.x = [la.Vector(t.N) for t in self.t]
        self.B = serve.BatchSolver(len(SIZES), sum(t.N for t in self.t))
        self.slots = [self.B.add(A, b, x, rtol=1e-10, max_it=2000)
                      for A, b, x in zip(self.A, self.b, self.x)]

    def load(self, seed):
        """kappa per cell and f per vertex of coefficient set `seed`."""
        rng = np.random.default_rng(seed)
        for t in self.t:
            nc = t.mesh.topology.cells.shape[0]
            nv = t.mesh.geometry.x.shape[0]
            t.a.pack(t.mesh, np.exp(0.5 * rng.standard_normal(nc)), "cell")
            t.L.pack(t.mesh, rng.standard_normal(nv), "vertex")

    def sequence(self):
        for A in self.A:
            A.set_value(0.0)
        fem.assemble_matrix_batched([t.a for t in self.t], [t.bc for t in self.t], self.A)
        for t, b in zip(self.t, self.b):
            b.set(0.0)
            fem.assemble_vector(t.L, b)
            fem.apply_lifting(b, [t.a], [[t.bc]])
            fem.set_bc(b, [t.bc])
        self.B.restart()
        self.B.step(ITERS, graph=False)

    def result(self):
        done, conv, its, rel = self.B.status(self.slots)
        return [x.array.copy() for x in self.x], done.copy(), conv.copy(), its.copy()


def test_batched_assemble_and_solve_replayed_from_one_graph():
    bt = Batch()
    # call by call, for three coefficient sets (the first run also builds
    # every plan and table the sequence needs)
    ref = {}
    for seed in (1, 2, 3):
        bt.load(seed)
        bt.sequence()
        ref[seed] = bt.result()
        xs, done, conv, its = ref[seed]
        assert done.all() and conv.all() and (its > 10).all()
    # the eager solutions against the single-system CG on the same A, b
    for A, b, x in zip(bt.A, bt.b, ref[3][0]):
        xv = la.Vector(b.array.size)
        solvers.cg(A, b, xv, rtol=1e-12)
        assert rel_err(x, xv.array) <= 1e-8
    g = graph.Graph()
    bt.load(1)
    with g.capture():
        bt.sequence()
    for seed in (2, 1, 3, 2):
        bt.load(seed)
        g.launch()
        xs, done, conv, its = bt.result()
        assert np.array_equal(its, ref[seed][3])
        assert done.all() and conv.all()
        for x, xr in zip(xs, ref[seed][0]):
            assert np.array_equal(x, xr)
        # the program's own eager work between replays
        scratch = la.Vector(bt.t[0].N)
        fem.assemble_vector(bt.t[0].L, scratch)
        assert np.isfinite(scratch.array).all()


def test_capture_refuses_host_steps_and_recovers():
    bt = Batch()
    bt.load(4)
    bt.sequence()
    # allocation inside a capture
    g = graph.Graph()
    with pytest.raises(DfxError) as e:
        with g.capture():
            la.Vector(100)
    assert e.value.code == _lib.DFX_ERR_CAPTURE
    # a solve that polls the host
    g2 = graph.Graph()
    with pytest.raises(DfxError):
        with g2.capture():
            solvers.cg(bt.A[0], bt.b[0], bt.x[0], rtol=1e-8)
    # a failed call inside the body: dfx_graph_end discards the capture
    # and reports the first failure
    _lib.call("dfx_graph_begin")
    with pytest.raises(DfxError):
        la.Vector(10)
    with pytest.raises(DfxError, match="graph capture discarded"):
        import ctypes as C
        h = C.c_void_p()
        _lib.call("dfx_graph_end", C.byref(h))
    # the library runs normally afterwards
    bt.sequence()
    xs, done, conv, its = bt.result()
    assert done.all() and conv.all()
    with pytest.raises(DfxError, match="no graph capture"):
        import ctypes as C
        _lib.call("dfx_graph_end", C.byref(C.c_void_p()))
