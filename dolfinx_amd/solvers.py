"""Device-resident Krylov solvers.

cg follows the reference CG (cpp/demo/poisson_matrix_free/main.cpp:84-132,
python/test/conftest.py:155-205; stops when ||r||^2/||r0||^2 < rtol^2 and
raises after max_it iterations), with an optional Jacobi preconditioner.
mixed_refinement wraps an FP32 inner CG in FP64 iterative refinement.
"""

from __future__ import annotations

import ctypes as C

from . import _lib
from .la import MatrixCSR, Vector


def cg(A: MatrixCSR, b: Vector, x: Vector, rtol=1e-8, max_it=10000, jacobi=True,
       check_every=32, graph=False, classic=False):
    """CG (Jacobi-preconditioned by default). A scalar system of at most
    65,536 rows is solved in one cooperative launch unless classic=True;
    otherwise three kernels per iteration, and graph=True replays each
    batch of check_every iterations from a captured HIP graph (same kernels
    and results; the capture costs about a millisecond per solve)."""
    opts = _lib.CGOpts(max_it=int(max_it), rtol=float(rtol), jacobi=1 if jacobi else 0,
                       check_every=int(check_every), graph=1 if graph else 0,
                       classic=1 if classic else 0)
    res = _lib.CGResult()
    _lib.call("dfx_cg_solve", A.handle, b.handle, x.handle, C.byref(opts), C.byref(res))
    return res


def mixed_refinement(A64: MatrixCSR, A32: MatrixCSR, b: Vector, x: Vector, rtol=1e-8,
                     inner_rtol=1e-4, max_outer=50, max_inner=20000, check_every=32,
                     restart=True, graph=False, fallback=True):
    """FP32 Jacobi-PCG inside FP64 iterative refinement: each inner solve
    reduces the residual by inner_rtol, then x += d in FP64 and the FP64
    true residual restarts the next inner solve. restart=False keeps the FP32
    search direction across refinements instead (reliable updates).
    If a refinement step fails to halve the FP64 residual norm (or the
    budgets run out), fallback=True finishes with FP64 Jacobi-PCG from the
    current x (res.fallback_iterations > 0); fallback=False raises."""
    opts = _lib.MixedOpts(max_outer=int(max_outer), max_inner=int(max_inner), rtol=float(rtol),
                          inner_rtol=float(inner_rtol), check_every=int(check_every),
                          restart=1 if restart else 0, graph=1 if graph else 0,
                          fallback=1 if fallback else 0)
    res = _lib.MixedResult()
    _lib.call("dfx_mixed_solve", A64.handle, A32.handle, b.handle, x.handle, C.byref(opts),
              C.byref(res))
    return res
