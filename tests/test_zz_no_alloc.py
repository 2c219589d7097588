"""dfx_cg_solve_noalloc (CG with its work vectors, state and reduction
partials in the library's fixed device array, no allocation) against
dfx_cg_solve, through the C ABI.

Two calls in a row, the second on a smaller system (the array keeps the
first call's contents, so the state must be cleared by the call). With the
SpMV set to two lanes per row both entry points run the same kernels:
iterates are bitwise equal. With the default SpMV, dfx_cg_solve uses its
sliced-ELL copy, or for these small systems its one-launch cooperative
CG (other reduction orders): solutions agree within 1e-8 relative L2 and
the iteration counts within 2; the bitwise comparisons ask dfx_cg_solve
for its three-kernel iteration (classic = 1). Sizes above the array's
maximum (1,048,576 scalar rows) are refused before any device work.
"""

import ctypes as C

import numpy as np
import pytest

from dolfinx_amd import _lib, la, problems
from dolfinx_amd._lib import DfxError

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _solve(name, A, b, rtol, jacobi=True, classic=0):
    x = la.Vector(b.array.shape[0], A.dtype)
    opts = _lib.CGOpts(max_it=20000, rtol=rtol, jacobi=1 if jacobi else 0, check_every=16,
                       classic=classic)
    res = _lib.CGResult()
    _lib.call(name, A.handle, b.handle, x.handle, C.byref(opts), C.byref(res))
    return x.array, res


def _system(n, cell, dtype, seed):
    p = problems.scalar_problem(n, cell, kappa="exp_normal", c=0.5, f="normal", seed=seed,
                                dtype=dtype)
    A, b = problems.assemble_system(p)
    return A, b


@pytest.mark.parametrize("dtype,rtol,tol", [(np.float64, 1e-10, 1e-8), (np.float32, 1e-5, 2e-4)])
def test_noalloc_cg_matches_cg_two_calls(dtype, rtol, tol):
    for n, cell in [((12, 12, 12), 0), ((6, 5, 4), 1)]:       # second call smaller
        A, b = _system(n, cell, dtype, seed=n[0])
        xa, ra = _solve("dfx_cg_solve", A, b, rtol)
        xn, rn = _solve("dfx_cg_solve_noalloc", A, b, rtol)
        assert ra.converged and rn.converged
        assert rel_err(xn, xa) <= tol
        assert abs(rn.iterations - ra.iterations) <= 2
        # same kernels: bitwise equal
        A.set_lanes(2)
        xa2, ra2 = _solve("dfx_cg_solve", A, b, rtol, classic=1)
        xn2, rn2 = _solve("dfx_cg_solve_noalloc", A, b, rtol)
        assert rn2.iterations == ra2.iterations
        assert np.array_equal(xn2, xa2)
        assert rn2.true_rel_residual == ra2.true_rel_residual


def test_noalloc_cg_plain_cg_repeated():
    A, b = _system((10, 9, 8), 0, np.float64, seed=3)
    A.set_lanes(2)
    xa, ra = _solve("dfx_cg_solve", A, b, 1e-10, jacobi=False, classic=1)
    for _ in range(2):
        xn, rn = _solve("dfx_cg_solve_noalloc", A, b, 1e-10, jacobi=False)
        assert rn.iterations == ra.iterations and np.array_equal(xn, xa)


def test_noalloc_cg_refuses_sizes_above_the_maximum():
    n = (1 << 20) + 1
    rp = np.arange(n + 1, dtype=np.int64)
    A = la.MatrixCSR(rp, np.arange(n, dtype=np.int32))
    b = la.Vector(n)
    x = la.Vector(n)
    opts = _lib.CGOpts(max_it=10, rtol=1e-8, jacobi=1, check_every=16)
    res = _lib.CGResult()
    with pytest.raises(DfxError, match="exceed the scratch maximum"):
        _lib.call("dfx_cg_solve_noalloc", A.handle, b.handle, x.handle, C.byref(opts),
                  C.byref(res))
