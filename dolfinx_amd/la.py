"""Device linear algebra (mirror of dolfinx.la for this path).

MatrixCSR and Vector hold HBM-resident data behind the C ABI; the host views
(``data``, ``array``) are copies made on request.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class SparsityPattern:
    """Finalised block sparsity pattern (la/SparsityPattern.h:25-180)."""

    def __init__(self, handle):
        self._h = handle
        nrows, nnz, bs = C.c_int32(), C.c_int64(), C.c_int32()
        _lib.call("dfx_sparsity_sizes", handle, C.byref(nrows), C.byref(nnz), C.byref(bs))
        self.num_rows = nrows.value
        self.block_size = bs.value
        self.row_ptr = np.zeros(nrows.value + 1, np.int64)
        self.cols = np.zeros(nnz.value, np.int32)
        _lib.call("dfx_sparsity_copy", handle, _lib.ptr(self.row_ptr), _lib.ptr(self.cols))

    @property
    def num_nonzeros(self):
        return int(self.row_ptr[-1])

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_sparsity_destroy(h)
            except Exception:
                pass


class MatrixCSR:
    """la::MatrixCSR with a device container (la/MatrixCSR.h:67-70)."""

    def __init__(self, row_ptr, cols, ncols=None, bs=1, dtype=np.float64):
        row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        cols = np.ascontiguousarray(cols, np.int32)
        nrows = len(row_ptr) - 1
        ncols = nrows if ncols is None else ncols
        self._h = C.c_void_p()
        _lib.call("dfx_csr_create", _lib.ptr(row_ptr), _lib.ptr(cols), nrows, ncols, bs,
                  _lib.code_of(dtype), C.byref(self._h))
        self.indptr = row_ptr
        self.indices = cols
        self.bs = bs
        self.shape = (nrows * bs, ncols * bs)
        self.dtype = np.dtype(dtype)

    @classmethod
    def from_sparsity(cls, sp: SparsityPattern, dtype=np.float64):
        return cls(sp.row_ptr, sp.cols, bs=sp.block_size, dtype=dtype)

    @classmethod
    def shared(cls, pattern: "MatrixCSR", dtype=np.float64):
        """A matrix with its own values on the device pattern of `pattern`
        (shared, not copied; `pattern` must outlive it)."""
        A = cls.__new__(cls)
        A._h = C.c_void_p()
        _lib.call("dfx_csr_create_shared", pattern.handle, _lib.code_of(dtype), C.byref(A._h))
        A._pattern = pattern
        A.indptr, A.indices, A.bs = pattern.indptr, pattern.indices, pattern.bs
        A.shape, A.dtype = pattern.shape, np.dtype(dtype)
        return A

    @property
    def handle(self):
        return self._h

    @property
    def nnz(self):
        return int(self.indptr[-1])

    @property
    def block_size(self):
        return [self.bs, self.bs]

    def set_value(self, v=0.0):
        if v != 0.0:
            raise ValueError("only zeroing is supported")
        _lib.call("dfx_csr_zero", self._h)

    @property
    def data(self):
        """Host copy of the values (nnz * bs * bs, row-major blocks)."""
        out = np.zeros(self.nnz * self.bs * self.bs, self.dtype)
        _lib.call("dfx_csr_get_values", self._h, _lib.ptr(out))
        return out

    @data.setter
    def data(self, values):
        v = np.ascontiguousarray(values, self.dtype)
        if v.size != self.nnz * self.bs * self.bs:
            raise ValueError("size mismatch")
        _lib.call("dfx_csr_set_values", self._h, _lib.ptr(v))

    def pattern(self):
        rp = np.zeros(len(self.indptr), np.int64)
        ci = np.zeros(self.nnz, np.int32)
        _lib.call("dfx_csr_get_pattern", self._h, _lib.ptr(rp), _lib.ptr(ci))
        return rp, ci

    def autotune(self):
        """Pick the SpMV lanes per row for this matrix by timing (kept for
        every later SpMV and CG on it). Returns (lanes, ms per SpMV)."""
        lanes, ms = C.c_int32(), C.c_float()
        _lib.call("dfx_csr_autotune", self._h, C.byref(lanes), C.byref(ms))
        return lanes.value, ms.value

    def add(self, rows, cols, values):
        """MatrixCSR::add of dense blocks (la/MatrixCSR.h:157-174): rows is
        (nblocks, m), cols (nblocks, n) block indices, values (nblocks,
        m bs, n bs); entries are added in block order, and a column outside
        the pattern raises "Entry not in sparsity"."""
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        if rows.ndim == 1:
            rows, cols = rows[None, :], cols[None, :]
        nb, m = rows.shape
        n = cols.shape[1]
        v = np.ascontiguousarray(values, self.dtype).reshape(nb, m * self.bs, n * self.bs)
        _lib.call("dfx_csr_add_blocks", self._h, int(nb), int(m), int(n), _lib.ptr(rows),
                  _lib.ptr(cols), _lib.ptr(v))

    def set_lanes(self, lanes):
        """SpMV kernel of mult: 0 default, 1 row-block stream, 2..64 lanes
        per row, -1 the solvers' sliced-ELL copy of the values as they are
        now (set again after the values change)."""
        _lib.call("dfx_csr_set_lanes", self._h, int(lanes))

    def mult(self, x, y):
        """y += A x (MatrixCSR::mult, la/MatrixCSR.h:877-946)."""
        _lib.call("dfx_spmv", self._h, x.handle, y.handle)

    def to_scipy(self):
        import scipy.sparse as sp

        if self.bs == 1:
            return sp.csr_matrix((self.data, self.indices, self.indptr), shape=self.shape)
        vals = self.data.reshape(-1, self.bs, self.bs)
        return sp.bsr_matrix((vals, self.indices, self.indptr), shape=self.shape).tocsr()

    def to_dense(self):
        return self.to_scipy().toarray()

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_csr_destroy(h)
            except Exception:
                pass


class Vector:
    """la::Vector with a device container (la/Vector.h:47-49)."""

    def __init__(self, n, dtype=np.float64):
        self._h = C.c_void_p()
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        _lib.call("dfx_vec_create", int(n), _lib.code_of(dtype), C.byref(self._h))

    @classmethod
    def from_array(cls, a, dtype=None):
        a = np.asarray(a)
        v = cls(a.size, a.dtype if dtype is None else dtype)
        v.array = a
        return v

    @property
    def handle(self):
        return self._h

    @property
    def array(self):
        """Host copy of the values."""
        out = np.zeros(self.size, self.dtype)
        _lib.call("dfx_vec_download", self._h, _lib.ptr(out))
        return out

    @array.setter
    def array(self, values):
        v = np.ascontiguousarray(values, self.dtype)
        if v.size != self.size:
            raise ValueError("size mismatch")
        _lib.call("dfx_vec_upload", self._h, _lib.ptr(v))

    def set(self, value):
        _lib.call("dfx_vec_set", self._h, float(value))

    def copy_from(self, other):
        _lib.call("dfx_vec_copy", self._h, other.handle)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            try:
                _lib.lib().dfx_vec_destroy(h)
            except Exception:
                pass


def inner_product(a: Vector, b: Vector) -> float:
    out = C.c_double()
    _lib.call("dfx_inner_product", a.handle, b.handle, C.byref(out))
    return out.value


def squared_norm(a: Vector) -> float:
    out = C.c_double()
    _lib.call("dfx_squared_norm", a.handle, C.byref(out))
    return out.value


def norm(a: Vector) -> float:
    return float(np.sqrt(squared_norm(a)))


def axpy(r: Vector, alpha: float, x: Vector, y: Vector) -> None:
    """r = alpha x + y (cpp/demo/poisson_matrix_free/main.cpp:65-69)."""
    _lib.call("dfx_axpy", r.handle, float(alpha), x.handle, y.handle)
