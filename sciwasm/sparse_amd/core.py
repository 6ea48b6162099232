"""Device context, device arrays and the sparse matrix objects.

The objects mirror the reference's JS object model
(seq/src/sparse32.js:29-198, par/src/sparse32.js:53-208): a COO/CSR/DIA/ELL
matrix is a set of arrays plus sizes; here the arrays live in HBM as
DeviceArray buffers instead of at byte offsets in a Wasm linear memory.
Every operation runs on the HIP library; there is no host compute path.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import call, c_vp, c_flt

_DT = {np.dtype(np.float32): _lib.F32, np.dtype(np.float64): _lib.F64}


def dtype_code(dt):
    try:
        return _DT[np.dtype(dt)]
    except KeyError:
        raise TypeError("values must be float32 or float64, got %s" % dt)


class Context:
    """A HIP device context with its own default-priority stream."""

    _default = None

    def __init__(self, device=0):
        h = c_vp()
        call("sable_ctx_create", device, None, ctypes.byref(h))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls):
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def sync(self):
        call("sable_ctx_sync", self.handle)

    def close(self):
        if self.handle:
            _lib.load().sable_ctx_destroy(self.handle)
            self.handle = None

    # ---- timing on the context stream
    def event(self):
        e = c_vp()
        call("sable_event_create", ctypes.byref(e))
        return e

    def record(self, ev):
        call("sable_event_record", self.handle, ev)

    @staticmethod
    def elapsed_ms(a, b):
        ms = c_flt()
        call("sable_event_elapsed_ms", a, b, ctypes.byref(ms))
        return ms.value


def _ctx(ctx):
    return ctx if ctx is not None else Context.default()


class DeviceArray:
    """A 1-D device buffer of a numpy dtype (freed when collected)."""

    def __init__(self, n, dtype, ctx=None):
        self.ctx = _ctx(ctx)
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = c_vp()
        call("sable_device_alloc", self.ctx.handle, max(self.nbytes, 1), ctypes.byref(p))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None, ctx=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype, ctx)
        if a.nbytes:
            call("sable_memcpy_h2d", d.ctx.handle, d.ptr, a.ctypes.data, a.nbytes)
        return d

    @classmethod
    def zeros(cls, n, dtype, ctx=None):
        d = cls(n, dtype, ctx)
        d.fill_bytes(0)
        return d

    def fill_bytes(self, byte):
        if self.nbytes:
            call("sable_fill_bytes", self.ctx.handle, self.ptr, byte, self.nbytes)

    def fill(self, value):
        if self.dtype == np.float32 or self.dtype == np.float64:
            call("sable_fill_value", self.ctx.handle, dtype_code(self.dtype), self.ptr, self.n, float(value))
        elif value == 0:
            self.fill_bytes(0)
        else:
            self.copy_from(np.full(self.n, value, self.dtype))

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if a.nbytes:
            call("sable_memcpy_h2d", self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes)

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.nbytes:
            call("sable_memcpy_d2h", self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes)
        return out

    def copy(self):
        d = DeviceArray(self.n, self.dtype, self.ctx)
        if self.nbytes:
            call("sable_memcpy_d2d", self.ctx.handle, d.ptr, self.ptr, self.nbytes)
        return d

    def free(self):
        if self.ptr:
            _lib.load().sable_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.n


# ------------------------------------------------------------------ objects
class sparse_COO_t:
    """seq/src/sparse32.js:29-63: row, col, val arrays of length nnz."""

    def __init__(self, row, col, val, N, nnz):
        self.row, self.col, self.val = row, col, val
        self.N, self.nnz = int(N), int(nnz)
        self._plan = None

    @property
    def dtype(self):
        return self.val.dtype

    def plan(self):
        """Cached row analysis for SpMV (rows checked once; call
        invalidate() after writing to the row array)."""
        if self._plan is None:
            from .spmv import CooPlan
            self._plan = CooPlan(self)
        return self._plan

    def invalidate(self):
        self._plan = None


class sparse_CSR_t:
    """seq/src/sparse32.js:67-108: row_ptr (N+1), col, val (nnz)."""

    def __init__(self, row, col, val, N, nnz):
        self.row, self.col, self.val = row, col, val
        self.N, self.nnz = int(N), int(nnz)
        self._plan = None
        # level-set analysis results (par/src/sparse32.js:98-115)
        self.nlevels = None
        self.level = None
        self.permutation = None

    @property
    def dtype(self):
        return self.val.dtype

    def plan(self, tune=None):
        """Cached device row-block schedule for SpMV."""
        if self._plan is None or tune is not None:
            from .spmv import CsrPlan
            self._plan = CsrPlan(self, tune)
        return self._plan


class sparse_CSC_t:
    """seq/src/sparse32.js:110-123: col_ptr (ncols+1), row, val (nnz)."""

    def __init__(self, col, row, val, ncols, nnz):
        self.col, self.row, self.val = col, row, val
        self.ncols, self.nnz = int(ncols), int(nnz)
        self.N = self.ncols

    @property
    def dtype(self):
        return self.val.dtype


class sparse_DIA_t:
    """seq/src/sparse32.js:125-155: offset (nd), data (nd*stride)."""

    def __init__(self, offset, data, ndiags, stride, N, nnz):
        self.offset, self.data = offset, data
        self.ndiags, self.stride, self.N, self.nnz = int(ndiags), int(stride), int(N), int(nnz)

    @property
    def dtype(self):
        return self.data.dtype


class sparse_ELL_t:
    """seq/src/sparse32.js:159-188: indices, data (ncols*N, column-major)."""

    def __init__(self, indices, data, ncols, N, nnz):
        self.indices, self.data = indices, data
        self.ncols, self.N, self.nnz = int(ncols), int(N), int(nnz)

    @property
    def dtype(self):
        return self.data.dtype


class sparse_vec_t:
    """seq/src/sparse32.js:190-198: a dense vector."""

    def __init__(self, vec, nelem):
        self.vec, self.vec_nelem = vec, int(nelem)

    def array(self):
        return self.vec.numpy()

    def size(self):
        return self.vec_nelem


# ------------------------------------------------------------- constructors
def coo_from_host(N, row, col, val, ctx=None):
    val = np.ascontiguousarray(val)
    dtype_code(val.dtype)
    return sparse_COO_t(DeviceArray.from_host(row, np.int32, ctx), DeviceArray.from_host(col, np.int32, ctx),
                        DeviceArray.from_host(val, None, ctx), N, len(val))


def csr_from_host(N, row_ptr, col, val, ctx=None):
    val = np.ascontiguousarray(val)
    dtype_code(val.dtype)
    return sparse_CSR_t(DeviceArray.from_host(row_ptr, np.int32, ctx), DeviceArray.from_host(col, np.int32, ctx),
                        DeviceArray.from_host(val, None, ctx), N, len(val))


def csc_from_host(N, col_ptr, row, val, ctx=None):
    val = np.ascontiguousarray(val)
    dtype_code(val.dtype)
    return sparse_CSC_t(DeviceArray.from_host(col_ptr, np.int32, ctx), DeviceArray.from_host(row, np.int32, ctx),
                        DeviceArray.from_host(val, None, ctx), N, len(val))


def dia_from_host(N, stride, offset, data, nnz, ctx=None):
    data = np.ascontiguousarray(data)
    dtype_code(data.dtype)
    return sparse_DIA_t(DeviceArray.from_host(offset, np.int32, ctx), DeviceArray.from_host(data, None, ctx),
                        len(offset), stride, N, nnz)


def ell_from_host(N, nc, indices, data, nnz, ctx=None):
    data = np.ascontiguousarray(data)
    dtype_code(data.dtype)
    return sparse_ELL_t(DeviceArray.from_host(indices, np.int32, ctx), DeviceArray.from_host(data, None, ctx),
                        nc, N, nnz)


def create_vec(nelem, val=0, dtype=np.float32, ctx=None):
    """seq/src/sparse32.js create_vec: a vector filled with val."""
    d = DeviceArray(nelem, dtype, ctx)
    d.fill(val)
    return sparse_vec_t(d, nelem)


def vec_from_host(a, dtype=None, ctx=None):
    a = np.ascontiguousarray(a, dtype=dtype)
    return sparse_vec_t(DeviceArray.from_host(a, None, ctx), a.size)


def clear_vec(y):
    """seq/src/sparse32.js:1319-1322."""
    y.vec.fill(0)


def init_x(x):
    """seq/src/sparse32.js init_x: x[i] = i."""
    x.vec.copy_from(np.arange(x.vec_nelem, dtype=np.float64).astype(x.vec.dtype))


def fletcher_sum(a):
    """seq/src/sparse32.js:1291-1305 on a host array (JS double arithmetic)."""
    s1 = 0.0
    s2 = 0.0
    for v in np.asarray(a, dtype=np.float64).tolist():
        s1 = s1 + v
        s1 = s1 - np.floor(s1 / 255.0) * 255.0
        s2 = s2 + s1
        s2 = s2 - np.floor(s2 / 255.0) * 255.0
    r = s2 * 256 + s1
    return r if np.isnan(r) else float(np.trunc(r))


def fletcher_sum_y(y):
    """seq/src/sparse32.js:1307-1311 (copies y to the host)."""
    return fletcher_sum(y.vec.numpy())


class DeviceView:
    """A non-owning window [offset, offset+n) of a DeviceArray (for the
    reference's pointer-offset row slices, e.g. rowptr + 4*start)."""

    def __init__(self, base, offset, n):
        assert 0 <= offset and offset + n <= base.n
        self.base, self.ctx, self.dtype, self.n = base, base.ctx, base.dtype, int(n)
        self.ptr = base.ptr + int(offset) * base.dtype.itemsize
        self.nbytes = self.n * self.dtype.itemsize

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.nbytes:
            call("sable_memcpy_d2h", self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes)
        return out
