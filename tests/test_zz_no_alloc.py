"""sable_spmv_csr_noalloc against sable_spmv_csr through the C-ABI.

The allocation-free entry point runs the same device-built plan and kernels
as sable_spmv_csr, with the plan in the library's fixed per-GPU work space,
so its y must equal sable_spmv_csr's bit for bit: on two calls in a row (the
second on a smaller matrix, over what the first left in the work space), at
the largest row count the work space holds, and with beta and repeated
products; a row count beyond it is refused before anything is launched."""
import numpy as np
import pytest

import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import SableError, _lib
from sciwasm.sparse_amd._lib import call
from sciwasm.sparse_amd.core import DeviceArray, dtype_code

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 23


def random_csr(N, ncols, mean_len, seed, dtype):
    rng = np.random.default_rng(seed)
    lens = np.minimum(rng.zipf(1.8, N), 4 * mean_len + 64).astype(np.int64)
    rp = np.zeros(N + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    col = rng.integers(0, ncols, nnz).astype(np.int32)
    val = rng.uniform(-1, 1, nnz).astype(dtype)
    return rp.astype(np.int32), col, val


def both(ctx, dtype, N, rp, col, val, x, y0, beta, inner_max):
    """(sable_spmv_csr's y, sable_spmv_csr_noalloc's y) on the same inputs."""
    R, C, V, X = (DeviceArray.from_host(a) for a in (rp, col, val, x))
    out = []
    for fn in ("sable_spmv_csr", "sable_spmv_csr_noalloc"):
        Y = DeviceArray.from_host(y0)
        print(fn, np.dtype(dtype).name, "N=%d nnz=%d beta=%g inner_max=%d" % (N, len(col), beta, inner_max),
              flush=True)
        call(fn, ctx.handle, dtype_code(dtype), N, R.ptr, C.ptr, V.ptr, X.ptr, Y.ptr, beta, inner_max)
        out.append(Y.numpy())
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_noalloc_equals_plan_free_twice(ctx, dtype):
    """A large matrix, then a smaller one: the second call's plan is built
    over the first call's work space contents and must not see them."""
    rng = np.random.default_rng(1)
    for N, ncols, mean_len, seed in ((300000, 300000, 12, 2), (7000, 9000, 5, 3)):
        rp, col, val = random_csr(N, ncols, mean_len, seed, dtype)
        x = rng.uniform(-1, 1, ncols).astype(dtype)
        y0 = np.zeros(N, dtype)
        want, got = both(ctx, dtype, N, rp, col, val, x, y0, 0.0, 1)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), N


def test_noalloc_beta_and_repeats(ctx):
    rng = np.random.default_rng(4)
    N = 50000
    rp, col, val = random_csr(N, N, 8, 5, np.float64)
    x = rng.uniform(-1, 1, N)
    y0 = rng.uniform(-1, 1, N)
    want, got = both(ctx, np.float64, N, rp, col, val, x, y0, 0.5, 3)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_noalloc_at_max_rows(ctx):
    """SABLE_NOALLOC_MAX_ROWS rows (one to three entries each): the plan's
    capacity fills the work space to its last block boundary."""
    rng = np.random.default_rng(6)
    N = MAX_ROWS
    lens = rng.integers(1, 4, N)
    rp = np.zeros(N + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, N, int(rp[-1])).astype(np.int32)
    val = rng.uniform(-1, 1, len(col))
    x = rng.uniform(-1, 1, N)
    want, got = both(ctx, np.float64, N, rp.astype(np.int32), col, val, x, np.zeros(N), 0.0, 1)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_noalloc_refuses_more_rows(ctx):
    R = DeviceArray.from_host(np.zeros(2, np.int32))
    with pytest.raises(SableError) as e:
        call("sable_spmv_csr_noalloc", ctx.handle, _lib.F64, MAX_ROWS + 1, R.ptr, R.ptr, R.ptr, R.ptr, R.ptr, 0.0, 1)
    assert e.value.name == "SABLE_EINVAL"
    assert "work space" in str(e.value)
