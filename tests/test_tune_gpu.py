"""The autotuner on the GPU: every candidate is timed, the winner is
cached per row-statistics class, and the schedule it builds computes the
CSR product within the CSR bound against the oracle."""
import numpy as np
import pytest

import oracle
import sciwasm.sparse_amd as sp
from sciwasm.sparse_amd import gen, DeviceArray

pytestmark = pytest.mark.gpu


def bound(rp, col, val, x, dtype):
    u = np.finfo(dtype).eps / 2
    k = np.diff(rp).astype(np.float64) + 1
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    s = np.bincount(rows, weights=np.abs(val.astype(np.float64) * x.astype(np.float64)[col]), minlength=len(rp) - 1)
    return 2 * k * u / (1 - k * u) * s + np.finfo(dtype).tiny


def run_schedule(runner, A, rp, c, v, dtype, seed):
    x = np.random.default_rng(seed).uniform(-1, 1, A.N).astype(dtype)
    y = DeviceArray.zeros(A.N, dtype)
    runner.run(DeviceArray.from_host(x), y, beta=0.0)
    yref = oracle.spmv_csr(A.N, rp, c, v, x, np.zeros(A.N, dtype))
    assert np.all(np.abs(y.numpy().astype(np.float64) - yref) <= bound(rp, c, v, x, dtype))


@pytest.mark.parametrize("allow_tiles", [True, False])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tune_picks_times_and_caches(ctx, dtype, allow_tiles):
    T = sp.Autotuner(ctx, reps=3, allow_tiles=allow_tiles)
    cases = [gen.laplacian_2d(128), gen.random_csr(20000, 9, seed=5), gen.zipf_csr(1 << 15, 1.5, 1024, seed=7)]
    for rp, c, v in cases:
        v = v.astype(dtype)
        A = sp.csr_from_host(len(rp) - 1, rp, c, v)
        key, best = T.tune(A, rp)
        assert best["ms"] > 0 and len(best["table"]) == len(T.candidates(A, rp))
        assert best["ms"] == min(r["ms"] for r in best["table"])
        assert allow_tiles or best["format"] != "tiles"
        # every candidate computes the product, not only the winner
        for cand in T.candidates(A, rp):
            run_schedule(T.build(A, cand), A, rp, c, v, dtype, seed=1)
        run_schedule(T.plan(A, rp), A, rp, c, v, dtype, seed=2)
    # a matrix of an already-tuned class reuses the choice without measuring
    n = len(T.cache)
    rp, c, v = gen.laplacian_2d(140)
    A = sp.csr_from_host(len(rp) - 1, rp, c, v.astype(dtype))
    key, best = T.tune(A, rp)
    assert len(T.cache) == n and best is T.cache[key]
