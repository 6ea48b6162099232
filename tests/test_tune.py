"""The SpMV autotuner: keys built from row-length statistics (host logic,
no GPU) and the candidate schedules it weighs."""
import numpy as np

from sciwasm.sparse_amd import gen
from sciwasm.sparse_amd.tune import Autotuner, row_stats_key


class _Shape:
    """The fields candidates() reads from a matrix."""

    def __init__(self, N, nnz, dtype=np.float64):
        self.N, self.nnz, self.dtype = N, nnz, np.dtype(dtype)


def test_key_groups_matrices_by_row_statistics():
    a, _, _ = gen.laplacian_2d(128)
    b, _, _ = gen.laplacian_2d(140)  # same log2 buckets of N, mean and max
    z, _, _ = gen.zipf_csr(1 << 14, 1.5, 1024, seed=3)
    assert row_stats_key(a) == row_stats_key(b)
    assert row_stats_key(a) != row_stats_key(z)
    assert row_stats_key(np.zeros(1, np.int64)).startswith("N0_mean0_max0")


def test_candidates_cover_formats():
    rp, _, _ = gen.laplacian_2d(64)  # 4096 rows of at most 5: ELL passes the 5x guard
    t = Autotuner(ctx=None, allow_tiles=True)
    c = t.candidates(_Shape(4096, int(rp[-1])), rp)
    fmts = [x["format"] for x in c]
    assert "csr" in fmts and "ell" in fmts and "panels" not in fmts
    tiles = [x for x in c if x["format"] == "tiles"]
    assert tiles == [{"format": "tiles", "tile_rows": 4096, "panel_cols": 4096}]
    assert any(x.get("vector_lanes") for x in c)
    # x larger than 2 MiB: panels and the 16384 x 8192 tiles
    big = 1 << 19
    c = t.candidates(_Shape(big, 5 * big), np.arange(0, 5 * big + 1, 5))
    assert {"format": "tiles", "tile_rows": 16384, "panel_cols": 8192} in c
    assert any(x["format"] == "panels" for x in c)
    # deterministic schedules only
    c = Autotuner(ctx=None, allow_tiles=False).candidates(_Shape(big, 5 * big), np.arange(0, 5 * big + 1, 5))
    assert all(x["format"] != "tiles" for x in c)
