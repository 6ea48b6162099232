"""sciwasm.sparse_amd -- MI355X (gfx950) HIP implementation of the
SciWasm.Sparse Level-II hot path, with a host API that mirrors the
reference's sparse32.js / sparse64.js call surface.

All compute runs in libsablesparse_hip.so (HIP kernels); this package is
the host-side object layer and binding.  Importing it does not touch the
GPU; the first operation creates the default device context.
"""
from ._lib import SableError, load as _load_library  # noqa: F401
from .core import (Context, DeviceArray, sparse_COO_t, sparse_CSC_t, sparse_CSR_t, sparse_DIA_t,  # noqa: F401
                   sparse_ELL_t, sparse_vec_t, coo_from_host, csc_from_host, csr_from_host, dia_from_host,
                   ell_from_host, create_vec, vec_from_host, clear_vec, init_x, fletcher_sum, fletcher_sum_y)
from .spmv import spmv, CsrPlan, CsrPanels, CsrTiles, CooPlan, CsrAuto  # noqa: F401
from .convert import coo_csr, coo_csc, num_cols, csr_ell, num_diags, csr_dia, transpose  # noqa: F401
from .spts import (SptsAnalysis, CSR_create_level_sets, CSR_create_level_sets_with_reorder,  # noqa: F401
                   spts_level_csr, spts_syncfree, spts_csc)
from .tune import Autotuner, BatchPlan, row_stats_key  # noqa: F401
from .serve import Server, BatchLevelPlan, batch_spmv, batch_spmv_tiles, batch_tile_plans, batch_spts, batch_spts_refined, lower_dominant  # noqa: F401
from .mm import MMInfo, read_matrix_MM_files, create_COO_from_MM, create_LCOO_from_MM, mmread  # noqa: F401
from .ewise import (expm1, log1p, sin, tan, pow, deg2rad, rad2deg, sign, abs, neg, sqrt, ceil, floor,  # noqa: F401
                    trunc, nearest, eliminate_zeros, count_nonzeros)
from .reorder import sort_rows_by_nnz, sort_y_rows_by_nnz, permute_rows, reorder_NN  # noqa: F401

__all__ = [
    "Context", "DeviceArray", "SableError",
    "sparse_COO_t", "sparse_CSC_t", "sparse_CSR_t", "sparse_DIA_t", "sparse_ELL_t", "sparse_vec_t",
    "coo_from_host", "csc_from_host", "csr_from_host", "dia_from_host", "ell_from_host",
    "create_vec", "vec_from_host", "clear_vec", "init_x", "fletcher_sum", "fletcher_sum_y",
    "spmv", "CsrPlan", "CsrPanels", "CsrTiles", "CooPlan", "CsrAuto", "Autotuner", "BatchPlan", "BatchLevelPlan", "row_stats_key",
    "coo_csr", "coo_csc", "num_cols", "csr_ell", "num_diags", "csr_dia", "transpose",
    "SptsAnalysis", "CSR_create_level_sets", "CSR_create_level_sets_with_reorder", "spts_level_csr",
    "spts_syncfree", "spts_csc", "Server", "batch_spmv", "batch_spmv_tiles", "batch_tile_plans", "batch_spts", "batch_spts_refined", "lower_dominant",
    "MMInfo", "read_matrix_MM_files", "create_COO_from_MM", "create_LCOO_from_MM", "mmread",
    "expm1", "log1p", "sin", "tan", "pow", "deg2rad", "rad2deg", "sign", "abs", "neg", "sqrt", "ceil", "floor",
    "trunc", "nearest", "eliminate_zeros", "count_nonzeros", "sort_rows_by_nnz", "sort_y_rows_by_nnz", "permute_rows",
    "reorder_NN",
]
