"""ctypes binding of libsablesparse_hip.so (the C ABI in include/sable_sparse.h).

The library is loaded from this package directory.  There is no fallback:
if the shared object is missing or fails to load, every call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsablesparse_hip.so")

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_flt = ctypes.c_float
c_vp = ctypes.c_void_p
c_sz = ctypes.c_size_t
c_vpp = ctypes.POINTER(ctypes.c_void_p)

F32 = 0
F64 = 1

ERRORS = {
    -1: "SABLE_EINVAL", -2: "SABLE_EREFUSED", -3: "SABLE_EMALFORMED", -4: "SABLE_ENOMEM",
    -5: "SABLE_EHIP", -6: "SABLE_EBUDGET", -7: "SABLE_ETIMEOUT", -8: "SABLE_ESINGULAR",
}


class SableError(RuntimeError):
    def __init__(self, code, fn, msg):
        super().__init__("%s from %s: %s" % (ERRORS.get(code, code), fn, msg))
        self.code = code
        self.name = ERRORS.get(code, str(code))


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", c_i64), ("mode", c_int), ("max_batch", c_int), ("slots", c_int),
                ("arena_bytes", c_i64), ("refine_iters", c_int), ("spts_lanes", c_int), ("group_rhs", c_int),
                ("batch_tiles", c_int), ("spts_sweep", c_int)]


class LoadSpec(ctypes.Structure):
    _fields_ = [("rate", c_dbl), ("duration_s", c_dbl), ("spts_fraction", c_dbl), ("seed", ctypes.c_uint64),
                ("max_requests", c_i64)]


class LoadStats(ctypes.Structure):
    _fields_ = [("submitted", c_i64), ("completed", c_i64), ("failed", c_i64), ("rejected", c_i64),
                ("batches", c_i64), ("elapsed_s", c_dbl), ("solves_per_s", c_dbl), ("p50_us", c_dbl),
                ("p99_us", c_dbl), ("mean_us", c_dbl), ("max_us", c_dbl), ("mean_batch", c_dbl)]


class CsrTune(ctypes.Structure):
    _fields_ = [("block_weight", c_int), ("row_weight", c_int), ("vector_lanes", c_int), ("vector_rows", c_int)]


class CsrAutoInfo(ctypes.Structure):
    _fields_ = [("schedule", c_int), ("param0", c_int), ("param1", c_int), ("ncols", c_i32), ("nnz", c_i64),
                ("extra_bytes", c_i64), ("ms", c_flt), ("from_cache", c_int)]


class MMInfo(ctypes.Structure):
    _fields_ = [("nrows", c_i64), ("ncols", c_i64), ("nentries", c_i64), ("nnz", c_i64), ("lower_nnz", c_i64),
                ("field", ctypes.c_int32), ("symmetric", ctypes.c_int32), ("body_offset", c_i64)]


# name -> argtypes; every function returns int unless listed in _RESTYPES.
SIGNATURES = {
    "sable_ctx_create": [c_int, c_vp, c_vpp],
    "sable_ctx_destroy": [c_vp],
    "sable_ctx_sync": [c_vp],
    "sable_ctx_stream": [c_vp],
    "sable_last_error": [],
    "sable_version": [],
    "sable_device_alloc": [c_vp, c_sz, c_vpp],
    "sable_device_free": [c_vp, c_vp],
    "sable_host_alloc": [c_sz, c_vpp],
    "sable_host_free": [c_vp],
    "sable_memcpy_h2d": [c_vp, c_vp, c_vp, c_sz],
    "sable_memcpy_d2h": [c_vp, c_vp, c_vp, c_sz],
    "sable_memcpy_d2d": [c_vp, c_vp, c_vp, c_sz],
    "sable_memcpy_h2d_staged": [c_vp, c_vp, c_vp, c_sz],
    "sable_memcpy_d2h_staged": [c_vp, c_vp, c_vp, c_sz],
    "sable_max_i32": [c_vp, c_vp, c_i64, ctypes.POINTER(c_i32)],
    "sable_spts_set_schedule": [c_vp, c_int, c_i32],
    "sable_spts_schedule_info": [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_i32), ctypes.POINTER(c_i32),
                                 ctypes.POINTER(c_i32), ctypes.POINTER(c_i64)],
    "sable_spts_tune": [c_vp, c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_i32), ctypes.POINTER(c_int),
                        ctypes.POINTER(ctypes.c_float)],
    "sable_csr_tiles_create_rect": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_vpp],
    "sable_csr_auto_create": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_int, c_vpp],
    "sable_csr_auto_destroy": [c_vp],
    "sable_csr_auto_info": [c_vp, c_vp],
    "sable_spmv_csr_auto": [c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_csr_auto_cache_clear": [],
    "sable_spmv_dia_range": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_dbl, c_int],
    "sable_spmv_ell_range": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_dbl, c_int],
    "sable_spmv_dia_rowmajor": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_int],
    "sable_spmv_ell_rowmajor": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_int, c_int],
    "sable_csr_ell_rowmajor": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_csr_dia_rowmajor": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_fill_bytes": [c_vp, c_vp, c_int, c_sz],
    "sable_fill_value": [c_vp, c_int, c_vp, c_i64, c_dbl],
    "sable_vec_add": [c_vp, c_int, c_vp, c_vp, c_i64],
    "sable_event_create": [c_vpp],
    "sable_event_destroy": [c_vp],
    "sable_event_record": [c_vp, c_vp],
    "sable_event_elapsed_ms": [c_vp, c_vp, ctypes.POINTER(c_flt)],
    "sable_csr_plan_create": [c_vp, c_i32, c_vp, ctypes.POINTER(CsrTune), c_vpp],
    "sable_csr_plan_destroy": [c_vp],
    "sable_csr_plan_info": [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), ctypes.POINTER(c_i32)],
    "sable_spmv_csr_plan": [c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_csr_kernel_select": [c_int],
    "sable_spmv_csr": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_spmv_csr_noalloc": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_csr_panels_create": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_i32, ctypes.POINTER(CsrTune), c_vpp],
    "sable_csr_panels_destroy": [c_vp],
    "sable_csr_panels_info": [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), ctypes.POINTER(c_i64)],
    "sable_spmv_csr_panels": [c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_spmv_coo": [c_vp, c_int, c_i32, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_csr_tiles_create": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_i32, c_i32, c_vpp],
    "sable_csr_tiles_destroy": [c_vp],
    "sable_csr_tiles_info": [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), ctypes.POINTER(c_i64)],
    "sable_spmv_csr_tiles": [c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_csr_tiles_select": [c_int],
    "sable_csr_tiles_batch_item": [c_vp, c_vp, c_vp, c_dbl, c_i64, c_vp],
    "sable_batch_spmv_tiles": [c_vp, c_int, c_int, c_vp, c_int, c_i64, c_i32, c_int],
    "sable_coo_plan_row_ptr": [c_vp, c_vpp],
    "sable_coo_plan_create": [c_vp, c_i32, c_i64, c_vp, c_vpp],
    "sable_coo_plan_destroy": [c_vp],
    "sable_coo_plan_info": [c_vp, ctypes.POINTER(c_int), ctypes.POINTER(c_i64)],
    "sable_spmv_coo_plan": [c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_spmv_dia": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_spmv_ell": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_dbl, c_int],
    "sable_coo_csr": [c_vp, c_int, c_i32, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_coo_csc": [c_vp, c_int, c_i32, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_coo_sort": [c_vp, c_int, c_i32, c_i64, c_vp, c_vp, c_vp],
    "sable_ew_values": [c_vp, c_int, c_int, c_dbl, c_i64, c_vp, c_vp],
    "sable_ew_dia": [c_vp, c_int, c_int, c_dbl, c_i32, c_i32, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp],
    "sable_ew_ell": [c_vp, c_int, c_int, c_dbl, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp],
    "sable_count_nonzeros": [c_vp, c_int, c_i64, c_vp, ctypes.POINTER(c_i64)],
    "sable_eliminate_zeros_coo": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_i64)],
    "sable_eliminate_zeros_csr": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_i64)],
    "sable_count_nonzeros_dia": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, ctypes.POINTER(c_i32),
                                 ctypes.POINTER(c_i64)],
    "sable_eliminate_zeros_dia": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_i32)],
    "sable_count_nonzeros_ell": [c_vp, c_int, c_i32, c_i32, c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i64)],
    "sable_eliminate_zeros_ell": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    "sable_csr_sort_rows_by_nnz": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                   ctypes.POINTER(c_i32)],
    "sable_permute_rows": [c_vp, c_int, c_i32, c_vp, c_int, c_vp, c_vp],
    "sable_csr_reorder_nn": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp],
    "sable_mm_header": [c_vp, c_sz, ctypes.POINTER(MMInfo)],
    "sable_mm_parse": [c_vp, c_vp, c_sz, c_int, c_vpp],
    "sable_mm_get_info": [c_vp, ctypes.POINTER(MMInfo)],
    "sable_mm_stats": [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)],
    "sable_mm_coo": [c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp],
    "sable_mm_destroy": [c_vp],
    "sable_csr_num_cols": [c_vp, c_i32, c_vp, ctypes.POINTER(c_i32)],
    "sable_csr_ell": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_csr_num_diags": [c_vp, c_i32, c_vp, c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)],
    "sable_csr_dia": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_transpose_csr": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_transpose_dia": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    "sable_count_transpose_cols_ell": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, ctypes.POINTER(c_i32)],
    "sable_transpose_ell": [c_vp, c_int, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp],
    "sable_spts_analyze": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_int, c_vpp],
    "sable_spts_destroy": [c_vp],
    "sable_spts_info": [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)],
    "sable_spts_set_lanes": [c_vp, c_int],
    "sable_spts_get_schedule": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "sable_spts_solve_level": [c_vp, c_vp, c_vp, c_vp],
    "sable_spts_solve_syncfree": [c_vp, c_vp, c_vp, c_vp],
    "sable_spts_solve_indegree": [c_vp, c_vp, c_vp, c_vp],
    "sable_batch_level_plan_create": [c_vp, c_int, c_vp, c_int, c_vpp],
    "sable_batch_level_plan_destroy": [c_vp],
    "sable_batch_level_plan_info": [c_vp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i64)],
    "sable_batch_level_plan_solve": [c_vp, c_vp, c_int, c_vp, c_int],
    "sable_batch_level_plan_solve_refined": [c_vp, c_vp, c_vp, c_int],
    "sable_spts_solve_combined": [c_vp, c_vp, c_vp, c_vp],
    "sable_spts_status": [c_vp, c_vp],
    "sable_spts_solve_refined": [c_vp, c_vp, c_vp, c_vp, c_int, c_int],
    "sable_spts_csc_analyze": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_int, c_vpp],
    "sable_spts_csc": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_int],
    "sable_spts_level_csr": [c_vp, c_int, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int],
    "sable_spts_syncfree_csr": [c_vp, c_int, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, ctypes.c_longlong],
    "sable_csr_plan_blocks": [c_vp, c_vpp, ctypes.POINTER(c_i64)],
    "sable_batch_spmv": [c_vp, c_int, c_int, c_vp, c_int, c_i64, c_int],
    "sable_batch_spts": [c_vp, c_int, c_int, c_vp, c_int, c_i64],
    "sable_batch_spts_refined": [c_vp, c_int, c_vp, c_int, c_i64, c_int],
    "sable_server_create": [c_vp, c_vp, c_vpp],
    "sable_server_destroy": [c_vp],
    "sable_server_add_matrix": [c_vp, c_i32, c_vp, c_vp, c_vp, c_int, ctypes.POINTER(c_i32)],
    "sable_server_submit": [c_vp, c_i32, c_int, c_vp, c_vp, ctypes.POINTER(c_i64)],
    "sable_server_wait": [c_vp, c_i64, ctypes.POINTER(c_dbl)],
    "sable_server_run_load": [c_vp, c_vp, c_vp],
    "sable_server_info": [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i32)],
    "sable_scan_exclusive_i32": [c_vp, c_vp, c_vp, c_i64],
    "sable_sort_pairs_u64": [c_vp, c_vp, c_vp, c_i64, c_int],
}

_RESTYPES = {
    "sable_ctx_destroy": None,
    "sable_csr_plan_destroy": None,
    "sable_coo_plan_destroy": None,
    "sable_csr_tiles_destroy": None,
    "sable_csr_auto_destroy": None,
    "sable_csr_auto_cache_clear": None,
    "sable_mm_destroy": None,
    "sable_spts_destroy": None,
    "sable_csr_panels_destroy": None,
    "sable_server_destroy": None,
    "sable_batch_level_plan_destroy": None,
    "sable_ctx_stream": c_vp,
    "sable_last_error": ctypes.c_char_p,
    "sable_version": ctypes.c_char_p,
}

_lib = None


def load():
    """Load the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsablesparse_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = L
    return L


def call(name, *args):
    """Call an int-returning entry point and raise SableError on failure."""
    L = load()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.sable_last_error()
        raise SableError(rc, name, msg.decode() if msg else "")
    return rc


def exported_symbols():
    return sorted(SIGNATURES)
