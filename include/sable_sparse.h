/*
 * sable_sparse.h -- C ABI of libsablesparse_hip, an MI355X (gfx950) HIP
 * implementation of the SciWasm.Sparse Level-II hot path: SpMV over
 * COO/CSR/DIA/ELL, lower sparse triangular solve (level-set and sync-free),
 * and the coo_csr / csr_dia / csr_ell conversions and transposes.
 *
 * Conventions
 *  - Every matrix/vector pointer is a DEVICE pointer (allocate with
 *    sable_device_alloc or any hipMalloc'd memory on the context's device).
 *    Index arrays are int32, as in the reference; element counts that can
 *    exceed 2^31 (nnz, stored slots) are int64.
 *  - Calls are asynchronous on the context's stream unless they return a
 *    host-side quantity (counts, widths), which synchronises.
 *  - Every function returns SABLE_OK (0) or a negative SABLE_E* code;
 *    sable_last_error() gives the message of the last failure on the
 *    calling thread.
 *  - dtype selects float32 (SABLE_F32) or float64 (SABLE_F64) values.
 *
 * Reference interface each entry point replaces is cited as
 * (file:line) relative to the reference repository, with
 *   seq/ = src/sequential/, par/ = src/parallel/.
 */
#ifndef SABLE_SPARSE_H
#define SABLE_SPARSE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SABLE_API __attribute__((visibility("default")))

enum {
    SABLE_OK = 0,
    SABLE_EINVAL = -1,     /* bad argument                                   */
    SABLE_EREFUSED = -2,   /* format guard refused (reference returns undefined) */
    SABLE_EMALFORMED = -3, /* malformed matrix: non-monotone row_ptr, index out
                              of range, missing/non-last diagonal            */
    SABLE_ENOMEM = -4,     /* device allocation failed                        */
    SABLE_EHIP = -5,       /* HIP runtime error                               */
    SABLE_EBUDGET = -6,    /* request exceeds the HBM budget of a server       */
    SABLE_ETIMEOUT = -7,   /* bounded wait expired                            */
    SABLE_ESINGULAR = -8   /* zero on the diagonal of a triangular solve      */
};

typedef enum { SABLE_F32 = 0, SABLE_F64 = 1 } sable_dtype_t;

typedef struct sable_ctx_s sable_ctx_t;

/* ------------------------------------------------------------ context */
/* Replaces the WebAssembly instance + linear memory set-up
 * (seq/src/sparse32.js:1-9, par/src/sparse32.js:3-30).  stream may be NULL
 * (the context then owns a default-priority stream). */
SABLE_API int sable_ctx_create(int device, void* hip_stream, sable_ctx_t** out);
SABLE_API void sable_ctx_destroy(sable_ctx_t* ctx);
SABLE_API int sable_ctx_sync(sable_ctx_t* ctx);
SABLE_API void* sable_ctx_stream(sable_ctx_t* ctx);
SABLE_API const char* sable_last_error(void);
SABLE_API const char* sable_version(void);

/* Device memory: replaces _malloc/_free of the L0 runtime
 * (seq/libs/matmachjs-lib.js, used at seq/src/sparse32.js:2174-2395). */
SABLE_API int sable_device_alloc(sable_ctx_t* ctx, size_t bytes, void** dptr);
SABLE_API int sable_device_free(sable_ctx_t* ctx, void* dptr);
SABLE_API int sable_memcpy_h2d(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
SABLE_API int sable_memcpy_d2h(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
SABLE_API int sable_memcpy_d2d(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
/* Fill with a byte value using a kernel (no hipMemset nodes). */
SABLE_API int sable_fill_bytes(sable_ctx_t* ctx, void* dst, int byte, size_t bytes);
/* Replaces clear_vec (seq/src/sparse32.js:1319-1322) and create_vec's fill. */
SABLE_API int sable_fill_value(sable_ctx_t* ctx, sable_dtype_t dtype, void* dst, int64_t n, double value);

/* Timing on the context stream (HIP events). */
SABLE_API int sable_event_create(void** ev);
SABLE_API int sable_event_destroy(void* ev);
SABLE_API int sable_event_record(sable_ctx_t* ctx, void* ev);
SABLE_API int sable_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms);

/* ------------------------------------------------------------ SpMV     */
/* All SpMV entry points compute  y = A.x + beta*y, repeated inner_max times.
 * beta = 1 with inner_max = n is the reference's y += A.x wrapper
 * semantics; beta = 0 is a plain product. */

/* CSR.  Adaptive row-block schedule computed on the device from row_ptr.
 * row_ptr may be a slice (row_ptr[0] != 0): col/val are indexed by the
 * absolute row_ptr values, as in the parallel worker slices.
 * Replaces spmv_csr / spmv_csr_wrapper (seq/src/sparse_opt_32.wat:227-288,
 * seq/src/sparse_opt_64.wat:239-300; par/src/sparse_opt_32.wat:1712-1771). */
typedef struct sable_csr_plan_s sable_csr_plan_t;
typedef struct {
    int block_weight;  /* nnz + row_weight*rows per work block (0 = default) */
    int row_weight;    /* cost of one row in nnz units (0 = default)         */
} sable_csr_tune_t;
SABLE_API int sable_csr_plan_create(sable_ctx_t* ctx, int32_t N, const int32_t* row_ptr,
                                    const sable_csr_tune_t* tune, sable_csr_plan_t** out);
SABLE_API void sable_csr_plan_destroy(sable_csr_plan_t* plan);
SABLE_API int sable_csr_plan_info(const sable_csr_plan_t* plan, int64_t* nblocks, int64_t* nnz,
                                  int32_t* max_block_nnz);
SABLE_API int sable_spmv_csr_plan(sable_ctx_t* ctx, const sable_csr_plan_t* plan, sable_dtype_t dtype,
                                  const int32_t* col, const void* val, const void* x, void* y,
                                  double beta, int inner_max);
SABLE_API int sable_spmv_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                             const int32_t* col, const void* val, const void* x, void* y,
                             double beta, int inner_max);

/* COO.  Row-sorted input runs a deterministic segmented reduction (row
 * heads found on the device, then the CSR block kernel); unsorted input is
 * reduced with device atomics.  Replaces spmv_coo / spmv_coo_wrapper
 * (seq/src/sparse_opt_32.wat:175-225) and the parallel
 * sparse_spmv_coo + sum driver (par/src/sparse32.js:2637-2693). */
SABLE_API int sable_spmv_coo(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz,
                             const int32_t* row, const int32_t* col, const void* val,
                             const void* x, void* y, double beta, int inner_max);

/* DIA, row-indexed storage: entry (i, i+offset[k]) at
 * data[k*stride + i - (offset[k] < 0 ? N - stride : 0)].  Per-row order is
 * the stored diagonal order, with a rounded multiply then add, so the
 * result is bit-identical to the reference.  Replaces spmv_dia /
 * spmv_dia_wrapper (seq/src/sparse_opt_32.wat:290-425) and spmv_dia_col
 * (par/src/sparse_opt_32.wat:9112-9246). */
SABLE_API int sable_spmv_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd,
                             int32_t stride, const int32_t* offset, const void* data,
                             const void* x, void* y, double beta, int inner_max);

/* ELL, column-major slots: data[k*N+i], indices[k*N+i].  Slot order per
 * row, rounded multiply then add: bit-identical to the reference.
 * Replaces spmv_ell / spmv_ell_wrapper (seq/src/sparse_opt_32.wat:427-498)
 * and spmv_ell_col (par/src/sparse_opt_32.wat:11055-11106). */
SABLE_API int sable_spmv_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                             const int32_t* indices, const void* data, const void* x, void* y,
                             double beta, int inner_max);

#ifdef __cplusplus
}
#endif
#endif /* SABLE_SPARSE_H */
