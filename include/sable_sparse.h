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
/* Pinned (page-locked) host memory, e.g. for a host-side linear memory
 * that is copied to and from the device. */
SABLE_API int sable_host_alloc(size_t bytes, void** out);
SABLE_API int sable_host_free(void* p);
SABLE_API int sable_memcpy_h2d(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
SABLE_API int sable_memcpy_d2h(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
SABLE_API int sable_memcpy_d2d(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
/* Copies between PAGEABLE host memory and the device through the context's
 * two pinned 4 MiB staging halves, the host copy of one half overlapping the
 * DMA of the other (the addon's write-tracked linear memory is pageable).
 * h2d returns once the host bytes have been read (the device copy is ordered
 * on the context stream); d2h returns with the bytes in place.  One caller at
 * a time per context. */
SABLE_API int sable_memcpy_h2d_staged(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
SABLE_API int sable_memcpy_d2h_staged(sable_ctx_t* ctx, void* dst, const void* src, size_t bytes);
/* Fill with a byte value using a kernel (no hipMemset nodes). */
SABLE_API int sable_fill_bytes(sable_ctx_t* ctx, void* dst, int byte, size_t bytes);
/* Replaces clear_vec (seq/src/sparse32.js:1319-1322) and create_vec's fill. */
SABLE_API int sable_fill_value(sable_ctx_t* ctx, sable_dtype_t dtype, void* dst, int64_t n, double value);
/* y[i] = y[i] + w[i] for i in [0, n): the reduction of per-worker partial
 * vectors after a parallel COO SpMV.  Replaces sum
 * (par/src/sparse_opt_32.wat:1420-1481, driven from par/src/sparse32.js:2675-2691). */
SABLE_API int sable_vec_add(sable_ctx_t* ctx, sable_dtype_t dtype, void* y, const void* w, int64_t n);

/* Largest element of a device int32 array (-1 for n = 0; synchronises):
 * the x and y lengths of a resident COO/ELL, found once per upload instead
 * of scanning the host arrays on every call. */
SABLE_API int sable_max_i32(sable_ctx_t* ctx, const int32_t* a, int64_t n, int32_t* out);

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
    int vector_lanes;  /* 0: row blocks; 2..64: the row-vector kernel with this many
                          lanes per row (for short, regular rows) */
    int vector_rows;   /* row-vector kernel: rows per lane group, 1, 2, 4, 8 or 16 (0 = 4) */
} sable_csr_tune_t;
SABLE_API int sable_csr_plan_create(sable_ctx_t* ctx, int32_t N, const int32_t* row_ptr,
                                    const sable_csr_tune_t* tune, sable_csr_plan_t** out);
SABLE_API void sable_csr_plan_destroy(sable_csr_plan_t* plan);
SABLE_API int sable_csr_plan_info(const sable_csr_plan_t* plan, int64_t* nblocks, int64_t* nnz,
                                  int32_t* max_block_nnz);
SABLE_API int sable_spmv_csr_plan(sable_ctx_t* ctx, const sable_csr_plan_t* plan, sable_dtype_t dtype,
                                  const int32_t* col, const void* val, const void* x, void* y,
                                  double beta, int inner_max);
/* Selects the CSR block kernel for all later launches (process-wide): 0 =
 * basic block kernel, one workgroup per row block (default); 1 = the kernel
 * with staged row pointers and early y loads, one workgroup per row block;
 * 2 = that kernel on a persistent grid of 8 workgroups per CU; 3 = the basic
 * kernel with non-temporal (streaming) col/val loads. */
SABLE_API int sable_csr_kernel_select(int kernel);
/* Plan-free CSR product: the row-block plan is built on the device on every
 * call (block weight raised when the default would need more than N/2 +
 * 1024 blocks) and a persistent grid runs it, so the call neither
 * synchronises nor reads anything back.  A row_ptr that decreases leaves y
 * unchanged (no error is reported without a synchronisation; the plan
 * entry points check it). */
SABLE_API int sable_spmv_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                             const int32_t* col, const void* val, const void* x, void* y,
                             double beta, int inner_max);

/* sable_spmv_csr without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (work under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * device-built row-block plan goes to a fixed work space in the library's
 * code object, one copy per GPU (SABLE_NOALLOC_WORK_BYTES, held until the
 * process ends), so N is at most SABLE_NOALLOC_MAX_ROWS (larger N:
 * SABLE_EINVAL, nothing launched).  Same kernels, plan and bits as
 * sable_spmv_csr.  Every call on one GPU shares that GPU's work space:
 * calls that overlap on one GPU (other streams or contexts) race. */
#define SABLE_NOALLOC_MAX_ROWS (1 << 23)
#define SABLE_NOALLOC_WORK_BYTES (256 + 4 * ((SABLE_NOALLOC_MAX_ROWS / 2) + 1024 + 1))
SABLE_API int sable_spmv_csr_noalloc(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                     const int32_t* col, const void* val, const void* x, void* y, double beta,
                                     int inner_max);

/* Column-panel blocked CSR (no reference counterpart; a schedule of the same
 * CSR product for matrices whose x does not fit one XCD's L2): the columns
 * are cut into panels of panel_cols columns (0 = 2 MiB of x) and the matrix
 * is copied panel-major; one launch per panel keeps every x gather of that
 * launch inside an L2-resident slice of x.  The plan owns its copy of the
 * matrix (the caller's arrays may be freed afterwards).  Same y = A.x +
 * beta*y semantics and tolerance as sable_spmv_csr (the per-row sum is
 * accumulated panel by panel). */
typedef struct sable_csr_panels_s sable_csr_panels_t;
SABLE_API int sable_csr_panels_create(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                      const int32_t* col, const void* val, int32_t panel_cols,
                                      const sable_csr_tune_t* tune, sable_csr_panels_t** out);
SABLE_API void sable_csr_panels_destroy(sable_csr_panels_t* p);
SABLE_API int sable_csr_panels_info(const sable_csr_panels_t* p, int32_t* npanels, int32_t* panel_cols, int64_t* bytes);
SABLE_API int sable_spmv_csr_panels(sable_ctx_t* ctx, const sable_csr_panels_t* p, const void* x, void* y, double beta,
                                    int inner_max);

/* Column-sorted tiles (no reference counterpart; a schedule of the same
 * CSR product for matrices whose columns are scattered over an x much larger
 * than one XCD's L2): rows are cut into tiles of tile_rows rows and columns
 * into panels of panel_cols columns (tile_rows a multiple of 64 and at most
 * 20480, panel_cols a power of two >= 64, tile_rows * panel_cols <= 2^32;
 * 0 = 16384 rows and 8192 columns); each (panel, tile) keeps its entries
 * sorted by column, so that consecutive lanes gather neighbouring x values,
 * and sums its rows in LDS in FP64 (products in the value type, each row's
 * sum rounded to it once).  The plan owns its copy of the
 * matrix (row_ptr[0] must be 0).  Same y = A.x + beta*y semantics and
 * tolerance as sable_spmv_csr; the summation order of a row depends on LDS
 * atomic order, so results are not bit-reproducible from run to run. */
typedef struct sable_csr_tiles_s sable_csr_tiles_t;
SABLE_API int sable_csr_tiles_create(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                     const int32_t* col, const void* val, int32_t tile_rows, int32_t panel_cols,
                                     sable_csr_tiles_t** out);
SABLE_API void sable_csr_tiles_destroy(sable_csr_tiles_t* p);
SABLE_API int sable_csr_tiles_info(const sable_csr_tiles_t* p, int32_t* row_tiles, int32_t* npanels, int64_t* bytes);
SABLE_API int sable_spmv_csr_tiles(sable_ctx_t* ctx, const sable_csr_tiles_t* p, const void* x, void* y, double beta,
                                   int inner_max);
/* The same plan for a matrix of N rows whose columns lie in [0, ncols)
 * (e.g. a row slice of a larger matrix); sable_csr_tiles_create is this
 * with ncols = N. */
SABLE_API int sable_csr_tiles_create_rect(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t ncols,
                                          const int32_t* row_ptr, const int32_t* col, const void* val,
                                          int32_t tile_rows, int32_t panel_cols, sable_csr_tiles_t** out);
/* Tile schedule for all later products (process-wide): 0 = one launch per
 * panel, y updated after each; 1 = one launch, each row tile walks all
 * panels with its row sums in LDS and writes y once (default). */
SABLE_API int sable_csr_tiles_select(int mode);

/* Many tile plans in one persistent launch (no reference counterpart; the
 * batched form of sable_spmv_csr_tiles for batches of matrices whose x
 * outgrows what a row block's random gathers can hit in the L1), with nrhs
 * in {1, 2, 4, 8} right-hand sides: x and y of every item are row-major
 * [N][nrhs], and x must be 16-byte aligned (device allocations are).  The row tiles of all items form one global work
 * list: item k owns units [unit_begin_k, unit_begin_k + ntr_k).
 * Workgroups claim units from a device counter; each one sums its tile's
 * rows in LDS over the item's panels and writes y = A.x + beta*y for those
 * rows.  Items are filled by sable_csr_tiles_batch_item, live in DEVICE
 * memory and are ordered by ascending unit_begin; all plans share `dtype`.
 * max_tile_rows (>= every item's tile_rows; max_tile_rows * nrhs FP64 row
 * sums must fit the 160 KiB LDS, i.e. at most 20480 / nrhs) sizes the LDS
 * of a workgroup.  An item whose tile_rows exceeds it
 * is skipped and the call fails with SABLE_EINVAL (the call synchronises to
 * read that flag).  workgroups = 0 picks 4 per CU. */
typedef struct {
    const int32_t* tile_ptr;
    const uint32_t* packed;
    const void* val;
    const void* x;
    void* y;
    int64_t unit_begin; /* first global unit (row tile) of this item */
    double beta;
    int32_t N;
    int32_t tile_rows;
    int32_t ntr;        /* row tiles = units of this item */
    int32_t npanels;
    int32_t panel_cols;
    int32_t row_bits;
} sable_batch_tiles_item_t;
SABLE_API int sable_csr_tiles_batch_item(const sable_csr_tiles_t* p, const void* x, void* y, double beta,
                                         int64_t unit_begin, sable_batch_tiles_item_t* out);
SABLE_API int sable_batch_spmv_tiles(sable_ctx_t* ctx, sable_dtype_t dtype, int nrhs,
                                     const sable_batch_tiles_item_t* d_items, int nitems, int64_t total_units,
                                     int32_t max_tile_rows, int workgroups);

/* Autotuned CSR product for a resident matrix: the schedule behind the
 * reference's spmv_csr / spmv_csr_wrapper (seq/src/sparse_opt_32.wat:227-288,
 * seq/src/sparse_opt_64.wat:239-300, par/src/sparse_opt_32.wat:1712-1771)
 * and spmv_coo when the Node-API addon keeps the matrix on the device.
 * Candidate schedules (row blocks at three weights, the row-vector kernel
 * with lanes around the mean row length, column-sorted tiles) are built and
 * timed with HIP events on the context stream and the fastest is kept; the
 * choice is cached per matrix class (row-length statistics, value type,
 * whether x fits one XCD's L2), so later matrices of a class skip the timing.
 * row_ptr may be a slice (row_ptr[0] != 0, col/val indexed by the absolute
 * values); ncols <= 0 measures the x length as the largest column + 1.
 * Creation synchronises; products only launch.  The caller's row_ptr, col
 * and val must stay valid and unchanged while the plan is used (a tile
 * schedule keeps its own copy of col and val).  Flags:
 * SABLE_AUTO_DETERMINISTIC keeps to the bit-reproducible schedules (no
 * tiles); SABLE_AUTO_NO_TIMING takes a rule of thumb instead of timing. */
enum { SABLE_AUTO_DETERMINISTIC = 1, SABLE_AUTO_NO_TIMING = 2 };
enum { SABLE_SCHED_BLOCKS = 0, SABLE_SCHED_VECTOR = 1, SABLE_SCHED_TILES = 2 };
typedef struct sable_csr_auto_s sable_csr_auto_t;
typedef struct {
    int schedule;        /* SABLE_SCHED_*                                          */
    int param0, param1;  /* blocks: block/row weight; vector: lanes/rows; tiles:
                            tile rows/panel columns                               */
    int32_t ncols;       /* x length                                              */
    int64_t nnz;
    int64_t extra_bytes; /* device bytes of the schedule's own copy of the matrix */
    float ms;            /* measured time of one product (0 when taken from the cache) */
    int from_cache;      /* 1: the class's cached choice, not timed again           */
} sable_csr_auto_info_t;
SABLE_API int sable_csr_auto_create(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t ncols,
                                    const int32_t* row_ptr, const int32_t* col, const void* val, int flags,
                                    sable_csr_auto_t** out);
SABLE_API void sable_csr_auto_destroy(sable_csr_auto_t* plan);
SABLE_API int sable_csr_auto_info(const sable_csr_auto_t* plan, sable_csr_auto_info_t* info);
SABLE_API int sable_spmv_csr_auto(sable_ctx_t* ctx, const sable_csr_auto_t* plan, const void* x, void* y,
                                  double beta, int inner_max);
/* Forgets every cached class choice (process-wide). */
SABLE_API void sable_csr_auto_cache_clear(void);

/* COO.  Row-sorted input runs a deterministic segmented reduction (row
 * heads found on the device, then the CSR block kernel); unsorted input is
 * reduced with device atomics.  Replaces spmv_coo / spmv_coo_wrapper
 * (seq/src/sparse_opt_32.wat:175-225) and the parallel
 * sparse_spmv_coo + sum driver (par/src/sparse32.js:2637-2693). */
SABLE_API int sable_spmv_coo(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz,
                             const int32_t* row, const int32_t* col, const void* val,
                             const void* x, void* y, double beta, int inner_max);
/* The same product with the row analysis done once: the plan checks the
 * rows (SABLE_EMALFORMED when one is out of range; synchronises once) and,
 * for row-sorted input, keeps the row heads and row blocks, so that
 * sable_spmv_coo_plan only launches.  `row` must not change while the
 * plan is used. */
typedef struct sable_coo_plan_s sable_coo_plan_t;
SABLE_API int sable_coo_plan_create(sable_ctx_t* ctx, int32_t N, int64_t nnz, const int32_t* row,
                                    sable_coo_plan_t** out);
SABLE_API void sable_coo_plan_destroy(sable_coo_plan_t* plan);
SABLE_API int sable_coo_plan_info(const sable_coo_plan_t* plan, int* sorted, int64_t* nblocks);
/* Device row heads (N + 1 entries, owned by the plan) of a row-sorted COO:
 * with them the COO's col and val arrays form a CSR, e.g. for a tile plan.
 * SABLE_EINVAL when the rows are not sorted.  No reference counterpart
 * (the reference's COO product has no plan). */
SABLE_API int sable_coo_plan_row_ptr(const sable_coo_plan_t* plan, const int32_t** d_row_ptr);
SABLE_API int sable_spmv_coo_plan(sable_ctx_t* ctx, const sable_coo_plan_t* plan, sable_dtype_t dtype,
                                  const int32_t* row, const int32_t* col, const void* val, const void* x, void* y,
                                  double beta, int inner_max);

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

/* Rows [r0, r1) only of the DIA and ELL products above (the other rows of y
 * are not touched), each row with the same order and bits.  The parallel
 * tree's column-major slices: spmv_dia_col_wrapper
 * (par/src/sparse_opt_32.wat:9112-9246), spmv_dia_col_basic_wrapper
 * (:8991-9111) and spmv_bdia_col_wrapper (:8757-8965), whose blocked and
 * basic sweeps add each row's diagonals in the same stored order; and
 * spmv_ell_col_wrapper (:11055-11106), spmv_ell_col_gs_wrapper
 * (:11291-11400) and spmv_bell_col_gs_wrapper (:11130-11290), whose SIMD
 * runs across consecutive rows, so each row adds its slots in order. */
SABLE_API int sable_spmv_dia_range(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                   const int32_t* offset, const void* data, int32_t r0, int32_t r1, const void* x,
                                   void* y, double beta, int inner_max);
SABLE_API int sable_spmv_ell_range(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                                   const int32_t* indices, const void* data, int32_t r0, int32_t r1, const void* x,
                                   void* y, double beta, int inner_max);

/* Row-major ("major: row") DIA and ELL of the parallel tree, y += A.x on
 * rows [r0, r1) (inner_max times).  DIA: data[i*nd + k] = a(i, i +
 * offset[k]), columns outside [0, N) skipped; replaces spmv_dia /
 * spmv_dia_wrapper (par/src/sparse_opt_32.wat:8213-8290).  ELL:
 * data/indices[i*nc + k]; order 0 adds the slots in order and leaves the
 * row at the first negative index, replacing spmv_ell / spmv_ell_wrapper
 * (:8292-8367); order 1 is the gs order of spmv_ell_row_gs_wrapper
 * (:8369-8756): for nc >= 4 the first nc % 4 slots go into the running sum,
 * the rest into four lane sums that are added to it in lane order.  Both
 * bit-identical to the reference. */
SABLE_API int sable_spmv_dia_rowmajor(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd,
                                      const int32_t* offset, const void* data, int32_t r0, int32_t r1,
                                      const void* x, void* y, int inner_max);
SABLE_API int sable_spmv_ell_rowmajor(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                                      const int32_t* indices, const void* data, int32_t r0, int32_t r1,
                                      const void* x, void* y, int order, int inner_max);


/* ------------------------------------------------------------ conversions */
/* All conversions are bit-identical to the reference algorithms.  The
 * reference's storage guards (nc*N < 2^27 and nc*N/nnz <= 5 for ELL,
 * nd*stride < 2^27 and nd*stride/nnz <= 5 for DIA) live in the caller, as
 * in the reference's JS; these entry points convert whatever they are given. */

/* COO -> CSR: rows and, within a row, columns ascending; equal (row, col)
 * pairs keep their input order (a stable sort; duplicates are not summed).
 * out_row_ptr has N+1 entries.  Replaces coo_csr (seq/src/sparse32.js:
 * 1674-1720, seq/src/sparse64.js:1255; par/src/sparse32.js:2096-2145). */
SABLE_API int sable_coo_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz, const int32_t* row,
                            const int32_t* col, const void* val, int32_t* out_row_ptr, int32_t* out_col,
                            void* out_val);

/* Longest row (synchronises).  Replaces num_cols (seq/src/sparse32.js:1436-1447). */
SABLE_API int sable_csr_num_cols(sable_ctx_t* ctx, int32_t N, const int32_t* row_ptr, int32_t* nc);

/* CSR -> column-major ELL of width nc (padding value 0, index 0).
 * Replaces csr_ell (seq/src/sparse32.js:1449-1484) and csr_ell_col
 * (par/src/sparse32.js:1796-1836). */
SABLE_API int sable_csr_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, const int32_t* row_ptr,
                            const int32_t* col, const void* val, int32_t* out_indices, void* out_data);

/* Number of occupied diagonals and the smallest |offset| among them
 * (synchronises).  Replaces num_diags (seq/src/sparse32.js:1486-1512;
 * seq/src/sparse64.js:1066-1092 uses stride = N - min|offset|). */
SABLE_API int sable_csr_num_diags(sable_ctx_t* ctx, int32_t N, const int32_t* row_ptr, const int32_t* col,
                                  int32_t* nd, int32_t* min_abs);

/* CSR -> DIA: out_offset ascending (nd entries), out_data[k*stride + i -
 * (offset<0 ? N-stride : 0)] = a(i, i+offset[k]), zero padding; a later
 * duplicate (i, c) overwrites an earlier one.  Replaces csr_dia
 * (seq/src/sparse32.js:1515-1574, seq/src/sparse64.js:1095-1157) and
 * csr_dia_col (par/src/sparse32.js:1976-2040). */
SABLE_API int sable_csr_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                            const int32_t* row_ptr, const int32_t* col, const void* val, int32_t* out_offset,
                            void* out_data);

/* Row-major ("major: row") conversions of the parallel tree: ELL slots at
 * i*nc + k (padding value 0, index 0), replacing csr_ell
 * (par/src/sparse32.js:1840-1880); DIA data[i*nd + k] with stride N and the
 * offsets ascending (a later duplicate overwrites), replacing csr_dia
 * (par/src/sparse32.js:1912-1973).  Bit-identical to the reference. */
SABLE_API int sable_csr_ell_rowmajor(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                                     const int32_t* row_ptr, const int32_t* col, const void* val,
                                     int32_t* out_indices, void* out_data);
SABLE_API int sable_csr_dia_rowmajor(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd,
                                     const int32_t* row_ptr, const int32_t* col, const void* val,
                                     int32_t* out_offset, void* out_data);

/* COO -> CSC: the counting scatter by column with a stable per-column sort
 * by row (duplicates kept in input order), i.e. coo_csr with the roles of
 * row and column exchanged.  Replaces coo_csc (seq/src/sparse32.js:1629-1671). */
SABLE_API int sable_coo_csc(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz, const int32_t* row,
                            const int32_t* col, const void* val, int32_t* out_col_ptr, int32_t* out_row,
                            void* out_val);

/* Sorts a COO in place by (row, col), stably (equal keys keep their order).
 * For distinct (row, col) keys this is the order of quick_sort_COO
 * (seq/src/sparse32.js:1577-1609), used by the in-place COO transpose
 * (seq/src/sparse32.js:285-295) and the MatrixMarket reader. */
SABLE_API int sable_coo_sort(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz, int32_t* row,
                             int32_t* col, void* val);

/* CSR transpose (counting sort by column, rows ascending within each new
 * row; values moved bitwise).  Replaces transpose_csr
 * (seq/src/sparse_opt_32.wat:6348-6477, caller seq/src/sparse32.js:1116-1120). */
SABLE_API int sable_transpose_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                  const int32_t* col, const void* val, int32_t* out_row_ptr, int32_t* out_col,
                                  void* out_val);

/* DIA transpose, stride == N: out_offset[k] = -offset[k] (descending),
 * each stored run re-indexed by its new row; other slots zero.  Replaces
 * transpose_dia (seq/src/sparse_opt_32.wat:6479-6572). */
SABLE_API int sable_transpose_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                  const int32_t* offset, const void* data, int32_t* out_offset, void* out_data);

/* Width of the transposed ELL: the largest per-column count of slots whose
 * value compares > 0 or < 0 (zeros, -0 and NaN are dropped) (synchronises).
 * Replaces count_transpose_cols_ell (seq/src/sparse_opt_32.wat:6574-6632). */
SABLE_API int sable_count_transpose_cols_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                                             const void* data, const int32_t* indices, int32_t* nc_out);

/* ELL transpose into width nc_out: new row c lists input rows j ascending
 * (then slots ascending); padding zero.  Replaces transpose_ell
 * (seq/src/sparse_opt_32.wat:6634-6701). */
SABLE_API int sable_transpose_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, int32_t nc_out,
                                  const void* data, const int32_t* indices, void* out_data, int32_t* out_indices);

/* ------------------------------------------------------------ element-wise ops */
/* Operation on each stored value (seq/src/sparse32.js:671-1105 out of
 * place, :200-660 in place).  expm1, log1p, sin, tan and pow run in double
 * precision and round to the value type, as the reference's WebAssembly
 * calls into JavaScript's Math (pow with Math.pow's NaN cases); the rest are
 * the WebAssembly instructions. */
typedef enum {
    SABLE_EW_EXPM1 = 0,
    SABLE_EW_LOG1P = 1,
    SABLE_EW_SIN = 2,
    SABLE_EW_TAN = 3,
    SABLE_EW_POW = 4,      /* param = exponent p */
    SABLE_EW_DEG2RAD = 5,  /* param = pi: v * (pi / 180), pi / 180 rounded in the value type */
    SABLE_EW_RAD2DEG = 6,  /* param = pi: v / (pi / 180) */
    SABLE_EW_SIGN = 7,     /* v > 0 ? 1 : -1 (the COO/CSR exports) */
    SABLE_EW_SIGN0 = 8,    /* v == 0 ? 0 : (v > 0 ? 1 : -1) (the DIA/ELL exports) */
    SABLE_EW_ABS = 9,
    SABLE_EW_NEG = 10,
    SABLE_EW_SQRT = 11,
    SABLE_EW_CEIL = 12,
    SABLE_EW_FLOOR = 13,
    SABLE_EW_TRUNC = 14,
    SABLE_EW_NEAREST = 15, /* round half to even */
    SABLE_EW_MULTIPLY = 16 /* param = scalar */
} sable_ew_op_t;

/* out[i] = op(in[i]) for i < n: COO/CSR values (n = nnz) or all ELL slots
 * (n = ncols*N).  in == out is the in-place form.  Replaces <op>_coo,
 * self_<op>_coo, <op>_ell, self_<op>_ell and self_multiply_scalar_coo
 * (seq/src/sparse_opt_32.wat:597-1699, 4977-6347; sparse_opt_64.wat). */
SABLE_API int sable_ew_values(sable_ctx_t* ctx, sable_dtype_t dtype, int op, double param, int64_t n, const void* in,
                              void* out);
/* DIA: op on the in-range slots of every diagonal only (rows max(0,-k) to
 * min(N, N-k)), restricted to rows [r0, r1) (0, N for the whole matrix);
 * padding slots of out are not written.  Replaces <op>_dia / self_<op>_dia
 * (seq/src/sparse_opt_32.wat:2137-4976) and the worker-range forms
 * (par/src/sparse_opt_32.wat, e.g. self_expm1_dia(id, offset, data,
 * start_row, end_row, ...)). */
SABLE_API int sable_ew_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int op, double param, int32_t N, int32_t nd,
                           int32_t stride, const int32_t* offset, int32_t r0, int32_t r1, const void* in, void* out);
/* ELL rows [r0, r1) of every slot column (column-major slots k*N + i).
 * Replaces the parallel tree's <op>_ell / self_<op>_ell worker forms
 * (par/src/sparse_opt_32.wat, (id, [scalar,] data, start_row, end_row,
 * ncols, N)); sable_ew_values covers the serial form. */
SABLE_API int sable_ew_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int op, double param, int32_t N, int32_t nc,
                           int32_t r0, int32_t r1, const void* in, void* out);

/* Number of values v with v > 0 || v < 0 (synchronises).  Replaces
 * count_nonzeros_coo (seq/src/sparse_opt_32.wat:500-521). */
SABLE_API int sable_count_nonzeros(sable_ctx_t* ctx, sable_dtype_t dtype, int64_t n, const void* val, int64_t* count);
/* COO without its zero (and NaN) values, order kept; outputs hold *count
 * entries (size them with sable_count_nonzeros).  Replaces copy_nonzeros_coo
 * (seq/src/sparse_opt_32.wat:523-546), eliminate_zeros (seq/src/sparse32.js:
 * 1140-1155). */
SABLE_API int sable_eliminate_zeros_coo(sable_ctx_t* ctx, sable_dtype_t dtype, int64_t nnz, const int32_t* row,
                                        const int32_t* col, const void* val, int32_t* out_row, int32_t* out_col,
                                        void* out_val, int64_t* count);
/* CSR without its zero values: out_row_ptr (N+1), out_col/out_val (*count).
 * Replaces copy_nonzeros_csr (seq/src/sparse_opt_32.wat:548-595; its outer
 * loop also visits a row N past the end, which is not repeated here). */
SABLE_API int sable_eliminate_zeros_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                        const int32_t* col, const void* val, int32_t* out_row_ptr, int32_t* out_col,
                                        void* out_val, int64_t* count);
/* DIA: diagonals with a nonzero in range, and nonzeros in range
 * (synchronises).  Replaces count_nonzeros_dia (:1948-2039). */
SABLE_API int sable_count_nonzeros_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                       const int32_t* offset, const void* data, int32_t* count_diags, int64_t* count);
/* DIA keeping the diagonals that hold a nonzero (offsets in order) in
 * count_diags*stride zero-filled slots; as the reference does, the kept
 * diagonal's nonzeros are packed in row order from the slot of its first
 * row (zeros inside a diagonal are squeezed out, moving later values up).
 * Replaces copy_nonzeros_dia (:2041-2135). */
SABLE_API int sable_eliminate_zeros_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                        const int32_t* offset, const void* data, int32_t* out_offset, void* out_data,
                                        int32_t* count_diags);
/* ELL: the largest per-row count of nonzero slots, and the total
 * (synchronises).  Replaces count_nonzeros_ell (:5821-5893). */
SABLE_API int sable_count_nonzeros_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, const void* data,
                                       int32_t* count_cols, int64_t* count);
/* ELL of width nc_out (= count_cols): each row's nonzero slots in slot
 * order, the rest zero.  Replaces copy_nonzeros_ell (:5895-5949). */
SABLE_API int sable_eliminate_zeros_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, int32_t nc_out,
                                        const void* data, const int32_t* indices, void* out_data,
                                        int32_t* out_indices);

/* ------------------------------------------------------------ row reordering */
/* Rows in ascending nnz order, stable: out_perm[i] = the old row placed at
 * new row i; out_row_ptr/out_col/out_val the permuted CSR (row_ptr[0] = 0);
 * out_nnz_row[r] = nnz of old row r (may be NULL); bins = the number of rows
 * with 0, <= 1, <= 2 and <= 3 entries (num_zero_rows ... num_three_rows),
 * which select the short-row SpMV variants (synchronises).  Replaces
 * sort_rows_by_nnz (par/src/sparse32.js:2148-2218; sparse64.js:1297-1367),
 * whose row_ptr stores go through a zero-length view and are lost; the
 * row_ptr here is the one that loop computes. */
SABLE_API int sable_csr_sort_rows_by_nnz(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                         const int32_t* col, const void* val, int32_t* out_row_ptr, int32_t* out_col,
                                         void* out_val, int32_t* out_perm, int32_t* out_nnz_row, int32_t* bins);
/* Locality ordering of a CSR whose rows are sorted by nnz: within each
 * bucket of <= 1024 rows of equal nnz, a greedy walk that appends the
 * unvisited row most similar (Jaccard similarity of the 16-column cache
 * lines it touches) to one of the last w rows placed, from each of those
 * rows' top-w candidate lists.  out_perm[i] = perm_in[r] for the input row r
 * placed at new row i (perm_in may be NULL); the permuted CSR goes to
 * out_row_ptr/out_col/out_val.  w in [1, 16] (synchronises).  Replaces
 * reorder_NN (par/src/sparse32.js:2501-2622), including its list rules; its
 * row_ptr stores go through a zero-length view as in sort_rows_by_nnz. */
SABLE_API int sable_csr_reorder_nn(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                   const int32_t* col, const void* val, const int32_t* perm_in, int w,
                                   int32_t* out_row_ptr, int32_t* out_col, void* out_val, int32_t* out_perm);
/* scatter = 1: out[perm[i]] = in[i] (results of the reordered matrix back to
 * the original row order; sort_y_rows_by_nnz, par/src/sparse32.js:2220-2235);
 * scatter = 0: out[i] = in[perm[i]]. */
SABLE_API int sable_permute_rows(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* perm, int scatter,
                                 const void* in, void* out);

/* ------------------------------------------------------------ MatrixMarket ingest */
/* Field of the banner (token 4 of "%%MatrixMarket matrix coordinate <field>
 * <symmetry>"); every field but pattern carries a value per entry (complex:
 * the real part, as the reference reads token 3 of each line). */
typedef enum {
    SABLE_MM_REAL = 0,
    SABLE_MM_INTEGER = 1,
    SABLE_MM_COMPLEX = 2,
    SABLE_MM_PATTERN = 3,
    SABLE_MM_OTHER = 4
} sable_mm_field_t;

/* mm_info of the reference (seq/src/sparse32.js:14-25). */
typedef struct {
    int64_t nrows, ncols, nentries; /* the size line */
    int64_t nnz;        /* entries of the full COO: symmetric off-diagonals
                         * twice, explicit zeros dropped, 0 -> nentries
                         * (calculate_actual_nnz, read_matrix_MM_files) */
    int64_t lower_nnz;  /* entries of the lower COO: nrows + strictly lower
                         * nonzeros (create_LCOO_from_MM's L_nnz) */
    int32_t field;      /* sable_mm_field_t */
    int32_t symmetric;  /* symmetry token == "symmetric" (others read as general) */
    int64_t body_offset; /* byte offset of the first line after the size line */
} sable_mm_info_t;

typedef struct sable_mm_s sable_mm_t;

/* Header only, on the host: banner, '%' comment lines, size line.
 * Replaces read_MM_header (seq/src/sparse32.js:1889-1911). */
SABLE_API int sable_mm_header(const char* text, size_t len, sable_mm_info_t* info);

/* Parses the whole text of a coordinate MatrixMarket file on the device.
 * text is a host pointer (copied to HBM and released before returning) or,
 * with text_on_device = 1, a device pointer that stays untouched.  Every
 * entry line is checked: tokens must be decimal ("row col [value]"), indices
 * inside the size line, and the entry count must match nentries
 * (SABLE_EMALFORMED names the byte offset and the line).  Values are rounded
 * to the nearest binary64 exactly as JavaScript's Number() does
 * (synchronises).  Replaces read_matrix_MM_files + calculate_actual_nnz
 * (seq/src/sparse32.js:1914-1981; seq/src/sparse64.js:1536-1564). */
SABLE_API int sable_mm_parse(sable_ctx_t* ctx, const char* text, size_t len, int text_on_device, sable_mm_t** out);
SABLE_API int sable_mm_get_info(const sable_mm_t* mm, sable_mm_info_t* info);
/* Lines holding an entry or a blank, and values decided by the exact
 * big-integer conversion (diagnostics). */
SABLE_API int sable_mm_stats(const sable_mm_t* mm, int64_t* lines, int64_t* exact_values);

/* The COO (0-based, sorted by (row, col)) into device arrays of info.nnz
 * entries (lower = 0) or info.lower_nnz entries (lower = 1).  lower = 0
 * replaces create_COO_from_MM (seq/src/sparse32.js:2100-2172): symmetric
 * off-diagonals expanded to (r, c) and (c, r), explicit zeros dropped,
 * pattern values 1; when every entry was zero the nentries outputs are zero.
 * lower = 1 replaces create_LCOO_from_MM (seq/src/sparse32.js:1983-2097;
 * par/src/spts32.js:14): strictly lower nonzeros (upper ones mirrored when
 * symmetric) plus a unit diagonal on every row.  Equal (row, col) keys keep
 * file order; the reference's quicksort leaves them in pivot order. */
SABLE_API int sable_mm_coo(sable_ctx_t* ctx, const sable_mm_t* mm, sable_dtype_t dtype, int lower, int32_t* row,
                           int32_t* col, void* val);
SABLE_API void sable_mm_destroy(sable_mm_t* mm);

/* ------------------------------------------------------------ triangular solve */
/* L.x = b for a lower-triangular CSR whose rows each end with the diagonal
 * (the reference's layout; par/src/spts32.js:14-136 builds L = tril(A,-1)+I
 * that way).  lanes = lanes per row (1, 4, 8, 16 or 64; 0 = chosen from the
 * mean row length) of the level graph and the natural-order sync-free
 * solve: they form a row's products in parallel and add them in stored
 * order.  Every schedule and lane count keeps the reference's per-row
 * operation order, so the level-set and sync-free solves are bit-identical
 * to the reference's. */
typedef struct sable_spts_s sable_spts_t;

/* Level-set analysis: checks the structure (SABLE_EMALFORMED) and the
 * diagonal (SABLE_ESINGULAR), computes level[i], level_ptr, the stable
 * level permutation and order, and the level-ordered CSR (all bit-identical
 * to CSR_create_level_sets_with_reorder, par/src/sparse32.js:2942-3181,
 * par/src/spts32.js:158-357).  The caller's arrays must outlive the
 * analysis (the sync-free solve and the refinement read them). */
SABLE_API int sable_spts_analyze(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                 const int32_t* col, const void* val, int lanes, sable_spts_t** out);
SABLE_API void sable_spts_destroy(sable_spts_t* a);
SABLE_API int sable_spts_info(const sable_spts_t* a, int32_t* nlevels, int32_t* max_width, int32_t* lanes);
SABLE_API int sable_spts_set_lanes(sable_spts_t* a, int lanes);
/* Copies of the schedule (any pointer may be NULL): h_level_ptr is a host
 * array of nlevels+1 entries, the others are device arrays. */
SABLE_API int sable_spts_get_schedule(sable_ctx_t* ctx, const sable_spts_t* a, int32_t* h_level_ptr,
                                      int32_t* d_permutation, int32_t* d_order, int32_t* d_row_ptr, int32_t* d_col,
                                      void* d_val);

/* Schedules of the level-set and sync-free solves of an analysis:
 *  SABLE_SPTS_BLOCKS (default): one launch.  Rows are cut into blocks of
 *    consecutive rows, claimed by workgroups in increasing order; inside a
 *    block the rows run in level order, one level per step, with the
 *    block's solution in LDS and a barrier per step, and a dependency on an
 *    earlier block is waited for on its published value.  One thread per row, entries in stored order, so the
 *    level solve's bits equal the reference level solve's and the sync-free
 *    solve's equal the reference sync-free solve's.  block_rows: a power of
 *    two in [64, 16384] (0 keeps the current one; the analysis starts with
 *    about N/256 rows, in [1024, 8192]).
 *  SABLE_SPTS_LEVEL_GRAPH: one launch per level over the level-ordered
 *    matrix, captured once in a hipGraph and replayed (the analysis's lane
 *    count per row); the sync-free solve: waves claim rows in natural order.
 *  SABLE_SPTS_LEVEL_LAUNCHES: the same per-level launches without a graph. */
enum { SABLE_SPTS_BLOCKS = 0, SABLE_SPTS_LEVEL_GRAPH = 1, SABLE_SPTS_LEVEL_LAUNCHES = 2 };
SABLE_API int sable_spts_set_schedule(sable_spts_t* a, int schedule, int32_t block_rows);
SABLE_API int sable_spts_schedule_info(const sable_spts_t* a, int* schedule, int32_t* block_rows,
                                       int32_t* nblocks, int32_t* nsteps, int64_t* slots);
/* Times the level-set solve under row blocks of 1024, 2048, ... 16384 rows
 * (up to twice N) and under the level graph at 1, 4, 8, 16 and 64 lanes per
 * row, on b = L*1 in buffers of its own, and keeps the fastest schedule and
 * lane count (the row-block schedule stays built at its fastest size either
 * way).  The solution depends on neither, bit for bit.  Outputs may be NULL.
 * Has no counterpart in the reference, whose solves take a fixed schedule. */
SABLE_API int sable_spts_tune(sable_ctx_t* ctx, sable_spts_t* a, int* schedule, int32_t* block_rows, int* lanes,
                              float* best_ms);

/* Level-set solve with b and x in natural order (distinct arrays), on the
 * analysis's schedule.  Replaces spts_level_csr
 * (par/src/sparse_opt_32.wat:11759-11984 with spts_csr :11679-11757), whose
 * workers run the levels of one call behind an in-call barrier, and the
 * level + sync-free combined solve (:12846-13353). */
SABLE_API int sable_spts_solve_level(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x);

/* Sync-free solve, natural order (b and x distinct), one launch, on the
 * analysis's schedule.  Replaces spts_csr_sync_free(_wrapper)
 * (par/src/spts_opt_32.wat:8-260). */
SABLE_API int sable_spts_solve_syncfree(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x);

/* Sync-free solve with per-row in-degree counters, natural order (b and x
 * distinct), one launch: row i waits until the count of its solved
 * dependencies reaches its in-degree, then forms its sum in stored order
 * and increments the count of every row that depends on it (the analysis
 * keeps those lists, built on first use).  The same bits as
 * sable_spts_solve_syncfree (the reference's spts_csr_sync_free,
 * par/src/spts_opt_32.wat:8-260), a different way of waiting. */
SABLE_API int sable_spts_solve_indegree(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x);

/* Level + sync-free combined solve, b and x in natural order: the
 * sync-free kernel over the analysis's level-ordered matrix, with the
 * level solve's arithmetic, so that the result equals sable_spts_solve_level
 * bit for bit at every lane count.  Replaces the combined solve
 * spts_csr_opt_level_sync_free_wrapper (par/src/sparse_opt_32.wat:12846-13353)
 * and spts_csr_level_sync_free_wrapper, whose arithmetic is spts_csr's
 * (:11679-11757). */
SABLE_API int sable_spts_solve_combined(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x);

/* Synchronises; SABLE_ETIMEOUT if a sync-free solve of this analysis hit
 * its spin limit since the last call. */
SABLE_API int sable_spts_status(sable_ctx_t* ctx, sable_spts_t* a);

/* FP32 solve plus `iters` steps of FP64 iterative refinement on an FP64
 * analysis: r = b - L.x in FP64, L32.d = r in FP32 (inner 0 = level-set,
 * 1 = sync-free), x += d in FP64.  inner 2 (iters must be 1): the FP32
 * level solve and its one refinement step fused into one sweep of the
 * analysis's schedule (row blocks, level graph or per-level launches), each
 * row forming x0 = L32 solve of (float)b, r = (float)(b - L.x0) with the
 * FP64 products chained by fma in stored order (diagonal last), d = L32
 * solve of r, and x = x0 + d: the same x on every schedule.  No reference
 * counterpart (the reference has no mixed-precision solve). */
SABLE_API int sable_spts_solve_refined(sable_ctx_t* ctx, sable_spts_t* a, const double* b, double* x, int iters,
                                       int inner);

/* Lower L given in CSC with the diagonal first in each column (the serial
 * tree's layout, seq/src/sparse32.js:1629-1671 builds it): the analysis
 * transposes it into an owned CSR (rows in ascending column order,
 * diagonal last) and analyses that, so every solve entry point above
 * applies.  SABLE_EMALFORMED when a column is empty or does not start with
 * its diagonal. */
SABLE_API int sable_spts_csc_analyze(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* col_ptr,
                                     const int32_t* row, const void* val, int lanes, sable_spts_t** out);
/* The column sweep y = L^-1 b, inner_max times (y overwritten each time):
 * per row the subtractions happen in ascending column order, then the
 * division, so the sync-free solve on the transposed CSR performs the
 * same operations in the same order and the result is bit-identical.  Replaces spts_csc / spts_csc_wrapper
 * (seq/src/sparse_opt_32.wat:11-173).  Allocates and synchronises. */
SABLE_API int sable_spts_csc(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* col_ptr,
                             const int32_t* row, const void* val, const void* b, void* y, int inner_max);

/* Reference-shaped level-set solve on an already level-ordered matrix with
 * level-ordered b and y; h_level_ptr on the host.  One launch per level.
 * Replaces spts_level_csr (par/src/sparse_opt_32.wat:11759-11984). */
SABLE_API int sable_spts_level_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nlevels,
                                   const int32_t* h_level_ptr, const int32_t* row_ptr, const int32_t* col,
                                   const void* val, const void* b, void* y, int lanes);

/* Reference-shaped sync-free solve on a natural-order matrix; checks the
 * structure first; every wait is bounded by spin_limit polls without
 * progress (0 = default); synchronises and returns SABLE_ETIMEOUT when the
 * bound was hit.  Replaces spts_csr_sync_free_wrapper
 * (par/src/spts_opt_32.wat:119-260). */
SABLE_API int sable_spts_syncfree_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                      const int32_t* col, const void* val, const void* b, void* x, int lanes,
                                      long long spin_limit);

/* ------------------------------------------------------------ batches */
/* No reference counterpart (the reference runs one matrix per call): many
 * independent matrices in one persistent launch.  Item arrays live in
 * DEVICE memory; items are ordered by ascending blk_begin / row_begin. */

/* Device row-block boundaries of a CSR plan (nblocks+1 entries) for
 * building batch items. */
SABLE_API int sable_csr_plan_blocks(const sable_csr_plan_t* plan, const int32_t** d_blk, int64_t* nblocks);

typedef struct {
    const int32_t* row_ptr;
    const int32_t* col;
    const void* val;
    const int32_t* blk;  /* the matrix's plan blocks (sable_csr_plan_blocks) */
    const void* x;       /* [N][nrhs] row-major */
    void* y;             /* [N][nrhs] row-major */
    int64_t blk_begin;   /* first global block of this item */
    double beta;
} sable_batch_spmv_item_t;

/* y_k = A_k.x_k + beta_k*y_k for every item, nrhs in {1, 2, 4, 8} right-hand
 * sides, over total_blocks = sum of the items' block counts; workgroups = 0
 * picks a persistent grid of 4 per CU.  x of every item must be 16-byte
 * aligned (device allocations are). */
SABLE_API int sable_batch_spmv(sable_ctx_t* ctx, sable_dtype_t dtype, int nrhs, const sable_batch_spmv_item_t* d_items,
                               int nitems, int64_t total_blocks, int workgroups);

typedef struct {
    const int32_t* row_ptr;  /* lower CSR, diagonal last in each row */
    const int32_t* col;
    const void* val;
    const void* b;
    void* x;
    int64_t row_begin;  /* first global row of this item */
    int32_t N;
    int32_t pad;
} sable_batch_spts_item_t;

/* Sync-free solve of every item in one launch (synchronises; SABLE_ETIMEOUT
 * when a wait exceeded its bound).  lanes per row (1, 4, 8, 16, 64) form a
 * row's products in parallel and subtract them in stored order: the
 * reference sync-free solve's bits at every lane count. */
SABLE_API int sable_batch_spts(sable_ctx_t* ctx, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* d_items,
                               int nitems, int64_t total_rows);

typedef struct {
    const int32_t* row_ptr;
    const int32_t* col;
    const double* val64;
    const float* val32;
    const double* b64;
    double* x64;
    float* r32;  /* N floats of work space */
    float* d32;  /* N floats of work space */
    int64_t row_begin;
    int32_t N;
    int32_t pad;
} sable_batch_refine_item_t;

/* Level-set solve of every item of a batch, level by level: one launch per
 * level of the deepest matrix, each solving that level of every item at
 * once (the batched form of the per-level launches of
 * sable_spts_solve_level; the reference level solve's bits for each item at
 * every lane count).  The plan reads only row_ptr, col and N of the host
 * item array given at creation (device pointers; checks the structure:
 * SABLE_EMALFORMED), computes every item's levels and keeps the row lists
 * and chunk tables; a solve takes a device item array of the same matrices
 * in the same order, with any b and x, and only launches.  No reference
 * counterpart (the reference solves one matrix per call). */
typedef struct sable_batch_level_plan_s sable_batch_level_plan_t;
SABLE_API int sable_batch_level_plan_create(sable_ctx_t* ctx, int lanes, const sable_batch_spts_item_t* h_items,
                                            int nitems, sable_batch_level_plan_t** out);
SABLE_API void sable_batch_level_plan_destroy(sable_batch_level_plan_t* p);
SABLE_API int sable_batch_level_plan_info(const sable_batch_level_plan_t* p, int32_t* levels, int64_t* chunks);
SABLE_API int sable_batch_level_plan_solve(sable_ctx_t* ctx, const sable_batch_level_plan_t* p, sable_dtype_t dtype,
                                           const sable_batch_spts_item_t* d_items, int nitems);
/* The FP32 solve of every item with one FP64 refinement step fused into
 * the same level sweep (one launch per level of the deepest matrix): per
 * row x0 = L32 solve of (float)b, r = (float)(b - L.x0) (FP64 products
 * chained by fma in stored order, diagonal last), d = L32 solve of r, x64 =
 * x0 + d, as sable_spts_solve_refined does with inner 2 (the FP32 values
 * are the FP64 values rounded as they are read).  Items (device memory, the
 * plan's matrices in its order) give val64; r32 must hold 2N floats of work
 * space; val32 and d32 are not read.  No reference counterpart. */
SABLE_API int sable_batch_level_plan_solve_refined(sable_ctx_t* ctx, const sable_batch_level_plan_t* p,
                                                   const sable_batch_refine_item_t* d_items, int nitems);

/* FP32 sync-free solve + `iters` FP64 refinement steps for every item
 * (synchronises). */
SABLE_API int sable_batch_spts_refined(sable_ctx_t* ctx, int lanes, const sable_batch_refine_item_t* d_items,
                                       int nitems, int64_t total_rows, int iters);

/* ------------------------------------------------------------ serving */
/* No reference counterpart: a request server over resident matrices.
 * Matrices and their SpMV plans (and FP32 copies) stay in HBM and are
 * charged against hbm_budget; requests queue on the host and are batched
 * continuously into `slots` in-flight batches (see mode). */
typedef struct sable_server_s sable_server_t;
typedef struct {
    int64_t hbm_budget;   /* bytes for matrices + batch arenas (0 = unlimited) */
    int mode;             /* 0 plain FP64, one launch per request; 1 batched FP64;
                             2 batched, FP32 SpMV and FP32 solve + FP64 refinement */
    int max_batch;        /* requests per batch (0 = 256) */
    int slots;            /* batches in flight (0 = 3) */
    int64_t arena_bytes;  /* device vector arena per slot (0 = 256 MiB) */
    int refine_iters;     /* refinement steps in mode 2 (< 0: 1); one step runs fused
                             into the level sweep of the solve */
    int spts_lanes;       /* lanes per row of the solves (0 = 4) */
    int group_rhs;        /* modes 1 and 2: SpMV requests on one matrix in a batch share
                             one multi-RHS product, up to this many (0 = 8, 1 = off) */
    int batch_tiles;      /* modes 1 and 2, matrices below 2^20 rows: 0 = row-block batch
                             (each row summed in a fixed order, so a result is the same
                             bits alone and in any mix); 1 = column-sorted tile plans in
                             the mode's precision, one batched tile launch per width
                             (faster; rows summed in LDS-atomic order, so results agree
                             within the CSR bound) */
    int spts_sweep;       /* modes 1 and 2, the batch's solves: 0 = level by level, one
                             launch per level of the deepest matrix for all solves at
                             once (mode 1: the reference level solve's bits); 1 = one
                             sync-free launch for all of them (mode 1: the reference
                             sync-free solve's bits) */
} sable_server_config_t;

SABLE_API int sable_server_create(sable_ctx_t* ctx, const sable_server_config_t* cfg, sable_server_t** out);
SABLE_API void sable_server_destroy(sable_server_t* s);
/* Host CSR arrays (row_ptr[0] = 0); kind 0 = general (SpMV), 1 = lower
 * triangular with the diagonal last (SpMV and solves).  SABLE_EBUDGET when
 * the matrix does not fit the remaining HBM budget. */
SABLE_API int sable_server_add_matrix(sable_server_t* s, int32_t N, const int32_t* row_ptr, const int32_t* col,
                                      const double* val, int kind, int32_t* matrix_id);
/* op 0: out = A.in; op 1: solve A.out = in.  Host vectors of N doubles;
 * `in` must stay valid until the request completes. */
SABLE_API int sable_server_submit(sable_server_t* s, int32_t matrix_id, int op, const double* in, double* out,
                                  int64_t* request_id);
/* Blocks until the request is done; returns its status (SABLE_EBUDGET when
 * it can never fit a batch arena). */
SABLE_API int sable_server_wait(sable_server_t* s, int64_t request_id, double* latency_us);

typedef struct {
    double rate;           /* requests per second (Poisson arrivals) */
    double duration_s;
    double spts_fraction;  /* share of lower solves (the rest are SpMV) */
    uint64_t seed;
    int64_t max_requests;  /* 0 = no cap */
} sable_load_t;
typedef struct {
    int64_t submitted, completed, failed, rejected, batches;
    double elapsed_s, solves_per_s, p50_us, p99_us, mean_us, max_us, mean_batch;
} sable_load_stats_t;
/* Open-loop load: requests arrive at exponential gaps, choose a matrix
 * uniformly (solves: among the lower ones), inputs U(-1,1) generated on the
 * device from a per-request seed and results left in HBM (nothing crosses
 * PCIe); latency runs from the scheduled arrival to the completion of the
 * request's batch. */
SABLE_API int sable_server_run_load(sable_server_t* s, const sable_load_t* load, sable_load_stats_t* stats);
SABLE_API int sable_server_info(sable_server_t* s, int64_t* hbm_used, int32_t* nmatrices);

/* ------------------------------------------------------------ building blocks */
/* No reference counterpart: device-wide primitives the conversions are
 * built from, exported for callers that assemble their own formats. */
/* out[0..n) = exclusive prefix sum of in[0..n); out[n] = total (out has n+1). */
SABLE_API int sable_scan_exclusive_i32(sable_ctx_t* ctx, const int32_t* in, int32_t* out, int64_t n);
/* Stable sort of (keys, vals) by key bits [0, end_bit), in place; key bits
 * from end_bit up travel with their keys unsorted.  At most 2^31 - 1 keys. */
SABLE_API int sable_sort_pairs_u64(sable_ctx_t* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int end_bit);

#ifdef __cplusplus
}
#endif
#endif /* SABLE_SPARSE_H */
