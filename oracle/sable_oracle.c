/*
 * sable_oracle.c -- CPU restatement of the SciWasm.Sparse Level-II hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled
 * into oracle/liboracle.so and may be loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and there only
 * as the checker (or the timed CPU baseline).  The shipped product
 * (libsablesparse_hip.so) never links or calls it.
 *
 * Each function restates one reference routine in the reference's own
 * operation order (built with -ffp-contract=off so a*b+c is a rounded
 * multiply followed by a rounded add, as in WebAssembly f32/f64 arithmetic).
 * The *_ld functions are the FP64 oracle: long-double accumulation used to
 * report the error of every floating-point result.
 *
 * Pinning (how each restatement is checked, see tests/test_oracle.py):
 *   coo_csr, csr_ell, csr_dia, transposes, level sets: brute-force dense
 *     reconstruction on random small matrices, closed forms (5-point and
 *     7-point Laplacians: DIA offsets, ELL width, level of grid point
 *     (i,j,k) = i+j+k and trinomial level sizes), invariants (monotone
 *     row_ptr, ascending columns, stable ties, bijective permutations,
 *     transpose(transpose(A)) = A).
 *   spmv_*: exact integer arithmetic cases (values and x small integers, so
 *     every rounding is exact) against a dense product; Laplacian closed
 *     forms; order-sensitive cases (1e16, 1, -1e16) that a reordered sum
 *     fails.
 *   spts_*: manufactured solutions with exactly representable integer data,
 *     summation-order-sensitive cases, and residual <= 1e-12 of the FP64
 *     oracle.
 *   fletcher_sum: hand-computed checksums.
 * The reference ships no golden vectors or fixtures for this path, so these
 * checks are the pin.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* coo_csr: seq/src/sparse32.js:1674-1720 (+ sort :1612-1627).          */
/* Row histogram, exclusive scan, scatter in input order, shift back,   */
/* then a per-row bubble sort by column that swaps only on '>' (so ties  */
/* keep input order) and has no early exit.                             */
/* ------------------------------------------------------------------ */
#define DEF_COO_CSR(T, SUF)                                                       \
EXPORT void oracle_coo_csr_##SUF(int N, int nnz, const int *row, const int *col,  \
                                 const T *val, int *csr_row, int *csr_col,       \
                                 T *csr_val)                                     \
{                                                                                 \
    memset(csr_row, 0, sizeof(int) * (size_t)(N + 1));                           \
    memset(csr_col, 0, sizeof(int) * (size_t)nnz);                               \
    memset(csr_val, 0, sizeof(T) * (size_t)nnz);                                 \
    for (int i = 0; i < nnz; i++) csr_row[row[i]]++;                             \
    int j = 0, j0;                                                                \
    for (int i = 0; i < N; i++) { j0 = csr_row[i]; csr_row[i] = j; j += j0; }     \
    for (int i = 0; i < nnz; i++) {                                               \
        j = csr_row[row[i]];                                                      \
        csr_col[j] = col[i];                                                      \
        csr_val[j] = val[i];                                                      \
        csr_row[row[i]]++;                                                        \
    }                                                                             \
    for (int i = N - 1; i > 0; i--) csr_row[i] = csr_row[i - 1];                 \
    csr_row[0] = 0;                                                               \
    csr_row[N] = nnz;                                                             \
    for (int r = 0; r < N; r++) {                                                 \
        int start = csr_row[r], end = csr_row[r + 1];                            \
        for (int p = 0; p < end - start - 1; p++)                                 \
            for (int q = start; q < end - p - 1; q++)                             \
                if (csr_col[q] > csr_col[q + 1]) {                                \
                    int tc = csr_col[q]; csr_col[q] = csr_col[q + 1];             \
                    csr_col[q + 1] = tc;                                          \
                    T tv = csr_val[q]; csr_val[q] = csr_val[q + 1];               \
                    csr_val[q + 1] = tv;                                          \
                }                                                                 \
    }                                                                             \
}
DEF_COO_CSR(float, f32)
DEF_COO_CSR(double, f64)

/* num_cols: seq/src/sparse32.js:1436-1447 -- the longest row. */
EXPORT int oracle_num_cols(int N, const int *rowptr)
{
    int max = 0;
    for (int i = 0; i < N; i++) {
        int t = rowptr[i + 1] - rowptr[i];
        if (max < t) max = t;
    }
    return max;
}

/* csr_ell: seq/src/sparse32.js:1449-1484.  Column-major slots,
 * data[k*N+i] / indices[k*N+i], padding value 0 and index 0.
 * The guard (nc*N < 2^27 && nc*N/nnz <= 5) lives in the caller. */
#define DEF_CSR_ELL(T, SUF)                                                       \
EXPORT void oracle_csr_ell_##SUF(int N, int nc, const int *rowptr,               \
                                 const int *col, const T *val, int *indices,     \
                                 T *data)                                        \
{                                                                                 \
    memset(indices, 0, sizeof(int) * (size_t)nc * N);                            \
    memset(data, 0, sizeof(T) * (size_t)nc * N);                                 \
    for (int i = 0; i < N; i++) {                                                 \
        int k = 0;                                                                \
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {                         \
            data[(size_t)k * N + i] = val[j];                                     \
            indices[(size_t)k * N + i] = col[j];                                  \
            k++;                                                                  \
        }                                                                         \
    }                                                                             \
}
DEF_CSR_ELL(float, f32)
DEF_CSR_ELL(double, f64)

/* num_diags: seq/src/sparse32.js:1486-1512 (stride N) and
 * seq/src/sparse64.js:1066-1092 (stride N - min|offset|).
 * Returns the number of occupied diagonals; *min_abs receives the
 * smallest |offset| among them (N-1 when none are occupied). */
EXPORT int oracle_num_diags(int N, const int *rowptr, const int *col, int *min_abs)
{
    int *ind = (int *)calloc((size_t)(2 * N > 1 ? 2 * N - 1 : 1), sizeof(int));
    int num_diag = 0;
    for (int i = 0; i < N; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            if (!ind[N + col[j] - i - 1]++) num_diag++;
    int diag_no = -((2 * N - 1) / 2);
    int mn = abs(diag_no);
    for (int i = 0; i < 2 * N - 1; i++) {
        if (ind[i] && mn > abs(diag_no)) mn = abs(diag_no);
        diag_no++;
    }
    free(ind);
    if (min_abs) *min_abs = mn;
    return num_diag;
}

/* csr_dia: seq/src/sparse32.js:1515-1574 (and sparse64.js:1095-1157).
 * offset[] ascending over occupied diagonals; entry (i, i+off_k) goes to
 * data[k*stride + i - move], move = N - stride for negative offsets.
 * A later duplicate (i, c) overwrites an earlier one.  Padding is zero
 * (sparse32.js zero-fills; sparse64.js leaves malloc contents, which this
 * restatement defines as zero). */
#define DEF_CSR_DIA(T, SUF)                                                       \
EXPORT void oracle_csr_dia_##SUF(int N, int nd, int stride, const int *rowptr,   \
                                 const int *col, const T *val, int *offset,      \
                                 T *data)                                        \
{                                                                                 \
    memset(data, 0, sizeof(T) * (size_t)nd * stride);                            \
    int *ind = (int *)calloc((size_t)(2 * N > 1 ? 2 * N - 1 : 1), sizeof(int));  \
    for (int i = 0; i < N; i++)                                                   \
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) ind[N + col[j] - i - 1]++;\
    int diag_no = -((2 * N - 1) / 2), index = 0;                                 \
    for (int i = 0; i < 2 * N - 1; i++) {                                         \
        if (ind[i]) offset[index++] = diag_no;                                    \
        diag_no++;                                                                \
    }                                                                             \
    free(ind);                                                                    \
    for (int i = 0; i < N; i++)                                                   \
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {                         \
            int c = col[j];                                                       \
            for (int k = 0; k < nd; k++) {                                        \
                int move = 0;                                                     \
                if (c - i == offset[k]) {                                         \
                    if (offset[k] < 0) move = N - stride;                         \
                    data[(size_t)k * stride + i - move] = val[j];                 \
                    break;                                                        \
                }                                                                 \
            }                                                                     \
        }                                                                         \
}
DEF_CSR_DIA(float, f32)
DEF_CSR_DIA(double, f64)

/* transpose_csr: seq/src/sparse_opt_32.wat:6348-6477.  Column histogram
 * into out_rowptr, exclusive scan, scatter in row order (so the new
 * columns ascend), shift back.  Values are moved as bit patterns
 * (esize bytes each). */
EXPORT void oracle_transpose_csr(int N, int nnz, int esize, const int *rowptr,
                                 const int *col, const void *val, int *out_rowptr,
                                 int *out_col, void *out_val)
{
    memset(out_rowptr, 0, sizeof(int) * (size_t)(N + 1));
    for (int i = 0; i < nnz; i++) out_rowptr[col[i]]++;
    int sum = 0;
    for (int i = 0; i < N; i++) { int c = out_rowptr[i]; out_rowptr[i] = sum; sum += c; }
    out_rowptr[N] = nnz;
    for (int i = 0; i < N; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            int addr = out_rowptr[col[j]];
            out_col[addr] = i;
            memcpy((char *)out_val + (size_t)addr * esize,
                   (const char *)val + (size_t)j * esize, (size_t)esize);
            out_rowptr[col[j]] = addr + 1;
        }
    int prev = 0;
    for (int i = 0; i <= N; i++) { int t = out_rowptr[i]; out_rowptr[i] = prev; prev = t; }
}

/* transpose_dia: seq/src/sparse_opt_32.wat:6479-6572.  out_offset[k] =
 * -offset[k] (so the order is descending); each diagonal's stored run is
 * shifted so that it is indexed by the new row.  Slots outside the run
 * are zero here (the reference leaves them as allocated). */
EXPORT void oracle_transpose_dia(int N, int nd, int stride, int esize,
                                 const int *offset, const void *data,
                                 int *out_offset, void *out_data)
{
    memset(out_data, 0, (size_t)esize * nd * stride);
    int exp1 = N - stride;
    for (int i = 0; i < nd; i++) {
        int k = offset[i];
        int base = i * stride - (k < 0 ? exp1 : 0);
        int in_n = k < 0 ? -k : 0;
        int in_end = (N < N - k) ? N : N - k;
        out_offset[i] = -k;
        int ko = -k;
        int obase = i * stride - (ko < 0 || k < 0 ? exp1 : 0);
        int out_n = ko < 0 ? -ko : 0;
        for (int t = 0; in_n + t < in_end; t++)
            memcpy((char *)out_data + (size_t)(obase + out_n + t) * esize,
                   (const char *)data + (size_t)(base + in_n + t) * esize, (size_t)esize);
    }
}

/* count_transpose_cols_ell + transpose_ell: seq/src/sparse_opt_32.wat
 * :6574-6701.  A slot counts when its value compares > 0 or < 0 (so zeros,
 * -0 and NaN are dropped); nnz_row[col] counts per new row; the new width
 * is the largest count.  Scatter walks input rows j ascending and their
 * slots ascending; padding of the output is zero here. */
#define DEF_TELL(T, SUF)                                                          \
EXPORT int oracle_count_transpose_cols_ell_##SUF(int N, int nc, const T *data,   \
                                                 const int *indices)             \
{                                                                                 \
    if (!(nc > 0 && N > 0)) return 0;                                             \
    int *cnt = (int *)calloc((size_t)N, sizeof(int));                             \
    int mx = 0;                                                                   \
    for (size_t e = 0; e < (size_t)nc * N; e++)                                   \
        if (data[e] > 0 || data[e] < 0) {                                         \
            int c = ++cnt[indices[e]];                                            \
            if (c > mx) mx = c;                                                   \
        }                                                                         \
    free(cnt);                                                                    \
    return mx;                                                                    \
}                                                                                 \
EXPORT void oracle_transpose_ell_##SUF(int N, int nc, int nc_out, const T *data, \
                                       const int *indices, T *out_data,          \
                                       int *out_indices)                         \
{                                                                                 \
    memset(out_data, 0, sizeof(T) * (size_t)nc_out * N);                          \
    memset(out_indices, 0, sizeof(int) * (size_t)nc_out * N);                     \
    if (!(nc > 0 && N > 0)) return;                                               \
    int *cnt = (int *)calloc((size_t)N, sizeof(int));                             \
    for (int j = 0; j < N; j++)                                                   \
        for (int s = 0; s < nc; s++) {                                            \
            size_t e = (size_t)s * N + j;                                         \
            if (data[e] > 0 || data[e] < 0) {                                     \
                int c = indices[e];                                               \
                size_t o = (size_t)cnt[c] * N + c;                                \
                out_data[o] = data[e];                                            \
                out_indices[o] = j;                                               \
                cnt[c]++;                                                         \
            }                                                                     \
        }                                                                         \
    free(cnt);                                                                    \
}
DEF_TELL(float, f32)
DEF_TELL(double, f64)

/* ------------------------------------------------------------------ */
/* Level sets: par/src/sparse32.js:2942-3008 and :3010-3181.            */
/* level[i] = 0 for a row with exactly one entry, else 1 + max(0, levels  */
/* of the columns of all entries but the last).  level_ptr is the prefix */
/* sum of level sizes; permutation[new] = old by a stable counting sort; */
/* order[old] = new.  Returns the number of levels.                     */
/* ------------------------------------------------------------------ */
EXPORT int oracle_level_sets(int N, const int *rowptr, const int *col,
                             int *level_of_row, int *level_ptr_out /* N+1 cap */,
                             int *permutation, int *order)
{
    int max_level = 0;
    for (int i = 0; i < N; i++) {
        if (rowptr[i + 1] - rowptr[i] == 1) {
            level_of_row[i] = 0;
        } else {
            int m = 0;
            for (int j = rowptr[i]; j < rowptr[i + 1] - 1; j++)
                if (level_of_row[col[j]] > m) m = level_of_row[col[j]];
            level_of_row[i] = 1 + m;
            if (level_of_row[i] > max_level) max_level = level_of_row[i];
        }
    }
    int tot = max_level + 1;
    int *freq = (int *)calloc((size_t)tot, sizeof(int));
    int *start = (int *)calloc((size_t)tot, sizeof(int));
    for (int i = 0; i < N; i++) freq[level_of_row[i]]++;
    level_ptr_out[0] = 0;
    start[0] = 0;
    for (int l = 1; l < tot; l++) start[l] = level_ptr_out[l] = level_ptr_out[l - 1] + freq[l - 1];
    level_ptr_out[tot] = level_ptr_out[tot - 1] + freq[tot - 1];
    for (int i = 0; i < N; i++) {
        int l = level_of_row[i];
        permutation[start[l]] = i;
        if (order) order[i] = start[l];
        start[l]++;
    }
    free(freq);
    free(start);
    return tot;
}

/* Reordered CSR of CSR_create_level_sets_with_reorder
 * (par/src/sparse32.js:3136-3150): new row i is old row permutation[i],
 * its columns renamed order[col], column order within the row unchanged. */
#define DEF_REORDER(T, SUF)                                                       \
EXPORT void oracle_level_reorder_##SUF(int N, const int *rowptr, const int *col, \
                                       const T *val, const int *permutation,     \
                                       const int *order, int *new_rowptr,        \
                                       int *new_col, T *new_val)                 \
{                                                                                 \
    int j = 0;                                                                    \
    new_rowptr[0] = 0;                                                            \
    for (int i = 0; i < N; i++) {                                                 \
        int k = rowptr[permutation[i]];                                           \
        int len = rowptr[permutation[i] + 1] - k;                                 \
        new_rowptr[i + 1] = new_rowptr[i] + len;                                  \
        while (len--) { new_col[j] = order[col[k]]; new_val[j++] = val[k++]; }   \
    }                                                                             \
}
DEF_REORDER(float, f32)
DEF_REORDER(double, f64)

/* ------------------------------------------------------------------ */
/* SpMV in reference order.  Semantics are y += A.x, inner_max times    */
/* (the *_wrapper exports).                                             */
/* ------------------------------------------------------------------ */
/* spmv_coo: seq/src/sparse_opt_32.wat:175-202: y[r] = v*x[c] + y[r], k ascending. */
#define DEF_SPMV(T, SUF)                                                          \
EXPORT void oracle_spmv_coo_##SUF(int nnz, const int *row, const int *col,       \
                                  const T *val, const T *x, T *y, int inner_max) \
{                                                                                 \
    for (int it = 0; it < inner_max; it++)                                        \
        for (int k = 0; k < nnz; k++) {                                           \
            T p = val[k] * x[col[k]];                                             \
            y[row[k]] = p + y[row[k]];                                            \
        }                                                                         \
}                                                                                 \
/* spmv_csr: seq/src/sparse_opt_32.wat:227-271: t = y[i]; t = v*x + t. */         \
EXPORT void oracle_spmv_csr_##SUF(int N, const int *rowptr, const int *col,      \
                                  const T *val, const T *x, T *y, int inner_max) \
{                                                                                 \
    for (int it = 0; it < inner_max; it++)                                        \
        for (int i = 0; i < N; i++) {                                             \
            int j = rowptr[i], e = rowptr[i + 1];                                 \
            if (j < e) {                                                          \
                T t = y[i];                                                       \
                for (; j < e; j++) { T p = val[j] * x[col[j]]; t = p + t; }       \
                y[i] = t;                                                         \
            }                                                                     \
        }                                                                         \
}                                                                                 \
/* spmv_dia: seq/src/sparse_opt_32.wat:290-407: diagonals in stored order,       \
 * rows n in [max(0,-off), min(N,N-off)), y[n] = d*x[n+off] + y[n]. */           \
EXPORT void oracle_spmv_dia_##SUF(int N, int nd, int stride, const int *offset,  \
                                  const T *data, const T *x, T *y, int inner_max)\
{                                                                                 \
    for (int it = 0; it < inner_max; it++)                                        \
        for (int k = 0; k < nd; k++) {                                            \
            int off = offset[k];                                                  \
            long base = (long)k * stride - (off < 0 ? N - stride : 0);            \
            int n0 = off < 0 ? -off : 0;                                          \
            int n1 = (N < N - off) ? N : N - off;                                 \
            for (int n = n0; n < n1; n++) {                                       \
                T p = data[base + n] * x[n + off];                                \
                y[n] = p + y[n];                                                  \
            }                                                                     \
        }                                                                         \
}                                                                                 \
/* spmv_ell: seq/src/sparse_opt_32.wat:427-479: slot-major sweep,                \
 * y[i] = d*x[idx] + y[i] (padding adds 0*x[0]). */                              \
EXPORT void oracle_spmv_ell_##SUF(int N, int nc, const int *indices,             \
                                  const T *data, const T *x, T *y, int inner_max)\
{                                                                                 \
    if (!(nc > 0 && N > 0)) return;                                               \
    for (int it = 0; it < inner_max; it++)                                        \
        for (int j = 0; j < nc; j++)                                              \
            for (int i = 0; i < N; i++) {                                         \
                size_t e = (size_t)j * N + i;                                     \
                T p = data[e] * x[indices[e]];                                    \
                y[i] = p + y[i];                                                  \
            }                                                                     \
}
DEF_SPMV(float, f32)
DEF_SPMV(double, f64)

/* Row-major ("major: row") storage of the parallel tree. */
#define DEF_ROWMAJOR(T, SUF)                                                      \
/* csr_ell, row-major: par/src/sparse32.js:1840-1880: data[i*nc+k],              \
 * indices[i*nc+k], padding value 0 and index 0. */                             \
EXPORT void oracle_csr_ell_row_##SUF(int N, int nc, const int *rowptr,           \
                                     const int *col, const T *val, int *indices, \
                                     T *data)                                    \
{                                                                                 \
    memset(indices, 0, sizeof(int) * (size_t)nc * N);                            \
    memset(data, 0, sizeof(T) * (size_t)nc * N);                                 \
    for (int i = 0; i < N; i++) {                                                 \
        int k = 0;                                                                \
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {                         \
            data[(size_t)i * nc + k] = val[j];                                    \
            indices[(size_t)i * nc + k] = col[j];                                 \
            k++;                                                                  \
        }                                                                         \
    }                                                                             \
}                                                                                 \
/* csr_dia, row-major: par/src/sparse32.js:1912-1973: offsets ascending over     \
 * the occupied diagonals (stride N), data[i*nd + k] = a(i, i+offset[k]) with    \
 * the first k whose offset matches; a later duplicate overwrites. */           \
EXPORT void oracle_csr_dia_row_##SUF(int N, int nd, const int *rowptr,           \
                                     const int *col, const T *val, int *offset,  \
                                     T *data)                                    \
{                                                                                 \
    memset(data, 0, sizeof(T) * (size_t)nd * N);                                 \
    int *ind = (int *)calloc((size_t)(2 * N > 1 ? 2 * N - 1 : 1), sizeof(int));  \
    for (int i = 0; i < N; i++)                                                   \
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) ind[N + col[j] - i - 1]++;\
    int diag_no = -((2 * N - 1) / 2), index = 0;                                 \
    for (int i = 0; i < 2 * N - 1; i++) {                                         \
        if (ind[i]) offset[index++] = diag_no;                                    \
        diag_no++;                                                                \
    }                                                                             \
    free(ind);                                                                    \
    for (int i = 0; i < N; i++)                                                   \
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)                           \
            for (int k = 0; k < nd; k++)                                          \
                if (col[j] - i == offset[k]) {                                    \
                    data[(size_t)i * nd + k] = val[j];                            \
                    break;                                                        \
                }                                                                 \
}                                                                                 \
/* spmv_dia, row-major: par/src/sparse_opt_32.wat:8213-8269: rows [r0, r1),      \
 * t = y[i]; for each diagonal k with c = offset[k] + i in [0, N):               \
 * t = d*x[c] + t; y[i] = t. */                                                  \
EXPORT void oracle_spmv_dia_row_##SUF(int N, int nd, const int *offset,          \
                                      const T *data, int r0, int r1, const T *x, \
                                      T *y, int inner_max)                       \
{                                                                                 \
    if (r0 >= r1 || nd <= 0) return;                                              \
    for (int it = 0; it < inner_max; it++)                                        \
        for (int i = r0; i < r1; i++) {                                           \
            T t = y[i];                                                           \
            for (int k = 0; k < nd; k++) {                                        \
                int c = offset[k] + i;                                            \
                if (c >= 0 && c < N) { T p = data[(size_t)i * nd + k] * x[c]; t = p + t; } \
            }                                                                     \
            y[i] = t;                                                             \
        }                                                                         \
}                                                                                 \
/* spmv_ell, row-major: par/src/sparse_opt_32.wat:8292-8346 (gs = 0): slots in   \
 * order from t = y[i], leaving the row at the first negative index; and         \
 * spmv_ell_row_gs :8369-8756 (gs = 1): for nc >= 4 the first nc % 4 slots go    \
 * into t, slot h + 4q + l into lane l (each lane from 0), and                   \
 * y[i] = (((t + l0) + l1) + l2) + l3; narrower rows take the plain order. */    \
EXPORT void oracle_spmv_ell_row_##SUF(int N, int nc, const int *indices,         \
                                      const T *data, int r0, int r1, const T *x, \
                                      T *y, int gs, int inner_max)               \
{                                                                                 \
    (void)N;                                                                      \
    if (r0 >= r1 || nc <= 0) return;                                              \
    for (int it = 0; it < inner_max; it++)                                        \
        for (int i = r0; i < r1; i++) {                                           \
            const int *ci = indices + (size_t)i * nc;                             \
            const T *d = data + (size_t)i * nc;                                   \
            T t = y[i];                                                           \
            if (!gs || nc < 4) {                                                  \
                for (int k = 0; k < nc && ci[k] >= 0; k++) { T p = d[k] * x[ci[k]]; t = p + t; } \
                y[i] = t;                                                         \
            } else {                                                              \
                int h = nc % 4;                                                   \
                for (int k = 0; k < h; k++) { T p = d[k] * x[ci[k]]; t = p + t; } \
                T l[4] = {0, 0, 0, 0};                                            \
                for (int k = h; k < nc; k += 4)                                   \
                    for (int q = 0; q < 4; q++) { T p = d[k + q] * x[ci[k + q]]; l[q] = p + l[q]; } \
                t = t + l[0]; t = t + l[1]; t = t + l[2]; t = t + l[3];           \
                y[i] = t;                                                         \
            }                                                                     \
        }                                                                         \
}
DEF_ROWMAJOR(float, f32)
DEF_ROWMAJOR(double, f64)

/* FP64 oracle for CSR SpMV: y_ld[i] = y0[i] + sum_j v*x in long double. */
EXPORT void oracle_spmv_csr_ld(int N, const int *rowptr, const int *col,
                               const double *val, const double *x, const double *y0,
                               double *y_out, double *absrow /* sum |v||x| (nullable) */)
{
    for (int i = 0; i < N; i++) {
        long double t = y0 ? (long double)y0[i] : 0.0L, a = 0.0L;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            t += (long double)val[j] * (long double)x[col[j]];
            a += fabsl((long double)val[j] * (long double)x[col[j]]);
        }
        y_out[i] = (double)t;
        if (absrow) absrow[i] = (double)a;
    }
}

/* ------------------------------------------------------------------ */
/* Lower triangular solve.  The diagonal is the last entry of each row. */
/* ------------------------------------------------------------------ */
/* Level-set order, spts_csr par/src/sparse_opt_32.wat:11679-11757 driven
 * by spts_level_csr :11759-11984: t = 0; t = t + v*y[c] over the
 * off-diagonals in stored order; y[i] = (b[i] - t)/diag.  Rows are processed
 * level by level (in row order within a level, which gives the same result
 * because rows of one level are independent). */
#define DEF_SPTS(T, SUF)                                                          \
EXPORT void oracle_spts_level_##SUF(int N, const int *rowptr, const int *col,    \
                                    const T *val, const T *b, T *y)              \
{                                                                                 \
    for (int i = 0; i < N; i++) {                                                 \
        int j = rowptr[i], end = rowptr[i + 1] - 1;                               \
        if (end > j) {                                                            \
            T t = 0;                                                              \
            for (; j < end; j++) { T p = val[j] * y[col[j]]; t = t + p; }         \
            y[i] = (b[i] - t) / val[end];                                         \
        } else {                                                                  \
            y[i] = b[i] / val[j];                                                 \
        }                                                                         \
    }                                                                             \
}                                                                                 \
/* Sync-free order, spts_csr_sync_free par/src/spts_opt_32.wat:8-117:            \
 * t = b[i]; t = t - v*y[c]; y[i] = t/diag. */                                    \
EXPORT void oracle_spts_syncfree_##SUF(int N, const int *rowptr, const int *col, \
                                       const T *val, const T *b, T *y)           \
{                                                                                 \
    for (int i = 0; i < N; i++) {                                                 \
        int j = rowptr[i], end = rowptr[i + 1] - 1;                               \
        T t = b[i];                                                               \
        for (; j < end; j++) { T p = val[j] * y[col[j]]; t = t - p; }             \
        y[i] = t / val[end];                                                      \
    }                                                                             \
}
DEF_SPTS(float, f32)
DEF_SPTS(double, f64)

/* Column-oriented forward substitution, seq/src/sparse_opt_32.wat:11-146
 * (spts_csc) and :149-173 (spts_csc_wrapper): L in CSC with the diagonal
 * first in each column; y = b; then for each column j in order:
 * y[j] = y[j] / diag_j, and for the column's other entries (in stored order)
 * y[row] = y[row] - val * y[j] (mul then sub).  The wat's 4-wide SIMD body
 * performs the same per-element operations. */
#define DEF_SPTS_CSC(T, SUF)                                                       \
EXPORT void oracle_spts_csc_##SUF(int N, const int *colptr, const int *row,       \
                                  const T *val, const T *b, T *y)                 \
{                                                                                 \
    for (int i = 0; i < N; i++) y[i] = b[i];                                      \
    for (int j = 0; j < N; j++) {                                                 \
        int k = colptr[j];                                                        \
        y[j] = y[j] / val[k];                                                     \
        for (k = k + 1; k < colptr[j + 1]; k++) {                                 \
            T p = val[k] * y[j];                                                  \
            y[row[k]] = y[row[k]] - p;                                            \
        }                                                                         \
    }                                                                             \
}
DEF_SPTS_CSC(float, f32)
DEF_SPTS_CSC(double, f64)

/* FP64 oracle forward substitution in long double. */
EXPORT void oracle_spts_ld(int N, const int *rowptr, const int *col,
                           const double *val, const double *b, double *y)
{
    long double *t = (long double *)malloc(sizeof(long double) * (size_t)(N > 0 ? N : 1));
    for (int i = 0; i < N; i++) {
        int end = rowptr[i + 1] - 1;
        long double s = (long double)b[i];
        for (int j = rowptr[i]; j < end; j++) s -= (long double)val[j] * t[col[j]];
        t[i] = s / (long double)val[end];
        y[i] = (double)t[i];
    }
    free(t);
}

/* Relative residual ||b - L x||_inf / (||L||_inf ||x||_inf + ||b||_inf) in
 * long double, for any square CSR (the lower-solve acceptance metric). */
EXPORT double oracle_rel_residual(int N, const int *rowptr, const int *col,
                                  const double *val, const double *x, const double *b)
{
    long double rmax = 0, lnorm = 0, xmax = 0, bmax = 0;
    for (int i = 0; i < N; i++) {
        long double s = (long double)b[i], rs = 0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            s -= (long double)val[j] * (long double)x[col[j]];
            rs += fabsl((long double)val[j]);
        }
        if (fabsl(s) > rmax) rmax = fabsl(s);
        if (rs > lnorm) lnorm = rs;
        if (fabsl((long double)x[i]) > xmax) xmax = fabsl((long double)x[i]);
        if (fabsl((long double)b[i]) > bmax) bmax = fabsl((long double)b[i]);
    }
    long double den = lnorm * xmax + bmax;
    return den > 0 ? (double)(rmax / den) : (double)rmax;
}

/* fletcher_sum_y: seq/src/sparse32.js:1291-1311.  JS numbers are doubles,
 * matlab_modulo(x,255) = x - floor(x/255)*255, parseInt truncates. */
static double mmod(double x, double y) { return x - floor(x / y) * y; }
EXPORT double oracle_fletcher_f32(int n, const float *y)
{
    double s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) { s1 = mmod(s1 + (double)y[i], 255); s2 = mmod(s2 + s1, 255); }
    double r = s2 * 256 + s1;
    return isnan(r) ? r : trunc(r);
}
EXPORT double oracle_fletcher_f64(int n, const double *y)
{
    double s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) { s1 = mmod(s1 + y[i], 255); s2 = mmod(s2 + s1, 255); }
    double r = s2 * 256 + s1;
    return isnan(r) ? r : trunc(r);
}

/* CPU baseline timing of the serial restatement (seq/tests/main32.js:122-134
 * method: warm-up calls, then timed calls; y is cleared before each call).
 * Returns the mean seconds per spmv_csr call over `reps` timed calls. */
#include <time.h>
static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
EXPORT double oracle_time_spmv_csr_f64(int nrows, const int *rowptr, const int *col,
                                       const double *val, const double *x, double *y,
                                       int warm, int reps)
{
    for (int w = 0; w < warm; w++) {
        memset(y, 0, sizeof(double) * (size_t)nrows);
        oracle_spmv_csr_f64(nrows, rowptr, col, val, x, y, 1);
    }
    double t = 0;
    for (int r = 0; r < reps; r++) {
        memset(y, 0, sizeof(double) * (size_t)nrows);
        double t0 = now_s();
        oracle_spmv_csr_f64(nrows, rowptr, col, val, x, y, 1);
        t += now_s() - t0;
    }
    return reps > 0 ? t / reps : 0.0;
}

/* Threaded CPU baseline: rows split into W contiguous ranges of about equal
 * nnz by the greedy rule of static_nnz (par/src/sparse32.js:2769-2821: a
 * worker takes rows while its running sum is below remaining_nnz /
 * remaining_workers), one persistent thread per range, a barrier before and
 * after each timed product (the reference's workers are persistent too and
 * report back once per product). */
#include <pthread.h>
typedef struct {
    int r0, r1, nrows;
    const int *rowptr, *col;
    const double *val, *x;
    double *y;
    int runs;
    pthread_barrier_t *bar;
} mt_job;

static void *mt_worker(void *arg)
{
    mt_job *j = (mt_job *)arg;
    for (int r = 0; r < j->runs; r++) {
        pthread_barrier_wait(j->bar);
        for (int i = j->r0; i < j->r1; i++) j->y[i] = 0.0;
        if (j->r1 > j->r0)
            oracle_spmv_csr_f64(j->r1 - j->r0, j->rowptr + j->r0, j->col, j->val, j->x, j->y + j->r0, 1);
        pthread_barrier_wait(j->bar);
    }
    return NULL;
}

EXPORT void oracle_static_nnz(int nrows, const int *rowptr, int workers, int *row_start, int *row_end)
{
    long long rem = (long long)rowptr[nrows] - rowptr[0];
    int rem_w = workers, next = 0;
    for (int w = 0; w < workers; w++) {
        if (next == nrows) { row_start[w] = row_end[w] = nrows; continue; }
        long long ideal = rem / rem_w, got = 0;
        row_start[w] = next;
        while (next < nrows && got < ideal) { got += rowptr[next + 1] - rowptr[next]; next++; }
        row_end[w] = next;
        rem -= got;
        rem_w--;
    }
    row_end[workers - 1] = nrows;
}

EXPORT double oracle_time_spmv_csr_f64_mt(int nrows, const int *rowptr, const int *col,
                                          const double *val, const double *x, double *y,
                                          int workers, int warm, int reps)
{
    if (workers < 1) workers = 1;
    int *rs = (int *)malloc(sizeof(int) * workers), *re = (int *)malloc(sizeof(int) * workers);
    mt_job *jobs = (mt_job *)malloc(sizeof(mt_job) * workers);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * workers);
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)workers + 1);
    oracle_static_nnz(nrows, rowptr, workers, rs, re);
    for (int w = 0; w < workers; w++) {
        mt_job j = {rs[w], re[w], nrows, rowptr, col, val, x, y, warm + reps, &bar};
        jobs[w] = j;
        pthread_create(&th[w], NULL, mt_worker, &jobs[w]);
    }
    double t = 0;
    for (int r = 0; r < warm + reps; r++) {
        double t0 = now_s();
        pthread_barrier_wait(&bar);
        pthread_barrier_wait(&bar);
        if (r >= warm) t += now_s() - t0;
    }
    for (int w = 0; w < workers; w++) pthread_join(th[w], NULL);
    pthread_barrier_destroy(&bar);
    free(rs); free(re); free(jobs); free(th);
    return reps > 0 ? t / reps : 0.0;
}

/* ================================================================ MatrixMarket
 * seq/src/sparse32.js:1889-2172 (FP64: seq/src/sparse64.js:1536-1760), as
 * the reference runs it: the text is cut into '\n'-terminated lines (the
 * loader keeps an unterminated last line back, :2470-2485), every line is
 * split on single spaces (String.split(" ")), and each token goes through
 * JavaScript's Number() and, for the index arrays, the Int32Array store
 * (ToInt32).  Values are held in a Float64Array, then stored to the COO's
 * Float32Array (32-bit tree) or Float64Array (64-bit tree).  The COO is
 * sorted by quick_sort_COO (:1577-1609) exactly as written; its arrays start
 * zeroed, as freshly grown WebAssembly memory does. */

typedef struct {
    const char* p;
    size_t n;
} mm_str;

/* Number(s) for the decimal literals this path meets (StringNumericLiteral
 * without the hexadecimal/binary/octal forms): whitespace trimmed, empty ->
 * 0, [+-]Infinity, otherwise strtod when the whole token is a decimal
 * literal, NaN when it is not. */
static double js_number(mm_str s) {
    size_t a = 0, b = s.n;
    while (a < b && (s.p[a] == ' ' || s.p[a] == '\t' || s.p[a] == '\r' || s.p[a] == '\n' || s.p[a] == '\v' ||
                     s.p[a] == '\f'))
        ++a;
    while (b > a && (s.p[b - 1] == ' ' || s.p[b - 1] == '\t' || s.p[b - 1] == '\r' || s.p[b - 1] == '\n' ||
                     s.p[b - 1] == '\v' || s.p[b - 1] == '\f'))
        --b;
    if (a == b) return 0.0;
    char buf[4096];
    size_t n = b - a;
    if (n >= sizeof(buf)) return NAN;
    memcpy(buf, s.p + a, n);
    buf[n] = 0;
    const char* q = buf;
    int neg = 0;
    if (*q == '+' || *q == '-') {
        neg = *q == '-';
        ++q;
    }
    if (strcmp(q, "Infinity") == 0) return neg ? -INFINITY : INFINITY;
    /* digits [. digits] | . digits, then [eE][+-]digits */
    const char* r = q;
    int dig = 0;
    while (*r >= '0' && *r <= '9') ++r, ++dig;
    if (*r == '.') {
        ++r;
        while (*r >= '0' && *r <= '9') ++r, ++dig;
    }
    if (!dig) return NAN;
    if (*r == 'e' || *r == 'E') {
        ++r;
        if (*r == '+' || *r == '-') ++r;
        if (!(*r >= '0' && *r <= '9')) return NAN;
        while (*r >= '0' && *r <= '9') ++r;
    }
    if (*r) return NAN;
    return strtod(buf, NULL);
}

/* ECMAScript ToInt32 (the Int32Array store). */
static int32_t js_to_int32(double d) {
    if (!isfinite(d)) return 0;
    d = trunc(d);
    double m = fmod(d, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    uint32_t u = (uint32_t)m;
    return (int32_t)u;
}

/* The k-th piece of line.split(" "); .p == NULL when there are fewer. */
static mm_str js_split_space(mm_str line, int k) {
    size_t a = 0;
    for (int i = 0; i < k; ++i) {
        const char* sp = memchr(line.p + a, ' ', line.n - a);
        if (!sp) {
            mm_str none = {NULL, 0};
            return none;
        }
        a = (size_t)(sp - line.p) + 1;
    }
    const char* sp = memchr(line.p + a, ' ', line.n - a);
    mm_str r = {line.p + a, sp ? (size_t)(sp - line.p) - a : line.n - a};
    return r;
}

static int mm_streq(mm_str s, const char* lit) {
    size_t n = strlen(lit);
    return s.p && s.n == n && memcmp(s.p, lit, n) == 0;
}

typedef struct {
    int pattern, symmetric;
    int64_t nrows, ncols, nentries, nnz;
    int32_t* row; /* mm_info.row/col/val: nentries each */
    int32_t* col;
    double* val;
} mm_info_t;

static void mm_info_free(mm_info_t* m) {
    free(m->row);
    free(m->col);
    free(m->val);
}

/* read_MM_header + read_matrix_MM_files + calculate_actual_nnz
 * (:1889-1981).  Returns 0, or -1 when the reference throws. */
static int oracle_mm_read_info(const char* text, size_t len, mm_info_t* m) {
    memset(m, 0, sizeof(*m));
    /* complete lines only */
    size_t nl = 0;
    for (size_t i = 0; i < len; ++i) nl += text[i] == '\n';
    mm_str* lines = (mm_str*)malloc((nl + 1) * sizeof(mm_str));
    size_t a = 0, k = 0;
    for (size_t i = 0; i < len; ++i)
        if (text[i] == '\n') {
            lines[k].p = text + a;
            lines[k].n = i - a;
            ++k;
            a = i + 1;
        }
    if (k == 0) {
        free(lines);
        return -1;
    }
    mm_str first = lines[0];
    m->pattern = mm_streq(js_split_space(first, 3), "pattern");
    m->symmetric = mm_streq(js_split_space(first, 4), "symmetric");
    size_t n = 0;
    while (n < k && lines[n].n > 0 && lines[n].p[0] == '%') ++n;
    if (n >= k) {
        free(lines);
        return -1;
    }
    mm_str info = lines[n++];
    m->nrows = (int64_t)js_number(js_split_space(info, 0));
    m->ncols = (int64_t)js_number(js_split_space(info, 1));
    m->nentries = (int64_t)js_number(js_split_space(info, 2));
    if (m->nentries > (1LL << 27) || m->nentries < 0) {
        free(lines);
        return -1;
    }
    const int64_t ne = m->nentries;
    m->row = (int32_t*)calloc((size_t)(ne ? ne : 1), 4);
    m->col = (int32_t*)calloc((size_t)(ne ? ne : 1), 4);
    m->val = (double*)calloc((size_t)(ne ? ne : 1), 8);
    int64_t nnz = 0;
    for (int64_t j = 0; n < k; ++n, ++j) {
        mm_str c0 = js_split_space(lines[n], 0), c1 = js_split_space(lines[n], 1), c2 = js_split_space(lines[n], 2);
        /* Number(undefined) is NaN */
        const double r = c0.p ? js_number(c0) : NAN, c = c1.p ? js_number(c1) : NAN;
        const int in = j < ne; /* typed-array stores past the end are dropped */
        if (in) {
            m->row[j] = js_to_int32(r);
            m->col[j] = js_to_int32(c);
        }
        if (m->symmetric) {
            if (!m->pattern) {
                if (in) m->val[j] = c2.p ? js_number(c2) : NAN;
                const double v = in ? m->val[j] : NAN; /* undefined compares false */
                if (v < 0 || v > 0) nnz += (in && m->row[j] == m->col[j]) ? 1 : 2;
            } else {
                /* undefined == undefined past the end */
                nnz += (!in || m->row[j] == m->col[j]) ? 1 : 2;
            }
        } else if (!m->pattern) {
            if (in) m->val[j] = c2.p ? js_number(c2) : NAN;
            const double v = in ? m->val[j] : NAN;
            if (v < 0 || v > 0) ++nnz;
        }
    }
    free(lines);
    if (nnz == 0) nnz = ne;
    if (nnz > (1LL << 28)) {
        mm_info_free(m);
        return -1;
    }
    m->nnz = nnz;
    return 0;
}

#define DEF_MM_COO(T, SUF)                                                                                        \
    /* quick_sort_COO (:1577-1609) */                                                                              \
    static void mm_qsort_##SUF(int32_t* row, int32_t* col, T* val, int64_t left, int64_t right) {                   \
        int64_t i = left, j = right;                                                                               \
        const int64_t mid = (left + right) / 2;                                                                    \
        const int32_t pivot = row[mid], pivot_col = col[mid];                                                      \
        while (i <= j) {                                                                                           \
            while (row[i] < pivot || (row[i] == pivot && col[i] < pivot_col)) ++i;                                 \
            while (row[j] > pivot || (row[j] == pivot && col[j] > pivot_col)) --j;                                 \
            if (i <= j) {                                                                                          \
                int32_t tr = row[i]; row[i] = row[j]; row[j] = tr;                                                 \
                int32_t tc = col[i]; col[i] = col[j]; col[j] = tc;                                                 \
                T tv = val[i]; val[i] = val[j]; val[j] = tv;                                                       \
                ++i;                                                                                               \
                --j;                                                                                               \
            }                                                                                                      \
        }                                                                                                          \
        if (left < j) mm_qsort_##SUF(row, col, val, left, j);                                                      \
        if (i < right) mm_qsort_##SUF(row, col, val, i, right);                                                    \
    }                                                                                                              \
    /* create_COO_from_MM (:2100-2172) or, lower = 1, create_LCOO_from_MM (:1983-2097) into arrays of *nnz      \
     * entries (call with row == NULL for the size).  Returns 0 or -1. */                                          \
    EXPORT int oracle_mm_coo_##SUF(const char* text, size_t len, int lower, int64_t* info6, int32_t* orow,          \
                                   int32_t* ocol, T* oval) {                                                       \
        mm_info_t m;                                                                                               \
        if (oracle_mm_read_info(text, len, &m)) return -1;                                                          \
        const int64_t ne = m.nentries;                                                                             \
        const int32_t* row = m.row;                                                                                \
        const int32_t* col = m.col;                                                                                \
        const double* val = m.val;                                                                                 \
        int64_t nnz = m.nnz;                                                                                       \
        if (lower) {                                                                                               \
            nnz = m.nrows;                                                                                         \
            for (int64_t n = 0; n < ne; ++n) {                                                                     \
                if (m.symmetric) {                                                                                 \
                    if (row[n] != col[n] && (m.pattern || val[n] > 0 || val[n] < 0)) ++nnz;                         \
                } else if (row[n] > col[n] && (m.pattern || val[n] > 0 || val[n] < 0)) {                            \
                    ++nnz;                                                                                         \
                }                                                                                                  \
            }                                                                                                      \
        }                                                                                                          \
        info6[0] = m.nrows; info6[1] = m.ncols; info6[2] = ne; info6[3] = nnz;                                     \
        info6[4] = m.pattern; info6[5] = m.symmetric;                                                              \
        if (!orow) {                                                                                               \
            mm_info_free(&m);                                                                                      \
            return 0;                                                                                              \
        }                                                                                                          \
        memset(orow, 0, (size_t)nnz * 4);                                                                          \
        memset(ocol, 0, (size_t)nnz * 4);                                                                          \
        memset(oval, 0, (size_t)nnz * sizeof(T));                                                                  \
        int64_t i = 0;                                                                                             \
        /* stores at i >= nnz are dropped like typed-array stores past the end */                                 \
        if (lower) {                                                                                               \
            for (int64_t n = 0; n < ne; ++n) {                                                                     \
                const int kept = m.pattern || val[n] < 0 || val[n] > 0;                                            \
                const T v = m.pattern ? (T)1.0 : (T)val[n];                                                        \
                if (!kept) continue;                                                                               \
                if (row[n] > col[n]) {                                                                             \
                    if (i < nnz) { orow[i] = row[n] - 1; ocol[i] = col[n] - 1; oval[i] = v; }                      \
                    ++i;                                                                                           \
                } else if (m.symmetric && row[n] < col[n]) {                                                       \
                    if (i < nnz) { orow[i] = col[n] - 1; ocol[i] = row[n] - 1; oval[i] = v; }                      \
                    ++i;                                                                                           \
                }                                                                                                  \
            }                                                                                                      \
            for (int64_t n = 0; n < m.nrows; ++n) {                                                                \
                if (i < nnz) { orow[i] = (int32_t)n; ocol[i] = (int32_t)n; oval[i] = (T)1.0; }                     \
                ++i;                                                                                               \
            }                                                                                                      \
        } else {                                                                                                   \
            for (int64_t n = 0; n < ne; ++n) {                                                                     \
                if (!m.pattern && !(val[n] < 0 || val[n] > 0)) continue;                                           \
                const T v = m.pattern ? (T)1.0 : (T)val[n];                                                        \
                if (i < nnz) { orow[i] = row[n] - 1; ocol[i] = col[n] - 1; oval[i] = v; }                          \
                if (m.symmetric && row[n] != col[n]) {                                                             \
                    if (i + 1 < nnz) { orow[i + 1] = col[n] - 1; ocol[i + 1] = row[n] - 1; oval[i + 1] = v; }      \
                    i += 2;                                                                                        \
                } else {                                                                                           \
                    ++i;                                                                                           \
                }                                                                                                  \
            }                                                                                                      \
        }                                                                                                          \
        if (nnz > 0) mm_qsort_##SUF(orow, ocol, oval, 0, nnz - 1);                                                 \
        mm_info_free(&m);                                                                                          \
        return 0;                                                                                                  \
    }

DEF_MM_COO(float, f32)
DEF_MM_COO(double, f64)

/* ================================================================ element-wise
 * seq/src/sparse_opt_32.wat / sparse_opt_64.wat exports behind
 * seq/src/sparse32.js:200-1207.  Op codes follow include/sable_sparse.h
 * (sable_ew_op_t).  Transcendental ops: JavaScript Math in double on the
 * widened value, rounded back (the module imports Math.expm1 ... Math.pow,
 * seq/src/sparse32.js:8); Math.pow differs from C pow for a NaN exponent
 * and for |x| == 1 with an infinite exponent (both NaN). */
static double js_pow(double x, double y) {
    if (isnan(y)) return NAN;
    if (fabs(x) == 1.0 && isinf(y)) return NAN;
    return pow(x, y);
}

#define DEF_EW(T, SUF)                                                                                            \
    static T ew_##SUF(int op, T param, T v) {                                                                     \
        switch (op) {                                                                                             \
            case 0: return (T)expm1((double)v);                                                                   \
            case 1: return (T)log1p((double)v);                                                                   \
            case 2: return (T)sin((double)v);                                                                     \
            case 3: return (T)tan((double)v);                                                                     \
            case 4: return (T)js_pow((double)v, (double)param);                                                   \
            case 5: return v * param;                 /* deg2rad: param = pi/180 in T */                         \
            case 6: return v / param;                 /* rad2deg */                                               \
            case 7: return v > 0 ? (T)1 : (T)-1;      /* sign, COO/CSR exports (:1137-1183) */                   \
            case 8: return v == 0 ? (T)0 : (v > 0 ? (T)1 : (T)-1); /* sign, DIA/ELL exports (:2879-2960) */      \
            case 9: return (T)fabs((double)v);                                                                    \
            case 10: return -v;                                                                                   \
            case 11: return (sizeof(T) == 4) ? (T)sqrtf((float)v) : (T)sqrt((double)v);                           \
            case 12: return (sizeof(T) == 4) ? (T)ceilf((float)v) : (T)ceil((double)v);                           \
            case 13: return (sizeof(T) == 4) ? (T)floorf((float)v) : (T)floor((double)v);                         \
            case 14: return (sizeof(T) == 4) ? (T)truncf((float)v) : (T)trunc((double)v);                         \
            case 15: return (sizeof(T) == 4) ? (T)rintf((float)v) : (T)rint((double)v); /* f32.nearest */        \
            case 16: return v * param;                /* self_multiply_scalar_coo (:788-830) */                  \
            default: return v;                                                                                    \
        }                                                                                                         \
    }                                                                                                             \
    static T ew_param_##SUF(int op, double param) {                                                               \
        /* pi / 180 rounded in T (f32.div / f64.div at :683-700) */                                               \
        if (op == 5 || op == 6) return (T)((T)param / (T)180);                                                    \
        return (T)param;                                                                                          \
    }                                                                                                             \
    /* <op>_coo, self_<op>_coo, <op>_ell, self_<op>_ell: every value */                                           \
    EXPORT void oracle_ew_values_##SUF(int op, double param, int64_t n, const T* in, T* out) {                    \
        const T p = ew_param_##SUF(op, param);                                                                    \
        for (int64_t i = 0; i < n; ++i) out[i] = ew_##SUF(op, p, in[i]);                                          \
    }                                                                                                             \
    /* <op>_dia, self_<op>_dia: rows max(0,-k) .. min(N, N-k) of each diagonal (:2137-2207) */                    \
    EXPORT void oracle_ew_dia_##SUF(int op, double param, int N, int nd, int stride, const int32_t* offset,         \
                                    const T* in, T* out) {                                                        \
        const T p = ew_param_##SUF(op, param);                                                                    \
        for (int d = 0; d < nd; ++d) {                                                                            \
            const int k = offset[d];                                                                              \
            const int64_t base = (int64_t)d * stride - (k < 0 ? (int64_t)(N - stride) : 0);                       \
            const int lo = k < 0 ? -k : 0, hi = k > 0 ? N - k : N;                                                \
            for (int r = lo; r < hi; ++r) out[base + r] = ew_##SUF(op, p, in[base + r]);                          \
        }                                                                                                         \
    }                                                                                                             \
    /* count_nonzeros_coo (:500-521): v > 0 || v < 0 */                                                            \
    EXPORT int64_t oracle_count_nonzeros_##SUF(int64_t n, const T* v) {                                           \
        int64_t c = 0;                                                                                            \
        for (int64_t i = 0; i < n; ++i) c += (v[i] > 0 || v[i] < 0);                                              \
        return c;                                                                                                 \
    }                                                                                                             \
    /* copy_nonzeros_coo (:523-546) */                                                                             \
    EXPORT void oracle_copy_nonzeros_coo_##SUF(int64_t n, const int32_t* row, const int32_t* col, const T* val,     \
                                               int32_t* orow, int32_t* ocol, T* oval) {                           \
        int64_t o = 0;                                                                                            \
        for (int64_t i = 0; i < n; ++i)                                                                           \
            if (val[i] > 0 || val[i] < 0) {                                                                       \
                orow[o] = row[i];                                                                                 \
                ocol[o] = col[i];                                                                                 \
                oval[o] = val[i];                                                                                 \
                ++o;                                                                                              \
            }                                                                                                     \
    }                                                                                                             \
    /* copy_nonzeros_csr (:548-595), rows 0..N-1 */                                                                \
    EXPORT void oracle_copy_nonzeros_csr_##SUF(int N, const int32_t* rp, const int32_t* col, const T* val,        \
                                               int32_t* orp, int32_t* ocol, T* oval) {                            \
        int64_t o = 0;                                                                                            \
        orp[0] = 0;                                                                                               \
        for (int i = 0; i < N; ++i) {                                                                             \
            for (int j = rp[i]; j < rp[i + 1]; ++j)                                                               \
                if (val[j] > 0 || val[j] < 0) {                                                                   \
                    ocol[o] = col[j];                                                                             \
                    oval[o] = val[j];                                                                             \
                    ++o;                                                                                          \
                }                                                                                                 \
            orp[i + 1] = (int32_t)o;                                                                              \
        }                                                                                                         \
    }                                                                                                             \
    /* count_nonzeros_dia (:1948-2039) -> out2 = {count_diags, count} */                                          \
    EXPORT void oracle_count_nonzeros_dia_##SUF(int N, int nd, int stride, const int32_t* offset, const T* data,   \
                                                int64_t* out2) {                                                  \
        int64_t cd = 0, nnz = 0;                                                                                  \
        for (int d = 0; d < nd; ++d) {                                                                            \
            const int k = offset[d];                                                                              \
            const int64_t base = (int64_t)d * stride - (k < 0 ? (int64_t)(N - stride) : 0);                       \
            const int lo = k < 0 ? -k : 0, hi = k > 0 ? N - k : N;                                                \
            int64_t c = 0;                                                                                        \
            for (int r = lo; r < hi; ++r) c += (data[base + r] > 0 || data[base + r] < 0);                        \
            if (c) {                                                                                              \
                nnz += c;                                                                                         \
                ++cd;                                                                                             \
            }                                                                                                     \
        }                                                                                                         \
        out2[0] = cd;                                                                                             \
        out2[1] = nnz;                                                                                            \
    }                                                                                                             \
    /* copy_nonzeros_dia (:2041-2135): the j-th kept diagonal's nonzeros, in row order, from slot               \
     * j*stride + (first row) onwards (for stride < N, k < 0, the slot of the first row in that layout);        \
     * out_data zero-filled first as eliminate_zeros does (seq/src/sparse32.js:1176-1179) */                     \
    EXPORT void oracle_copy_nonzeros_dia_##SUF(int N, int nd, int stride, const int32_t* offset, const T* data,    \
                                               int32_t* ooff, T* odata, int64_t out_slots) {                      \
        for (int64_t i = 0; i < out_slots; ++i) odata[i] = 0;                                                     \
        int64_t j = 0;                                                                                            \
        for (int d = 0; d < nd; ++d) {                                                                            \
            const int k = offset[d];                                                                              \
            const int64_t shift = (k < 0 ? (int64_t)(N - stride) : 0);                                            \
            const int64_t base = (int64_t)d * stride - shift;                                                     \
            const int lo = k < 0 ? -k : 0, hi = k > 0 ? N - k : N;                                                \
            int64_t o = j * stride - shift + lo, c = 0;                                                           \
            for (int r = lo; r < hi; ++r)                                                                         \
                if (data[base + r] > 0 || data[base + r] < 0) {                                                   \
                    odata[o++] = data[base + r];                                                                  \
                    ++c;                                                                                          \
                }                                                                                                 \
            if (c) ooff[j++] = k;                                                                                 \
        }                                                                                                         \
    }                                                                                                             \
    /* count_nonzeros_ell (:5821-5893) -> out2 = {count_cols, count} */                                           \
    EXPORT void oracle_count_nonzeros_ell_##SUF(int N, int nc, const T* data, int64_t* out2) {                    \
        int64_t cc = 0, nnz = 0;                                                                                  \
        for (int i = 0; i < N; ++i) {                                                                             \
            int64_t c = 0;                                                                                        \
            for (int k = 0; k < nc; ++k) c += (data[(int64_t)k * N + i] > 0 || data[(int64_t)k * N + i] < 0);     \
            if (c > cc) cc = c;                                                                                   \
            nnz += c;                                                                                             \
        }                                                                                                         \
        out2[0] = cc;                                                                                             \
        out2[1] = nnz;                                                                                            \
    }                                                                                                             \
    /* copy_nonzeros_ell (:5895-5949) into zero-filled nc_out-wide arrays (seq/src/sparse32.js:1196-1203) */      \
    EXPORT void oracle_copy_nonzeros_ell_##SUF(int N, int nc, int nc_out, const T* data, const int32_t* idx,       \
                                               T* odata, int32_t* oidx) {                                         \
        for (int64_t i = 0; i < (int64_t)nc_out * N; ++i) {                                                       \
            odata[i] = 0;                                                                                         \
            oidx[i] = 0;                                                                                          \
        }                                                                                                         \
        for (int i = 0; i < N; ++i) {                                                                             \
            int o = 0;                                                                                            \
            for (int k = 0; k < nc; ++k) {                                                                        \
                const T v = data[(int64_t)k * N + i];                                                             \
                if (v > 0 || v < 0) {                                                                             \
                    odata[(int64_t)o * N + i] = v;                                                                \
                    oidx[(int64_t)o * N + i] = idx[(int64_t)k * N + i];                                           \
                    ++o;                                                                                          \
                }                                                                                                 \
            }                                                                                                     \
        }                                                                                                         \
    }

DEF_EW(float, f32)
DEF_EW(double, f64)

/* ================================================================ row reordering
 * sort_rows_by_nnz (par/src/sparse32.js:2148-2218): counting sort of the
 * rows by nnz, stable; the new row_ptr is the one its loop computes (the
 * reference stores it through a zero-length view).  bins4 = cumulative
 * counts of rows with 0, <=1, <=2, <=3 entries. */
#define DEF_SORT_ROWS(T, SUF)                                                                                      \
    EXPORT void oracle_sort_rows_by_nnz_##SUF(int N, const int32_t* rp, const int32_t* col, const T* val,          \
                                              int32_t* orp, int32_t* ocol, T* oval, int32_t* perm,               \
                                              int32_t* nnz_row, int32_t* bins4) {                                 \
        /* frequencies up to the longest row (the reference sizes them N + 1, which drops rows longer  \
         * than N, possible only with duplicate entries, and reads undefined bins when N < 3) */      \
        int32_t maxl = 3;                                                                                         \
        for (int i = 0; i < N; ++i) {                                                                             \
            nnz_row[i] = rp[i + 1] - rp[i];                                                                       \
            if (nnz_row[i] > maxl) maxl = nnz_row[i];                                                             \
        }                                                                                                         \
        int32_t* freq = (int32_t*)calloc((size_t)maxl + 1, 4);                                                    \
        int32_t* start = (int32_t*)calloc((size_t)maxl + 1, 4);                                                   \
        for (int i = 0; i < N; ++i) freq[nnz_row[i]]++;                                                           \
        start[0] = 0;                                                                                             \
        for (int i = 1; i <= maxl; ++i) start[i] = start[i - 1] + freq[i - 1];                                    \
        bins4[0] = freq[0];                                                                                       \
        bins4[1] = bins4[0] + freq[1];                                                                            \
        bins4[2] = bins4[1] + freq[2];                                                                            \
        bins4[3] = bins4[2] + freq[3];                                                                            \
        for (int i = 0; i < N; ++i) perm[start[nnz_row[i]]++] = i;                                                \
        int64_t j = 0;                                                                                            \
        orp[0] = 0;                                                                                               \
        for (int i = 0; i < N; ++i) {                                                                             \
            int64_t k = rp[perm[i]];                                                                              \
            int t = nnz_row[perm[i]];                                                                             \
            orp[i + 1] = orp[i] + t;                                                                              \
            while (t != 0) {                                                                                      \
                ocol[j] = col[k];                                                                                 \
                oval[j++] = val[k++];                                                                             \
                t--;                                                                                              \
            }                                                                                                     \
        }                                                                                                         \
        free(freq);                                                                                               \
        free(start);                                                                                              \
    }                                                                                                             \
    /* sort_y_rows_by_nnz (:2220-2235) */                                                                          \
    EXPORT void oracle_sort_y_rows_##SUF(int N, const int32_t* perm, const T* y, T* out) {                        \
        for (int i = 0; i < N; ++i) out[perm[i]] = y[i];                                                          \
    }

DEF_SORT_ROWS(float, f32)
DEF_SORT_ROWS(double, f64)

/* reorder_NN (par/src/sparse32.js:2501-2622) with reorder_subset
 * (:2439-2492), find_next_row (:2330-2419) and assign_w_min_elems
 * (:2275-2327), over a CSR whose rows are sorted by nnz.  out_perm[i] = the
 * row of the input placed at new row i (the caller composes it with the
 * sort's permutation).  Similarities are computed in double and kept in
 * float32 lists, compared against the double value, as the reference's
 * Float32Array lists are. */
static void nn_assign(float* jac, int32_t* nid, int32_t next, double max, int w) {
    int i = 0;
    while (i < w && jac[i] >= 0 && (double)jac[i] >= max) i++;
    if (i == w) return;
    if (jac[i] == -1) {
        jac[i] = (float)max;
        nid[i] = next;
        return;
    }
    int j = w - 1;
    while (jac[j] == -1) j--;
    while (j != i) {
        jac[j] = jac[j - 1];
        nid[j] = nid[j - 1];
        j--;
    }
    jac[i] = (float)max;
    nid[i] = next;
}

/* last_used has N/16 + 1 lines; a typed-array store past its end is dropped
 * and a load past it gives undefined, which equals no row */
static void nn_find(int index, int8_t* visited, const int32_t* nlines, float* jac, int32_t* nid, int32_t* last_used,
                    const int32_t* rp, const int32_t* col, int N, int w, int start, int end, int32_t* next_out,
                    double* max_out) {
    const int L = N / 16 + 1;
    if (visited[index] == 3) {
        visited[index] = 1;
        for (int k = 0; k < w; ++k) {
            jac[k] = -1;
            nid[k] = N;
        }
        for (int j = rp[index]; j < rp[index + 1]; ++j)
            if (col[j] >= 0 && col[j] / 16 < L && last_used[col[j] / 16] != index) last_used[col[j] / 16] = index;
        for (int i = start; i < end; ++i) {
            if (visited[i] != 0) continue;
            int j = rp[i], alines = 0;
            while (j < rp[i + 1]) {
                if (col[j] >= 0 && col[j] / 16 < L && last_used[col[j] / 16] == index) alines++;
                j++;
                while (j < rp[i + 1] && col[j - 1] / 16 == col[j] / 16) j++;
            }
            const int dissim = nlines[index] + nlines[i] - alines;
            nn_assign(jac, nid, i, (double)alines / (double)dissim, w);
        }
    }
    int32_t next = N;
    double max = 0;
    for (int k = 0; k < w; ++k)
        if (jac[k] >= 0 && nid[k] < N && visited[nid[k]] == 0) {
            next = nid[k];
            max = jac[k];
            break;
        }
    *next_out = next;
    *max_out = max;
}

EXPORT void oracle_reorder_nn(int N, const int32_t* rp, const int32_t* col, int w, int32_t* perm) {
    int8_t* visited = (int8_t*)calloc((size_t)N + 1, 1);
    int32_t* nlines = (int32_t*)calloc((size_t)N + 1, 4);
    int32_t maxl = 0;
    for (int i = 0; i < N; ++i)
        if (rp[i + 1] - rp[i] > maxl) maxl = rp[i + 1] - rp[i];
    int32_t* freq = (int32_t*)calloc((size_t)maxl + 1, 4);
    for (int i = 0; i < N; ++i) {
        perm[i] = i;
        freq[rp[i + 1] - rp[i]]++;
        if (rp[i + 1] == rp[i]) visited[i] = 2;
    }
    for (int i = 0; i < N; ++i) {
        if (visited[i]) continue;
        int j = rp[i];
        while (j < rp[i + 1]) {
            nlines[i]++;
            j++;
            while (j < rp[i + 1] && col[j - 1] / 16 == col[j] / 16) j++;
        }
    }
    int index = 0;
    while (index < N && visited[index] != 0) index++;
    const int B = 1024;
    const int L = N / 16 + 1;
    int32_t* last_used = (int32_t*)malloc((size_t)L * 4);
    float* jac = (float*)malloc((size_t)w * w * sizeof(float));
    int32_t* nid = (int32_t*)malloc((size_t)w * w * 4);
    int32_t* win = (int32_t*)malloc((size_t)w * 4);
    int start = index;
    for (int c = 1; c <= maxl; ++c) {
        while (freq[c] > 0) {
            const int take = freq[c] >= B ? B : freq[c];
            const int end = start + take;
            freq[c] -= take;
            /* reorder_subset(start, end) */
            for (int k = 0; k < L; ++k) last_used[k] = -1;
            for (int k = 0; k < w * w; ++k) {
                jac[k] = 0;
                nid[k] = 0;
            }
            int idx = start, pos = start, nnb = 0, count = 0;
            while (idx != N) {
                visited[idx] = 3;
                win[count++] = idx;
                if (nnb < w) nnb++;
                if (count == w) count = 0;
                perm[pos++] = idx;
                int32_t next = N;
                double max = 0;
                for (int s = 0; s < nnb; ++s) {
                    int32_t nx;
                    double mx;
                    nn_find(win[s], visited, nlines, jac + (size_t)s * w, nid + (size_t)s * w, last_used, rp, col, N,
                            w, start, end, &nx, &mx);
                    if (max <= mx && nx != N) {
                        max = mx;
                        next = nx;
                    }
                }
                idx = next;
            }
            start = end;
        }
    }
    free(visited);
    free(nlines);
    free(freq);
    free(last_used);
    free(jac);
    free(nid);
    free(win);
}
