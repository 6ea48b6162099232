// reorder.hip -- row reordering of a CSR matrix on the device.
//
// Reference: par/src/sparse32.js (and sparse64.js:1297-1380)
//   sort_rows_by_nnz     :2148-2218  rows in ascending nnz order, stable
//                                    (a counting sort), the permutation
//                                    (new row i = old row perm[i]), nnz per
//                                    old row, and the cumulative counts of
//                                    rows with 0, <=1, <=2, <=3 entries that
//                                    select the short-row SpMV variants
//   sort_y_rows_by_nnz   :2220-2235  y_new[perm[i]] = y[i]
//
// The reference writes the new row_ptr through a view sized by the
// undefined A_csr_new.nrows, i.e. a zero-length view, so its row_ptr is
// never stored; the row_ptr here is the one its loop computes.
#include "common.h"
#include "prims.h"

#include <string.h>

#include <algorithm>

using namespace sable;

namespace {

constexpr int RB = 256;

unsigned gs(int64_t n) { return grid_for(n, RB); }

// nnz of every row; keys = nnz (for the stable sort), payload = row.
__global__ void k_row_lengths(int32_t N, const int32_t* __restrict__ rp, int32_t* __restrict__ len,
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ rows, int* __restrict__ bins) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    if (i < N) {
        const int32_t l = rp[i + 1] - rp[i];
        len[i] = l;
        keys[i] = (uint64_t)(uint32_t)l;
        rows[i] = (uint32_t)i;
        b0 = l == 0;
        b1 = l == 1;
        b2 = l == 2;
        b3 = l == 3;
    }
    b0 = wave_sum(b0);
    b1 = wave_sum(b1);
    b2 = wave_sum(b2);
    b3 = wave_sum(b3);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        if (b0) atomicAdd(&bins[0], b0);
        if (b1) atomicAdd(&bins[1], b1);
        if (b2) atomicAdd(&bins[2], b2);
        if (b3) atomicAdd(&bins[3], b3);
    }
}

__global__ void k_perm_out(int32_t N, const uint32_t* __restrict__ rows, const int32_t* __restrict__ len,
                           int32_t* __restrict__ perm, int32_t* __restrict__ newlen) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int32_t r = (int32_t)rows[i];
    perm[i] = r;
    newlen[i] = len[r];
}

// One wave per new row: copies old row perm[i]'s entries.
template <typename T>
__global__ void k_copy_rows(int32_t N, const int32_t* __restrict__ perm, const int32_t* __restrict__ rp,
                            const int32_t* __restrict__ col, const T* __restrict__ val,
                            const int32_t* __restrict__ nrp, int32_t* __restrict__ ocol, T* __restrict__ oval) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) / WAVE;
    const int64_t base = rp[0];
    for (int64_t i = w0; i < N; i += nw) {
        const int32_t r = perm[i];
        const int64_t s = rp[r] - base, n = rp[r + 1] - rp[r], d = nrp[i];
        for (int64_t k = lane; k < n; k += WAVE) {
            ocol[d + k] = col[s + k];
            oval[d + k] = val[s + k];
        }
    }
}

template <typename T>
__global__ void k_scatter_rows(int32_t N, const int32_t* __restrict__ perm, const T* __restrict__ y,
                               T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[perm[i]] = y[i];
}

template <typename T>
__global__ void k_gather_rows(int32_t N, const int32_t* __restrict__ perm, const T* __restrict__ x,
                              T* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out[i] = x[perm[i]];
}

template <typename T>
int sort_rows_t(sable_ctx_t* ctx, int32_t N, const int32_t* rp, const int32_t* col, const void* val, int32_t* orp,
                int32_t* ocol, void* oval, int32_t* perm, int32_t* nnz_row, int32_t* bins_out) {
    hipStream_t st = ctx->stream;
    int32_t ends[2];
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, rp, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipMemcpyAsync((char*)ctx->pinned + 4, rp + N, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    memcpy(ends, ctx->pinned, 8);
    const int64_t nnz = (int64_t)ends[1] - ends[0];
    SABLE_CHECK(nnz >= 0, SABLE_EMALFORMED, "row_ptr[N] < row_ptr[0]");
    const size_t need = Carver::round(4 * (size_t)N) * 2 + Carver::round(8 * (size_t)N) +
                        Carver::round(4 * (size_t)N) + 256 + std::max(sort_tmp_bytes(N), scan_tmp_bytes(N + 1));
    void* s;
    int rc = ctx_scratch(ctx, need, &s);
    if (rc) return rc;
    Carver cv(s);
    int32_t* len = nnz_row ? nnz_row : cv.take<int32_t>(N);
    int32_t* nlen = cv.take<int32_t>(N);
    uint64_t* keys = cv.take<uint64_t>(N);
    uint32_t* rows = cv.take<uint32_t>(N);
    int* bins = cv.take<int>(4);
    void* tmp = cv.p + cv.used;
    SABLE_HIP(hipMemsetAsync(bins, 0, 16, st));
    hipLaunchKernelGGL(k_row_lengths, dim3(gs(N)), dim3(RB), 0, st, N, rp, len, keys, rows, bins);
    SABLE_LAUNCH_CHECK();
    rc = sort_pairs_u64(st, keys, rows, N, bits_for(nnz), tmp);
    if (rc) return rc;
    hipLaunchKernelGGL(k_perm_out, dim3(gs(N)), dim3(RB), 0, st, N, rows, len, perm, nlen);
    SABLE_LAUNCH_CHECK();
    rc = scan_exclusive_i32(st, nlen, orp, N, orp + N, tmp);  // orp[N] = total
    if (rc) return rc;
    hipLaunchKernelGGL(k_copy_rows<T>, dim3(gs((int64_t)std::min<int64_t>(N, 1 << 20) * WAVE)), dim3(RB), 0, st, N,
                       perm, rp, col, (const T*)val, orp, ocol, (T*)oval);
    SABLE_LAUNCH_CHECK();
    int h[4];
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, bins, 16, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    memcpy(h, ctx->pinned, 16);
    // cumulative, as num_zero_rows / num_one_rows / num_two_rows / num_three_rows
    bins_out[0] = h[0];
    bins_out[1] = h[0] + h[1];
    bins_out[2] = h[0] + h[1] + h[2];
    bins_out[3] = h[0] + h[1] + h[2] + h[3];
    return SABLE_OK;
}

}  // namespace

SABLE_API int sable_csr_sort_rows_by_nnz(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                         const int32_t* col, const void* val, int32_t* out_row_ptr, int32_t* out_col,
                                         void* out_val, int32_t* out_perm, int32_t* out_nnz_row, int32_t* bins) {
    SABLE_CHECK(ctx && row_ptr && out_row_ptr && out_perm && bins, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0, SABLE_EINVAL, "negative size");
    memset(bins, 0, 4 * sizeof(int32_t));
    if (N == 0) return SABLE_OK;
    if (dtype == SABLE_F32)
        return sort_rows_t<float>(ctx, N, row_ptr, col, val, out_row_ptr, out_col, out_val, out_perm, out_nnz_row,
                                  bins);
    if (dtype == SABLE_F64)
        return sort_rows_t<double>(ctx, N, row_ptr, col, val, out_row_ptr, out_col, out_val, out_perm, out_nnz_row,
                                   bins);
    return fail(SABLE_EINVAL, "bad dtype");
}

SABLE_API int sable_permute_rows(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* perm, int scatter,
                                 const void* in, void* out) {
    SABLE_CHECK(ctx && perm && in && out && in != out, SABLE_EINVAL, "bad argument");
    SABLE_CHECK(N >= 0, SABLE_EINVAL, "negative size");
    if (N == 0) return SABLE_OK;
    if (dtype == SABLE_F32) {
        if (scatter)
            hipLaunchKernelGGL(k_scatter_rows<float>, dim3(gs(N)), dim3(RB), 0, ctx->stream, N, perm,
                               (const float*)in, (float*)out);
        else
            hipLaunchKernelGGL(k_gather_rows<float>, dim3(gs(N)), dim3(RB), 0, ctx->stream, N, perm, (const float*)in,
                               (float*)out);
    } else if (dtype == SABLE_F64) {
        if (scatter)
            hipLaunchKernelGGL(k_scatter_rows<double>, dim3(gs(N)), dim3(RB), 0, ctx->stream, N, perm,
                               (const double*)in, (double*)out);
        else
            hipLaunchKernelGGL(k_gather_rows<double>, dim3(gs(N)), dim3(RB), 0, ctx->stream, N, perm,
                               (const double*)in, (double*)out);
    } else {
        return fail(SABLE_EINVAL, "bad dtype");
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}
