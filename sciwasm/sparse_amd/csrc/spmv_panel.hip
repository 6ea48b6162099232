// spmv_panel.hip -- column-panel blocked CSR SpMV for matrices whose x does
// not fit the L2 of one XCD.
//
// With uniform random columns over a large x, nearly every x[col] gather of
// the row-block kernel misses the 4 MB L2 of its XCD and pulls a 128-byte
// line across the fabric for 8 useful bytes.  The panel plan splits the
// columns into panels of `panel_cols` columns (default: 2 MB of x) and
// re-lays the CSR out panel-major: panel p holds, for every row, the row's
// entries whose column falls in p, in their original (ascending) order.  One
// launch per panel then keeps every gather of that launch inside an
// L2-resident slice of x; the launches accumulate into y (beta = 1 after the
// first panel, and rows with no entries in a panel are not touched).
//
// All P row pointers share one array S of P*N+1 offsets: panel p's row
// pointer is S + p*N (its N+1 entries overlap the next panel's first).
#include "common.h"
#include "prims.h"

#include <cstring>
#include <vector>

struct sable_csr_panels_s {
    sable_ctx_t* ctx = nullptr;
    sable_dtype_t dtype = SABLE_F64;
    int32_t N = 0;
    int64_t nnz = 0;
    int32_t panel_cols = 0;
    int32_t npanels = 0;
    int32_t* S = nullptr;       // P*N+1 panel row pointers
    int32_t* col = nullptr;     // panel-major copy
    void* val = nullptr;        // panel-major copy
    int32_t* blk = nullptr;     // concatenated row-block boundaries of all panels
    std::vector<int64_t> blk_off;  // first boundary of panel p inside blk
    std::vector<int64_t> nblocks;
    int lds_cap = 0;
    size_t bytes = 0;
};

namespace sable {

constexpr int PB = 256;

// counts[p*N + i] = entries of row i in panel p
__global__ void k_panel_count(int32_t N, int32_t W, int32_t P, const int32_t* __restrict__ rp,
                              const int32_t* __restrict__ col, int32_t* __restrict__ counts, int* __restrict__ err) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        for (int32_t p = 0; p < P; p++) counts[(int64_t)p * N + i] = 0;
        for (int64_t j = rp[i]; j < rp[i + 1]; j++) {
            const int c = col[j];
            if (c < 0 || c >= N) {
                atomicOr(err, 1);
                continue;
            }
            counts[(int64_t)(c / W) * N + i]++;
        }
    }
}

template <typename T>
__global__ void k_panel_scatter(int32_t N, int32_t W, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                const T* __restrict__ val, const int32_t* __restrict__ S, int32_t* __restrict__ ocol,
                                T* __restrict__ oval) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t cur_p = -1;
        int64_t dst = 0;
        for (int64_t j = rp[i]; j < rp[i + 1]; j++) {
            const int c = col[j];
            if (c < 0 || c >= N) continue;
            const int32_t p = c / W;
            if (p != cur_p) {
                cur_p = p;
                dst = S[(int64_t)p * N + i];
            }
            ocol[dst] = c;
            oval[dst] = val[j];
            dst++;
        }
    }
}

static unsigned gsz(int64_t n) {
    unsigned g = grid_for(n, PB);
    return g > 16384 ? 16384 : g;
}

static void free_panels(sable_csr_panels_s* p) {
    if (!p) return;
    void* ptrs[] = {p->S, p->col, p->val, p->blk};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_csr_panels_create(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                      const int32_t* col, const void* val, int32_t panel_cols,
                                      const sable_csr_tune_t* tune, sable_csr_panels_t** out) {
    SABLE_CHECK(ctx && row_ptr && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N > 0, SABLE_EINVAL, "N must be positive");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    *out = nullptr;
    hipStream_t st = ctx->stream;
    const size_t vs = dtype == SABLE_F64 ? 8 : 4;
    int32_t W = panel_cols > 0 ? panel_cols : (int32_t)((2u << 20) / vs);
    if (W > N) W = N;
    const int32_t P = (int32_t)(((int64_t)N + W - 1) / W);
    SABLE_CHECK((int64_t)P * N < 0x7fffffffLL, SABLE_EINVAL, "too many panels for int32 offsets");
    int32_t ends[2];
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, row_ptr, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipMemcpyAsync((char*)ctx->pinned + 4, row_ptr + N, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    memcpy(ends, ctx->pinned, 8);
    SABLE_CHECK(ends[0] == 0 && ends[1] >= 0, SABLE_EINVAL, "row_ptr[0] must be 0");
    const int64_t nnz = ends[1];
    auto* p = new sable_csr_panels_s();
    p->ctx = ctx;
    p->dtype = dtype;
    p->N = N;
    p->nnz = nnz;
    p->panel_cols = W;
    p->npanels = P;
    const int64_t SN = (int64_t)P * N + 1;
    int T = (tune && tune->block_weight > 0) ? tune->block_weight : 1024;
    int C = (tune && tune->row_weight > 0) ? tune->row_weight : 2;
    bool ok = hipMalloc(&p->S, 4 * (size_t)SN) == hipSuccess &&
              hipMalloc(&p->col, 4 * (size_t)(nnz ? nnz : 1)) == hipSuccess &&
              hipMalloc(&p->val, vs * (size_t)(nnz ? nnz : 1)) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        free_panels(p);
        return fail(SABLE_ENOMEM, "panel plan allocation failed");
    }
    void* s;
    int rc = ctx_scratch(ctx, Carver::round(4 * (size_t)SN) + 512 + scan_tmp_bytes(SN), &s);
    if (rc) {
        free_panels(p);
        return rc;
    }
    Carver cv(s);
    int32_t* counts = cv.take<int32_t>(SN);
    int* err = cv.take<int>(4);
    void* tmp = cv.p + cv.used;
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_panel_count, dim3(gsz(N)), dim3(PB), 0, st, N, W, P, row_ptr, col, counts, err);
    SABLE_LAUNCH_CHECK();
    rc = scan_exclusive_i32(st, counts, p->S, SN - 1, p->S + (SN - 1), tmp);
    if (rc) {
        free_panels(p);
        return rc;
    }
    if (dtype == SABLE_F64)
        hipLaunchKernelGGL(k_panel_scatter<double>, dim3(gsz(N)), dim3(PB), 0, st, N, W, row_ptr, col,
                           (const double*)val, p->S, p->col, (double*)p->val);
    else
        hipLaunchKernelGGL(k_panel_scatter<float>, dim3(gsz(N)), dim3(PB), 0, st, N, W, row_ptr, col,
                           (const float*)val, p->S, p->col, (float*)p->val);
    SABLE_LAUNCH_CHECK();
    // panel boundaries on the host for the block counts
    std::vector<int32_t> starts(P + 1);
    for (int32_t q = 0; q <= P; q++) {
        SABLE_HIP(hipMemcpyAsync(&((int32_t*)ctx->pinned)[q % 512], p->S + (int64_t)q * N, 4, hipMemcpyDeviceToHost, st));
        if (q % 512 == 511 || q == P) {
            SABLE_HIP(hipStreamSynchronize(st));
            for (int32_t k = q - (q % 512); k <= q; k++) starts[k] = ((int32_t*)ctx->pinned)[k % 512];
        }
    }
    int e = 0;
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    e = *(int*)ctx->pinned;
    if (e) {
        free_panels(p);
        return fail(SABLE_EMALFORMED, "column index outside [0, N)");
    }
    p->blk_off.resize(P + 1);
    p->nblocks.resize(P);
    int64_t tot = 0;
    for (int32_t q = 0; q < P; q++) {
        p->blk_off[q] = tot;
        p->nblocks[q] = plan_nblocks((int64_t)starts[q + 1] - starts[q], N, T, C);
        tot += p->nblocks[q] + 1;
    }
    p->blk_off[P] = tot;
    if (hipMalloc(&p->blk, 4 * (size_t)(tot + 1)) != hipSuccess) {
        (void)hipGetLastError();
        free_panels(p);
        return fail(SABLE_ENOMEM, "panel plan allocation failed");
    }
    for (int32_t q = 0; q < P; q++) {
        rc = plan_fill_stream(st, N, p->S + (int64_t)q * N, T, C, p->nblocks[q], p->blk + p->blk_off[q], err);
        if (rc) {
            free_panels(p);
            return rc;
        }
    }
    p->lds_cap = plan_lds_cap(T);
    p->bytes = 4 * (size_t)SN + (4 + vs) * (size_t)nnz + 4 * (size_t)tot;
    SABLE_HIP(hipStreamSynchronize(st));
    *out = p;
    return SABLE_OK;
}

SABLE_API void sable_csr_panels_destroy(sable_csr_panels_t* p) {
    if (!p) return;
    (void)hipStreamSynchronize(p->ctx->stream);
    free_panels(p);
}

SABLE_API int sable_csr_panels_info(const sable_csr_panels_t* p, int32_t* npanels, int32_t* panel_cols, int64_t* bytes) {
    SABLE_CHECK(p, SABLE_EINVAL, "NULL plan");
    if (npanels) *npanels = p->npanels;
    if (panel_cols) *panel_cols = p->panel_cols;
    if (bytes) *bytes = (int64_t)p->bytes;
    return SABLE_OK;
}

SABLE_API int sable_spmv_csr_panels(sable_ctx_t* ctx, const sable_csr_panels_t* p, const void* x, void* y, double beta,
                                    int inner_max) {
    SABLE_CHECK(ctx && p && x && y, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(inner_max >= 0, SABLE_EINVAL, "negative inner_max");
    for (int it = 0; it < inner_max; it++) {
        for (int32_t q = 0; q < p->npanels; q++) {
            const double b = q == 0 ? beta : 1.0;
            if (q > 0 && p->nblocks[q] == 0) continue;
            int rc = csr_launch(ctx->stream, p->dtype, p->blk + p->blk_off[q], p->nblocks[q], p->S + (int64_t)q * p->N,
                                p->col, p->val, x, y, b, p->lds_cap);
            if (rc) return rc;
        }
    }
    return SABLE_OK;
}

}  // extern "C"
