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
//   reorder_NN           :2501-2622  greedy window ordering of each bucket
//                                    of <= 1024 equal-nnz rows by the
//                                    Jaccard similarity of their 16-column
//                                    cache-line sets (reorder_subset
//                                    :2439-2492, find_next_row :2330-2419,
//                                    assign_w_min_elems :2275-2327)
//
// The reference writes the new row_ptr through a view sized by the
// undefined A_csr_new.nrows, i.e. a zero-length view, so its row_ptr is
// never stored; the row_ptr here is the one its loop computes.
#include "common.h"
#include "prims.h"

#include <string.h>

#include <algorithm>
#include <vector>

using namespace sable;

namespace {

constexpr int RB = 256;

unsigned gs(int64_t n) { return grid_for(n, RB); }

// nnz of every row; keys = nnz (for the stable sort), payload = row.  The
// counts of rows of 0-3 entries are summed per workgroup (grid-stride, then
// waves, then LDS) and added with one atomic per counter and workgroup.
__global__ __launch_bounds__(RB) void k_row_lengths(int32_t N, const int32_t* __restrict__ rp,
                                                    int32_t* __restrict__ len, uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ rows, int* __restrict__ bins) {
    __shared__ int part[RB / WAVE][4];
    int b[4] = {0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t l = rp[i + 1] - rp[i];
        len[i] = l;
        keys[i] = (uint64_t)(uint32_t)l;
        rows[i] = (uint32_t)i;
#pragma unroll
        for (int q = 0; q < 4; q++) b[q] += l == q;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) b[q] = wave_sum(b[q]);
    const int w = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) part[w][q] = b[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        int t = 0;
        for (int k = 0; k < RB / WAVE; k++) t += part[k][threadIdx.x];
        if (t) atomicAdd(&bins[threadIdx.x], t);
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

__global__ void k_iota(int32_t N, int32_t* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) p[i] = (int32_t)i;
}

__global__ void k_row_len_of(int32_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ perm,
                             int32_t* __restrict__ len) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) len[i] = rp[perm[i] + 1] - rp[perm[i]];
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

// ---------------------------------------------------------------- reorder_NN
constexpr int NN_B = 1024;  // rows per bucket (the reference's B)
constexpr int NN_W = 16;    // largest window
constexpr int NN_RUNS = 12;  // rows of at most this many entries keep their line runs in LDS

// Runs of consecutive entries in the same 16-column line, per row.
__global__ void k_nlines(int32_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                         int32_t* __restrict__ nlines) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int32_t n = 0;
    for (int32_t j = rp[i]; j < rp[i + 1];) {
        ++n;
        ++j;
        while (j < rp[i + 1] && (col[j - 1] >> 4) == (col[j] >> 4)) ++j;
    }
    nlines[i] = n;
}

// assign_w_min_elems on a list held in registers: the first position whose
// entry is -1 or below m takes m; the entries from there to the last valid
// one move down by one, the last valid one being overwritten (as the
// reference's loop does, also when the list is not full).
__device__ __forceinline__ void nn_insert(float (&lj)[NN_W], int32_t (&ln)[NN_W], int32_t cand, double m, int w) {
    int ins = w, last = -1;
#pragma unroll
    for (int k = NN_W - 1; k >= 0; --k)
        if (k < w && !(lj[k] >= 0.f && (double)lj[k] >= m)) ins = k;
#pragma unroll
    for (int k = 0; k < NN_W; ++k)
        if (k < w && lj[k] != -1.f) last = k;
    if (ins >= w) return;
    if (lj[0] == -1.f && ins == 0) {  // empty list
        lj[0] = (float)m;
        ln[0] = cand;
        return;
    }
    bool empty_slot = false;
#pragma unroll
    for (int k = 0; k < NN_W; ++k)
        if (k == ins) empty_slot = lj[k] == -1.f;
    if (!empty_slot) {
#pragma unroll
        for (int k = NN_W - 1; k >= 1; --k)
            if (k > ins && k <= last) {
                lj[k] = lj[k - 1];
                ln[k] = ln[k - 1];
            }
    }
#pragma unroll
    for (int k = 0; k < NN_W; ++k)
        if (k == ins) {
            lj[k] = (float)m;
            ln[k] = cand;
        }
}

// One workgroup per bucket [bstart, bend) of the nnz-sorted rows.
__global__ __launch_bounds__(RB) void k_reorder_nn(int32_t N, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ nlines,
                                                   const int32_t* __restrict__ bstart,
                                                   const int32_t* __restrict__ bend, int32_t nbuckets, int w,
                                                   int32_t* __restrict__ lu_pool, int64_t L,
                                                   int32_t* __restrict__ perm) {
    __shared__ int8_t vis[NN_B];
    __shared__ double cj[NN_B];
    __shared__ float jac[NN_W][NN_W];
    __shared__ int32_t nid[NN_W][NN_W];
    __shared__ int32_t s_next;
    // line runs of every row of a bucket of short rows (nnz <= NN_RUNS):
    // run r of row i is runs[i * NN_RUNS + r], -1 for a line >= L
    __shared__ int32_t runs[NN_B * NN_RUNS];
    __shared__ uint8_t nruns[NN_B];
    const int tid = threadIdx.x;
    int32_t* lu = lu_pool + (int64_t)blockIdx.x * L;
    for (int b = blockIdx.x; b < nbuckets; b += gridDim.x) {
        const int32_t start = bstart[b], end = bend[b], n = end - start;
        // the rows of a bucket have equal nnz
        const bool in_lds = n > 0 && rp[start + 1] - rp[start] <= NN_RUNS;
        for (int i = tid; i < n; i += RB) {
            vis[i] = 0;
            if (in_lds) {
                const int32_t c = start + i, e = rp[c + 1];
                int r = 0;
                for (int32_t j = rp[c]; j < e;) {
                    const int32_t ln = col[j] >> 4;
                    runs[i * NN_RUNS + r++] = (ln >= 0 && ln < L) ? ln : -1;
                    ++j;
                    while (j < e && (col[j - 1] >> 4) == (col[j] >> 4)) ++j;
                }
                nruns[i] = (uint8_t)r;
            }
        }
        __syncthreads();
        int32_t index = start, pos = start;
        int nnb = 0, count = 0;
        while (index != N) {
            const int slot = count;
            if (tid == 0) {
                vis[index - start] = 3;
                perm[pos] = index;
            }
            ++pos;
            if (++count == w) count = 0;
            if (nnb < w) ++nnb;
            // find_next_row for the new window row: its lines, then the
            // similarity of every unvisited row of the bucket
            // lines past N/16 (columns >= N) are not tracked, as the
            // reference's Int32Array(N/16 + 1) drops those stores
            if (in_lds) {
                // the window row's valid lines are its own runs in LDS; a
                // candidate's run counts when its line is among them (the
                // line table's lu[ln] == index)
                __syncthreads();
                const int ii = index - start, ni = nruns[ii];
                for (int i = tid; i < n; i += RB) {
                    if (vis[i] != 0) continue;
                    const int nc = nruns[i];
                    int32_t al = 0;
                    for (int r = 0; r < nc; ++r) {
                        const int32_t ln = runs[i * NN_RUNS + r];
                        if (ln < 0) continue;
                        for (int q = 0; q < ni; ++q)
                            if (runs[ii * NN_RUNS + q] == ln) {
                                ++al;
                                break;
                            }
                    }
                    cj[i] = (double)al / (double)(ni + nc - al);
                }
            } else {
                for (int32_t j = rp[index] + tid; j < rp[index + 1]; j += RB) {
                    const int32_t ln = col[j] >> 4;
                    if (ln >= 0 && ln < L) lu[ln] = index;
                }
                __threadfence_block();
                __syncthreads();
                for (int i = tid; i < n; i += RB) {
                    if (vis[i] != 0) continue;
                    const int32_t c = start + i, e = rp[c + 1];
                    int32_t al = 0;
                    for (int32_t j = rp[c]; j < e;) {
                        const int32_t ln = col[j] >> 4;
                        if (ln >= 0 && ln < L && lu[ln] == index) ++al;
                        ++j;
                        while (j < e && (col[j - 1] >> 4) == (col[j] >> 4)) ++j;
                    }
                    cj[i] = (double)al / (double)(nlines[index] + nlines[c] - al);
                }
            }
            __syncthreads();
            // The candidates are folded into the list in row order by the
            // first wave, every lane holding the same list.  A candidate is
            // a no-op exactly when the list holds w valid (>= 0) entries,
            // all >= its score (nn_insert then finds no position); a ballot
            // skips those, and the others are inserted one by one, in order.
            if (tid < WAVE) {
                float lj[NN_W];
                int32_t ln[NN_W];
#pragma unroll
                for (int k = 0; k < NN_W; ++k) {
                    lj[k] = -1.f;
                    ln[k] = N;
                }
                auto changes = [&](bool cand, double m) {
                    bool full = true;
                    float thr = 0.f;
#pragma unroll
                    for (int k = 0; k < NN_W; ++k)
                        if (k < w) {
                            full = full && lj[k] >= 0.f;
                            thr = (k == 0 || lj[k] < thr) ? lj[k] : thr;
                        }
                    return cand && !(full && (double)thr >= m);
                };
                for (int base = 0; base < n; base += WAVE) {
                    const int i = base + tid;
                    const bool cand = i < n && vis[i] == 0;
                    const double m = cand ? cj[i] : 0.0;
                    uint64_t mask = __ballot(changes(cand, m));
                    while (mask != 0ull) {
                        const int j = __ffsll((unsigned long long)mask) - 1;
                        const double mj = __shfl(m, j, WAVE);
                        nn_insert(lj, ln, start + base + j, mj, w);
                        mask &= mask - 1ull;
                        mask &= __ballot(changes(cand, m));
                    }
                }
                if (tid == 0) {
#pragma unroll
                    for (int k = 0; k < NN_W; ++k) {
                        jac[slot][k] = lj[k];
                        nid[slot][k] = ln[k];
                    }
                    vis[index - start] = 1;
                }
                // next row: each slot's first unvisited candidate, then the
                // slot with the largest score, a later slot winning a tie
                // (lane s2 < nnb <= NN_W takes slot s2; LDS accesses of one
                // wave complete in order, so lane 0's writes are seen)
                int32_t nx = N;
                double m = 0;
                if (tid < nnb) {
                    for (int k = 0; k < w; ++k) {
                        const float v = jac[tid][k];
                        const int32_t q = nid[tid][k];
                        if (v >= 0.f && q < N && vis[q - start] == 0) {
                            nx = q;
                            m = v;
                            break;
                        }
                    }
                }
                bool ok = nx != N;
                int s2 = tid;
#pragma unroll
                for (int off = NN_W / 2; off > 0; off >>= 1) {
                    const double m2 = __shfl_xor(m, off, WAVE);
                    const int32_t nx2 = __shfl_xor(nx, off, WAVE);
                    const int s22 = __shfl_xor(s2, off, WAVE);
                    const bool ok2 = __shfl_xor((int)ok, off, WAVE) != 0;
                    if (ok2 && (!ok || m2 > m || (m2 == m && s22 > s2))) {
                        m = m2;
                        nx = nx2;
                        s2 = s22;
                        ok = true;
                    }
                }
                if (tid == 0) s_next = ok ? nx : N;
            }
            __syncthreads();
            index = s_next;
            __syncthreads();
        }
    }
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
    hipLaunchKernelGGL(k_row_lengths, dim3(std::min(gs(N), 2048u)), dim3(RB), 0, st, N, rp, len, keys, rows, bins);
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

SABLE_API int sable_csr_reorder_nn(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                   const int32_t* col, const void* val, const int32_t* perm_in, int w,
                                   int32_t* out_row_ptr, int32_t* out_col, void* out_val, int32_t* out_perm) {
    SABLE_CHECK(ctx && row_ptr && out_row_ptr && out_perm, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(w >= 1 && w <= NN_W, SABLE_EINVAL, "window w=%d outside [1, %d]", w, NN_W);
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (N == 0) return SABLE_OK;
    hipStream_t st = ctx->stream;
    // buckets from the row-length histogram, as reorder_NN forms them
    std::vector<int32_t> hrp((size_t)N + 1);
    SABLE_HIP(hipMemcpyAsync(hrp.data(), row_ptr, 4 * ((size_t)N + 1), hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    SABLE_CHECK(hrp[0] == 0, SABLE_EINVAL, "row_ptr[0] must be 0");
    int32_t maxl = 0, first = -1;
    for (int32_t i = 0; i < N; ++i) {
        const int32_t l = hrp[i + 1] - hrp[i];
        SABLE_CHECK(l >= 0, SABLE_EMALFORMED, "row_ptr decreases at row %d", i);
        if (l > maxl) maxl = l;
        if (first < 0 && l > 0) first = i;
    }
    std::vector<int64_t> freq((size_t)maxl + 1, 0);
    for (int32_t i = 0; i < N; ++i) freq[hrp[i + 1] - hrp[i]]++;
    std::vector<int32_t> bs, be;
    int64_t start = first < 0 ? N : first;
    for (int32_t c = 1; c <= maxl; ++c)
        while (freq[c] > 0) {
            const int64_t take = std::min<int64_t>(freq[c], NN_B);
            bs.push_back((int32_t)start);
            be.push_back((int32_t)(start + take));
            freq[c] -= take;
            start += take;
        }
    const int32_t nb = (int32_t)bs.size();
    const int64_t L = (int64_t)N / 16 + 1;
    const int nwg = (int)std::max<int64_t>(1, std::min<int64_t>(nb, std::min<int64_t>(512, 4 * (int64_t)ctx->num_cu)));
    const size_t need = Carver::round(4 * (size_t)N) * 3 + Carver::round(4 * (size_t)std::max(nb, 1)) * 2 +
                        Carver::round(4 * (size_t)nwg * L) + 256 + scan_tmp_bytes(N);
    void* s;
    int rc = ctx_scratch(ctx, need, &s);
    if (rc) return rc;
    Carver cv(s);
    int32_t* nl = cv.take<int32_t>(N);
    int32_t* pnn = cv.take<int32_t>(N);
    int32_t* nlen = cv.take<int32_t>(N);
    int32_t* dbs = cv.take<int32_t>(std::max(nb, 1));
    int32_t* dbe = cv.take<int32_t>(std::max(nb, 1));
    int32_t* lu = cv.take<int32_t>((size_t)nwg * L);
    void* tmp = cv.p + cv.used;
    // identity, then each bucket's walk overwrites its own positions
    hipLaunchKernelGGL(k_iota, dim3(gs(N)), dim3(RB), 0, st, N, pnn);
    SABLE_LAUNCH_CHECK();
    if (nb > 0) {
        SABLE_HIP(hipMemcpyAsync(dbs, bs.data(), 4 * (size_t)nb, hipMemcpyHostToDevice, st));
        SABLE_HIP(hipMemcpyAsync(dbe, be.data(), 4 * (size_t)nb, hipMemcpyHostToDevice, st));
        SABLE_HIP(hipMemsetAsync(lu, 0xff, 4 * (size_t)nwg * L, st));
        hipLaunchKernelGGL(k_nlines, dim3(gs(N)), dim3(RB), 0, st, N, row_ptr, col, nl);
        SABLE_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_reorder_nn, dim3(nwg), dim3(RB), 0, st, N, row_ptr, col, nl, dbs, dbe, nb, w, lu, L,
                           pnn);
        SABLE_LAUNCH_CHECK();
    }
    // rows of the input in the new order
    hipLaunchKernelGGL(k_row_len_of, dim3(gs(N)), dim3(RB), 0, st, N, row_ptr, pnn, nlen);
    SABLE_LAUNCH_CHECK();
    rc = scan_exclusive_i32(st, nlen, out_row_ptr, N, out_row_ptr + N, tmp);
    if (rc) return rc;
    const unsigned gc = gs((int64_t)std::min<int64_t>(N, 1 << 20) * WAVE);
    if (dtype == SABLE_F32)
        hipLaunchKernelGGL(k_copy_rows<float>, dim3(gc), dim3(RB), 0, st, N, pnn, row_ptr, col, (const float*)val,
                           out_row_ptr, out_col, (float*)out_val);
    else
        hipLaunchKernelGGL(k_copy_rows<double>, dim3(gc), dim3(RB), 0, st, N, pnn, row_ptr, col, (const double*)val,
                           out_row_ptr, out_col, (double*)out_val);
    SABLE_LAUNCH_CHECK();
    // permutation of the original rows
    if (perm_in)
        hipLaunchKernelGGL(k_gather_rows<int32_t>, dim3(gs(N)), dim3(RB), 0, st, N, pnn, perm_in, out_perm);
    else
        SABLE_HIP(hipMemcpyAsync(out_perm, pnn, 4 * (size_t)N, hipMemcpyDeviceToDevice, st));
    SABLE_LAUNCH_CHECK();
    SABLE_HIP(hipStreamSynchronize(st));  // the bucket vectors are host memory of this call
    return SABLE_OK;
}
