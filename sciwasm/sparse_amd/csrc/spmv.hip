// spmv.hip -- SpMV over CSR, COO, DIA and ELL for gfx950.
//
// CSR uses a row-block schedule computed on the device: rows are cut into
// blocks of roughly equal weight  nnz + row_weight*rows  (a monotone key, so
// block boundaries are found in parallel, one thread per row).  A block whose
// nonzeros fit the LDS buffer is "streamed": the workgroup reads val/col
// fully coalesced, gathers x, stores the products in LDS, then reduces each
// row with L lanes (L = largest power of two with L*rows <= 256) and a
// shuffle tree.  A block that does not fit (it holds a long row) reduces
// each of its rows with the whole workgroup.  Workgroup ids are remapped so
// that each XCD processes a contiguous range of blocks (its L2 then sees a
// contiguous band of rows and, for banded matrices, of x).
#include "common.h"

#include <cstdlib>
#include <cstring>

struct sable_coo_plan_s {
    int32_t N = 0;
    int64_t nnz = 0;
    bool sorted = false;       // rows non-decreasing: segmented (CSR block) path
    int32_t* rp = nullptr;     // sorted: row heads, N + 1
    int32_t* blk = nullptr;    // sorted: row blocks
    int64_t nblocks = 0;
    int lds_cap = 0;
};

struct sable_csr_plan_s {
    sable_ctx_t* ctx = nullptr;
    int32_t N = 0;
    int64_t nnz = 0;
    const int32_t* row_ptr = nullptr;
    int32_t* blk = nullptr;  // nblocks + 1 row boundaries
    int64_t nblocks = 0;
    int block_weight = 0;
    int row_weight = 0;
    int lds_cap = 0;  // elements staged per block
    int vector_lanes = 0;  // > 0: the row-vector kernel with this many lanes per row
    int vector_rows = 4;   // rows per lane group of the row-vector kernel
    bool owns_blk = false;
};

namespace sable {

constexpr int BS = 256;
constexpr int DEFAULT_BLOCK_WEIGHT = 1024;
constexpr int DEFAULT_ROW_WEIGHT = 2;
constexpr int NUM_XCD = 8;

// Physical workgroup p runs on XCD p % 8; give XCD x a contiguous range of
// logical blocks.
__device__ __forceinline__ int64_t xcd_remap(int64_t p, int64_t nb) {
    int64_t q = nb / NUM_XCD, r = nb % NUM_XCD;
    int64_t x = p % NUM_XCD, idx = p / NUM_XCD;
    return x * q + (x < r ? x : r) + idx;
}

// One thread per row i in [0, N]: write blk[b] = first row whose key reaches
// b*T.  key(i) = (row_ptr[i] - row_ptr[0]) + C*i.  Flags a non-monotone
// row_ptr in *err.
__global__ void k_plan_fill(int32_t N, const int32_t* __restrict__ rp, int64_t C, int64_t T,
                            int64_t B, int32_t* __restrict__ blk, int* __restrict__ err) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    int64_t base = rp[0];
    int64_t kprev = -1;
    if (i > 0) kprev = ((int64_t)rp[i - 1] - base) + C * (i - 1);
    int64_t blo = kprev < 0 ? 0 : kprev / T + 1;
    int64_t bhi;
    if (i < N) {
        int64_t a = rp[i], e = rp[i + 1];
        if (e < a) atomicOr(err, 1);
        bhi = ((a - base) + C * i) / T;
        if (bhi > B) bhi = B;
    } else {
        bhi = B;
    }
    for (int64_t b = blo; b <= bhi; b++) blk[b] = (int32_t)i;
}

// The block count of a plan built on the device without reading nnz on the
// host: weight W = nnz + C*N, block weight Teff = max(T, ceil(W / cap)) so
// that at most cap blocks are needed, B = ceil(W / Teff).  Every thread of
// the fill kernel and of the persistent kernel derives the same B from
// row_ptr[0] and row_ptr[N].
__device__ __forceinline__ void dev_plan_shape(int32_t N, const int32_t* __restrict__ rp, int64_t C, int64_t T,
                                               int64_t cap, int64_t* Teff, int64_t* B) {
    const int64_t nnz = (int64_t)rp[N] - rp[0];
    const int64_t W = (nnz > 0 ? nnz : 0) + C * N;
    int64_t t = (W + cap - 1) / cap;
    if (t < T) t = T;
    *Teff = t;
    *B = W <= 0 ? 0 : (W + t - 1) / t;
}

__global__ void k_devplan_init(int64_t* shape, int64_t C, int64_t T, int64_t cap, int* err) {
    if (threadIdx.x == 0) {
        shape[0] = C;
        shape[1] = T;
        shape[2] = cap;
        *err = 0;
    }
}

// k_plan_fill with the shape computed on the device; blk holds cap + 1
// entries.  A row_ptr that decreases anywhere, or ends below its start,
// raises *err (the product kernel then writes nothing).
__global__ void k_plan_fill_dev(int32_t N, const int32_t* __restrict__ rp, int64_t C, int64_t T, int64_t cap,
                                int32_t* __restrict__ blk, int* __restrict__ err) {
    int64_t Teff, B;
    dev_plan_shape(N, rp, C, T, cap, &Teff, &B);
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    if (i == 0 && rp[N] < rp[0]) atomicOr(err, 1);
    int64_t base = rp[0];
    int64_t kprev = -1;
    if (i > 0) kprev = ((int64_t)rp[i - 1] - base) + C * (i - 1);
    int64_t blo = kprev < 0 ? 0 : kprev / Teff + 1;
    int64_t bhi;
    if (i < N) {
        int64_t a = rp[i], e = rp[i + 1];
        if (e < a) atomicOr(err, 1);
        bhi = ((a - base) + C * i) / Teff;
        if (bhi > B) bhi = B;
    } else {
        bhi = B;
    }
    for (int64_t b = blo; b <= bhi; b++) blk[b] = (int32_t)i;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    v = wave_sum(v);
    int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < BS / WAVE; k++) s += red[k];
    }
    return s;
}

template <typename T>
__device__ __forceinline__ void write_y(T* y, int64_t r, T s, T beta, bool nonempty) {
    if (beta == T(0)) {
        y[r] = s;
    } else if (beta == T(1)) {
        if (nonempty) y[r] = y[r] + s;  // the reference does not store empty rows
    } else {
        y[r] = beta * y[r] + s;
    }
}

// Streaming loads of the matrix arrays: with NT the col/val loads carry the
// non-temporal hint, so the L2 keeps x (the only reused data) instead of
// matrix lines that are read exactly once.
template <bool NT, typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <typename T, int U, bool NT>
__global__ __launch_bounds__(BS) void k_csr_block(const int32_t* __restrict__ blk, int64_t nblocks,
                                                  const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const T* __restrict__ val, const T* __restrict__ x,
                                                  T* __restrict__ y, T beta, int lds_cap) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* prod = reinterpret_cast<T*>(smem_raw);
    __shared__ T red[BS / WAVE];
    const int64_t b = xcd_remap(blockIdx.x, nblocks);
    const int r0 = blk[b], r1 = blk[b + 1];
    if (r0 >= r1) return;
    const int tid = threadIdx.x;
    const int64_t nz0 = rp[r0];
    const int nnzb = (int)((int64_t)rp[r1] - nz0);
    const int R = r1 - r0;
    if (nnzb <= lds_cap) {
        const int32_t* cb = col + nz0;
        const T* vb = val + nz0;
        // All loads of an unrolled step are issued before any is used, with
        // out-of-range slots predicated off, so a short block still keeps U
        // gathers per thread in flight.
        for (int k = tid; k < nnzb; k += U * BS) {
            int c[U];
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int kk = k + u * BS;
                c[u] = kk < nnzb ? ld_stream<NT>(cb + kk) : 0;
                v[u] = kk < nnzb ? ld_stream<NT>(vb + kk) : T(0);
            }
            T xv[U];
#pragma unroll
            for (int u = 0; u < U; u++) xv[u] = (k + u * BS < nnzb) ? x[c[u]] : T(0);
#pragma unroll
            for (int u = 0; u < U; u++)
                if (k + u * BS < nnzb) prod[k + u * BS] = v[u] * xv[u];
        }
        __syncthreads();
        int L = WAVE;
        while (L > 1 && L * R > BS) L >>= 1;
        const int groups = BS / L, g = tid / L, l = tid & (L - 1);
        for (int rb = 0; rb < R; rb += groups) {
            const int r = r0 + rb + g;
            T s = 0;
            int a = 0, e = 0;
            if (r < r1) {
                a = (int)((int64_t)rp[r] - nz0);
                e = (int)((int64_t)rp[r + 1] - nz0);
                for (int kk = a + l; kk < e; kk += L) s += prod[kk];
            }
            s = group_sum(s, L);
            if (l == 0 && r < r1) write_y(y, r, s, beta, e > a);
        }
    } else {
        for (int r = r0; r < r1; r++) {
            const int64_t a = rp[r], e = rp[r + 1];
            T s = 0;
            int64_t k = a + tid;
            for (; k + (U - 1) * BS < e; k += U * BS) {
                int c[U];
                T v[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    c[u] = ld_stream<NT>(col + k + u * BS);
                    v[u] = ld_stream<NT>(val + k + u * BS);
                }
#pragma unroll
                for (int u = 0; u < U; u++) s += v[u] * x[c[u]];
            }
            for (; k < e; k += BS) s += ld_stream<NT>(val + k) * x[ld_stream<NT>(col + k)];
            s = block_sum(s, red);
            if (tid == 0) write_y(y, r, s, beta, e > a);
        }
    }
}

// Row-vector kernel for matrices of short, regular rows: G lanes per row
// (G >= the typical row length), each group handling R rows, R*G loads in
// flight per group before any is used; no LDS and no block barrier, so a
// row costs three dependent memory round trips (row_ptr, col/val, x).
// Consecutive groups of a wave take consecutive rows, so the col/val loads
// of a wave are contiguous.
template <typename T, int G, int R, bool NT>
__global__ __launch_bounds__(BS) void k_csr_vec(int32_t N, const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ col, const T* __restrict__ val,
                                                const T* __restrict__ x, T* __restrict__ y, T beta) {
    const int64_t gg = ((int64_t)blockIdx.x * BS + threadIdx.x) / G;
    const int l = threadIdx.x & (G - 1);
    const int64_t NG = (int64_t)gridDim.x * (BS / G);
    int a[R], e[R];
    T s[R];
    int maxlen = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int64_t i = gg + k * NG;
        a[k] = i < N ? rp[i] : 0;
        e[k] = i < N ? rp[i + 1] : 0;
        s[k] = T(0);
        maxlen = max(maxlen, e[k] - a[k]);
    }
    for (int c = l; c < maxlen; c += G) {
        int cc[R];
        T vv[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int j = a[k] + c;
            cc[k] = j < e[k] ? ld_stream<NT>(col + j) : 0;
            vv[k] = j < e[k] ? ld_stream<NT>(val + j) : T(0);
        }
        T xv[R];
#pragma unroll
        for (int k = 0; k < R; k++) xv[k] = (a[k] + c < e[k]) ? x[cc[k]] : T(0);
#pragma unroll
        for (int k = 0; k < R; k++) s[k] += vv[k] * xv[k];
    }
#pragma unroll
    for (int k = 0; k < R; k++) s[k] = group_sum(s[k], G);
    if (l == 0) {
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int64_t i = gg + k * NG;
            if (i < N) write_y(y, i, s[k], beta, e[k] > a[k]);
        }
    }
}

template <typename T, int R>
static void vec_launch_r(hipStream_t st, int G, int32_t N, const int32_t* rp, const int32_t* col, const void* val,
                         const void* x, void* y, double beta) {
    const int64_t groups = ((int64_t)N + R - 1) / R;
    const unsigned grid = (unsigned)((groups * G + BS - 1) / BS);
#define VEC_LAUNCH(GG)                                                                                     \
    hipLaunchKernelGGL((k_csr_vec<T, GG, R, true>), dim3(grid), dim3(BS), 0, st, N, rp, col, (const T*)val, \
                       (const T*)x, (T*)y, (T)beta)
    switch (G) {
        case 2: VEC_LAUNCH(2); break;
        case 4: VEC_LAUNCH(4); break;
        case 8: VEC_LAUNCH(8); break;
        case 16: VEC_LAUNCH(16); break;
        case 32: VEC_LAUNCH(32); break;
        default: VEC_LAUNCH(64); break;
    }
#undef VEC_LAUNCH
}

template <typename T>
static void vec_launch_t(hipStream_t st, int G, int R, int32_t N, const int32_t* rp, const int32_t* col,
                         const void* val, const void* x, void* y, double beta) {
    switch (R) {
        case 1: vec_launch_r<T, 1>(st, G, N, rp, col, val, x, y, beta); break;
        case 2: vec_launch_r<T, 2>(st, G, N, rp, col, val, x, y, beta); break;
        case 8: vec_launch_r<T, 8>(st, G, N, rp, col, val, x, y, beta); break;
        case 16: vec_launch_r<T, 16>(st, G, N, rp, col, val, x, y, beta); break;
        default: vec_launch_r<T, 4>(st, G, N, rp, col, val, x, y, beta); break;
    }
}

// Persistent variant of the block kernel.  A grid of about 8 workgroups
// per CU walks the blocks (XCD x owns the contiguous x-th eighth of them;
// its workgroups stride through it), so no workgroup is dispatched per
// block.  Per block the dependent memory round trips are cut to three: the
// next block's boundaries are prefetched while the current one runs, and
// the block's row-pointer slice and the y values (beta != 0) are loaded in
// the same step as col/val.
constexpr int RPS_CAP = 520;

// dshape != nullptr: the plan was built on the device (k_plan_fill_dev);
// dshape = {row weight, block weight, cap} and the block count is derived
// from row_ptr, with nothing done when *err is set.
template <typename T, int U>
__global__ __launch_bounds__(BS) void k_csr_persist(const int32_t* __restrict__ blk, int64_t nblocks,
                                                    const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col,
                                                    const T* __restrict__ val, const T* __restrict__ x,
                                                    T* __restrict__ y, T beta, int lds_cap,
                                                    int32_t N = 0, const int64_t* __restrict__ dshape = nullptr,
                                                    const int* __restrict__ derr = nullptr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* prod = reinterpret_cast<T*>(smem_raw);
    __shared__ T red[BS / WAVE];
    __shared__ int32_t rps[RPS_CAP + 1];
    const int tid = threadIdx.x;
    if (dshape) {
        if (*derr) return;
        int64_t Teff;
        dev_plan_shape(N, rp, dshape[0], dshape[1], dshape[2], &Teff, &nblocks);
    }
    const int64_t G = gridDim.x;
    const int64_t xcd = blockIdx.x % NUM_XCD, idx = blockIdx.x / NUM_XCD;
    const int64_t gx = G / NUM_XCD + (xcd < G % NUM_XCD ? 1 : 0);  // workgroups on this XCD
    const int64_t b_lo = nblocks * xcd / NUM_XCD, b_hi = nblocks * (xcd + 1) / NUM_XCD;
    int64_t b = b_lo + idx;
    int nr0 = 0, nr1 = 0;
    if (b < b_hi) {
        nr0 = blk[b];
        nr1 = blk[b + 1];
    }
    for (; b < b_hi; b += gx) {
        const int r0 = nr0, r1 = nr1;
        // prefetch the boundaries of this workgroup's next block
        if (b + gx < b_hi) {
            nr0 = blk[b + gx];
            nr1 = blk[b + gx + 1];
        }
        if (r0 >= r1) continue;
        const int64_t nz0 = rp[r0];
        const int nnzb = (int)((int64_t)rp[r1] - nz0);
        const int R = r1 - r0;
        if (nnzb <= lds_cap) {
            int L = WAVE;
            while (L > 1 && L * R > BS) L >>= 1;
            const int groups = BS / L, g = tid / L, l = tid & (L - 1);
            const bool stage_rp = R < RPS_CAP;
            if (stage_rp)
                for (int k = tid; k <= R; k += BS) rps[k] = rp[r0 + k];
            // y for the (single) row of each group, when every group has at most one row
            T yold = T(0);
            const bool pre_y = beta != T(0) && R <= groups && l == 0 && g < R;
            if (pre_y) yold = y[r0 + g];
            const int32_t* cb = col + nz0;
            const T* vb = val + nz0;
            for (int k = tid; k < nnzb; k += U * BS) {
                int c[U];
                T v[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int kk = k + u * BS;
                    c[u] = kk < nnzb ? cb[kk] : 0;
                    v[u] = kk < nnzb ? vb[kk] : T(0);
                }
                T xv[U];
#pragma unroll
                for (int u = 0; u < U; u++) xv[u] = (k + u * BS < nnzb) ? x[c[u]] : T(0);
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (k + u * BS < nnzb) prod[k + u * BS] = v[u] * xv[u];
            }
            __syncthreads();
            for (int rb = 0; rb < R; rb += groups) {
                const int r = r0 + rb + g;
                T s = 0;
                int a = 0, e = 0;
                if (r < r1) {
                    if (stage_rp) {
                        a = rps[rb + g] - (int)(rps[0]);
                        e = rps[rb + g + 1] - (int)(rps[0]);
                    } else {
                        a = (int)((int64_t)rp[r] - nz0);
                        e = (int)((int64_t)rp[r + 1] - nz0);
                    }
                    for (int kk = a + l; kk < e; kk += L) s += prod[kk];
                }
                s = group_sum(s, L);
                if (l == 0 && r < r1) {
                    if (beta == T(0)) {
                        y[r] = s;
                    } else if (e > a || beta != T(1)) {
                        const T yo = pre_y ? yold : y[r];
                        y[r] = (beta == T(1)) ? yo + s : beta * yo + s;
                    }
                }
            }
            __syncthreads();  // prod and rps are reused by the next block
        } else {
            for (int r = r0; r < r1; r++) {
                const int64_t a = rp[r], e = rp[r + 1];
                T s = 0;
                for (int64_t k = a + tid; k < e; k += U * BS) {
                    int c[U];
                    T v[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int64_t kk = k + u * BS;
                        c[u] = kk < e ? col[kk] : 0;
                        v[u] = kk < e ? val[kk] : T(0);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++)
                        if (k + u * BS < e) s += v[u] * x[c[u]];
                }
                s = block_sum(s, red);
                if (tid == 0) write_y(y, r, s, beta, e > a);
                __syncthreads();
            }
        }
    }
}

// ------------------------------------------------------------------ DIA
// One thread per row; diagonals in stored order; rounded mul then add.
template <typename T>
__global__ __launch_bounds__(BS) void k_dia(int32_t N, int32_t nd, int32_t stride,
                                            const int32_t* __restrict__ offset,
                                            const T* __restrict__ data, const T* __restrict__ x,
                                            T* __restrict__ y, T beta, int32_t r0, int32_t r1) {
    __shared__ int32_t offs[1024];
    const int ndl = nd < 1024 ? nd : 1024;
    for (int k = threadIdx.x; k < ndl; k += blockDim.x) offs[k] = offset[k];
    __syncthreads();
    const int64_t shiftneg = (int64_t)N - stride;
    constexpr int U = 8;
    for (int64_t n = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < r1;
         n += (int64_t)gridDim.x * blockDim.x) {
        T t = beta == T(1) ? y[n] : (beta == T(0) ? T(0) : beta * y[n]);
        // U diagonals per step: all loads of a step are issued before the
        // products are added in diagonal order.  Entries outside the stored
        // run of a diagonal load a clamped address and are skipped, so the
        // loads are straight-line code (a guarded load makes every later
        // wait conservative); the diagonals' data are read once
        // (non-temporal), x is reused by neighbouring rows.
        for (int k0 = 0; k0 < nd; k0 += U) {
            T d[U], xv[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int k = k0 + u;
                const int off = k < nd ? (k < 1024 ? offs[k] : offset[k]) : 0;
                const int64_t c = n + off;
                const int64_t slot = n - (off < 0 ? shiftneg : 0);
                ok[u] = k < nd && c >= 0 && c < N && slot >= 0 && slot < stride;
                const int64_t kc = k < nd ? k : nd - 1;
                const int64_t sc = slot < 0 ? 0 : (slot >= stride ? stride - 1 : slot);
                const int64_t cc = c < 0 ? 0 : (c >= N ? N - 1 : c);
                d[u] = __builtin_nontemporal_load(data + kc * stride + sc);
                xv[u] = x[cc];
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (ok[u]) t = add_rn(mul_rn(d[u], xv[u]), t);
        }
        y[n] = t;
    }
}

// ------------------------------------------------------------------ ELL
// One thread per row; slots in order; rounded mul then add (padding slots
// add 0*x[0], as in the reference).
template <typename T>
__global__ __launch_bounds__(BS) void k_ell(int32_t N, int32_t nc, const int32_t* __restrict__ idx,
                                            const T* __restrict__ data, const T* __restrict__ x,
                                            T* __restrict__ y, T beta, int32_t r0, int32_t r1) {
    for (int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r1;
         i += (int64_t)gridDim.x * blockDim.x) {
        T t = beta == T(1) ? y[i] : (beta == T(0) ? T(0) : beta * y[i]);
        // U slots per step, all loads issued before the products are added
        // in slot order (4 keeps the number of column-major streams a wave
        // touches at once small).  Slots past the width load the last slot
        // and are not added, so the loads are straight-line code; the
        // matrix is read once (non-temporal).
        constexpr int U = 4;
        for (int j = 0; j < nc; j += U) {
            int c[U];
            T d[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int64_t e = (int64_t)min(j + u, nc - 1) * N + i;
                c[u] = __builtin_nontemporal_load(idx + e);
                d[u] = __builtin_nontemporal_load(data + e);
            }
            T xv[U];
#pragma unroll
            for (int u = 0; u < U; u++) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; u++)
                if (j + u < nc) t = add_rn(mul_rn(d[u], xv[u]), t);
        }
        y[i] = t;
    }
}

// ------------------------------------------------------- row-major layouts
// The parallel tree's "major: row" storage (par/src/sparse32.js:1840-1880,
// 1912-1973): row i's slots at [i*w, (i+1)*w).  One thread per row of
// [r0, r1); t starts from y[i] and every slot adds with a rounded multiply
// then add, in the reference's order:
//  * DIA (par/src/sparse_opt_32.wat:8213-8269): diagonals in stored order,
//    skipping columns outside [0, N);
//  * ELL (:8292-8346): slots in order, leaving the row at the first
//    negative index;
//  * ELL, gs order (:8369-8756), for width >= 4: the first width % 4 slots
//    into t, the rest into four lane sums l0..l3 (slot s into lane
//    (s - width % 4) % 4, each lane from 0), then y = (((t + l0) + l1) + l2)
//    + l3; narrower rows take the plain order.
template <typename T>
__global__ __launch_bounds__(BS) void k_dia_rm(int32_t N, int32_t nd, const int32_t* __restrict__ offset,
                                               const T* __restrict__ data, const T* __restrict__ x,
                                               T* __restrict__ y, int32_t r0, int32_t r1) {
    __shared__ int32_t offs[1024];
    const int ndl = nd < 1024 ? nd : 1024;
    for (int k = threadIdx.x; k < ndl; k += blockDim.x) offs[k] = offset[k];
    __syncthreads();
    for (int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r1;
         i += (int64_t)gridDim.x * blockDim.x) {
        T t = y[i];
        const T* d = data + i * nd;
        for (int k = 0; k < nd; k++) {
            const int64_t c = i + (k < 1024 ? offs[k] : offset[k]);
            if (c >= 0 && c < N) t = add_rn(mul_rn(d[k], x[c]), t);
        }
        y[i] = t;
    }
}

template <typename T, bool GS>
__global__ __launch_bounds__(BS) void k_ell_rm(int32_t nc, const int32_t* __restrict__ idx,
                                               const T* __restrict__ data, const T* __restrict__ x,
                                               T* __restrict__ y, int32_t r0, int32_t r1) {
    for (int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r1;
         i += (int64_t)gridDim.x * blockDim.x) {
        T t = y[i];
        const int32_t* ci = idx + i * nc;
        const T* d = data + i * nc;
        if (!GS || nc < 4) {
            for (int k = 0; k < nc; k++) {
                const int32_t c = ci[k];
                if (c < 0) break;
                t = add_rn(mul_rn(d[k], x[c]), t);
            }
            y[i] = t;
        } else {
            const int h = nc & 3;
            for (int k = 0; k < h; k++) t = add_rn(mul_rn(d[k], x[ci[k]]), t);
            T l[4] = {T(0), T(0), T(0), T(0)};
            for (int k = h; k < nc; k += 4) {
#pragma unroll
                for (int q = 0; q < 4; q++) l[q] = add_rn(mul_rn(d[k + q], x[ci[k + q]]), l[q]);
            }
            y[i] = add_rn(add_rn(add_rn(add_rn(t, l[0]), l[1]), l[2]), l[3]);
        }
    }
}

// ------------------------------------------------------------------ COO
// Flags rows out of order (bit 1) or out of [0, N) (bit 2).  Each thread
// collects its flags over a grid-stride loop and each wave ORs them into
// err once: rows out of order are the normal case of a shuffled COO, and
// one atomic per entry on one address would serialise the whole check.
__global__ void k_coo_check(int64_t nnz, int32_t N, const int32_t* __restrict__ row, int* __restrict__ err) {
    int f = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int r = row[k];
        if (r < 0 || r >= N) f |= 2;
        if (k + 1 < nnz && row[k + 1] < r) f |= 1;
    }
    const int w = (__ballot(f & 1) ? 1 : 0) | (__ballot(f & 2) ? 2 : 0);
    if ((threadIdx.x & (WAVE - 1)) == 0 && w) atomicOr(err, w);
}

// Row heads of a row-sorted COO become a row pointer: thread k in [0, nnz]
// writes rp[r] = k for every row r in (row[k-1], row[k]].
__global__ void k_coo_heads(int64_t nnz, int32_t N, const int32_t* __restrict__ row, int32_t* __restrict__ rp) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nnz) return;
    const int lo = k == 0 ? 0 : row[k - 1] + 1;
    const int hi = k == nnz ? N : row[k];
    for (int r = lo; r <= hi; r++) rp[r] = (int32_t)k;
}

template <typename T>
__global__ void k_scale(int32_t N, T* __restrict__ y, T beta) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = beta == T(0) ? T(0) : beta * y[i];
}

template <typename T>
__global__ void k_coo_atomic(int64_t nnz, const int32_t* __restrict__ row, const int32_t* __restrict__ col,
                             const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&y[row[k]], val[k] * x[col[k]]);
}

// ---------------------------------------------------------------- host side
static void tune_defaults(const sable_csr_tune_t* t, int* T, int* C) {
    *T = (t && t->block_weight > 0) ? t->block_weight : DEFAULT_BLOCK_WEIGHT;
    *C = (t && t->row_weight > 0) ? t->row_weight : DEFAULT_ROW_WEIGHT;
}

static int64_t nblocks_for(int64_t nnz, int32_t N, int T, int C) {
    int64_t W = nnz + (int64_t)C * N;
    return W <= 0 ? 0 : (W + T - 1) / T;
}

static int lds_cap_for(int T) {
    int cap = 2 * T;
    if (cap < 512) cap = 512;
    if (cap > 8192) cap = 8192;
    return cap;
}

// Reads row_ptr[0] and row_ptr[N] (synchronises).
static int read_nnz(sable_ctx_t* ctx, int32_t N, const int32_t* rp, int64_t* nnz) {
    int32_t ends[2];
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, rp, 4, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipMemcpyAsync((char*)ctx->pinned + 4, rp + N, 4, hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(ends, ctx->pinned, 8);
    SABLE_CHECK(ends[1] >= ends[0] && ends[0] >= 0, SABLE_EMALFORMED,
                "row_ptr[N]=%d < row_ptr[0]=%d", ends[1], ends[0]);
    *nnz = (int64_t)ends[1] - ends[0];
    return SABLE_OK;
}

// Fill blk (device, nblocks+1) and check monotonicity; err is a device int.
static int plan_fill(sable_ctx_t* ctx, int32_t N, const int32_t* rp, int T, int C, int64_t B,
                     int32_t* blk, int* err, bool check) {
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_plan_fill, dim3(grid_for((int64_t)N + 1, BS)), dim3(BS), 0, ctx->stream, N, rp,
                       (int64_t)C, (int64_t)T, B, blk, err);
    SABLE_LAUNCH_CHECK();
    if (check) {
        int64_t e = 0;
        SABLE_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        SABLE_HIP(hipStreamSynchronize(ctx->stream));
        e = *(int*)ctx->pinned;
        SABLE_CHECK(e == 0, SABLE_EMALFORMED, "row_ptr is not monotone");
    }
    return SABLE_OK;
}

int64_t plan_nblocks(int64_t nnz, int32_t N, int T, int C) { return nblocks_for(nnz, N, T, C); }
int plan_lds_cap(int T) { return lds_cap_for(T); }
int plan_fill_stream(hipStream_t st, int32_t N, const int32_t* rp, int T, int C, int64_t B, int32_t* blk, int* err) {
    if (N <= 0) return SABLE_OK;
    hipLaunchKernelGGL(k_plan_fill, dim3(grid_for((int64_t)N + 1, BS)), dim3(BS), 0, st, N, rp, (int64_t)C, (int64_t)T,
                       B, blk, err);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

int g_num_cu = 256;
// 0: k_csr_block, one workgroup per block; 1: k_csr_persist with one
// workgroup per block; 2: k_csr_persist with ~8 workgroups per CU;
// 3: k_csr_block with non-temporal col/val loads.
int g_csr_kernel = 0;

template <typename T>
static void csr_launch_t(hipStream_t st, const int32_t* blk, int64_t nb, const int32_t* rp, const int32_t* col,
                         const void* val, const void* x, void* y, double beta, int lds_cap) {
    const size_t shm = (size_t)lds_cap * sizeof(T);
    if (g_csr_kernel == 1 || g_csr_kernel == 2) {
        // at least one workgroup per XCD, so that every XCD's block range is covered
        int64_t grid = g_csr_kernel == 2 ? (int64_t)g_num_cu * 8 : nb;
        if (grid > nb) grid = nb;
        if (grid < NUM_XCD) grid = NUM_XCD;
        hipLaunchKernelGGL((k_csr_persist<T, 8>), dim3((unsigned)grid), dim3(BS), shm, st, blk, nb, rp, col,
                           (const T*)val, (const T*)x, (T*)y, (T)beta, lds_cap);
    } else if (g_csr_kernel == 3) {
        hipLaunchKernelGGL((k_csr_block<T, 8, true>), dim3((unsigned)nb), dim3(BS), shm, st, blk, nb, rp, col,
                           (const T*)val, (const T*)x, (T*)y, (T)beta, lds_cap);
    } else {
        hipLaunchKernelGGL((k_csr_block<T, 8, false>), dim3((unsigned)nb), dim3(BS), shm, st, blk, nb, rp, col,
                           (const T*)val, (const T*)x, (T*)y, (T)beta, lds_cap);
    }
}

int csr_launch(hipStream_t st, sable_dtype_t dt, const int32_t* blk, int64_t nb, const int32_t* rp,
               const int32_t* col, const void* val, const void* x, void* y, double beta, int lds_cap) {
    if (nb <= 0) return SABLE_OK;
    SABLE_CHECK(nb < 0x7fffffffLL, SABLE_EINVAL, "too many row blocks");
    if (dt == SABLE_F32)
        csr_launch_t<float>(st, blk, nb, rp, col, val, x, y, beta, lds_cap);
    else if (dt == SABLE_F64)
        csr_launch_t<double>(st, blk, nb, rp, col, val, x, y, beta, lds_cap);
    else
        return fail(SABLE_EINVAL, "bad dtype");
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

int csr_plan_view(const sable_csr_plan_t* p, const int32_t** blk, int64_t* nb, int* lds_cap) {
    *blk = p->blk;
    *nb = p->nblocks;
    *lds_cap = p->lds_cap;
    return SABLE_OK;
}

static int dispatch_csr(sable_ctx_t* ctx, sable_dtype_t dt, const int32_t* blk, int64_t nb, const int32_t* rp,
                        const int32_t* col, const void* val, const void* x, void* y, double beta,
                        int inner_max, int lds_cap) {
    g_num_cu = ctx->num_cu;
    for (int it = 0; it < inner_max; it++) {
        int rc = csr_launch(ctx->stream, dt, blk, nb, rp, col, val, x, y, beta, lds_cap);
        if (rc) return rc;
    }
    return SABLE_OK;
}

// Unsorted COO: y = beta*y first, then one atomic add per entry (the order
// of accumulation is not deterministic).
static int coo_atomic_launch(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz, const int32_t* row,
                             const int32_t* col, const void* val, const void* x, void* y, double beta,
                             int inner_max) {
    unsigned g = grid_for(N, BS);
    if (g > 16384) g = 16384;
    unsigned gz = grid_for(nnz, BS);
    if (gz > 65536) gz = 65536;
    for (int it = 0; it < inner_max; it++) {
        if (beta != 1.0) {
            if (dtype == SABLE_F32)
                hipLaunchKernelGGL(k_scale<float>, dim3(g), dim3(BS), 0, ctx->stream, N, (float*)y, (float)beta);
            else
                hipLaunchKernelGGL(k_scale<double>, dim3(g), dim3(BS), 0, ctx->stream, N, (double*)y, beta);
        }
        if (nnz > 0) {
            if (dtype == SABLE_F32)
                hipLaunchKernelGGL(k_coo_atomic<float>, dim3(gz), dim3(BS), 0, ctx->stream, nnz, row, col,
                                   (const float*)val, (const float*)x, (float*)y);
            else
                hipLaunchKernelGGL(k_coo_atomic<double>, dim3(gz), dim3(BS), 0, ctx->stream, nnz, row, col,
                                   (const double*)val, (const double*)x, (double*)y);
        }
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

// Work space of a device-built plan: the shape {C, T, cap}, the error flag
// and cap + 1 block boundaries.
constexpr size_t devplan_bytes(int64_t cap) { return 256 + sizeof(int32_t) * (size_t)(cap + 1); }
// Block capacity from N alone: matrices whose default-weight plan would
// need more blocks get proportionally heavier blocks.
constexpr int64_t devplan_cap(int32_t N) { return (int64_t)N / 2 + 1024; }

// y = A.x + beta*y, inner_max times, with the block plan built on the
// device into `work` (devplan_bytes(cap) bytes): no host synchronisation
// and no allocation.  A malformed row_ptr leaves y unchanged.
int devplan_spmv(hipStream_t st, int num_cu, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                 const int32_t* col, const void* val, const void* x, void* y, double beta, int inner_max,
                 void* work, int64_t cap) {
    int T, C;
    tune_defaults(nullptr, &T, &C);
    int64_t* shape = (int64_t*)work;
    int* err = (int*)((char*)work + 64);
    int32_t* blk = (int32_t*)((char*)work + 256);
    // shape and flag are written by a kernel (no memset node in captured work)
    hipLaunchKernelGGL(k_devplan_init, dim3(1), dim3(64), 0, st, shape, (int64_t)C, (int64_t)T, cap, err);
    hipLaunchKernelGGL(k_plan_fill_dev, dim3(grid_for((int64_t)N + 1, BS)), dim3(BS), 0, st, N, row_ptr, (int64_t)C,
                       (int64_t)T, cap, blk, err);
    SABLE_LAUNCH_CHECK();
    const int lds_cap = lds_cap_for(T);
    int64_t grid = (int64_t)num_cu * 8;
    if (grid > cap) grid = cap;
    if (grid < NUM_XCD) grid = NUM_XCD;
    for (int it = 0; it < inner_max; it++) {
        if (dtype == SABLE_F32)
            hipLaunchKernelGGL((k_csr_persist<float, 8>), dim3((unsigned)grid), dim3(BS), (size_t)lds_cap * 4, st, blk,
                               (int64_t)0, row_ptr, col, (const float*)val, (const float*)x, (float*)y, (float)beta,
                               lds_cap, N, (const int64_t*)shape, (const int*)err);
        else
            hipLaunchKernelGGL((k_csr_persist<double, 8>), dim3((unsigned)grid), dim3(BS), (size_t)lds_cap * 8, st,
                               blk, (int64_t)0, row_ptr, col, (const double*)val, (const double*)x, (double*)y, beta,
                               lds_cap, N, (const int64_t*)shape, (const int*)err);
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

// The allocation-free product's work space: the device-built plan of a
// matrix of up to SABLE_NOALLOC_MAX_ROWS rows, one copy per GPU for the
// life of the process.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): it uses the
// device address that hipGetSymbolAddress returns on each device.
static_assert(devplan_bytes(devplan_cap(SABLE_NOALLOC_MAX_ROWS)) <= SABLE_NOALLOC_WORK_BYTES,
              "allocation-free work space too small for its row bound");
static __device__ __align__(256) unsigned char scratch[(SABLE_NOALLOC_WORK_BYTES + 255) / 256 * 256];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

int spmv_noalloc_setup(int device, void** out) {
    *out = nullptr;
    SABLE_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, SABLE_EINVAL, "device %d: at most %d devices", device,
                NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            return fail(SABLE_EHIP, "hipGetSymbolAddress(work space) on device %d: %s", device,
                        hipGetErrorString(e));
        }
        g_noalloc_ptr[device] = p;
    }
    *out = g_noalloc_ptr[device];
    return SABLE_OK;
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_csr_plan_create(sable_ctx_t* ctx, int32_t N, const int32_t* row_ptr,
                                    const sable_csr_tune_t* tune, sable_csr_plan_t** out) {
    SABLE_CHECK(ctx && out && row_ptr, SABLE_EINVAL, "sable_csr_plan_create: NULL argument");
    SABLE_CHECK(N >= 0, SABLE_EINVAL, "negative N");
    *out = nullptr;
    int T, C;
    tune_defaults(tune, &T, &C);
    int64_t nnz = 0;
    int rc = read_nnz(ctx, N, row_ptr, &nnz);
    if (rc) return rc;
    int64_t B = nblocks_for(nnz, N, T, C);
    auto* p = new sable_csr_plan_s();
    p->ctx = ctx;
    p->N = N;
    p->nnz = nnz;
    p->row_ptr = row_ptr;
    p->nblocks = B;
    p->block_weight = T;
    p->row_weight = C;
    p->lds_cap = lds_cap_for(T);
    if (tune && tune->vector_lanes > 0) {
        const int G = tune->vector_lanes;
        if (G != 2 && G != 4 && G != 8 && G != 16 && G != 32 && G != 64) {
            delete p;
            return fail(SABLE_EINVAL, "vector_lanes must be 2, 4, 8, 16, 32 or 64");
        }
        p->vector_lanes = G;
        const int R = tune->vector_rows > 0 ? tune->vector_rows : 4;
        if (R != 1 && R != 2 && R != 4 && R != 8 && R != 16) {
            delete p;
            return fail(SABLE_EINVAL, "vector_rows must be 1, 2, 4, 8 or 16");
        }
        p->vector_rows = R;
    }
    if (hipMalloc(&p->blk, sizeof(int32_t) * (size_t)(B + 1) + 256) != hipSuccess) {
        (void)hipGetLastError();
        delete p;
        return fail(SABLE_ENOMEM, "plan allocation failed");
    }
    p->owns_blk = true;
    int* err = (int*)((char*)p->blk + sizeof(int32_t) * (size_t)(B + 1));
    err = (int*)(((uintptr_t)err + 3) & ~(uintptr_t)3);
    if (N > 0) {
        rc = plan_fill(ctx, N, row_ptr, T, C, B, p->blk, err, true);
        if (rc) {
            (void)hipFree(p->blk);
            delete p;
            return rc;
        }
    }
    *out = p;
    return SABLE_OK;
}

SABLE_API void sable_csr_plan_destroy(sable_csr_plan_t* p) {
    if (!p) return;
    if (p->owns_blk && p->blk) {
        (void)hipStreamSynchronize(p->ctx->stream);
        (void)hipFree(p->blk);
    }
    delete p;
}

SABLE_API int sable_csr_plan_info(const sable_csr_plan_t* p, int64_t* nblocks, int64_t* nnz, int32_t* lds_cap) {
    SABLE_CHECK(p, SABLE_EINVAL, "NULL plan");
    if (nblocks) *nblocks = p->nblocks;
    if (nnz) *nnz = p->nnz;
    if (lds_cap) *lds_cap = p->lds_cap;
    return SABLE_OK;
}

SABLE_API int sable_csr_kernel_select(int kernel) {
    SABLE_CHECK(kernel >= 0 && kernel <= 3, SABLE_EINVAL, "kernel must be 0, 1, 2 or 3");
    g_csr_kernel = kernel;
    return SABLE_OK;
}

SABLE_API int sable_csr_plan_blocks(const sable_csr_plan_t* p, const int32_t** d_blk, int64_t* nblocks) {
    SABLE_CHECK(p && d_blk && nblocks, SABLE_EINVAL, "NULL argument");
    *d_blk = p->blk;
    *nblocks = p->nblocks;
    return SABLE_OK;
}

SABLE_API int sable_spmv_csr_plan(sable_ctx_t* ctx, const sable_csr_plan_t* p, sable_dtype_t dtype,
                                  const int32_t* col, const void* val, const void* x, void* y, double beta,
                                  int inner_max) {
    SABLE_CHECK(ctx && p, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(inner_max >= 0, SABLE_EINVAL, "negative inner_max");
    if (p->N == 0 || inner_max == 0) return SABLE_OK;
    if (p->vector_lanes > 0) {
        SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
        for (int it = 0; it < inner_max; it++) {
            if (dtype == SABLE_F32)
                vec_launch_t<float>(ctx->stream, p->vector_lanes, p->vector_rows, p->N, p->row_ptr, col, val, x, y,
                                    beta);
            else
                vec_launch_t<double>(ctx->stream, p->vector_lanes, p->vector_rows, p->N, p->row_ptr, col, val, x, y,
                                     beta);
            SABLE_LAUNCH_CHECK();
        }
        return SABLE_OK;
    }
    return dispatch_csr(ctx, dtype, p->blk, p->nblocks, p->row_ptr, col, val, x, y, beta, inner_max, p->lds_cap);
}

SABLE_API int sable_spmv_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                             const int32_t* col, const void* val, const void* x, void* y, double beta,
                             int inner_max) {
    SABLE_CHECK(ctx && row_ptr, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (N == 0 || inner_max == 0) return SABLE_OK;
    const int64_t cap = devplan_cap(N);
    void* s;
    int rc = ctx_scratch(ctx, devplan_bytes(cap), &s);
    if (rc) return rc;
    return devplan_spmv(ctx->stream, ctx->num_cu, dtype, N, row_ptr, col, val, x, y, beta, inner_max, s, cap);
}

SABLE_API int sable_spmv_csr_noalloc(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                     const int32_t* col, const void* val, const void* x, void* y, double beta,
                                     int inner_max) {
    SABLE_CHECK(ctx && row_ptr, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    SABLE_CHECK(N <= SABLE_NOALLOC_MAX_ROWS, SABLE_EINVAL,
                "sable_spmv_csr_noalloc: N = %d rows exceeds the %d its fixed work space holds", N,
                SABLE_NOALLOC_MAX_ROWS);
    SABLE_CHECK(ctx->noalloc, SABLE_EINVAL, "sable_spmv_csr_noalloc: no work space on device %d", ctx->device);
    if (N == 0 || inner_max == 0) return SABLE_OK;
    return devplan_spmv(ctx->stream, ctx->num_cu, dtype, N, row_ptr, col, val, x, y, beta, inner_max, ctx->noalloc,
                        devplan_cap(N));
}

SABLE_API int sable_spmv_coo(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int64_t nnz, const int32_t* row,
                             const int32_t* col, const void* val, const void* x, void* y, double beta,
                             int inner_max) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "NULL ctx");
    SABLE_CHECK(N >= 0 && nnz >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (N == 0 || inner_max == 0) return SABLE_OK;
    int T, C;
    tune_defaults(nullptr, &T, &C);
    int64_t B = nblocks_for(nnz, N, T, C);
    size_t rp_bytes = (sizeof(int32_t) * ((size_t)N + 1) + 255) & ~(size_t)255;
    size_t blk_bytes = (sizeof(int32_t) * ((size_t)B + 1) + 255) & ~(size_t)255;
    void* s;
    int rc = ctx_scratch(ctx, rp_bytes + blk_bytes + 256, &s);
    if (rc) return rc;
    int32_t* rp = (int32_t*)s;
    int32_t* blk = (int32_t*)((char*)s + rp_bytes);
    int* err = (int*)((char*)s + rp_bytes + blk_bytes);
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
    if (nnz > 0) {
        unsigned g = grid_for(nnz, BS);
        if (g > 2048) g = 2048;
        hipLaunchKernelGGL(k_coo_check, dim3(g), dim3(BS), 0, ctx->stream, nnz, N, row, err);
        SABLE_LAUNCH_CHECK();
    }
    int64_t e;
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    e = *(int*)ctx->pinned;
    SABLE_CHECK(!(e & 2), SABLE_EMALFORMED, "COO row index out of range");
    if (e & 1) return coo_atomic_launch(ctx, dtype, N, nnz, row, col, val, x, y, beta, inner_max);
    hipLaunchKernelGGL(k_coo_heads, dim3(grid_for(nnz + 1, BS)), dim3(BS), 0, ctx->stream, nnz, N, row, rp);
    SABLE_LAUNCH_CHECK();
    rc = plan_fill(ctx, N, rp, T, C, B, blk, err, false);
    if (rc) return rc;
    return dispatch_csr(ctx, dtype, blk, B, rp, col, val, x, y, beta, inner_max, lds_cap_for(T));
}

SABLE_API int sable_coo_plan_create(sable_ctx_t* ctx, int32_t N, int64_t nnz, const int32_t* row,
                                    sable_coo_plan_t** out) {
    SABLE_CHECK(ctx && out && (row || nnz == 0), SABLE_EINVAL, "sable_coo_plan_create: NULL argument");
    SABLE_CHECK(N >= 0 && nnz >= 0, SABLE_EINVAL, "negative size");
    *out = nullptr;
    auto* p = new sable_coo_plan_s();
    p->N = N;
    p->nnz = nnz;
    int T, C;
    tune_defaults(nullptr, &T, &C);
    p->lds_cap = lds_cap_for(T);
    int* err = nullptr;
    int rc = SABLE_OK;
    if (hipMalloc(&err, 256) != hipSuccess) {
        (void)hipGetLastError();
        delete p;
        return fail(SABLE_ENOMEM, "COO plan allocation failed");
    }
    auto done = [&](int code) {
        (void)hipFree(err);
        if (code) sable_coo_plan_destroy(p);
        else *out = p;
        return code;
    };
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
    if (nnz > 0) {
        unsigned g = grid_for(nnz, BS);
        if (g > 2048) g = 2048;
        hipLaunchKernelGGL(k_coo_check, dim3(g), dim3(BS), 0, ctx->stream, nnz, N, row, err);
        if (hipGetLastError() != hipSuccess) return done(fail(SABLE_EHIP, "COO check launch failed"));
    }
    if (hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return done(fail(SABLE_EHIP, "COO check failed"));
    const int e = *(int*)ctx->pinned;
    if (e & 2) return done(fail(SABLE_EMALFORMED, "COO row index out of range"));
    p->sorted = !(e & 1);
    if (p->sorted && N > 0) {
        p->nblocks = nblocks_for(nnz, N, T, C);
        if (hipMalloc(&p->rp, 4 * ((size_t)N + 1)) != hipSuccess ||
            hipMalloc(&p->blk, 4 * ((size_t)p->nblocks + 1)) != hipSuccess) {
            (void)hipGetLastError();
            return done(fail(SABLE_ENOMEM, "COO plan allocation failed"));
        }
        hipLaunchKernelGGL(k_coo_heads, dim3(grid_for(nnz + 1, BS)), dim3(BS), 0, ctx->stream, nnz, N, row, p->rp);
        if (hipGetLastError() != hipSuccess) return done(fail(SABLE_EHIP, "COO heads launch failed"));
        rc = plan_fill(ctx, N, p->rp, T, C, p->nblocks, p->blk, err, false);
        if (rc) return done(rc);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return done(fail(SABLE_EHIP, "COO plan failed"));
    }
    return done(SABLE_OK);
}

SABLE_API void sable_coo_plan_destroy(sable_coo_plan_t* p) {
    if (!p) return;
    if (p->rp) (void)hipFree(p->rp);
    if (p->blk) (void)hipFree(p->blk);
    delete p;
}

SABLE_API int sable_coo_plan_info(const sable_coo_plan_t* p, int* sorted, int64_t* nblocks) {
    SABLE_CHECK(p, SABLE_EINVAL, "NULL plan");
    if (sorted) *sorted = p->sorted ? 1 : 0;
    if (nblocks) *nblocks = p->nblocks;
    return SABLE_OK;
}

SABLE_API int sable_coo_plan_row_ptr(const sable_coo_plan_t* p, const int32_t** d_row_ptr) {
    SABLE_CHECK(p && d_row_ptr, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(p->sorted && p->rp, SABLE_EINVAL, "the COO rows are not sorted: no row heads");
    *d_row_ptr = p->rp;
    return SABLE_OK;
}

SABLE_API int sable_spmv_coo_plan(sable_ctx_t* ctx, const sable_coo_plan_t* p, sable_dtype_t dtype,
                                  const int32_t* row, const int32_t* col, const void* val, const void* x, void* y,
                                  double beta, int inner_max) {
    SABLE_CHECK(ctx && p, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(inner_max >= 0, SABLE_EINVAL, "negative inner_max");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (p->N == 0 || inner_max == 0) return SABLE_OK;
    if (p->sorted) return dispatch_csr(ctx, dtype, p->blk, p->nblocks, p->rp, col, val, x, y, beta, inner_max, p->lds_cap);
    return coo_atomic_launch(ctx, dtype, p->N, p->nnz, row, col, val, x, y, beta, inner_max);
}

SABLE_API int sable_spmv_dia_range(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                                   const int32_t* offset, const void* data, int32_t r0, int32_t r1, const void* x,
                                   void* y, double beta, int inner_max) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "NULL ctx");
    SABLE_CHECK(N >= 0 && nd >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(stride >= 0 && stride <= N, SABLE_EINVAL, "stride %d outside [0, N=%d]", stride, N);
    SABLE_CHECK(r0 >= 0 && r0 <= r1 && r1 <= N, SABLE_EINVAL, "row range [%d, %d) outside [0, %d)", r0, r1, N);
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (r0 == r1 || inner_max == 0) return SABLE_OK;
    unsigned g = grid_for(r1 - r0, BS);
    // with stride 0 no diagonal stores an entry (and the kernel's clamped
    // loads need stride >= 1): y = beta*y
    const int32_t ndk = stride > 0 ? nd : 0;
    for (int it = 0; it < inner_max; it++) {
        if (dtype == SABLE_F32)
            hipLaunchKernelGGL(k_dia<float>, dim3(g), dim3(BS), 0, ctx->stream, N, ndk, stride, offset,
                               (const float*)data, (const float*)x, (float*)y, (float)beta, r0, r1);
        else
            hipLaunchKernelGGL(k_dia<double>, dim3(g), dim3(BS), 0, ctx->stream, N, ndk, stride, offset,
                               (const double*)data, (const double*)x, (double*)y, beta, r0, r1);
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

SABLE_API int sable_spmv_dia(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd, int32_t stride,
                             const int32_t* offset, const void* data, const void* x, void* y, double beta,
                             int inner_max) {
    return sable_spmv_dia_range(ctx, dtype, N, nd, stride, offset, data, 0, N, x, y, beta, inner_max);
}

SABLE_API int sable_spmv_ell_range(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                                   const int32_t* indices, const void* data, int32_t r0, int32_t r1, const void* x,
                                   void* y, double beta, int inner_max) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "NULL ctx");
    SABLE_CHECK(N >= 0 && nc >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(r0 >= 0 && r0 <= r1 && r1 <= N, SABLE_EINVAL, "row range [%d, %d) outside [0, %d)", r0, r1, N);
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    // The reference returns without touching y when nc or N is not positive.
    if (r0 == r1 || nc == 0 || inner_max == 0) return SABLE_OK;
    unsigned g = grid_for(r1 - r0, BS);
    for (int it = 0; it < inner_max; it++) {
        if (dtype == SABLE_F32)
            hipLaunchKernelGGL(k_ell<float>, dim3(g), dim3(BS), 0, ctx->stream, N, nc, indices, (const float*)data,
                               (const float*)x, (float*)y, (float)beta, r0, r1);
        else
            hipLaunchKernelGGL(k_ell<double>, dim3(g), dim3(BS), 0, ctx->stream, N, nc, indices, (const double*)data,
                               (const double*)x, (double*)y, beta, r0, r1);
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

SABLE_API int sable_spmv_ell(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc, const int32_t* indices,
                             const void* data, const void* x, void* y, double beta, int inner_max) {
    return sable_spmv_ell_range(ctx, dtype, N, nc, indices, data, 0, N, x, y, beta, inner_max);
}

SABLE_API int sable_spmv_dia_rowmajor(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nd,
                                      const int32_t* offset, const void* data, int32_t r0, int32_t r1,
                                      const void* x, void* y, int inner_max) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "NULL ctx");
    SABLE_CHECK(N >= 0 && nd >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(r0 >= 0 && r1 <= N, SABLE_EINVAL, "row range [%d, %d) outside [0, %d)", r0, r1, N);
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    // the reference returns at once for an empty range or no diagonals
    if (r0 >= r1 || nd == 0 || inner_max == 0) return SABLE_OK;
    unsigned g = grid_for(r1 - r0, BS);
    for (int it = 0; it < inner_max; it++) {
        if (dtype == SABLE_F32)
            hipLaunchKernelGGL(k_dia_rm<float>, dim3(g), dim3(BS), 0, ctx->stream, N, nd, offset, (const float*)data,
                               (const float*)x, (float*)y, r0, r1);
        else
            hipLaunchKernelGGL(k_dia_rm<double>, dim3(g), dim3(BS), 0, ctx->stream, N, nd, offset,
                               (const double*)data, (const double*)x, (double*)y, r0, r1);
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

SABLE_API int sable_spmv_ell_rowmajor(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nc,
                                      const int32_t* indices, const void* data, int32_t r0, int32_t r1,
                                      const void* x, void* y, int order, int inner_max) {
    SABLE_CHECK(ctx, SABLE_EINVAL, "NULL ctx");
    SABLE_CHECK(N >= 0 && nc >= 0 && inner_max >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(r0 >= 0 && r1 <= N, SABLE_EINVAL, "row range [%d, %d) outside [0, %d)", r0, r1, N);
    SABLE_CHECK(order == 0 || order == 1, SABLE_EINVAL, "order must be 0 (plain) or 1 (gs)");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (r0 >= r1 || nc == 0 || inner_max == 0) return SABLE_OK;
    unsigned g = grid_for(r1 - r0, BS);
    for (int it = 0; it < inner_max; it++) {
        if (dtype == SABLE_F32) {
            if (order)
                hipLaunchKernelGGL((k_ell_rm<float, true>), dim3(g), dim3(BS), 0, ctx->stream, nc, indices,
                                   (const float*)data, (const float*)x, (float*)y, r0, r1);
            else
                hipLaunchKernelGGL((k_ell_rm<float, false>), dim3(g), dim3(BS), 0, ctx->stream, nc, indices,
                                   (const float*)data, (const float*)x, (float*)y, r0, r1);
        } else {
            if (order)
                hipLaunchKernelGGL((k_ell_rm<double, true>), dim3(g), dim3(BS), 0, ctx->stream, nc, indices,
                                   (const double*)data, (const double*)x, (double*)y, r0, r1);
            else
                hipLaunchKernelGGL((k_ell_rm<double, false>), dim3(g), dim3(BS), 0, ctx->stream, nc, indices,
                                   (const double*)data, (const double*)x, (double*)y, r0, r1);
        }
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

}  // extern "C"
