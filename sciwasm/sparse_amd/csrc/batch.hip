// batch.hip -- many independent matrices in one persistent launch.
//
// SpMV: every matrix brings its own row-block schedule (sable_csr_plan_t).
// The batch's blocks form one global work list (item k owns global blocks
// [blk_begin_k, blk_begin_k + nblocks_k)); a persistent grid of workgroups
// claims them BATCH_CH at a time from a device counter, finds each block's
// item by binary search and runs the block with the single-matrix CSR
// block arithmetic (products staged in LDS, each row summed in order).
// Items are ordered by the caller (the server puts blocks of long rows
// first).
// Up to 8 right-hand sides share each matrix read: x and y are row-major
// [N][nrhs], so one gather of x[col] fetches all right-hand sides.
//
// SpTS: the batch's rows form one global row space (item k owns rows
// [row_begin_k, row_begin_k + N_k)); waves claim consecutive global rows
// from a ticket, so the sync-free forward-progress argument of the single
// solve holds for the whole batch.  The refined variant adds batched FP64
// residual and update kernels.
#include "common.h"
#include "batch.h"
#include "refine.h"

struct sable_batch_level_plan_s {
    sable_ctx_t* ctx = nullptr;
    int lanes = 4;
    int nitems = 0;
    int64_t rows = 0;
    int64_t nchunks = 0;
    int32_t* d_rows = nullptr;           // every item's rows listed by level
    sable::RowChunk* d_chunks = nullptr;  // the chunks of every level, level after level
    sable::LevelPlan plan;               // level l: chunks [chunk_lo[l], chunk_lo[l + 1])
};

namespace sable {

constexpr int BB = 256;

__device__ __forceinline__ int find_item(const int64_t* __restrict__ begin, int n, int64_t g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (begin[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// LDS staging per workgroup: 2048 nonzeros x nrhs products, at most 8192.
template <int NR>
struct Lds {
    static constexpr int ELEMS = (2048 * NR < 8192) ? 2048 * NR : 8192;
};

// The NR values of row c of a row-major [N][NR] x: 16-byte vector loads when
// a row is a multiple of 16 bytes (rows are then 16-byte aligned, since x is),
// else one 8- or 4-byte load.
template <typename T, int NR>
__device__ __forceinline__ void load_xrow(const T* __restrict__ x, int c, T (&out)[NR]) {
    constexpr int BYTES = NR * (int)sizeof(T);
    const char* p = (const char*)(x + (int64_t)c * NR);
    if constexpr (BYTES % 16 == 0) {
#pragma unroll
        for (int q = 0; q < BYTES / 16; q++) {
            const uint4 w = *(const uint4*)(p + 16 * q);
            __builtin_memcpy((char*)out + 16 * q, &w, 16);
        }
    } else if constexpr (BYTES == 8) {
        const uint2 w = *(const uint2*)p;
        __builtin_memcpy(out, &w, 8);
    } else {
        out[0] = *(const T*)p;
    }
}

// Loads in flight per thread in the staged path: fewer for wide rows of x,
// whose values already fill the registers.
template <typename T, int NR>
struct Unroll {
    static constexpr int U = NR * (int)sizeof(T) >= 32 ? 2 : (NR * (int)sizeof(T) >= 16 ? 4 : 8);
};

// Pipelined path for a block whose entries fit one step (every thread
// loads at most U of them and the products fit the LDS staging buffer):
// the entries (and the first row extents each thread reduces) are loaded
// into a Stage, the products are formed into LDS, and the next block's
// Stage is loaded before this block's rows are reduced, so that its loads
// are in flight during the barrier and the reduction.
template <typename T, int NR>
struct Stage {
    static constexpr int U = Unroll<T, NR>::U;
    int c[U];
    T v[U];
    int a, e;  // entry range (block-relative) of the thread's first reduced row
};

template <typename T, int NR>
__device__ __forceinline__ bool one_step(int nnzb) {
    return nnzb <= Stage<T, NR>::U * BB && nnzb * NR <= Lds<NR>::ELEMS;
}

// lanes per row of the reduction: the widest power of two <= WAVE that
// gives every row of the block a group of lanes in one pass, else 1
__device__ __forceinline__ int reduce_lanes(int R) {
    int L = WAVE;
    while (L > 1 && L * R > BB) L >>= 1;
    return L;
}

template <typename T, int NR>
__device__ __forceinline__ void stage_load(Stage<T, NR>& st, const sable_batch_spmv_item_t& it, int r0, int r1,
                                           int64_t nz0, int nnzb) {
    const int tid = threadIdx.x;
    const int32_t* __restrict__ col = it.col;
    const T* __restrict__ val = (const T*)it.val;
#pragma unroll
    for (int u = 0; u < Stage<T, NR>::U; u++) {
        const int kk = tid + u * BB;
        st.c[u] = kk < nnzb ? col[nz0 + kk] : 0;
        st.v[u] = kk < nnzb ? val[nz0 + kk] : T(0);
    }
    const int L = reduce_lanes(r1 - r0);
    const int row = r0 + tid / L;
    st.a = row < r1 ? (int)((int64_t)it.row_ptr[row] - nz0) : 0;
    st.e = row < r1 ? (int)((int64_t)it.row_ptr[row + 1] - nz0) : 0;
}

template <typename T, int NR>
__device__ __forceinline__ void stage_prod(const Stage<T, NR>& st, const sable_batch_spmv_item_t& it, int nnzb,
                                           T* prod) {
    constexpr int U = Stage<T, NR>::U;
    const int tid = threadIdx.x;
    const T* __restrict__ x = (const T*)it.x;
    T xr[U][NR];
#pragma unroll
    for (int u = 0; u < U; u++)
        if (tid + u * BB < nnzb) load_xrow<T, NR>(x, st.c[u], xr[u]);
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int kk = tid + u * BB;
        if (kk < nnzb) {
#pragma unroll
            for (int r = 0; r < NR; r++) prod[kk * NR + r] = st.v[u] * xr[u][r];
        }
    }
}

template <typename T, int NR>
__device__ __forceinline__ void stage_reduce(const sable_batch_spmv_item_t& it, int r0, int r1, int64_t nz0, int a0,
                                             int e0, const T* prod) {
    const int tid = threadIdx.x;
    const int R = r1 - r0;
    const int L = reduce_lanes(R);
    const int groups = BB / L, g = tid / L, l = tid & (L - 1);
    T* __restrict__ y = (T*)it.y;
    const T beta = (T)it.beta;
    for (int rb = 0; rb < R; rb += groups) {
        const int row = r0 + rb + g;
        T s[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) s[r] = T(0);
        int a = 0, e = 0;
        if (row < r1) {
            if (rb == 0) {
                a = a0;
                e = e0;
            } else {
                a = (int)((int64_t)it.row_ptr[row] - nz0);
                e = (int)((int64_t)it.row_ptr[row + 1] - nz0);
            }
            for (int kk = a + l; kk < e; kk += L)
#pragma unroll
                for (int r = 0; r < NR; r++) s[r] += prod[kk * NR + r];
        }
#pragma unroll
        for (int r = 0; r < NR; r++) s[r] = group_sum(s[r], L);
        if (l == 0 && row < r1) {
#pragma unroll
            for (int r = 0; r < NR; r++) {
                T* yp = &y[(int64_t)row * NR + r];
                if (beta == T(0)) *yp = s[r];
                else if (beta == T(1)) { if (e > a) *yp = *yp + s[r]; }
                else *yp = beta * *yp + s[r];
            }
        }
    }
}

// One row block [r0, r1) of an item whose entries are [nz0, nz0 + nnzb).
template <typename T, int NR>
__device__ __forceinline__ void block_spmv(const sable_batch_spmv_item_t& it, int r0, int r1, int64_t nz0, int nnzb,
                                           T* prod, T* red) {
    constexpr int LDS_ELEMS = Lds<NR>::ELEMS;
    const int32_t* __restrict__ rp = it.row_ptr;
    const int32_t* __restrict__ col = it.col;
    const T* __restrict__ val = (const T*)it.val;
    const T* __restrict__ x = (const T*)it.x;
    T* __restrict__ y = (T*)it.y;
    const T beta = (T)it.beta;
    if (r0 >= r1) return;
    const int tid = threadIdx.x;
    const int R = r1 - r0;
    if (nnzb * NR <= LDS_ELEMS) {
        constexpr int U = Unroll<T, NR>::U;
        // all loads of a step are issued before any product is formed, with
        // out-of-range slots predicated off
        for (int k = tid; k < nnzb; k += U * BB) {
            int c[U];
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int kk = k + u * BB;
                c[u] = kk < nnzb ? col[nz0 + kk] : 0;
                v[u] = kk < nnzb ? val[nz0 + kk] : T(0);
            }
            T xr[U][NR];
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (k + u * BB < nnzb) load_xrow<T, NR>(x, c[u], xr[u]);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int kk = k + u * BB;
                if (kk < nnzb) {
#pragma unroll
                    for (int r = 0; r < NR; r++) prod[kk * NR + r] = v[u] * xr[u][r];
                }
            }
        }
        __syncthreads();
        int L = WAVE;
        while (L > 1 && L * R > BB) L >>= 1;
        const int groups = BB / L, g = tid / L, l = tid & (L - 1);
        for (int rb = 0; rb < R; rb += groups) {
            const int row = r0 + rb + g;
            T s[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) s[r] = T(0);
            int a = 0, e = 0;
            if (row < r1) {
                a = (int)((int64_t)rp[row] - nz0);
                e = (int)((int64_t)rp[row + 1] - nz0);
                for (int kk = a + l; kk < e; kk += L)
#pragma unroll
                    for (int r = 0; r < NR; r++) s[r] += prod[kk * NR + r];
            }
#pragma unroll
            for (int r = 0; r < NR; r++) s[r] = group_sum(s[r], L);
            if (l == 0 && row < r1) {
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    T* yp = &y[(int64_t)row * NR + r];
                    if (beta == T(0)) *yp = s[r];
                    else if (beta == T(1)) { if (e > a) *yp = *yp + s[r]; }
                    else *yp = beta * *yp + s[r];
                }
            }
        }
        __syncthreads();  // prod is reused by the next block of this workgroup
    } else if (R > 1) {
        // several rows whose products do not fit the LDS: a group of L
        // lanes per row, L the power of two (at most a wave) at or above
        // the block's mean row length
        int L = 1;
        while (L < WAVE && L * R < nnzb) L <<= 1;
        const int grp = tid / L, lane = tid & (L - 1);
        for (int row = r0 + grp; row < r1; row += BB / L) {
            const int64_t a = rp[row], e = rp[row + 1];
            T s[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) s[r] = T(0);
            for (int64_t k = a + lane; k < e; k += L) {
                const int c = col[k];
                const T v = val[k];
                T xr[NR];
                load_xrow<T, NR>(x, c, xr);
#pragma unroll
                for (int r = 0; r < NR; r++) s[r] += v * xr[r];
            }
#pragma unroll
            for (int r = 0; r < NR; r++) s[r] = group_sum(s[r], L);
            if (lane == 0) {
#pragma unroll
                for (int r = 0; r < NR; r++) {
                    T* yp = &y[(int64_t)row * NR + r];
                    if (beta == T(0)) *yp = s[r];
                    else if (beta == T(1)) { if (e > a) *yp = *yp + s[r]; }
                    else *yp = beta * *yp + s[r];
                }
            }
        }
    } else {
        // one long row: the whole workgroup
        for (int row = r0; row < r1; row++) {
            const int64_t a = rp[row], e = rp[row + 1];
            T s[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) s[r] = T(0);
            for (int64_t k = a + tid; k < e; k += BB) {
                const int c = col[k];
                const T v = val[k];
                T xr[NR];
                load_xrow<T, NR>(x, c, xr);
#pragma unroll
                for (int r = 0; r < NR; r++) s[r] += v * xr[r];
            }
#pragma unroll
            for (int r = 0; r < NR; r++) {
                T v = wave_sum(s[r]);
                __syncthreads();
                if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = v;
                __syncthreads();
                if (tid == 0) {
                    T t = 0;
                    for (int w = 0; w < BB / WAVE; w++) t += red[w];
                    T* yp = &y[(int64_t)row * NR + r];
                    if (beta == T(0)) *yp = t;
                    else if (beta == T(1)) { if (e > a) *yp = *yp + t; }
                    else *yp = beta * *yp + t;
                }
            }
        }
        __syncthreads();
    }
}

// The items' block begins are searched in LDS (each workgroup copies them
// once) when they fit beside the staging buffer within the default 64 KiB
// of LDS per workgroup (lds_items = nitems); otherwise (lds_items = 0)
// items[] is searched in global memory.
constexpr int BATCH_LDS_ITEMS = 4096;
// Blocks claimed per counter increment: the first BATCH_CH threads resolve
// one block each (owning item, row range, entry range) in parallel, so the
// chain of dependent loads behind a block is paid once per claim.
constexpr int BATCH_CH = 8;

template <typename T, int NR>
__global__ __launch_bounds__(BB) void k_batch_spmv(const sable_batch_spmv_item_t* __restrict__ items, int nitems,
                                                   int64_t total, unsigned long long* __restrict__ counter,
                                                   int lds_items) {
    __shared__ __align__(16) T prod[Lds<NR>::ELEMS];
    __shared__ T red[BB / WAVE];
    __shared__ long long next;
    __shared__ sable_batch_spmv_item_t s_it[BATCH_CH];
    __shared__ int s_r0[BATCH_CH], s_r1[BATCH_CH], s_nnz[BATCH_CH];
    __shared__ int64_t s_nz0[BATCH_CH];
    extern __shared__ __align__(16) int64_t begins[];  // lds_items entries
    const bool in_lds = lds_items == nitems;
    if (in_lds)
        for (int k = threadIdx.x; k < nitems; k += BB) begins[k] = items[k].blk_begin;
    // the next claim is made while the current one runs
    long long ahead = 0;
    if (threadIdx.x == 0) ahead = (long long)atomicAdd(counter, 1ull);
    __syncthreads();
    while (true) {
        if (threadIdx.x == 0) next = ahead;
        __syncthreads();
        const int64_t g0 = next * BATCH_CH;
        if (g0 >= total) break;
        if (threadIdx.x == 0) ahead = (long long)atomicAdd(counter, 1ull);
        const int cnt = (int)min((int64_t)BATCH_CH, total - g0);
        if (threadIdx.x < cnt) {
            const int64_t g = g0 + threadIdx.x;
            // items[].blk_begin is ascending: the owning item is the last
            // one that begins at or before g
            int lo = 0, hi = nitems - 1;
            if (in_lds) {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (begins[mid] <= g) lo = mid;
                    else hi = mid - 1;
                }
            } else {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (items[mid].blk_begin <= g) lo = mid;
                    else hi = mid - 1;
                }
            }
            const sable_batch_spmv_item_t it = items[lo];
            const int64_t b = g - it.blk_begin;
            const int r0 = it.blk[b], r1 = it.blk[b + 1];
            const int64_t nz0 = it.row_ptr[r0];
            s_it[threadIdx.x] = it;
            s_r0[threadIdx.x] = r0;
            s_r1[threadIdx.x] = r1;
            s_nz0[threadIdx.x] = nz0;
            s_nnz[threadIdx.x] = (int)((int64_t)it.row_ptr[r1] - nz0);
        }
        __syncthreads();
        Stage<T, NR> st;
        bool staged = false;  // st holds block q's loads (uniform over the workgroup)
        for (int q = 0; q < cnt; q++) {
            const int r0 = s_r0[q], r1 = s_r1[q], nnzb = s_nnz[q];
            const int64_t nz0 = s_nz0[q];
            if (r0 >= r1) continue;
            if (!one_step<T, NR>(nnzb)) {
                block_spmv<T, NR>(s_it[q], r0, r1, nz0, nnzb, prod, red);
                continue;
            }
            const sable_batch_spmv_item_t it = s_it[q];
            if (!staged) stage_load<T, NR>(st, it, r0, r1, nz0, nnzb);
            stage_prod<T, NR>(st, it, nnzb, prod);
            const int a0 = st.a, e0 = st.e;
            staged = false;
            if (q + 1 < cnt && s_r0[q + 1] < s_r1[q + 1] && one_step<T, NR>(s_nnz[q + 1])) {
                stage_load<T, NR>(st, s_it[q + 1], s_r0[q + 1], s_r1[q + 1], s_nz0[q + 1], s_nnz[q + 1]);
                staged = true;
            }
            __syncthreads();
            stage_reduce<T, NR>(it, r0, r1, nz0, a0, e0, prod);
            __syncthreads();  // prod is rewritten by the next block
        }
        __syncthreads();  // the claim's descriptors are rewritten by the next claim
    }
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// ------------------------------------------------------------- batched SpTS
constexpr uint64_t BSENT64 = 0x7FF4DEADBEEF0001ull;
constexpr uint32_t BSENT32 = 0x7FA0BEEFu;

__device__ __forceinline__ double bsent(double) { return __longlong_as_double((long long)BSENT64); }
__device__ __forceinline__ float bsent(float) { return __uint_as_float(BSENT32); }
__device__ __forceinline__ bool is_bsent(double v) { return (uint64_t)__double_as_longlong(v) == BSENT64; }
__device__ __forceinline__ bool is_bsent(float v) { return __float_as_uint(v) == BSENT32; }

template <typename T>
__global__ void k_batch_sentinel(const sable_batch_spts_item_t* __restrict__ items, int nitems, int64_t total,
                                 unsigned* __restrict__ ticket) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].row_begin <= g) lo = mid;
            else hi = mid - 1;
        }
        ((T*)items[lo].x)[g - items[lo].row_begin] = bsent(T(0));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0u;
}

template <typename T, int G>
__global__ __launch_bounds__(BB) void k_batch_syncfree(const sable_batch_spts_item_t* __restrict__ items, int nitems,
                                                       int64_t total, unsigned* __restrict__ ticket,
                                                       int* __restrict__ err, long long spin_limit) {
    constexpr int R = WAVE / G;
    const int lane = threadIdx.x & (WAVE - 1);
    const int l = lane & (G - 1);
    const int grp = lane / G;
    const uint64_t gmask = (G == 64) ? ~0ull : (((1ull << G) - 1ull) << (grp * G));
    while (true) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(ticket, (unsigned)R);
        base = __shfl(base, 0, WAVE);
        if ((int64_t)base >= total) break;
        const int64_t g = (int64_t)base + grp;
        bool published = g >= total;
        const int32_t* rp = nullptr;
        const int32_t* col = nullptr;
        const T* val = nullptr;
        T* x = nullptr;
        // j0: the group's next chunk of G entries; p: this lane's product of
        // it once its x has been published
        int32_t i = 0, j0 = 0, e = 0;
        T s = T(0), p = T(0);
        bool have = false;
        if (!published) {
            int lo = 0, hi = nitems - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (items[mid].row_begin <= g) lo = mid;
                else hi = mid - 1;
            }
            rp = items[lo].row_ptr;
            col = items[lo].col;
            val = (const T*)items[lo].val;
            x = (T*)items[lo].x;
            i = (int32_t)(g - items[lo].row_begin);
            j0 = rp[i];
            e = rp[i + 1] - 1;
            s = ((const T*)items[lo].b)[i];
        }
        long long spins = 0;
        while (__ballot(!published) != 0ull) {
            bool progressed = false;
            if (!published) {
                if (G == 1) {
                    while (j0 < e) {
                        const T v = __hip_atomic_load(&x[col[j0]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (is_bsent(v)) break;
                        s = sub_rn(s, mul_rn(val[j0], v));
                        j0++;
                        progressed = true;
                    }
                } else {
                    while (j0 < e) {
                        const int32_t j = j0 + l;
                        if (!have && j < e) {
                            const T v = __hip_atomic_load(&x[col[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if (!is_bsent(v)) {
                                p = mul_rn(val[j], v);
                                have = true;
                            }
                        }
                        // group-uniform: the whole group leaves or goes on
                        if ((__ballot(have || j >= e) & gmask) != gmask) break;
                        const int n = min(G, e - j0);
                        s = group_ordered_sum<true, G>(s, p, n);
                        j0 += G;
                        have = false;
                        progressed = true;
                    }
                }
            }
            const bool ldone = published || j0 >= e;
            const uint64_t m = __ballot(ldone);
            const bool gdone = (m & gmask) == gmask;
            if (!published && gdone) {
                if (l == 0) {
                    const T yv = div_rn(s, val[e]);
                    __hip_atomic_store(&x[i], yv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                published = true;
                progressed = true;
            }
            if (__ballot(progressed) == 0ull) {
                if (++spins > spin_limit) {
                    if (lane == 0) atomicOr(err, 4);
                    if (!published && l == 0)
                        __hip_atomic_store(&x[i], T(__builtin_nan("")), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    published = true;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// r32 = b64 - L64 x64 for every item (G lanes per row); first pass (x = 0)
// just converts b.
template <int G>
__global__ __launch_bounds__(BB) void k_batch_residual(const sable_batch_refine_item_t* __restrict__ items,
                                                       int nitems, int64_t total, int first) {
    const int64_t gid = ((int64_t)blockIdx.x * BB + threadIdx.x) / G;
    const int l = threadIdx.x & (G - 1);
    const bool valid = gid < total;
    int lo = 0;
    if (valid) {
        int hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].row_begin <= gid) lo = mid;
            else hi = mid - 1;
        }
    }
    double s = 0.0;
    int32_t i = 0;
    if (valid) {
        const sable_batch_refine_item_t& it = items[lo];
        i = (int32_t)(gid - it.row_begin);
        if (!first)
            for (int32_t j = it.row_ptr[i] + l; j < it.row_ptr[i + 1]; j += G) s = fma(it.val64[j], it.x64[it.col[j]], s);
    }
    if (G > 1) s = group_sum(s, G);
    if (valid && l == 0) {
        const sable_batch_refine_item_t& it = items[lo];
        it.r32[i] = (float)(it.b64[i] - s);
    }
}

__global__ void k_batch_update(const sable_batch_refine_item_t* __restrict__ items, int nitems, int64_t total,
                               int first) {
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nitems - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].row_begin <= g) lo = mid;
            else hi = mid - 1;
        }
        const sable_batch_refine_item_t& it = items[lo];
        const int64_t i = g - it.row_begin;
        it.x64[i] = first ? (double)it.d32[i] : it.x64[i] + (double)it.d32[i];
    }
}

// Sync-free item view of a refine item: solve L32 d = r.
__global__ void k_refine_to_spts(const sable_batch_refine_item_t* __restrict__ in, sable_batch_spts_item_t* __restrict__ out,
                                 int nitems) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nitems; k += gridDim.x * blockDim.x) {
        sable_batch_spts_item_t o;
        o.row_ptr = in[k].row_ptr;
        o.col = in[k].col;
        o.val = in[k].val32;
        o.b = in[k].r32;
        o.x = in[k].d32;
        o.row_begin = in[k].row_begin;
        o.N = in[k].N;
        o.pad = 0;
        out[k] = o;
    }
}

static unsigned gsz(int64_t n) {
    unsigned g = grid_for(n, BB);
    return g > 16384 ? 16384 : g;
}

// One level of a batched level-set solve, one workgroup per chunk of the
// level's rows: G lanes per row; t = 0 + v0*x0 + v1*x1 + ... over the row's
// off-diagonal entries in stored order (the lanes form G products at once,
// then every lane of the group adds them in order), x[i] = (b[i] - t) /
// diag: the level-set solve's bits at every G.
template <typename T, int G>
__device__ __forceinline__ void level_chunk(const sable_batch_spts_item_t* __restrict__ items, const RowChunk& c,
                                            T* stage) {
    const int g = threadIdx.x / G;
    const int l = threadIdx.x & (G - 1);
    const bool valid = g < c.count;
    const sable_batch_spts_item_t& it = items[c.item];
    T s = T(0);
    int32_t i = 0, a = 0, e = 0;
    if (valid) {
        i = c.rows ? c.rows[g] : c.first + g;
        a = it.row_ptr[i];
        e = it.row_ptr[i + 1] - 1;
    }
    const int32_t* __restrict__ col = it.col;
    const T* __restrict__ val = (const T*)it.val;
    const T* __restrict__ x = (const T*)it.x;
    // a and e are the same on the lanes of a group, so is every trip count
    for (int32_t j0 = a; j0 < e; j0 += G) {
        const int32_t j = j0 + l;
        const T p = j < e ? mul_rn(val[j], x[col[j]]) : T(0);
        if (G == 1) {
            s = add_rn(s, p);
        } else {
            const int n = min(G, e - j0);
            s = group_ordered_sum_lds<false, G>(s, p, n, stage);
        }
    }
    if (valid && l == 0) {
        const T d = ((const T*)it.val)[e];
        ((T*)it.x)[i] = div_rn(sub_rn(((const T*)it.b)[i], s), d);
    }
}

template <typename T, int G>
__global__ __launch_bounds__(BB) void k_batch_level(const sable_batch_spts_item_t* __restrict__ items,
                                                    const RowChunk* __restrict__ chunks) {
    __shared__ T stage[BB];
    level_chunk<T, G>(items, chunks[blockIdx.x], stage);
}

int batch_level_launch(hipStream_t st, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* items,
                       const RowChunk* d_level_chunks, const LevelPlan& plan) {
    for (int lv = 0; lv < plan.levels(); lv++) {
        const int nch = plan.chunk_lo[lv + 1] - plan.chunk_lo[lv];
        if (nch <= 0) continue;
        const RowChunk* ch = d_level_chunks + plan.chunk_lo[lv];
#define LV_LAUNCH(T, G) hipLaunchKernelGGL((k_batch_level<T, G>), dim3((unsigned)nch), dim3(BB), 0, st, items, ch)
#define LV_SWITCH(T)                        \
        switch (lanes) {                    \
            case 1: LV_LAUNCH(T, 1); break;  \
            case 4: LV_LAUNCH(T, 4); break;  \
            case 8: LV_LAUNCH(T, 8); break;  \
            case 16: LV_LAUNCH(T, 16); break; \
            default: LV_LAUNCH(T, 64); break; \
        }
        if (dtype == SABLE_F64) {
            LV_SWITCH(double)
        } else {
            LV_SWITCH(float)
        }
#undef LV_SWITCH
#undef LV_LAUNCH
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// Chunked residual r32 = b64 - L64 x64 (first pass: r32 = b64) and update
// x64 += d32 (first pass: x64 = d32), one workgroup per chunk of rows.
template <int G>
__device__ __forceinline__ void residual_chunk(const sable_batch_refine_item_t* __restrict__ items, const RowChunk& c,
                                               int first) {
    const int g = threadIdx.x / G;
    const int l = threadIdx.x & (G - 1);
    const bool valid = g < c.count;
    const sable_batch_refine_item_t& it = items[c.item];
    const int32_t i = c.first + g;
    double s = 0.0;
    if (valid && !first)
        for (int32_t j = it.row_ptr[i] + l; j < it.row_ptr[i + 1]; j += G) s = fma(it.val64[j], it.x64[it.col[j]], s);
    if (G > 1) s = group_sum(s, G);
    if (valid && l == 0) it.r32[i] = (float)(it.b64[i] - s);
}

__device__ __forceinline__ void update_chunk(const sable_batch_refine_item_t* __restrict__ items, const RowChunk& c,
                                             int first) {
    if ((int)threadIdx.x >= c.count) return;
    const sable_batch_refine_item_t& it = items[c.item];
    const int32_t i = c.first + threadIdx.x;
    it.x64[i] = first ? (double)it.d32[i] : it.x64[i] + (double)it.d32[i];
}

template <int G>
__global__ __launch_bounds__(BB) void k_chunk_residual(const sable_batch_refine_item_t* __restrict__ items,
                                                       const RowChunk* __restrict__ chunks, int first) {
    residual_chunk<G>(items, chunks[blockIdx.x], first);
}

__global__ __launch_bounds__(BB) void k_chunk_update(const sable_batch_refine_item_t* __restrict__ items,
                                                     const RowChunk* __restrict__ chunks, int first) {
    update_chunk(items, chunks[blockIdx.x], first);
}

int batch_spts_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* items,
                      int nitems, int64_t total_rows, unsigned* ticket, int* err, long long spin) {
    if (total_rows <= 0) return SABLE_OK;
    const int64_t waves = (total_rows * lanes + WAVE - 1) / WAVE;
    int64_t blocks = (waves + (BB / WAVE) - 1) / (BB / WAVE);
    const int64_t cap = (int64_t)num_cu * 8;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    // rows are claimed from a 32-bit ticket, at most WAVE per wave past the
    // end before every wave sees it exhausted: the last claim must not wrap
    const int64_t overshoot = blocks * (BB / WAVE) * (int64_t)WAVE;
    SABLE_CHECK(total_rows + overshoot < ((int64_t)1 << 32), SABLE_EINVAL,
                "batched solve of %lld rows: the 32-bit row ticket holds fewer than %lld rows",
                (long long)total_rows, (long long)(((int64_t)1 << 32) - overshoot));
#define SF_LAUNCH(T, G)                                                                                          \
    hipLaunchKernelGGL((k_batch_syncfree<T, G>), dim3((unsigned)blocks), dim3(BB), 0, st, items, nitems, total_rows, \
                       ticket, err, spin)
#define SF_SWITCH(T)                                   \
    hipLaunchKernelGGL(k_batch_sentinel<T>, dim3(gsz(total_rows)), dim3(BB), 0, st, items, nitems, total_rows, ticket); \
    switch (lanes) {                                   \
        case 1: SF_LAUNCH(T, 1); break;                \
        case 4: SF_LAUNCH(T, 4); break;                \
        case 8: SF_LAUNCH(T, 8); break;                \
        case 16: SF_LAUNCH(T, 16); break;              \
        default: SF_LAUNCH(T, 64); break;              \
    }
    if (dtype == SABLE_F64) {
        SF_SWITCH(double)
    } else {
        SF_SWITCH(float)
    }
#undef SF_SWITCH
#undef SF_LAUNCH
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

int batch_spmv_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int nrhs, const sable_batch_spmv_item_t* items,
                      int nitems, int64_t total_blocks, unsigned long long* counter, int wgs) {
    if (total_blocks <= 0 || nitems <= 0) return SABLE_OK;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, counter);
    int64_t grid = wgs > 0 ? wgs : (int64_t)num_cu * 4;
    const int64_t claims = (total_blocks + BATCH_CH - 1) / BATCH_CH;
    if (grid > claims) grid = claims;
#define BS_LAUNCH(T, NR)                                                                                          \
    do {                                                                                                           \
        const size_t fixed = sizeof(T) * (Lds<NR>::ELEMS + BB / WAVE) + 1024;                                     \
        const size_t want = sizeof(int64_t) * (size_t)nitems;                                                     \
        const bool fits = nitems <= BATCH_LDS_ITEMS && fixed + want <= 65536;                                     \
        hipLaunchKernelGGL((k_batch_spmv<T, NR>), dim3((unsigned)grid), dim3(BB), fits ? want : 0, st, items,     \
                           nitems, total_blocks, counter, fits ? nitems : 0);                                      \
    } while (0)
#define BS_SWITCH(T)                        \
    switch (nrhs) {                         \
        case 1: BS_LAUNCH(T, 1); break;     \
        case 2: BS_LAUNCH(T, 2); break;     \
        case 4: BS_LAUNCH(T, 4); break;     \
        default: BS_LAUNCH(T, 8); break;    \
    }
    if (dtype == SABLE_F64) {
        BS_SWITCH(double)
    } else {
        BS_SWITCH(float)
    }
#undef BS_SWITCH
#undef BS_LAUNCH
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

int batch_refine_launch(hipStream_t st, int num_cu, int lanes, const sable_batch_refine_item_t* items,
                        sable_batch_spts_item_t* spts_view, int nitems, int64_t total_rows, int iters,
                        unsigned* ticket, int* err, long long spin) {
    if (total_rows <= 0) return SABLE_OK;
    hipLaunchKernelGGL(k_refine_to_spts, dim3(1), dim3(BB), 0, st, items, spts_view, nitems);
    const unsigned rgrid = (unsigned)((total_rows * lanes + BB - 1) / BB);
    for (int it = 0; it <= iters; it++) {
        const int first = it == 0;
        switch (lanes) {
            case 1: hipLaunchKernelGGL((k_batch_residual<1>), dim3(rgrid), dim3(BB), 0, st, items, nitems, total_rows, first); break;
            case 4: hipLaunchKernelGGL((k_batch_residual<4>), dim3(rgrid), dim3(BB), 0, st, items, nitems, total_rows, first); break;
            case 8: hipLaunchKernelGGL((k_batch_residual<8>), dim3(rgrid), dim3(BB), 0, st, items, nitems, total_rows, first); break;
            case 16: hipLaunchKernelGGL((k_batch_residual<16>), dim3(rgrid), dim3(BB), 0, st, items, nitems, total_rows, first); break;
            default: hipLaunchKernelGGL((k_batch_residual<64>), dim3(rgrid), dim3(BB), 0, st, items, nitems, total_rows, first); break;
        }
        int rc = batch_spts_launch(st, num_cu, SABLE_F32, lanes, spts_view, nitems, total_rows, ticket, err, spin);
        if (rc) return rc;
        hipLaunchKernelGGL(k_batch_update, dim3(gsz(total_rows)), dim3(BB), 0, st, items, nitems, total_rows, first);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// One level of the batched FP32 solve with one FP64 refinement step fused
// into the same sweep (refine.h), one workgroup per chunk of the level's
// rows: the (x0, d) pair of row i is kept at float offset 2i of the item's
// work space from r32 (r32 and d32 must then be one region of 2N floats, as
// the server lays them out), x goes to x64.
template <int G>
__global__ __launch_bounds__(BB) void k_batch_level_ref(const sable_batch_refine_item_t* __restrict__ items,
                                                        const RowChunk* __restrict__ chunks) {
    __shared__ RefEntry stage[BB];
    const RowChunk c = chunks[blockIdx.x];
    const int g = threadIdx.x / G;
    const bool valid = g < c.count;
    const sable_batch_refine_item_t& it = items[c.item];
    const int32_t i = valid ? (c.rows ? c.rows[g] : c.first + g) : 0;
    ref_row<G>(i, valid, it.row_ptr, it.col, it.val64, it.b64, (float2*)it.r32, it.x64, stage);
}

int batch_refine_level_launch(hipStream_t st, int lanes, const sable_batch_refine_item_t* items,
                              sable_batch_spts_item_t* spts_view, int nitems, int iters,
                              const RowChunk* d_level_chunks, const LevelPlan& plan, const RowChunk* d_row_chunks,
                              int nrow_chunks) {
    if (nitems <= 0) return SABLE_OK;
    if (iters == 1) {
        // one sweep: each row's FP32 solve, residual, correction and update
        for (int lv = 0; lv < plan.levels(); lv++) {
            const int nch = plan.chunk_lo[lv + 1] - plan.chunk_lo[lv];
            if (nch <= 0) continue;
            const RowChunk* ch = d_level_chunks + plan.chunk_lo[lv];
            switch (lanes) {
                case 1: hipLaunchKernelGGL((k_batch_level_ref<1>), dim3((unsigned)nch), dim3(BB), 0, st, items, ch); break;
                case 4: hipLaunchKernelGGL((k_batch_level_ref<4>), dim3((unsigned)nch), dim3(BB), 0, st, items, ch); break;
                case 8: hipLaunchKernelGGL((k_batch_level_ref<8>), dim3((unsigned)nch), dim3(BB), 0, st, items, ch); break;
                case 16: hipLaunchKernelGGL((k_batch_level_ref<16>), dim3((unsigned)nch), dim3(BB), 0, st, items, ch); break;
                default: hipLaunchKernelGGL((k_batch_level_ref<64>), dim3((unsigned)nch), dim3(BB), 0, st, items, ch); break;
            }
        }
        SABLE_LAUNCH_CHECK();
        return SABLE_OK;
    }
    if (nrow_chunks <= 0) return SABLE_OK;
    hipLaunchKernelGGL(k_refine_to_spts, dim3(1), dim3(BB), 0, st, items, spts_view, nitems);
    for (int it = 0; it <= iters; it++) {
        const int first = it == 0;
        const dim3 g((unsigned)nrow_chunks);
        switch (lanes) {
            case 1: hipLaunchKernelGGL((k_chunk_residual<1>), g, dim3(BB), 0, st, items, d_row_chunks, first); break;
            case 4: hipLaunchKernelGGL((k_chunk_residual<4>), g, dim3(BB), 0, st, items, d_row_chunks, first); break;
            case 8: hipLaunchKernelGGL((k_chunk_residual<8>), g, dim3(BB), 0, st, items, d_row_chunks, first); break;
            case 16: hipLaunchKernelGGL((k_chunk_residual<16>), g, dim3(BB), 0, st, items, d_row_chunks, first); break;
            default: hipLaunchKernelGGL((k_chunk_residual<64>), g, dim3(BB), 0, st, items, d_row_chunks, first); break;
        }
        int rc = batch_level_launch(st, SABLE_F32, lanes, spts_view, d_level_chunks, plan);
        if (rc) return rc;
        hipLaunchKernelGGL(k_chunk_update, g, dim3(BB), 0, st, items, d_row_chunks, first);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_batch_spmv(sable_ctx_t* ctx, sable_dtype_t dtype, int nrhs, const sable_batch_spmv_item_t* d_items,
                               int nitems, int64_t total_blocks, int workgroups) {
    SABLE_CHECK(ctx && (d_items || nitems == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(nrhs == 1 || nrhs == 2 || nrhs == 4 || nrhs == 8, SABLE_EINVAL, "nrhs must be 1, 2, 4 or 8");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (nitems == 0 || total_blocks == 0) return SABLE_OK;
    void* s;
    int rc = ctx_scratch(ctx, 256, &s);
    if (rc) return rc;
    return batch_spmv_launch(ctx->stream, ctx->num_cu, dtype, nrhs, d_items, nitems, total_blocks,
                             (unsigned long long*)s, workgroups);
}

SABLE_API int sable_batch_spts(sable_ctx_t* ctx, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* d_items,
                               int nitems, int64_t total_rows) {
    SABLE_CHECK(ctx && (d_items || nitems == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(lanes == 1 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 64, SABLE_EINVAL, "bad lanes");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (nitems == 0 || total_rows == 0) return SABLE_OK;
    void* s;
    int rc = ctx_scratch(ctx, 512, &s);
    if (rc) return rc;
    int* err = (int*)((char*)s + 256);
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
    rc = batch_spts_launch(ctx->stream, ctx->num_cu, dtype, lanes, d_items, nitems, total_rows, (unsigned*)s, err,
                           1ll << 24);
    if (rc) return rc;
    int flag;
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    flag = *(int*)ctx->pinned;
    SABLE_CHECK(!flag, SABLE_ETIMEOUT, "batched sync-free solve exceeded its spin limit");
    return SABLE_OK;
}

SABLE_API int sable_batch_spts_refined(sable_ctx_t* ctx, int lanes, const sable_batch_refine_item_t* d_items,
                                       int nitems, int64_t total_rows, int iters) {
    SABLE_CHECK(ctx && (d_items || nitems == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(lanes == 1 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 64, SABLE_EINVAL, "bad lanes");
    SABLE_CHECK(iters >= 0, SABLE_EINVAL, "negative iters");
    if (nitems == 0 || total_rows == 0) return SABLE_OK;
    void* s;
    int rc = ctx_scratch(ctx, 1024 + sizeof(sable_batch_spts_item_t) * (size_t)nitems, &s);
    if (rc) return rc;
    int* err = (int*)((char*)s + 256);
    auto* view = (sable_batch_spts_item_t*)((char*)s + 1024);
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), ctx->stream));
    rc = batch_refine_launch(ctx->stream, ctx->num_cu, lanes, d_items, view, nitems, total_rows, iters, (unsigned*)s,
                             err, 1ll << 24);
    if (rc) return rc;
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    SABLE_CHECK(!*(int*)ctx->pinned, SABLE_ETIMEOUT, "batched sync-free solve exceeded its spin limit");
    return SABLE_OK;
}

// ------------------------------------------------ batched level schedule
SABLE_API int sable_batch_level_plan_create(sable_ctx_t* ctx, int lanes, const sable_batch_spts_item_t* h_items,
                                            int nitems, sable_batch_level_plan_t** out) {
    SABLE_CHECK(ctx && out && (h_items || nitems == 0) && nitems >= 0, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(lanes == 1 || lanes == 4 || lanes == 8 || lanes == 16 || lanes == 64, SABLE_EINVAL,
                "lanes must be 1, 4, 8, 16 or 64");
    *out = nullptr;
    hipStream_t st = ctx->stream;
    const int32_t CH = BB / lanes;
    // levels of every item (par/src/sparse32.js:2952-2974: 0 for a row with
    // only its diagonal, else 1 + the largest level among its columns), rows
    // listed by level, ascending within a level
    std::vector<std::vector<int32_t>> lps((size_t)nitems);
    std::vector<int32_t> rows_all;
    std::vector<int64_t> row_off((size_t)nitems + 1, 0);
    std::vector<int32_t> rp, col, level, next;
    for (int k = 0; k < nitems; k++) {
        const sable_batch_spts_item_t& it = h_items[k];
        const int32_t N = it.N;
        SABLE_CHECK(N >= 0 && (N == 0 || (it.row_ptr && it.col)), SABLE_EINVAL, "item %d: bad matrix", k);
        row_off[(size_t)k + 1] = row_off[k] + N;
        if (N == 0) continue;
        rp.resize((size_t)N + 1);
        SABLE_HIP(hipMemcpyAsync(rp.data(), it.row_ptr, 4 * ((size_t)N + 1), hipMemcpyDeviceToHost, st));
        SABLE_HIP(hipStreamSynchronize(st));
        SABLE_CHECK(rp[0] == 0 && rp[N] >= N, SABLE_EMALFORMED, "item %d: row_ptr must start at 0", k);
        col.resize((size_t)rp[N]);
        SABLE_HIP(hipMemcpyAsync(col.data(), it.col, 4 * (size_t)rp[N], hipMemcpyDeviceToHost, st));
        SABLE_HIP(hipStreamSynchronize(st));
        level.resize(N);
        int32_t top = 0;
        for (int32_t i = 0; i < N; i++) {
            SABLE_CHECK(rp[i + 1] > rp[i] && col[(size_t)rp[i + 1] - 1] == i, SABLE_EMALFORMED,
                        "item %d, row %d: the last entry is not the diagonal", k, i);
            int32_t m = -1;
            for (int32_t j = rp[i]; j < rp[i + 1] - 1; j++) {
                SABLE_CHECK(col[j] >= 0 && col[j] < i, SABLE_EMALFORMED, "item %d, row %d: column %d is not below "
                            "the diagonal", k, i, col[j]);
                m = std::max(m, level[col[j]]);
            }
            level[i] = m + 1;
            top = std::max(top, level[i]);
        }
        std::vector<int32_t>& lp = lps[k];
        lp.assign((size_t)top + 2, 0);
        for (int32_t i = 0; i < N; i++) lp[(size_t)level[i] + 1]++;
        for (int32_t l = 0; l <= top; l++) lp[(size_t)l + 1] += lp[l];
        next.assign(lp.begin(), lp.end() - 1);
        const size_t base = rows_all.size();
        rows_all.resize(base + N);
        for (int32_t i = 0; i < N; i++) rows_all[base + next[level[i]]++] = i;
    }
    auto* p = new sable_batch_level_plan_s();
    p->ctx = ctx;
    p->lanes = lanes;
    p->nitems = nitems;
    p->rows = (int64_t)rows_all.size();
    if (!rows_all.empty() && hipMalloc(&p->d_rows, 4 * rows_all.size()) != hipSuccess) {
        (void)hipGetLastError();
        delete p;
        return fail(SABLE_ENOMEM, "batched level plan: row lists");
    }
    // chunks, level by level: level l lists, item by item, the rows of level
    // l of every item that has one, CH rows (one workgroup) per chunk
    int top = 0;
    for (auto& lp : lps) top = std::max(top, (int)lp.size() - 1);
    std::vector<RowChunk> ch;
    p->plan.chunk_lo.assign((size_t)top + 1, 0);
    for (int lv = 0; lv < top; lv++) {
        p->plan.chunk_lo[lv] = (int32_t)ch.size();
        for (int k = 0; k < nitems; k++) {
            const std::vector<int32_t>& lp = lps[k];
            if (lv + 1 >= (int)lp.size()) continue;
            for (int32_t a = lp[lv]; a < lp[(size_t)lv + 1]; a += CH)
                ch.push_back(RowChunk{p->d_rows + row_off[k] + a, k, 0, std::min(CH, lp[(size_t)lv + 1] - a), 0});
        }
    }
    p->plan.chunk_lo[top] = (int32_t)ch.size();
    p->nchunks = (int64_t)ch.size();
    if (!ch.empty() && hipMalloc(&p->d_chunks, sizeof(RowChunk) * ch.size()) != hipSuccess) {
        (void)hipGetLastError();
        sable_batch_level_plan_destroy(p);
        return fail(SABLE_ENOMEM, "batched level plan: chunk table");
    }
    hipError_t e = hipSuccess;
    if (!rows_all.empty())
        e = hipMemcpyAsync(p->d_rows, rows_all.data(), 4 * rows_all.size(), hipMemcpyHostToDevice, st);
    if (!e && !ch.empty()) e = hipMemcpyAsync(p->d_chunks, ch.data(), sizeof(RowChunk) * ch.size(),
                                              hipMemcpyHostToDevice, st);
    if (!e) e = hipStreamSynchronize(st);
    if (e) {
        sable_batch_level_plan_destroy(p);
        return fail(SABLE_EHIP, "batched level plan upload: %s", hipGetErrorString(e));
    }
    *out = p;
    return SABLE_OK;
}

SABLE_API void sable_batch_level_plan_destroy(sable_batch_level_plan_t* p) {
    if (!p) return;
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->d_rows) (void)hipFree(p->d_rows);
    if (p->d_chunks) (void)hipFree(p->d_chunks);
    delete p;
}

SABLE_API int sable_batch_level_plan_info(const sable_batch_level_plan_t* p, int32_t* levels, int64_t* chunks) {
    SABLE_CHECK(p, SABLE_EINVAL, "NULL plan");
    if (levels) *levels = p->plan.levels();
    if (chunks) *chunks = p->nchunks;
    return SABLE_OK;
}

SABLE_API int sable_batch_level_plan_solve(sable_ctx_t* ctx, const sable_batch_level_plan_t* p, sable_dtype_t dtype,
                                           const sable_batch_spts_item_t* d_items, int nitems) {
    SABLE_CHECK(ctx && p && (d_items || nitems == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == p->ctx, SABLE_EINVAL, "plan belongs to another context");
    SABLE_CHECK(nitems == p->nitems, SABLE_EINVAL, "%d items for a plan of %d", nitems, p->nitems);
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (p->nchunks == 0) return SABLE_OK;
    return batch_level_launch(ctx->stream, dtype, p->lanes, d_items, p->d_chunks, p->plan);
}

SABLE_API int sable_batch_level_plan_solve_refined(sable_ctx_t* ctx, const sable_batch_level_plan_t* p,
                                                   const sable_batch_refine_item_t* d_items, int nitems) {
    SABLE_CHECK(ctx && p && (d_items || nitems == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == p->ctx, SABLE_EINVAL, "plan belongs to another context");
    SABLE_CHECK(nitems == p->nitems, SABLE_EINVAL, "%d items for a plan of %d", nitems, p->nitems);
    if (p->nchunks == 0) return SABLE_OK;
    // one refinement step: the fused sweep (no residual or update launches,
    // so no spts view or row chunks are needed)
    return batch_refine_level_launch(ctx->stream, p->lanes, d_items, nullptr, nitems, 1, p->d_chunks, p->plan,
                                     nullptr, 0);
}

}  // extern "C"
