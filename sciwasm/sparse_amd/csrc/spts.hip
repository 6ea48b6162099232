// spts.hip -- lower sparse triangular solve L.x = b for gfx950.
//
// Matrices are lower-triangular CSR with the diagonal stored LAST in each
// row (the reference's convention, par/src/sparse32.js:2956-2974 and
// par/src/sparse_opt_32.wat:11714-11737).
//
// * Level-set analysis (host C++, the reference's algorithm, bit-exact):
//   level[i] = 0 for a one-entry row, else 1 + max level of its other
//   columns; level_ptr = prefix sum of level sizes; permutation by a stable
//   counting sort; the reordered CSR (rows in level order, columns renamed)
//   is built on the device.
// * Level-set solve: one launch per level over the level's contiguous rows,
//   G lanes per row; the launches are captured once into a hipGraph and
//   replayed.  With G = 1 the per-row order is the reference's
//   t = 0 + v0*y0 + v1*y1 ...; y = (b - t)/d with separately rounded ops.
// * Sync-free solve: waves claim consecutive rows from a ticket counter (so
//   every row a wave waits on belongs to a wave that is already running).
//   Readiness is carried by the solution itself: x starts as a signalling-NaN
//   sentinel that arithmetic never produces, and each row publishes its
//   value with one agent-scope atomic store; consumers poll with
//   agent-scope atomic loads.  A wave loops uniformly until all of its rows
//   are published, so rows that depend on rows of the same wave make
//   progress.  A bounded spin count turns a stuck solve into an error.
//   With G = 1 the order is the reference's t = b; t = t - v*y ...; y = t/d.
// * FP32 solve + FP64 iterative refinement: r = b - L.x in FP64 from the
//   FP64 values, correction solved in FP32, x += d in FP64.
#include "common.h"
#include "prims.h"

#include <algorithm>
#include <cstring>
#include <vector>

struct sable_spts_s {
    sable_ctx_t* ctx = nullptr;
    sable_dtype_t dtype = SABLE_F64;
    int32_t N = 0;
    int64_t nnz = 0;
    int lanes = 1;
    // natural-order matrix (caller-owned) for the sync-free solve and residual
    const int32_t* rp = nullptr;
    const int32_t* col = nullptr;
    const void* val = nullptr;
    // level schedule
    int32_t nlevels = 0;
    int32_t max_width = 0;
    std::vector<int32_t> h_level_ptr;
    int32_t* d_perm = nullptr;
    int32_t* d_order = nullptr;
    int32_t* d_rp = nullptr;   // reordered
    int32_t* d_col = nullptr;  // reordered
    void* d_val = nullptr;     // reordered, dtype
    void* d_bl = nullptr;      // level-ordered rhs
    void* d_yl = nullptr;      // level-ordered solution
    // FP32 copies for the refined solve (created on first use)
    float* d_val32 = nullptr;    // natural order
    float* d_lval32 = nullptr;   // level order
    void* d_work = nullptr;      // r32, d32 for refinement
    unsigned* d_ticket = nullptr;
    int* d_err = nullptr;
    // captured level sweeps: [0] analysis dtype, [1] FP32 copy (refinement)
    int graph_lanes[2] = {0, 0};
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    // CSR copy owned by the analysis (built from a CSC input), else null
    void* owned[3] = {nullptr, nullptr, nullptr};
};

namespace sable {

constexpr int TB = 256;
constexpr uint64_t SENT64 = 0x7FF4DEADBEEF0001ull;  // signalling NaN payload
constexpr uint32_t SENT32 = 0x7FA0BEEFu;

template <typename T>
struct Sent;
template <>
struct Sent<double> {
    static __device__ __forceinline__ double value() { return __longlong_as_double((long long)SENT64); }
    static __device__ __forceinline__ bool is(double v) { return (uint64_t)__double_as_longlong(v) == SENT64; }
};
template <>
struct Sent<float> {
    static __device__ __forceinline__ float value() { return __uint_as_float(SENT32); }
    static __device__ __forceinline__ bool is(float v) { return __float_as_uint(v) == SENT32; }
};

template <typename T>
__device__ __forceinline__ T load_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void store_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------- level-set
// Rows [r0, r0+nrows) of a level; G lanes per row.
template <typename T, int G>
__global__ __launch_bounds__(TB) void k_level(int32_t r0, int32_t nrows, const int32_t* __restrict__ rp,
                                              const int32_t* __restrict__ col, const T* __restrict__ val,
                                              const T* __restrict__ bl, T* __restrict__ yl) {
    const int64_t gt = (int64_t)blockIdx.x * TB + threadIdx.x;
    const int64_t g = gt / G;
    const int l = threadIdx.x & (G - 1);
    const bool valid = g < nrows;
    const int32_t i = r0 + (int32_t)(valid ? g : 0);
    T s = T(0);
    int32_t e = 0;
    if (valid) {
        const int32_t a = rp[i];
        e = rp[i + 1] - 1;
        for (int32_t j = a + l; j < e; j += G) s = add_rn(s, mul_rn(val[j], yl[col[j]]));
    }
    if (G > 1) s = group_sum(s, G);
    if (valid && l == 0) yl[i] = div_rn(sub_rn(bl[i], s), val[e]);
}

template <typename T>
__global__ void k_gather_perm(int32_t N, const int32_t* __restrict__ perm, const T* __restrict__ src,
                              T* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

template <typename T>
__global__ void k_scatter_perm(int32_t N, const int32_t* __restrict__ perm, const T* __restrict__ src,
                               T* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        dst[perm[i]] = src[i];
}

// Reordered CSR: new row i = old row perm[i], columns renamed through order.
template <typename T>
__global__ void k_reorder(int32_t N, const int32_t* __restrict__ perm, const int32_t* __restrict__ order,
                          const int32_t* __restrict__ rp, const int32_t* __restrict__ col, const T* __restrict__ val,
                          const int32_t* __restrict__ nrp, int32_t* __restrict__ ncol, T* __restrict__ nval) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = perm[i];
        const int32_t a = rp[o], len = rp[o + 1] - a, d = nrp[i];
        for (int32_t t = 0; t < len; t++) {
            ncol[d + t] = order[col[a + t]];
            nval[d + t] = val[a + t];
        }
    }
}

// Zero diagonals (flag 1).
template <typename T>
__global__ void k_check_diag(int32_t N, const int32_t* __restrict__ rp, const T* __restrict__ val, int* err) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = rp[i + 1] - 1;
        if (e < rp[i] || !(val[e] != T(0))) atomicOr(err, 1);
    }
}

// Structure check for the raw sync-free entry: each row non-empty, the last
// column is the row itself and the others are < row (flag 2).
// CSC lower triangle as the column sweep reads it: every column starts with
// its diagonal (seq/src/sparse_opt_32.wat:52-57 divides by the column's
// first value).
__global__ void k_check_csc_diag_first(int32_t N, const int32_t* __restrict__ cp, const int32_t* __restrict__ row,
                                       int* err) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += (int64_t)gridDim.x * blockDim.x)
        if (cp[j + 1] <= cp[j] || row[cp[j]] != (int32_t)j) atomicOr(err, 1);
}

__global__ void k_check_lower(int32_t N, const int32_t* __restrict__ rp, const int32_t* __restrict__ col, int* err) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = rp[i], e = rp[i + 1] - 1;
        if (e < a || col[e] != i) {
            atomicOr(err, 2);
            continue;
        }
        for (int32_t j = a; j < e; j++)
            if (col[j] < 0 || col[j] >= i) {
                atomicOr(err, 2);
                break;
            }
    }
}

// ------------------------------------------------------------- sync-free
template <typename T>
__global__ void k_sentinel(int32_t N, T* __restrict__ x, unsigned* __restrict__ ticket) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = Sent<T>::value();
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0u;
}

// LVL: the level solve's arithmetic (t = 0, t += val * x over the
// off-diagonals, x = (b - t) / diag) also for G = 1, where the plain
// sync-free solve follows the reference's t = b, t -= val * x.  With LVL
// every lane count gives the level solve's result bit for bit.
template <typename T, int G, bool LVL = false>
__global__ __launch_bounds__(TB) void k_syncfree(int32_t N, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col, const T* __restrict__ val,
                                                 const T* __restrict__ b, T* x, unsigned* __restrict__ ticket,
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
        if (base >= (unsigned)N) break;
        const int64_t i = (int64_t)base + grp;
        bool published = i >= N;
        int32_t j = 0, e = 0;
        T s = T(0);
        T bi = T(0);
        if (!published) {
            j = rp[i] + l;
            e = rp[i + 1] - 1;
            bi = b[i];
            if (G == 1 && !LVL) s = bi;
        }
        long long spins = 0;
        while (__ballot(!published) != 0ull) {
            bool progressed = false;
            if (!published) {
                while (j < e) {
                    const T v = load_agent(&x[col[j]]);
                    if (Sent<T>::is(v)) break;
                    if (G == 1 && !LVL)
                        s = sub_rn(s, mul_rn(val[j], v));
                    else
                        s = add_rn(s, mul_rn(val[j], v));
                    j += G;
                    progressed = true;
                }
            }
            const bool ldone = published || j >= e;
            const uint64_t m = __ballot(ldone);
            const bool gdone = (m & gmask) == gmask;
            T tot = s;
            if (G > 1) tot = group_sum(s, G);
            if (!published && gdone) {
                if (l == 0) {
                    const T y = (G == 1 && !LVL) ? div_rn(tot, val[e]) : div_rn(sub_rn(bi, tot), val[e]);
                    store_agent(&x[i], y);
                }
                published = true;
                progressed = true;
            }
            if (__ballot(progressed) == 0ull) {
                if (++spins > spin_limit) {
                    if (lane == 0) atomicOr(err, 4);
                    // give up: publish NaN so that waiting rows finish too
                    if (!published && l == 0) store_agent(&x[i], T(__builtin_nan("")));
                    published = true;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// --------------------------------------------------------- refinement
// r32[i] = (float)(b[i] - sum_j L_ij x[j]) in FP64, G lanes per row.
template <int G>
__global__ __launch_bounds__(TB) void k_residual(int32_t N, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col, const double* __restrict__ val,
                                                 const double* __restrict__ x, const double* __restrict__ b,
                                                 float* __restrict__ r32) {
    const int64_t g = ((int64_t)blockIdx.x * TB + threadIdx.x) / G;
    const int l = threadIdx.x & (G - 1);
    const bool valid = g < N;
    double s = 0.0;
    if (valid)
        for (int32_t j = rp[g] + l; j < rp[g + 1]; j += G) s = fma(val[j], x[col[j]], s);
    if (G > 1) s = group_sum(s, G);
    if (valid && l == 0) r32[g] = (float)(b[g] - s);
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ o) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        o[i] = (float)a[i];
}

__global__ void k_set_from_f32(int32_t N, const float* __restrict__ d, double* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (double)d[i];
}

__global__ void k_add_f32(int32_t N, const float* __restrict__ d, double* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += (double)d[i];
}

// ------------------------------------------------------------------ host
static unsigned gsz(int64_t n) {
    unsigned g = grid_for(n, TB);
    return g > 16384 ? 16384 : g;
}

static int pick_lanes(int64_t nnz, int32_t N) {
    const double mean = N > 0 ? (double)nnz / N : 1.0;  // entries per row incl. diagonal
    if (mean <= 5.0) return 1;
    if (mean <= 12.0) return 4;
    if (mean <= 40.0) return 8;
    return 16;
}

template <typename T>
static void launch_level(hipStream_t st, int G, int32_t r0, int32_t nrows, const int32_t* rp, const int32_t* col,
                         const T* val, const T* bl, T* yl) {
    const unsigned grid = (unsigned)(((int64_t)nrows * G + TB - 1) / TB);
    switch (G) {
        case 1: hipLaunchKernelGGL((k_level<T, 1>), dim3(grid), dim3(TB), 0, st, r0, nrows, rp, col, val, bl, yl); break;
        case 4: hipLaunchKernelGGL((k_level<T, 4>), dim3(grid), dim3(TB), 0, st, r0, nrows, rp, col, val, bl, yl); break;
        case 8: hipLaunchKernelGGL((k_level<T, 8>), dim3(grid), dim3(TB), 0, st, r0, nrows, rp, col, val, bl, yl); break;
        case 16: hipLaunchKernelGGL((k_level<T, 16>), dim3(grid), dim3(TB), 0, st, r0, nrows, rp, col, val, bl, yl); break;
        default: hipLaunchKernelGGL((k_level<T, 64>), dim3(grid), dim3(TB), 0, st, r0, nrows, rp, col, val, bl, yl); break;
    }
}

template <typename T>
static int levels_on_stream(hipStream_t st, int G, const std::vector<int32_t>& lp, const int32_t* rp,
                            const int32_t* col, const T* val, const T* bl, T* yl) {
    const int nl = (int)lp.size() - 1;
    for (int l = 0; l < nl; l++) {
        const int32_t r0 = lp[l], nr = lp[l + 1] - lp[l];
        if (nr <= 0) continue;
        launch_level<T>(st, G, r0, nr, rp, col, val, bl, yl);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

template <typename T>
static void launch_syncfree(hipStream_t st, int G, int num_cu, int32_t N, const int32_t* rp, const int32_t* col,
                            const T* val, const T* b, T* x, unsigned* ticket, int* err, long long spin,
                            bool lvl = false) {
    hipLaunchKernelGGL(k_sentinel<T>, dim3(gsz(N)), dim3(TB), 0, st, N, x, ticket);
    const int64_t waves_needed = ((int64_t)N * G + WAVE - 1) / WAVE;
    int64_t blocks = (waves_needed + (TB / WAVE) - 1) / (TB / WAVE);
    const int64_t cap = (int64_t)num_cu * 8;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
#define SF_LAUNCH(GG, L)                                                                                    \
    hipLaunchKernelGGL((k_syncfree<T, GG, L>), dim3((unsigned)blocks), dim3(TB), 0, st, N, rp, col, val, b, x, \
                       ticket, err, spin)
    if (lvl) {
        switch (G) {
            case 1: SF_LAUNCH(1, true); break;
            case 4: SF_LAUNCH(4, true); break;
            case 8: SF_LAUNCH(8, true); break;
            case 16: SF_LAUNCH(16, true); break;
            default: SF_LAUNCH(64, true); break;
        }
    } else {
        switch (G) {
            case 1: SF_LAUNCH(1, false); break;
            case 4: SF_LAUNCH(4, false); break;
            case 8: SF_LAUNCH(8, false); break;
            case 16: SF_LAUNCH(16, false); break;
            default: SF_LAUNCH(64, false); break;
        }
    }
#undef SF_LAUNCH
}

int syncfree_launch(hipStream_t st, int num_cu, sable_dtype_t dt, int lanes, int32_t N, const int32_t* rp,
                    const int32_t* col, const void* val, const void* b, void* x, unsigned* ticket, int* err,
                    long long spin) {
    if (N <= 0) return SABLE_OK;
    if (dt == SABLE_F64)
        launch_syncfree<double>(st, lanes, num_cu, N, rp, col, (const double*)val, (const double*)b, (double*)x, ticket,
                                err, spin);
    else
        launch_syncfree<float>(st, lanes, num_cu, N, rp, col, (const float*)val, (const float*)b, (float*)x, ticket,
                               err, spin);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

static int read_flag(sable_ctx_t* ctx, int* dflag, int* out) {
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, dflag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SABLE_HIP(hipStreamSynchronize(ctx->stream));
    *out = *(int*)ctx->pinned;
    return SABLE_OK;
}

static int valid_lanes(int G) { return G == 1 || G == 4 || G == 8 || G == 16 || G == 64; }

constexpr long long DEFAULT_SPIN = 1ll << 24;

static void free_analysis(sable_spts_s* a) {
    if (!a) return;
    for (int k = 0; k < 2; k++) {
        if (a->exec[k]) (void)hipGraphExecDestroy(a->exec[k]);
        if (a->graph[k]) (void)hipGraphDestroy(a->graph[k]);
    }
    void* ptrs[] = {a->d_perm, a->d_order, a->d_rp, a->d_col, a->d_val, a->d_bl, a->d_yl,
                    a->d_val32, a->d_lval32, a->d_work, a->d_ticket, a->d_err,
                    a->owned[0], a->owned[1], a->owned[2]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete a;
}

template <typename T>
static int capture_levels(sable_spts_s* a, bool use32) {
    const int k = use32 ? 1 : 0;
    if (a->exec[k] && a->graph_lanes[k] == a->lanes) return SABLE_OK;
    if (a->exec[k]) {
        (void)hipGraphExecDestroy(a->exec[k]);
        a->exec[k] = nullptr;
    }
    if (a->graph[k]) {
        (void)hipGraphDestroy(a->graph[k]);
        a->graph[k] = nullptr;
    }
    hipStream_t cs;
    SABLE_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(SABLE_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    const T* val = use32 ? (const T*)a->d_lval32 : (const T*)a->d_val;
    levels_on_stream<T>(cs, a->lanes, a->h_level_ptr, a->d_rp, a->d_col, val, (const T*)a->d_bl, (T*)a->d_yl);
    e = hipStreamEndCapture(cs, &a->graph[k]);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return fail(SABLE_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&a->exec[k], a->graph[k], nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(SABLE_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    a->graph_lanes[k] = a->lanes;
    return SABLE_OK;
}

// Level solve on level-ordered buffers bl -> yl (graph replay).
template <typename T>
static int solve_level_ordered(sable_spts_s* a, bool use32) {
    int rc = capture_levels<T>(a, use32);
    if (rc) return rc;
    SABLE_HIP(hipGraphLaunch(a->exec[use32 ? 1 : 0], a->ctx->stream));
    return SABLE_OK;
}

static int ensure_f32(sable_spts_s* a) {
    if (a->d_val32) return SABLE_OK;
    SABLE_CHECK(a->dtype == SABLE_F64, SABLE_EINVAL, "refined solve needs an FP64 analysis");
    hipStream_t st = a->ctx->stream;
    SABLE_HIP(hipMalloc(&a->d_val32, sizeof(float) * (size_t)(a->nnz > 0 ? a->nnz : 1)));
    SABLE_HIP(hipMalloc(&a->d_lval32, sizeof(float) * (size_t)(a->nnz > 0 ? a->nnz : 1)));
    SABLE_HIP(hipMalloc(&a->d_work, sizeof(float) * 2 * (size_t)(a->N > 0 ? a->N : 1) + 256));
    const int32_t* base = a->rp;
    int32_t r0;
    SABLE_HIP(hipMemcpyAsync(a->ctx->pinned, base, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    memcpy(&r0, a->ctx->pinned, 4);
    hipLaunchKernelGGL(k_to_f32, dim3(gsz(a->nnz)), dim3(TB), 0, st, a->nnz, (const double*)a->val + r0, a->d_val32);
    hipLaunchKernelGGL(k_to_f32, dim3(gsz(a->nnz)), dim3(TB), 0, st, a->nnz, (const double*)a->d_val, a->d_lval32);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_spts_analyze(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                 const int32_t* col, const void* val, int lanes, sable_spts_t** out) {
    SABLE_CHECK(ctx && out && row_ptr && col && val, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N > 0, SABLE_EINVAL, "N must be positive");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    SABLE_CHECK(lanes == 0 || valid_lanes(lanes), SABLE_EINVAL, "lanes must be 0, 1, 4, 8, 16 or 64");
    *out = nullptr;
    hipStream_t st = ctx->stream;
    std::vector<int32_t> rp(N + 1);
    SABLE_HIP(hipMemcpyAsync(rp.data(), row_ptr, sizeof(int32_t) * (N + 1), hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    const int32_t base = rp[0];
    SABLE_CHECK(base == 0, SABLE_EINVAL, "row_ptr[0] must be 0");
    for (int32_t i = 0; i < N; i++)
        SABLE_CHECK(rp[i + 1] > rp[i], SABLE_EMALFORMED, "row %d is empty or row_ptr is not increasing", i);
    const int64_t nnz = rp[N];
    std::vector<int32_t> c(nnz);
    SABLE_HIP(hipMemcpyAsync(c.data(), col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    // structure: diagonal last, other columns strictly lower
    for (int32_t i = 0; i < N; i++) {
        SABLE_CHECK(c[rp[i + 1] - 1] == i, SABLE_EMALFORMED, "row %d: the last entry is not the diagonal", i);
        for (int32_t j = rp[i]; j < rp[i + 1] - 1; j++)
            SABLE_CHECK(c[j] >= 0 && c[j] < i, SABLE_EMALFORMED, "row %d: column %d is not below the diagonal", i,
                        c[j]);
    }
    // levels (par/src/sparse32.js:2952-2974)
    std::vector<int32_t> level(N);
    int32_t max_level = 0;
    for (int32_t i = 0; i < N; i++) {
        if (rp[i + 1] - rp[i] == 1) {
            level[i] = 0;
        } else {
            int32_t m = 0;
            for (int32_t j = rp[i]; j < rp[i + 1] - 1; j++) m = std::max(m, level[c[j]]);
            level[i] = 1 + m;
            max_level = std::max(max_level, level[i]);
        }
    }
    const int32_t tot = max_level + 1;
    std::vector<int32_t> freq(tot, 0), lp(tot + 1, 0), start(tot, 0), perm(N), order(N), nrp(N + 1);
    for (int32_t i = 0; i < N; i++) freq[level[i]]++;
    for (int32_t l = 1; l <= tot; l++) lp[l] = lp[l - 1] + freq[l - 1];
    for (int32_t l = 0; l < tot; l++) start[l] = lp[l];
    for (int32_t i = 0; i < N; i++) {
        const int32_t l = level[i];
        perm[start[l]] = i;
        order[i] = start[l];
        start[l]++;
    }
    nrp[0] = 0;
    for (int32_t i = 0; i < N; i++) nrp[i + 1] = nrp[i] + (rp[perm[i] + 1] - rp[perm[i]]);

    auto* a = new sable_spts_s();
    a->ctx = ctx;
    a->dtype = dtype;
    a->N = N;
    a->nnz = nnz;
    a->lanes = lanes ? lanes : pick_lanes(nnz, N);
    a->rp = row_ptr;
    a->col = col;
    a->val = val;
    a->nlevels = tot;
    a->h_level_ptr = lp;
    for (int32_t l = 0; l < tot; l++) a->max_width = std::max(a->max_width, freq[l]);
    const size_t vs = dtype == SABLE_F64 ? 8 : 4;
    bool ok = hipMalloc(&a->d_perm, 4 * (size_t)N) == hipSuccess && hipMalloc(&a->d_order, 4 * (size_t)N) == hipSuccess &&
              hipMalloc(&a->d_rp, 4 * (size_t)(N + 1)) == hipSuccess && hipMalloc(&a->d_col, 4 * (size_t)nnz) == hipSuccess &&
              hipMalloc(&a->d_val, vs * (size_t)nnz) == hipSuccess && hipMalloc(&a->d_bl, vs * (size_t)N) == hipSuccess &&
              hipMalloc(&a->d_yl, vs * (size_t)N) == hipSuccess && hipMalloc(&a->d_ticket, 256) == hipSuccess &&
              hipMalloc(&a->d_err, 256) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        free_analysis(a);
        return fail(SABLE_ENOMEM, "analysis allocation failed");
    }
    hipError_t e1 = hipMemcpyAsync(a->d_perm, perm.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st);
    hipError_t e2 = hipMemcpyAsync(a->d_order, order.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st);
    hipError_t e3 = hipMemcpyAsync(a->d_rp, nrp.data(), 4 * (size_t)(N + 1), hipMemcpyHostToDevice, st);
    hipError_t e4 = hipMemsetAsync(a->d_err, 0, 256, st);
    if (e1 || e2 || e3 || e4) {
        free_analysis(a);
        return fail(SABLE_EHIP, "analysis upload failed");
    }
    if (dtype == SABLE_F64) {
        hipLaunchKernelGGL(k_reorder<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, a->d_order, row_ptr, col,
                           (const double*)val, a->d_rp, a->d_col, (double*)a->d_val);
        hipLaunchKernelGGL(k_check_diag<double>, dim3(gsz(N)), dim3(TB), 0, st, N, row_ptr, (const double*)val, a->d_err);
    } else {
        hipLaunchKernelGGL(k_reorder<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, a->d_order, row_ptr, col,
                           (const float*)val, a->d_rp, a->d_col, (float*)a->d_val);
        hipLaunchKernelGGL(k_check_diag<float>, dim3(gsz(N)), dim3(TB), 0, st, N, row_ptr, (const float*)val, a->d_err);
    }
    int flag = 0;
    int rc = read_flag(ctx, a->d_err, &flag);  // also waits for the uploads of the host vectors above
    if (rc || flag) {
        free_analysis(a);
        return rc ? rc : fail(SABLE_ESINGULAR, "zero on the diagonal");
    }
    *out = a;
    return SABLE_OK;
}

SABLE_API void sable_spts_destroy(sable_spts_t* a) {
    if (!a) return;
    (void)hipStreamSynchronize(a->ctx->stream);
    free_analysis(a);
}

SABLE_API int sable_spts_info(const sable_spts_t* a, int32_t* nlevels, int32_t* max_width, int32_t* lanes) {
    SABLE_CHECK(a, SABLE_EINVAL, "NULL analysis");
    if (nlevels) *nlevels = a->nlevels;
    if (max_width) *max_width = a->max_width;
    if (lanes) *lanes = a->lanes;
    return SABLE_OK;
}

SABLE_API int sable_spts_set_lanes(sable_spts_t* a, int lanes) {
    SABLE_CHECK(a && valid_lanes(lanes), SABLE_EINVAL, "lanes must be 1, 4, 8, 16 or 64");
    a->lanes = lanes;
    return SABLE_OK;
}

SABLE_API int sable_spts_get_schedule(sable_ctx_t* ctx, const sable_spts_t* a, int32_t* h_level_ptr,
                                      int32_t* d_permutation, int32_t* d_order, int32_t* d_row_ptr, int32_t* d_col,
                                      void* d_val) {
    SABLE_CHECK(ctx && a, SABLE_EINVAL, "NULL argument");
    hipStream_t st = ctx->stream;
    const size_t vs = a->dtype == SABLE_F64 ? 8 : 4;
    if (h_level_ptr) memcpy(h_level_ptr, a->h_level_ptr.data(), 4 * a->h_level_ptr.size());
    if (d_permutation) SABLE_HIP(hipMemcpyAsync(d_permutation, a->d_perm, 4 * (size_t)a->N, hipMemcpyDeviceToDevice, st));
    if (d_order) SABLE_HIP(hipMemcpyAsync(d_order, a->d_order, 4 * (size_t)a->N, hipMemcpyDeviceToDevice, st));
    if (d_row_ptr) SABLE_HIP(hipMemcpyAsync(d_row_ptr, a->d_rp, 4 * (size_t)(a->N + 1), hipMemcpyDeviceToDevice, st));
    if (d_col) SABLE_HIP(hipMemcpyAsync(d_col, a->d_col, 4 * (size_t)a->nnz, hipMemcpyDeviceToDevice, st));
    if (d_val) SABLE_HIP(hipMemcpyAsync(d_val, a->d_val, vs * (size_t)a->nnz, hipMemcpyDeviceToDevice, st));
    return SABLE_OK;
}

SABLE_API int sable_spts_solve_level(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x) {
    SABLE_CHECK(ctx && a && b && x, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == a->ctx, SABLE_EINVAL, "analysis belongs to another context");
    hipStream_t st = ctx->stream;
    const int32_t N = a->N;
    if (a->dtype == SABLE_F64) {
        hipLaunchKernelGGL(k_gather_perm<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, (const double*)b,
                           (double*)a->d_bl);
        int rc = solve_level_ordered<double>(a, false);
        if (rc) return rc;
        hipLaunchKernelGGL(k_scatter_perm<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm,
                           (const double*)a->d_yl, (double*)x);
    } else {
        hipLaunchKernelGGL(k_gather_perm<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, (const float*)b,
                           (float*)a->d_bl);
        int rc = solve_level_ordered<float>(a, false);
        if (rc) return rc;
        hipLaunchKernelGGL(k_scatter_perm<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, (const float*)a->d_yl,
                           (float*)x);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_spts_solve_syncfree(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x) {
    SABLE_CHECK(ctx && a && b && x, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == a->ctx, SABLE_EINVAL, "analysis belongs to another context");
    if (a->dtype == SABLE_F64)
        launch_syncfree<double>(ctx->stream, a->lanes, ctx->num_cu, a->N, a->rp, a->col, (const double*)a->val,
                                (const double*)b, (double*)x, a->d_ticket, a->d_err, DEFAULT_SPIN);
    else
        launch_syncfree<float>(ctx->stream, a->lanes, ctx->num_cu, a->N, a->rp, a->col, (const float*)a->val,
                               (const float*)b, (float*)x, a->d_ticket, a->d_err, DEFAULT_SPIN);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// The reference's level + sync-free combination: the sync-free kernel over
// the analysis's level-ordered matrix (waves take rows in level order, so
// a wave's rows mostly share a level and rarely wait on each other), with
// the level solve's arithmetic.
SABLE_API int sable_spts_solve_combined(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x) {
    SABLE_CHECK(ctx && a && b && x, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == a->ctx, SABLE_EINVAL, "analysis belongs to another context");
    hipStream_t st = ctx->stream;
    const int32_t N = a->N;
    if (a->dtype == SABLE_F64) {
        hipLaunchKernelGGL(k_gather_perm<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, (const double*)b,
                           (double*)a->d_bl);
        launch_syncfree<double>(st, a->lanes, ctx->num_cu, N, a->d_rp, a->d_col, (const double*)a->d_val,
                                (const double*)a->d_bl, (double*)a->d_yl, a->d_ticket, a->d_err, DEFAULT_SPIN, true);
        hipLaunchKernelGGL(k_scatter_perm<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm,
                           (const double*)a->d_yl, (double*)x);
    } else {
        hipLaunchKernelGGL(k_gather_perm<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, (const float*)b,
                           (float*)a->d_bl);
        launch_syncfree<float>(st, a->lanes, ctx->num_cu, N, a->d_rp, a->d_col, (const float*)a->d_val,
                               (const float*)a->d_bl, (float*)a->d_yl, a->d_ticket, a->d_err, DEFAULT_SPIN, true);
        hipLaunchKernelGGL(k_scatter_perm<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, (const float*)a->d_yl,
                           (float*)x);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// Synchronises and reports (then clears) a timed-out sync-free solve.
SABLE_API int sable_spts_status(sable_ctx_t* ctx, sable_spts_t* a) {
    SABLE_CHECK(ctx && a, SABLE_EINVAL, "NULL argument");
    int flag = 0;
    int rc = read_flag(ctx, a->d_err, &flag);
    if (rc) return rc;
    if (flag) {
        SABLE_HIP(hipMemsetAsync(a->d_err, 0, sizeof(int), ctx->stream));
        return fail(SABLE_ETIMEOUT, "a sync-free solve exceeded its spin limit");
    }
    return SABLE_OK;
}

SABLE_API int sable_spts_solve_refined(sable_ctx_t* ctx, sable_spts_t* a, const double* b, double* x, int iters,
                                       int inner) {
    SABLE_CHECK(ctx && a && b && x, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(iters >= 0 && (inner == 0 || inner == 1), SABLE_EINVAL, "bad iters/inner");
    int rc = ensure_f32(a);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const int32_t N = a->N;
    float* r32 = (float*)a->d_work;
    float* d32 = r32 + ((N + 63) / 64) * 64;
    const int G = a->lanes;
    const unsigned rgrid = (unsigned)(((int64_t)N * G + TB - 1) / TB);
    for (int it = 0; it <= iters; it++) {
        // r = b - L x (x = 0 on the first pass)
        if (it == 0) {
            hipLaunchKernelGGL(k_to_f32, dim3(gsz(N)), dim3(TB), 0, st, (int64_t)N, b, r32);
        } else {
            switch (G) {
                case 1: hipLaunchKernelGGL((k_residual<1>), dim3(rgrid), dim3(TB), 0, st, N, a->rp, a->col, (const double*)a->val, x, b, r32); break;
                case 4: hipLaunchKernelGGL((k_residual<4>), dim3(rgrid), dim3(TB), 0, st, N, a->rp, a->col, (const double*)a->val, x, b, r32); break;
                case 8: hipLaunchKernelGGL((k_residual<8>), dim3(rgrid), dim3(TB), 0, st, N, a->rp, a->col, (const double*)a->val, x, b, r32); break;
                case 16: hipLaunchKernelGGL((k_residual<16>), dim3(rgrid), dim3(TB), 0, st, N, a->rp, a->col, (const double*)a->val, x, b, r32); break;
                default: hipLaunchKernelGGL((k_residual<64>), dim3(rgrid), dim3(TB), 0, st, N, a->rp, a->col, (const double*)a->val, x, b, r32); break;
            }
        }
        // L32 d = r (FP32)
        if (inner == 1) {
            launch_syncfree<float>(st, G, ctx->num_cu, N, a->rp, a->col, a->d_val32, r32, d32, a->d_ticket, a->d_err,
                                   DEFAULT_SPIN);
        } else {
            hipLaunchKernelGGL(k_gather_perm<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, r32, (float*)a->d_bl);
            // the level buffers are sized for the analysis dtype (FP64), large enough for FP32
            rc = solve_level_ordered<float>(a, true);
            if (rc) return rc;
            hipLaunchKernelGGL(k_scatter_perm<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm,
                               (const float*)a->d_yl, d32);
        }
        if (it == 0)
            hipLaunchKernelGGL(k_set_from_f32, dim3(gsz(N)), dim3(TB), 0, st, N, d32, x);
        else
            hipLaunchKernelGGL(k_add_f32, dim3(gsz(N)), dim3(TB), 0, st, N, d32, x);
        SABLE_LAUNCH_CHECK();
    }
    return SABLE_OK;
}

// Reference-shaped level-set solve (par/src/sparse_opt_32.wat:11759-11984):
// the matrix is already in level order, b and y are level-ordered and
// h_level_ptr (nlevels+1 entries) is on the host.  One launch per level.
SABLE_API int sable_spts_csc_analyze(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* col_ptr,
                                     const int32_t* row, const void* val, int lanes, sable_spts_t** out) {
    SABLE_CHECK(ctx && out && col_ptr && row && val, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N > 0, SABLE_EINVAL, "N must be positive");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    *out = nullptr;
    hipStream_t st = ctx->stream;
    int32_t ends[2];
    SABLE_HIP(hipMemcpyAsync(ends, col_ptr, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipMemcpyAsync(ends + 1, col_ptr + N, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    SABLE_CHECK(ends[0] == 0 && ends[1] >= N, SABLE_EMALFORMED, "col_ptr must start at 0 and hold the N diagonals");
    const int64_t nnz = ends[1];
    const size_t vs = dtype == SABLE_F64 ? 8 : 4;
    void *rp = nullptr, *cl = nullptr, *vl = nullptr;
    int* err = nullptr;
    auto release = [&]() {
        void* ps[] = {rp, cl, vl, err};
        for (void* q : ps)
            if (q) (void)hipFree(q);
    };
    if (hipMalloc(&rp, 4 * ((size_t)N + 1)) != hipSuccess || hipMalloc(&cl, 4 * (size_t)nnz) != hipSuccess ||
        hipMalloc(&vl, vs * (size_t)nnz) != hipSuccess || hipMalloc(&err, 256) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return fail(SABLE_ENOMEM, "CSC analysis allocation failed");
    }
    int rc = SABLE_OK, flag = 0;
    if (hipMemsetAsync(err, 0, sizeof(int), st) != hipSuccess) rc = fail(SABLE_EHIP, "memset failed");
    if (!rc) {
        hipLaunchKernelGGL(k_check_csc_diag_first, dim3(gsz(N)), dim3(TB), 0, st, N, col_ptr, row, err);
        rc = read_flag(ctx, err, &flag);
        if (!rc && flag) rc = fail(SABLE_EMALFORMED, "a CSC column is empty or does not start with its diagonal");
    }
    // the CSR of L is the transpose of the CSR view of the CSC arrays (rows
    // in ascending column order, diagonal last)
    if (!rc) rc = sable_transpose_csr(ctx, dtype, N, col_ptr, row, val, (int32_t*)rp, (int32_t*)cl, vl);
    if (!rc) rc = sable_spts_analyze(ctx, dtype, N, (const int32_t*)rp, (const int32_t*)cl, vl, lanes, out);
    (void)hipFree(err);
    err = nullptr;
    if (rc) {
        release();
        return rc;
    }
    (*out)->owned[0] = rp;
    (*out)->owned[1] = cl;
    (*out)->owned[2] = vl;
    return SABLE_OK;
}

SABLE_API int sable_spts_csc(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* col_ptr,
                             const int32_t* row, const void* val, const void* b, void* y, int inner_max) {
    SABLE_CHECK(inner_max >= 0, SABLE_EINVAL, "negative inner_max");
    if (inner_max == 0 || N <= 0) return SABLE_OK;
    sable_spts_t* a = nullptr;
    int rc = sable_spts_csc_analyze(ctx, dtype, N, col_ptr, row, val, 1, &a);
    if (rc) return rc;
    for (int it = 0; it < inner_max && !rc; it++) rc = sable_spts_solve_syncfree(ctx, a, b, y);
    if (!rc) rc = sable_spts_status(ctx, a);
    sable_spts_destroy(a);
    return rc;
}

SABLE_API int sable_spts_level_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t nlevels,
                                   const int32_t* h_level_ptr, const int32_t* row_ptr, const int32_t* col,
                                   const void* val, const void* b, void* y, int lanes) {
    SABLE_CHECK(ctx && h_level_ptr, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0 && nlevels >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(valid_lanes(lanes), SABLE_EINVAL, "lanes must be 1, 4, 8, 16 or 64");
    if (nlevels == 0 || N == 0) return SABLE_OK;
    SABLE_CHECK(h_level_ptr[0] == 0 && h_level_ptr[nlevels] == N, SABLE_EINVAL, "level_ptr must span [0, N]");
    std::vector<int32_t> lp(h_level_ptr, h_level_ptr + nlevels + 1);
    if (dtype == SABLE_F64)
        return levels_on_stream<double>(ctx->stream, lanes, lp, row_ptr, col, (const double*)val, (const double*)b,
                                        (double*)y);
    if (dtype == SABLE_F32)
        return levels_on_stream<float>(ctx->stream, lanes, lp, row_ptr, col, (const float*)val, (const float*)b,
                                       (float*)y);
    return fail(SABLE_EINVAL, "bad dtype");
}

// Reference-shaped sync-free solve (par/src/spts_opt_32.wat:8-260): natural
// order, the structure is checked first, waits are bounded by spin_limit
// polls without progress (0 = default).  Synchronises; SABLE_ETIMEOUT when
// the limit was hit.
SABLE_API int sable_spts_syncfree_csr(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                      const int32_t* col, const void* val, const void* b, void* x, int lanes,
                                      long long spin_limit) {
    SABLE_CHECK(ctx && row_ptr, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N >= 0, SABLE_EINVAL, "negative size");
    SABLE_CHECK(valid_lanes(lanes), SABLE_EINVAL, "lanes must be 1, 4, 8, 16 or 64");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    if (N == 0) return SABLE_OK;
    void* s;
    int rc = ctx_scratch(ctx, 512, &s);
    if (rc) return rc;
    unsigned* ticket = (unsigned*)s;
    int* err = (int*)((char*)s + 256);
    hipStream_t st = ctx->stream;
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_check_lower, dim3(gsz(N)), dim3(TB), 0, st, N, row_ptr, col, err);
    int flag = 0;
    rc = read_flag(ctx, err, &flag);
    if (rc) return rc;
    SABLE_CHECK(!flag, SABLE_EMALFORMED, "not a lower-triangular CSR with the diagonal last in each row");
    const long long spin = spin_limit > 0 ? spin_limit : DEFAULT_SPIN;
    if (dtype == SABLE_F64)
        launch_syncfree<double>(st, lanes, ctx->num_cu, N, row_ptr, col, (const double*)val, (const double*)b,
                                (double*)x, ticket, err, spin);
    else
        launch_syncfree<float>(st, lanes, ctx->num_cu, N, row_ptr, col, (const float*)val, (const float*)b, (float*)x,
                               ticket, err, spin);
    SABLE_LAUNCH_CHECK();
    rc = read_flag(ctx, err, &flag);
    if (rc) return rc;
    SABLE_CHECK(!flag, SABLE_ETIMEOUT, "sync-free solve exceeded its spin limit");
    return SABLE_OK;
}

}  // extern "C"
