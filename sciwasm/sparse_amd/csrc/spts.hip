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
#include "refine.h"
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
    // captured level sweeps: [0] analysis dtype, [1] FP32 copy (refinement),
    // [2] the refined solve fused into one sweep
    int graph_lanes[3] = {0, 0, 0};
    hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    // CSR copy owned by the analysis (built from a CSC input), else null
    void* owned[3] = {nullptr, nullptr, nullptr};
    // Row-block schedule (k_blocked): rows cut into blocks of bl_rows; in a
    // block, rows ordered by level and cut into steps of at most 256
    // rows; each step's off-diagonals stored step-major (entry j of the
    // step's row t at base + j*rows + t, padding column -1).
    int schedule = SABLE_SPTS_BLOCKS;
    int32_t bl_rows = 0;   // 0: not built
    int32_t bl_nblk = 0, bl_nsteps = 0, bl_max_steps = 0, bl_width = 0;
    int64_t bl_slots = 0;
    int32_t* d_bl_blkstep = nullptr;  // nblk + 1 step ranges
    int4* d_bl_steps = nullptr;       // {first schedule row, rows, width, entry base}
    int32_t* d_bl_row = nullptr;      // N: row of each schedule position
    void* d_bl_diag = nullptr;        // N diagonal values in schedule order (dtype)
    int32_t* d_bl_col = nullptr;      // slots
    void* d_bl_val = nullptr;         // slots (dtype)
    float* d_bl_diag32 = nullptr;     // FP32 copies for the refined solve
    float* d_bl_val32 = nullptr;
    // in-degree sync-free solve (built on first use): the rows that depend
    // on each row (strictly lower part by column), per-row in-degree, and
    // per-row counters of published dependencies
    int32_t* d_dep_ptr = nullptr;  // N + 1
    int32_t* d_dep_row = nullptr;  // nnz - N
    int32_t* d_indeg = nullptr;    // N
    int32_t* d_cnt = nullptr;      // N
    // host structure and levels kept to rebuild the block schedule
    std::vector<int32_t> h_rp, h_col, h_level;
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
// Rows [r0, r0+nrows) of a level; G lanes per row.  The lanes load and
// multiply G entries of the row at once (entries j0 + l), and the sum then
// takes those products in stored order: every lane of the group adds all of
// them, read across the group, so the sum, and y, are the one-lane result
// bit for bit whatever G.
template <typename T, int G>
__global__ __launch_bounds__(TB) void k_level(int32_t r0, int32_t nrows, const int32_t* __restrict__ rp,
                                              const int32_t* __restrict__ col, const T* __restrict__ val,
                                              const T* __restrict__ bl, T* __restrict__ yl) {
    __shared__ T stage[TB];
    const int64_t gt = (int64_t)blockIdx.x * TB + threadIdx.x;
    const int64_t g = gt / G;
    const int l = threadIdx.x & (G - 1);
    const bool valid = g < nrows;
    const int32_t i = r0 + (int32_t)(valid ? g : 0);
    T s = T(0);
    int32_t a = 0, e = 0;
    if (valid) {
        a = rp[i];
        e = rp[i + 1] - 1;
    }
    // a and e are the same on the lanes of a group, so is every trip count
    for (int32_t j0 = a; j0 < e; j0 += G) {
        const int32_t j = j0 + l;
        const T p = j < e ? mul_rn(val[j], yl[col[j]]) : T(0);
        if (G == 1) {
            s = add_rn(s, p);
        } else {
            const int n = min(G, e - j0);
            s = group_ordered_sum_lds<false, G>(s, p, n, stage);
        }
    }
    if (valid && l == 0) yl[i] = div_rn(sub_rn(bl[i], s), val[e]);
}

// A run of consecutive narrow levels in one launch of one workgroup: the
// levels one after another with a barrier between them, each in passes of
// 1024 / G rows, with k_level's arithmetic (the same bits).  The run's level
// bounds travel in the kernel arguments.
constexpr int RUN_MAX = 64;
constexpr int RUN_TB = 1024;
struct LevelRun {
    int32_t n;                 // levels in the run
    int32_t lp[RUN_MAX + 1];   // their bounds in the level-ordered rows
};

template <typename T, int G>
__global__ __launch_bounds__(RUN_TB) void k_level_run(LevelRun run, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ col, const T* __restrict__ val,
                                                      const T* __restrict__ bl, T* yl) {
    __shared__ T stage[RUN_TB];
    constexpr int RPP = RUN_TB / G;
    const int g = threadIdx.x / G;
    const int l = threadIdx.x & (G - 1);
    for (int k = 0; k < run.n; k++) {
        const int32_t r1 = run.lp[k + 1];
        for (int32_t base = run.lp[k]; base < r1; base += RPP) {
            const int32_t i = base + g;
            const bool valid = i < r1;
            T s = T(0);
            int32_t a = 0, e = 0;
            if (valid) {
                a = rp[i];
                e = rp[i + 1] - 1;
            }
            for (int32_t j0 = a; j0 < e; j0 += G) {
                const int32_t j = j0 + l;
                const T p = j < e ? mul_rn(val[j], yl[col[j]]) : T(0);
                if (G == 1) {
                    s = add_rn(s, p);
                } else {
                    const int n = min(G, e - j0);
                    s = group_ordered_sum_lds<false, G>(s, p, n, stage);
                }
            }
            if (valid && l == 0) yl[i] = div_rn(sub_rn(bl[i], s), val[e]);
        }
        __syncthreads();
    }
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
// off-diagonals, x = (b - t) / diag); else the reference sync-free solve's
// t = b, t -= val * x, x = t / diag.  G lanes per row take the row's entries
// G at a time: each lane waits for its own x and forms its product, and once
// the group holds all G products they are added to t in stored order (by
// every lane of the group, read across it), so every lane count gives the
// one-lane result bit for bit.
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
        // j0: the group's next chunk of G entries (the same on its lanes)
        int32_t j0 = 0, e = 0;
        T s = T(0);
        T bi = T(0);
        // this lane's product of the chunk, once its x has been published
        T p = T(0);
        bool have = false;
        if (!published) {
            j0 = rp[i];
            e = rp[i + 1] - 1;
            bi = b[i];
            if (!LVL) s = bi;
        }
        long long spins = 0;
        while (__ballot(!published) != 0ull) {
            bool progressed = false;
            if (!published) {
                if (G == 1) {
                    while (j0 < e) {
                        const T v = load_agent(&x[col[j0]]);
                        if (Sent<T>::is(v)) break;
                        s = LVL ? add_rn(s, mul_rn(val[j0], v)) : sub_rn(s, mul_rn(val[j0], v));
                        j0++;
                        progressed = true;
                    }
                } else {
                    while (j0 < e) {
                        const int32_t j = j0 + l;
                        if (!have && j < e) {
                            const T v = load_agent(&x[col[j]]);
                            if (!Sent<T>::is(v)) {
                                p = mul_rn(val[j], v);
                                have = true;
                            }
                        }
                        // group-uniform: the whole group leaves or goes on
                        if ((__ballot(have || j >= e) & gmask) != gmask) break;
                        const int n = min(G, e - j0);
                        s = group_ordered_sum<!LVL, G>(s, p, n);
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
                    const T y = LVL ? div_rn(sub_rn(bi, s), val[e]) : div_rn(s, val[e]);
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

// ------------------------------------------------- in-degree sync-free
// The sync-free solve with per-row in-degree counters: row i waits until
// cnt[i], the number of its off-diagonal dependencies already solved,
// reaches indeg[i]; it then forms t = b - v0*x0 - v1*x1 - ... in stored
// order (G lanes form G products at once, added in order), x = t / d, and
// pushes: one increment of cnt[r] for every row r that depends on it.  The
// push follows the completed store of x (the wave waits for it), and x and
// the counters move through agent-scope atomics, so a row that sees its
// count complete reads published values.  Rows are claimed in increasing
// order; a wave serves every row of its claim that is ready in each pass,
// so a row waiting on another row of the same wave does not stall it.
__global__ void k_indeg_reset(int32_t N, int32_t* __restrict__ cnt, unsigned* __restrict__ ticket) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        cnt[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0u;
}

template <typename T, int G>
__global__ __launch_bounds__(TB) void k_syncfree_indeg(int32_t N, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ col, const T* __restrict__ val,
                                                       const int32_t* __restrict__ dep_ptr,
                                                       const int32_t* __restrict__ dep_row,
                                                       const int32_t* __restrict__ indeg, int32_t* cnt,
                                                       const T* __restrict__ b, T* x, unsigned* __restrict__ ticket,
                                                       int* __restrict__ err, long long spin_limit) {
    constexpr int R = WAVE / G;
    const int lane = threadIdx.x & (WAVE - 1);
    const int l = lane & (G - 1);
    const int grp = lane / G;
    while (true) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(ticket, (unsigned)R);
        base = __shfl(base, 0, WAVE);
        if (base >= (unsigned)N) break;
        const int64_t i = (int64_t)base + grp;
        bool done = i >= N;
        const int32_t need = done ? 0 : indeg[i];
        long long spins = 0;
        while (__ballot(!done) != 0ull) {
            // the group's lanes read the same counter: `ready` is group-uniform
            const bool ready = !done && load_agent(&cnt[i]) >= need;
            if (ready) {
                const int32_t a0 = rp[i], e = rp[i + 1] - 1;
                T s = b[i];
                for (int32_t j0 = a0; j0 < e; j0 += G) {
                    const int32_t j = j0 + l;
                    const T p = j < e ? mul_rn(val[j], load_agent(&x[col[j]])) : T(0);
                    if (G == 1) s = sub_rn(s, p);
                    else s = group_ordered_sum<true, G>(s, p, min(G, e - j0));
                }
                if (l == 0) store_agent(&x[i], div_rn(s, val[e]));
                // the store has completed before any dependent is told
                __builtin_amdgcn_s_waitcnt(0);
                for (int32_t k = dep_ptr[i] + l; k < dep_ptr[i + 1]; k += G)
                    __hip_atomic_fetch_add(&cnt[dep_row[k]], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                done = true;
            }
            if (__ballot(ready) == 0ull) {
                if (++spins > spin_limit) {
                    if (lane == 0) atomicOr(err, 4);
                    // give up: publish NaN and push, so that waiting rows finish too
                    if (!done) {
                        if (l == 0) store_agent(&x[i], T(__builtin_nan("")));
                        __builtin_amdgcn_s_waitcnt(0);
                        for (int32_t k = dep_ptr[i] + l; k < dep_ptr[i + 1]; k += G)
                            __hip_atomic_fetch_add(&cnt[dep_row[k]], 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                    }
                    done = true;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
}

// ------------------------------------------------------- row blocks
// One workgroup per block of consecutive rows (claimed in increasing order
// from a ticket, so every block a workgroup waits on is held by a running
// workgroup).  Inside a block the rows run step by step in level order (one
// level per step, at most SB rows), with the block's solution in LDS and a
// barrier between steps; a dependency on an earlier block is read from x,
// whose entries hold
// a sentinel until they are published (one agent-scope store per row).  The
// next step's rows and entries are loaded while the current step runs.  One
// thread per row, entries in stored order: SF = false is the level solve's
// arithmetic t = 0 + v0*y0 + ..., y = (b - t)/d; SF = true the sync-free
// one, t = b - v0*y0 - ..., y = t/d.
constexpr int SB = 256;  // rows per step, threads per workgroup
// P (a template parameter): entries per row loaded ahead of their step, the
// schedule's widest row rounded up to 4, 8 or 16; wider rows take the rest P
// at a time in their step.

template <typename T, int P>
struct BlRow {
    int32_t rid;
    int32_t c[P];
    T v[P];
    T d;
};

template <typename T, int P>
__device__ __forceinline__ void bl_load(BlRow<T, P>& r, const int4 st, const int32_t* __restrict__ srow,
                                        const T* __restrict__ diag, const int32_t* __restrict__ scol,
                                        const T* __restrict__ sval) {
    const int t = threadIdx.x;
    r.rid = -1;
    if (t < st.y) {
        const int32_t k = st.x + t;
        r.rid = srow[k];
        r.d = diag[k];
#pragma unroll
        for (int j = 0; j < P; j++) {
            const bool in = j < st.z;
            const int64_t e = (int64_t)st.w + (int64_t)j * st.y + t;
            r.c[j] = in ? scol[e] : -1;
            r.v[j] = in ? sval[e] : T(0);
        }
    }
}

// The x values of one group of P entries: local columns from LDS; columns
// of earlier blocks with plain loads first, all issued together (a value
// already published is final wherever it is cached, so only the sentinel
// can be stale), then agent-scope polls, again all together, for the ones
// that read the sentinel (a spin limit ends the wait with NaN and raises
// err).  Padding (column -1) reads 0.
template <typename T>
__device__ __forceinline__ T load_cached(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, int P>
__device__ __forceinline__ void bl_fetch(const int32_t (&c)[P], T (&xv)[P], const T* xl, int32_t r0, T* x,
                                         int* err, long long spin) {
    uint32_t pend = 0;
#pragma unroll
    for (int j = 0; j < P; j++) {
        if (c[j] >= r0) {
            xv[j] = xl[c[j] - r0];
        } else if (c[j] >= 0) {
            xv[j] = load_cached(&x[c[j]]);
            pend |= 1u << j;
        } else {
            xv[j] = T(0);
        }
    }
    long long n = 0;
    while (true) {
#pragma unroll
        for (int j = 0; j < P; j++)
            if (((pend >> j) & 1u) && !Sent<T>::is(xv[j])) pend &= ~(1u << j);
        if (!pend) break;
        if (++n > spin) {
            atomicOr(err, 4);
#pragma unroll
            for (int j = 0; j < P; j++)
                if ((pend >> j) & 1u) xv[j] = T(__builtin_nan(""));
            break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int j = 0; j < P; j++)
            if ((pend >> j) & 1u) xv[j] = load_agent(&x[c[j]]);
    }
}

template <typename T, bool SF>
__device__ __forceinline__ T bl_acc(T t, T v, T y) {
    return SF ? sub_rn(t, mul_rn(v, y)) : add_rn(t, mul_rn(v, y));
}

template <typename T, bool SF, int P>
__global__ __launch_bounds__(SB) void k_blocked(int32_t nblk, int32_t B, const int32_t* __restrict__ blkstep,
                                                const int4* __restrict__ steps, const int32_t* __restrict__ srow,
                                                const T* __restrict__ diag, const int32_t* __restrict__ scol,
                                                const T* __restrict__ sval, const T* __restrict__ b, T* x,
                                                unsigned* __restrict__ ticket, int* __restrict__ err,
                                                long long spin) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* xl = reinterpret_cast<T*>(smem_raw);
    __shared__ unsigned s_blk;
    while (true) {
        __syncthreads();  // the previous block's xl and s_blk are no longer read
        if (threadIdx.x == 0) s_blk = atomicAdd(ticket, 1u);
        __syncthreads();
        const unsigned blk = s_blk;
        if (blk >= (unsigned)nblk) break;
        const int32_t r0 = (int32_t)blk * B;
        const int32_t s0 = blkstep[blk], s1 = blkstep[blk + 1];
        BlRow<T, P> cur, nxt;
        int4 st = s0 < s1 ? steps[s0] : make_int4(0, 0, 0, 0);
        if (s0 < s1) bl_load(cur, st, srow, diag, scol, sval);
        for (int32_t s = s0; s < s1; s++) {
            const int4 stn = s + 1 < s1 ? steps[s + 1] : make_int4(0, 0, 0, 0);
            if (s + 1 < s1) bl_load(nxt, stn, srow, diag, scol, sval);
            if (cur.rid >= 0) {
                const T bi = b[cur.rid];
                T xv[P];
                bl_fetch(cur.c, xv, xl, r0, x, err, spin);
                T t = SF ? bi : T(0);
#pragma unroll
                for (int j = 0; j < P; j++)
                    if (cur.c[j] >= 0) t = bl_acc<T, SF>(t, cur.v[j], xv[j]);
                // entries past the first P, P at a time
                for (int32_t j0 = P; j0 < st.z; j0 += P) {
                    int32_t c[P];
                    T v[P];
#pragma unroll
                    for (int j = 0; j < P; j++) {
                        const bool in = j0 + j < st.z;
                        const int64_t e = (int64_t)st.w + (int64_t)(j0 + j) * st.y + threadIdx.x;
                        c[j] = in ? scol[e] : -1;
                        v[j] = in ? sval[e] : T(0);
                    }
                    bl_fetch(c, xv, xl, r0, x, err, spin);
#pragma unroll
                    for (int j = 0; j < P; j++)
                        if (c[j] >= 0) t = bl_acc<T, SF>(t, v[j], xv[j]);
                }
                const T y = SF ? div_rn(t, cur.d) : div_rn(sub_rn(bi, t), cur.d);
                xl[cur.rid - r0] = y;
                store_agent(&x[cur.rid], y);
            }
            __syncthreads();
            cur = nxt;
            st = stn;
        }
    }
}

template <typename T>
__global__ void k_bl_gather(int64_t n, const int32_t* __restrict__ idx, const T* __restrict__ val,
                            T* __restrict__ out) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = idx[s];
        out[s] = j >= 0 ? val[j] : T(0);
    }
}

// b[i] = sum_j L_ij (a right-hand side whose solution is all ones, for
// timing solves on a matrix without the caller's vectors).
template <typename T>
__global__ void k_rowsum(int32_t N, const int32_t* __restrict__ rp, const T* __restrict__ val, T* __restrict__ b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        T s = T(0);
        for (int32_t j = rp[i]; j < rp[i + 1]; j++) s += val[j];
        b[i] = s;
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

// ------------------------------------------------- fused refinement
// The FP32 solve and one FP64 refinement step in one sweep.  Row i needs
// only rows of earlier levels, for all three of its passes:
//   x0_i = ((float)b_i - (0 + w_0 x0_c0 + w_1 x0_c1 + ...)) / w_d      FP32, w = (float)L
//   s    = fma(L_d, x0_i, fma(... fma(L_0, x0_c0, 0)))                  FP64, stored order
//   r_i  = (float)(b_i - s)
//   d_i  = (r_i - (0 + w_0 d_c0 + ...)) / w_d                           FP32
//   x_i  = (double)x0_i + (double)d_i
// so the sweep that forms x0_i can form r_i and d_i right after it: the
// refined solve costs one latency-bound sweep instead of two sweeps, a
// residual and an update.  x0 is the FP32 level solve's x bit for bit; every
// fused schedule gives the same x.

template <int G>
__global__ __launch_bounds__(TB) void k_level_ref(int32_t r0, int32_t nrows, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col, const double* __restrict__ v64,
                                                  const double* __restrict__ b, float2* XD,
                                                  double* __restrict__ xo) {
    __shared__ RefEntry stage[TB];
    const int64_t g = ((int64_t)blockIdx.x * TB + threadIdx.x) / G;
    const bool valid = g < nrows;
    ref_row<G>(r0 + (int32_t)(valid ? g : 0), valid, rp, col, v64, b, XD, xo, stage);
}

template <int G>
__global__ __launch_bounds__(RUN_TB) void k_level_ref_run(LevelRun run, const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ col,
                                                          const double* __restrict__ v64,
                                                          const double* __restrict__ b, float2* XD,
                                                          double* __restrict__ xo) {
    __shared__ RefEntry stage[RUN_TB];
    constexpr int RPP = RUN_TB / G;
    const int g = threadIdx.x / G;
    for (int k = 0; k < run.n; k++) {
        const int32_t r1 = run.lp[k + 1];
        for (int32_t base = run.lp[k]; base < r1; base += RPP) {
            const int32_t i = base + g;
            ref_row<G>(i < r1 ? i : base, i < r1, rp, col, v64, b, XD, xo, stage);
        }
        __syncthreads();
    }
}

// Row blocks (k_blocked's schedule and waits) for the fused refined solve:
// a row publishes x0 and d together, as one 64-bit word (x0 in the low
// half), in W, which holds the word sentinel until then; the block's words
// sit in LDS.  xo is the FP64 result.
constexpr uint64_t WSENT = ((uint64_t)SENT32 << 32) | (uint64_t)SENT32;

__device__ __forceinline__ uint64_t pack_x0_d(float x0, float d) {
    return (uint64_t)__float_as_uint(x0) | ((uint64_t)__float_as_uint(d) << 32);
}
__device__ __forceinline__ float word_x0(uint64_t w) { return __uint_as_float((uint32_t)w); }
__device__ __forceinline__ float word_d(uint64_t w) { return __uint_as_float((uint32_t)(w >> 32)); }

template <int P>
struct RfRow {
    int32_t rid;
    int32_t c[P];
    double v[P];
    double v_d;
};

template <int P>
__device__ __forceinline__ void rf_load(RfRow<P>& r, const int4 st, const int32_t* __restrict__ srow,
                                        const double* __restrict__ diag64, const int32_t* __restrict__ scol,
                                        const double* __restrict__ sval64) {
    const int t = threadIdx.x;
    r.rid = -1;
    if (t < st.y) {
        const int32_t k = st.x + t;
        r.rid = srow[k];
        r.v_d = diag64[k];
#pragma unroll
        for (int j = 0; j < P; j++) {
            const bool in = j < st.z;
            const int64_t e = (int64_t)st.w + (int64_t)j * st.y + t;
            r.c[j] = in ? scol[e] : -1;
            r.v[j] = in ? sval64[e] : 0.0;
        }
    }
}

// The words of one group of P entries, as bl_fetch reads x values.
template <int P>
__device__ __forceinline__ void rf_fetch(const int32_t (&c)[P], uint64_t (&wv)[P], const uint64_t* wl, int32_t r0,
                                         uint64_t* W, int* err, long long spin) {
    uint32_t pend = 0;
#pragma unroll
    for (int j = 0; j < P; j++) {
        if (c[j] >= r0) {
            wv[j] = wl[c[j] - r0];
        } else if (c[j] >= 0) {
            wv[j] = load_cached(&W[c[j]]);
            pend |= 1u << j;
        } else {
            wv[j] = 0;
        }
    }
    long long n = 0;
    while (true) {
#pragma unroll
        for (int j = 0; j < P; j++)
            if (((pend >> j) & 1u) && wv[j] != WSENT) pend &= ~(1u << j);
        if (!pend) break;
        if (++n > spin) {
            atomicOr(err, 4);
            const uint32_t q = __float_as_uint(__builtin_nanf(""));
#pragma unroll
            for (int j = 0; j < P; j++)
                if ((pend >> j) & 1u) wv[j] = ((uint64_t)q << 32) | q;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int j = 0; j < P; j++)
            if ((pend >> j) & 1u) wv[j] = load_agent(&W[c[j]]);
    }
}

// The FP32 value of an entry is its FP64 value rounded.
template <int P>
__device__ __forceinline__ void rf_acc(const int32_t (&c)[P], const double (&v)[P], const uint64_t (&wv)[P],
                                       float& s0, float& s2, double& s) {
#pragma unroll
    for (int j = 0; j < P; j++)
        if (c[j] >= 0) {
            const float w = (float)v[j];
            const float x0 = word_x0(wv[j]);
            s0 = add_rn(s0, mul_rn(w, x0));
            s2 = add_rn(s2, mul_rn(w, word_d(wv[j])));
            s = fma(v[j], (double)x0, s);
        }
}

template <int P>
__global__ __launch_bounds__(SB) void k_blocked_ref(int32_t nblk, int32_t B, const int32_t* __restrict__ blkstep,
                                                    const int4* __restrict__ steps, const int32_t* __restrict__ srow,
                                                    const double* __restrict__ diag64,
                                                    const int32_t* __restrict__ scol,
                                                    const double* __restrict__ sval64, const double* __restrict__ b,
                                                    uint64_t* W, double* __restrict__ xo,
                                                    unsigned* __restrict__ ticket, int* __restrict__ err,
                                                    long long spin) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint64_t* wl = reinterpret_cast<uint64_t*>(smem_raw);
    __shared__ unsigned s_blk;
    while (true) {
        __syncthreads();  // the previous block's wl and s_blk are no longer read
        if (threadIdx.x == 0) s_blk = atomicAdd(ticket, 1u);
        __syncthreads();
        const unsigned blk = s_blk;
        if (blk >= (unsigned)nblk) break;
        const int32_t r0 = (int32_t)blk * B;
        const int32_t s0i = blkstep[blk], s1i = blkstep[blk + 1];
        RfRow<P> cur, nxt;
        int4 st = s0i < s1i ? steps[s0i] : make_int4(0, 0, 0, 0);
        if (s0i < s1i) rf_load(cur, st, srow, diag64, scol, sval64);
        for (int32_t q = s0i; q < s1i; q++) {
            const int4 stn = q + 1 < s1i ? steps[q + 1] : make_int4(0, 0, 0, 0);
            if (q + 1 < s1i) rf_load(nxt, stn, srow, diag64, scol, sval64);
            if (cur.rid >= 0) {
                const double bi = b[cur.rid];
                uint64_t wv[P];
                rf_fetch(cur.c, wv, wl, r0, W, err, spin);
                float s0 = 0.f, s2 = 0.f;
                double s = 0.0;
                rf_acc(cur.c, cur.v, wv, s0, s2, s);
                // entries past the first P, P at a time
                for (int32_t j0 = P; j0 < st.z; j0 += P) {
                    int32_t c[P];
                    double v[P];
#pragma unroll
                    for (int j = 0; j < P; j++) {
                        const bool in = j0 + j < st.z;
                        const int64_t e = (int64_t)st.w + (int64_t)(j0 + j) * st.y + threadIdx.x;
                        c[j] = in ? scol[e] : -1;
                        v[j] = in ? sval64[e] : 0.0;
                    }
                    rf_fetch(c, wv, wl, r0, W, err, spin);
                    rf_acc(c, v, wv, s0, s2, s);
                }
                const float wd = (float)cur.v_d;
                const float x0 = div_rn(sub_rn((float)bi, s0), wd);
                s = fma(cur.v_d, (double)x0, s);
                const float d = div_rn(sub_rn((float)(bi - s), s2), wd);
                const uint64_t word = pack_x0_d(x0, d);
                wl[cur.rid - r0] = word;
                store_agent(&W[cur.rid], word);
                xo[cur.rid] = (double)x0 + (double)d;
            }
            __syncthreads();
            cur = nxt;
            st = stn;
        }
    }
}

__global__ void k_fill_words(int32_t N, uint64_t* __restrict__ W, unsigned* __restrict__ ticket) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        W[i] = WSENT;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0u;
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
static void launch_run(hipStream_t st, int G, const LevelRun& run, const int32_t* rp, const int32_t* col, const T* val,
                       const T* bl, T* yl) {
#define RUN_LAUNCH(GG) \
    hipLaunchKernelGGL((k_level_run<T, GG>), dim3(1), dim3(RUN_TB), 0, st, run, rp, col, val, bl, yl)
    switch (G) {
        case 1: RUN_LAUNCH(1); break;
        case 4: RUN_LAUNCH(4); break;
        case 8: RUN_LAUNCH(8); break;
        case 16: RUN_LAUNCH(16); break;
        default: RUN_LAUNCH(64); break;
    }
#undef RUN_LAUNCH
}

// The levels in order: a level of at most two passes of one 1024-thread
// workgroup (2048 / G rows) joins a run of such levels solved in one
// launch (up to RUN_MAX levels), every other level takes its own launch.
template <typename T>
static int levels_on_stream(hipStream_t st, int G, const std::vector<int32_t>& lp, const int32_t* rp,
                            const int32_t* col, const T* val, const T* bl, T* yl) {
    const int nl = (int)lp.size() - 1;
    const int32_t narrow = 2 * (RUN_TB / G);
    LevelRun run;
    run.n = 0;
    auto flush = [&]() {
        if (run.n == 1) launch_level<T>(st, G, run.lp[0], run.lp[1] - run.lp[0], rp, col, val, bl, yl);
        else if (run.n > 1) launch_run<T>(st, G, run, rp, col, val, bl, yl);
        run.n = 0;
    };
    for (int l = 0; l < nl; l++) {
        const int32_t r0 = lp[l], nr = lp[l + 1] - lp[l];
        if (nr <= 0) continue;
        if (nr <= narrow) {
            if (run.n == 0) run.lp[0] = r0;
            run.lp[++run.n] = r0 + nr;
            if (run.n == RUN_MAX) flush();
        } else {
            flush();
            launch_level<T>(st, G, r0, nr, rp, col, val, bl, yl);
        }
    }
    flush();
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// The fused refined solve over level-ordered buffers, one launch per wide
// level and one workgroup per run of narrow levels (as levels_on_stream).
template <int G>
static void launch_level_ref(hipStream_t st, int32_t r0, int32_t nr, const LevelRun* run, const int32_t* rp,
                             const int32_t* col, const double* v64, const double* bl, float2* XD, double* yl) {
    if (run)
        hipLaunchKernelGGL((k_level_ref_run<G>), dim3(1), dim3(RUN_TB), 0, st, *run, rp, col, v64, bl, XD, yl);
    else
        hipLaunchKernelGGL((k_level_ref<G>), dim3((unsigned)(((int64_t)nr * G + TB - 1) / TB)), dim3(TB), 0, st, r0,
                           nr, rp, col, v64, bl, XD, yl);
}

static int levels_ref_on_stream(hipStream_t st, int G, const std::vector<int32_t>& lp, const int32_t* rp,
                                const int32_t* col, const double* v64, const double* bl, float2* XD, double* yl) {
    const int nl = (int)lp.size() - 1;
    const int32_t narrow = 2 * (RUN_TB / G);
    LevelRun run;
    run.n = 0;
    auto go = [&](int32_t r0, int32_t nr, const LevelRun* rr) {
        switch (G) {
            case 1: launch_level_ref<1>(st, r0, nr, rr, rp, col, v64, bl, XD, yl); break;
            case 4: launch_level_ref<4>(st, r0, nr, rr, rp, col, v64, bl, XD, yl); break;
            case 8: launch_level_ref<8>(st, r0, nr, rr, rp, col, v64, bl, XD, yl); break;
            case 16: launch_level_ref<16>(st, r0, nr, rr, rp, col, v64, bl, XD, yl); break;
            default: launch_level_ref<64>(st, r0, nr, rr, rp, col, v64, bl, XD, yl); break;
        }
    };
    auto flush = [&]() {
        if (run.n == 1) go(run.lp[0], run.lp[1] - run.lp[0], nullptr);
        else if (run.n > 1) go(0, 0, &run);
        run.n = 0;
    };
    for (int l = 0; l < nl; l++) {
        const int32_t r0 = lp[l], nr = lp[l + 1] - lp[l];
        if (nr <= 0) continue;
        if (nr <= narrow) {
            if (run.n == 0) run.lp[0] = r0;
            run.lp[++run.n] = r0 + nr;
            if (run.n == RUN_MAX) flush();
        } else {
            flush();
            go(r0, nr, nullptr);
        }
    }
    flush();
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

// Rows per block: about 256 blocks for the matrix, between 1024 and 8192.
static int32_t default_block_rows(int32_t N) {
    const int64_t want = ((int64_t)N + 255) / 256;
    int32_t B = 1024;
    while (B < want && B < 8192) B *= 2;
    return B;
}

static void free_blocked(sable_spts_s* a) {
    void* ptrs[] = {a->d_bl_blkstep, a->d_bl_steps, a->d_bl_row, a->d_bl_diag, a->d_bl_col, a->d_bl_val,
                    a->d_bl_diag32, a->d_bl_val32};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    a->d_bl_blkstep = nullptr;
    a->d_bl_steps = nullptr;
    a->d_bl_row = nullptr;
    a->d_bl_diag = nullptr;
    a->d_bl_col = nullptr;
    a->d_bl_val = nullptr;
    a->d_bl_diag32 = nullptr;
    a->d_bl_val32 = nullptr;
    a->bl_rows = a->bl_nblk = a->bl_nsteps = a->bl_max_steps = a->bl_width = 0;
    a->bl_slots = 0;
}

// The row-block schedule of k_blocked for blocks of B rows, from the host
// copy of the structure; the values are gathered on the device from the
// caller's natural-order array.
static int build_blocked(sable_spts_s* a, int32_t B) {
    SABLE_CHECK(B >= 64 && B <= 16384 && (B & (B - 1)) == 0, SABLE_EINVAL,
                "block rows must be a power of two in [64, 16384]");
    free_blocked(a);
    const int32_t N = a->N;
    const std::vector<int32_t>& rp = a->h_rp;
    const std::vector<int32_t>& c = a->h_col;
    const std::vector<int32_t>& lev = a->h_level;
    const int32_t nblk = (int32_t)(((int64_t)N + B - 1) / B);
    std::vector<int32_t> blkstep((size_t)nblk + 1), srow(N), didx(N), eidx, scol, ord, first;
    std::vector<int4> steps;
    int64_t slots = 0;
    int32_t k = 0, max_steps = 0, max_width = 0;
    for (int32_t blk = 0; blk < nblk; blk++) {
        const int32_t r0 = blk * B, r1 = (int32_t)std::min<int64_t>(N, (int64_t)r0 + B);
        // Rows of the block in level order (stable), one level per step: a
        // row's dependencies inside the block then lie in earlier steps, and
        // a step waits only for rows of lower levels, so every block moves
        // through the levels in step with the others instead of waiting for
        // the whole of an earlier block.
        ord.resize((size_t)(r1 - r0));
        for (int32_t i = r0; i < r1; i++) ord[i - r0] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int32_t p, int32_t q) { return lev[p] < lev[q]; });
        first.clear();
        for (int32_t p = 0; p < r1 - r0; p++)
            if (p == 0 || lev[ord[p]] != lev[ord[p - 1]]) first.push_back(p);
        first.push_back(r1 - r0);
        const int32_t L = (int32_t)first.size() - 1;
        blkstep[blk] = (int32_t)steps.size();
        for (int32_t l = 0; l < L; l++)
            for (int32_t p = first[l]; p < first[l + 1]; p += SB) {
                const int32_t n = std::min(SB, first[l + 1] - p);
                int32_t width = 0;
                for (int32_t t = 0; t < n; t++) width = std::max(width, rp[ord[p + t] + 1] - 1 - rp[ord[p + t]]);
                SABLE_CHECK(slots + (int64_t)width * n < 0x7fffffffLL, SABLE_EINVAL,
                            "row-block schedule: more than 2^31 entry slots");
                steps.push_back(make_int4(k, n, width, (int)slots));
                max_width = std::max(max_width, width);
                eidx.resize((size_t)(slots + (int64_t)width * n));
                scol.resize(eidx.size());
                for (int32_t t = 0; t < n; t++) {
                    const int32_t i = ord[p + t], a0 = rp[i], len = rp[i + 1] - 1 - a0;
                    srow[k + t] = i;
                    didx[k + t] = rp[i + 1] - 1;
                    for (int32_t j = 0; j < width; j++) {
                        const size_t sl = (size_t)(slots + (int64_t)j * n + t);
                        eidx[sl] = j < len ? a0 + j : -1;
                        scol[sl] = j < len ? c[a0 + j] : -1;
                    }
                }
                slots += (int64_t)width * n;
                k += n;
            }
        max_steps = std::max(max_steps, (int32_t)steps.size() - blkstep[blk]);
    }
    blkstep[nblk] = (int32_t)steps.size();
    const size_t vs = a->dtype == SABLE_F64 ? 8 : 4;
    const size_t ns = std::max<size_t>(steps.size(), 1), sl = std::max<size_t>((size_t)slots, 1);
    int32_t* d_eidx = nullptr;
    int32_t* d_didx = nullptr;
    bool ok = hipMalloc(&a->d_bl_blkstep, 4 * ((size_t)nblk + 1)) == hipSuccess &&
              hipMalloc(&a->d_bl_steps, sizeof(int4) * ns) == hipSuccess &&
              hipMalloc(&a->d_bl_row, 4 * (size_t)N) == hipSuccess && hipMalloc(&a->d_bl_diag, vs * N) == hipSuccess &&
              hipMalloc(&a->d_bl_col, 4 * sl) == hipSuccess && hipMalloc(&a->d_bl_val, vs * sl) == hipSuccess &&
              hipMalloc(&d_eidx, 4 * sl) == hipSuccess && hipMalloc(&d_didx, 4 * (size_t)N) == hipSuccess;
    auto done = [&](int rc) {
        if (d_eidx) (void)hipFree(d_eidx);
        if (d_didx) (void)hipFree(d_didx);
        if (rc) free_blocked(a);
        return rc;
    };
    if (!ok) {
        (void)hipGetLastError();
        return done(fail(SABLE_ENOMEM, "row-block schedule allocation failed"));
    }
    hipStream_t st = a->ctx->stream;
    hipError_t e = hipMemcpyAsync(a->d_bl_blkstep, blkstep.data(), 4 * blkstep.size(), hipMemcpyHostToDevice, st);
    if (!e && !steps.empty())
        e = hipMemcpyAsync(a->d_bl_steps, steps.data(), sizeof(int4) * steps.size(), hipMemcpyHostToDevice, st);
    if (!e) e = hipMemcpyAsync(a->d_bl_row, srow.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st);
    if (!e) e = hipMemcpyAsync(d_didx, didx.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st);
    if (!e && slots) e = hipMemcpyAsync(a->d_bl_col, scol.data(), 4 * (size_t)slots, hipMemcpyHostToDevice, st);
    if (!e && slots) e = hipMemcpyAsync(d_eidx, eidx.data(), 4 * (size_t)slots, hipMemcpyHostToDevice, st);
    if (e) return done(fail(SABLE_EHIP, "row-block schedule upload: %s", hipGetErrorString(e)));
    if (a->dtype == SABLE_F64) {
        hipLaunchKernelGGL(k_bl_gather<double>, dim3(gsz(N)), dim3(TB), 0, st, (int64_t)N, d_didx,
                           (const double*)a->val, (double*)a->d_bl_diag);
        if (slots)
            hipLaunchKernelGGL(k_bl_gather<double>, dim3(gsz(slots)), dim3(TB), 0, st, slots, d_eidx,
                               (const double*)a->val, (double*)a->d_bl_val);
    } else {
        hipLaunchKernelGGL(k_bl_gather<float>, dim3(gsz(N)), dim3(TB), 0, st, (int64_t)N, d_didx,
                           (const float*)a->val, (float*)a->d_bl_diag);
        if (slots)
            hipLaunchKernelGGL(k_bl_gather<float>, dim3(gsz(slots)), dim3(TB), 0, st, slots, d_eidx,
                               (const float*)a->val, (float*)a->d_bl_val);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return done(fail(SABLE_EHIP, "row-block schedule build failed"));
    a->bl_rows = B;
    a->bl_nblk = nblk;
    a->bl_nsteps = (int32_t)steps.size();
    a->bl_max_steps = max_steps;
    a->bl_width = max_width;
    a->bl_slots = slots;
    return done(SABLE_OK);
}

// FP32 copies of the block schedule's values (refined solve).
static int ensure_bl32(sable_spts_s* a) {
    if (a->d_bl_val32) return SABLE_OK;
    SABLE_CHECK(a->dtype == SABLE_F64 && a->bl_rows, SABLE_EINVAL, "no FP64 row-block schedule");
    const size_t sl = std::max<size_t>((size_t)a->bl_slots, 1);
    SABLE_HIP(hipMalloc(&a->d_bl_diag32, 4 * (size_t)a->N));
    SABLE_HIP(hipMalloc(&a->d_bl_val32, 4 * sl));
    hipStream_t st = a->ctx->stream;
    hipLaunchKernelGGL(k_to_f32, dim3(gsz(a->N)), dim3(TB), 0, st, (int64_t)a->N, (const double*)a->d_bl_diag,
                       a->d_bl_diag32);
    if (a->bl_slots)
        hipLaunchKernelGGL(k_to_f32, dim3(gsz(a->bl_slots)), dim3(TB), 0, st, a->bl_slots,
                           (const double*)a->d_bl_val, a->d_bl_val32);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// x = L^-1 b over the row-block schedule (b and x natural order, distinct).
template <typename T, bool SF>
static int launch_blocked(hipStream_t st, int num_cu, sable_spts_s* a, const T* diag, const T* val, const T* b,
                          T* x) {
    hipLaunchKernelGGL(k_sentinel<T>, dim3(gsz(a->N)), dim3(TB), 0, st, a->N, x, a->d_ticket);
    const size_t shm = (size_t)a->bl_rows * sizeof(T);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(a->bl_nblk, (int64_t)num_cu * 2));
    // a dependency is waited for up to 2^20 polls (about a second): every
    // row a block waits on belongs to an earlier block held by a running
    // workgroup, so a wait that long is an error, reported by status
    constexpr long long BL_SPIN = 1ll << 20;
#define BL_LAUNCH(PP)                                                                                          \
    do {                                                                                                       \
        if (shm > 65536 - 1024)                                                                                \
            SABLE_HIP(hipFuncSetAttribute((const void*)k_blocked<T, SF, PP>,                                   \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));              \
        hipLaunchKernelGGL((k_blocked<T, SF, PP>), dim3((unsigned)grid), dim3(SB), shm, st, a->bl_nblk,         \
                           a->bl_rows, a->d_bl_blkstep, a->d_bl_steps, a->d_bl_row, diag, a->d_bl_col, val, b, x, \
                           a->d_ticket, a->d_err, BL_SPIN);                                                     \
    } while (0)
    if (a->bl_width <= 4)
        BL_LAUNCH(4);
    else if (a->bl_width <= 8)
        BL_LAUNCH(8);
    else
        BL_LAUNCH(16);
#undef BL_LAUNCH
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// The fused refined solve over the row-block schedule: b and x natural
// order (FP64), the row words in the analysis's work space.
static int launch_blocked_ref(hipStream_t st, int num_cu, sable_spts_s* a, const double* b, double* x) {
    uint64_t* W = (uint64_t*)a->d_work;
    hipLaunchKernelGGL(k_fill_words, dim3(gsz(a->N)), dim3(TB), 0, st, a->N, W, a->d_ticket);
    const size_t shm = (size_t)a->bl_rows * sizeof(uint64_t);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(a->bl_nblk, (int64_t)num_cu * 2));
    constexpr long long BL_SPIN = 1ll << 20;
#define BLR_LAUNCH(PP)                                                                                          \
    do {                                                                                                        \
        if (shm > 65536 - 1024)                                                                                 \
            SABLE_HIP(hipFuncSetAttribute((const void*)k_blocked_ref<PP>,                                       \
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));               \
        hipLaunchKernelGGL((k_blocked_ref<PP>), dim3((unsigned)grid), dim3(SB), shm, st, a->bl_nblk, a->bl_rows,  \
                           a->d_bl_blkstep, a->d_bl_steps, a->d_bl_row, (const double*)a->d_bl_diag, a->d_bl_col, \
                           (const double*)a->d_bl_val, b, W, x, a->d_ticket, a->d_err, BL_SPIN);                  \
    } while (0)
    if (a->bl_width <= 4)
        BLR_LAUNCH(4);
    else if (a->bl_width <= 8)
        BLR_LAUNCH(8);
    else
        BLR_LAUNCH(16);
#undef BLR_LAUNCH
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// The dependents of every row (the strictly lower part by column, rows
// ascending) and the in-degrees, from the host copy of the structure.
static int ensure_indeg(sable_spts_s* a) {
    if (a->d_dep_ptr) return SABLE_OK;
    const int32_t N = a->N;
    const std::vector<int32_t>& rp = a->h_rp;
    const std::vector<int32_t>& c = a->h_col;
    const int64_t m = a->nnz - N;
    std::vector<int32_t> dptr((size_t)N + 1, 0), drow((size_t)std::max<int64_t>(m, 1)), indeg(N);
    for (int32_t i = 0; i < N; i++) {
        indeg[i] = rp[i + 1] - 1 - rp[i];
        for (int32_t j = rp[i]; j < rp[i + 1] - 1; j++) dptr[c[j] + 1]++;
    }
    for (int32_t j = 0; j < N; j++) dptr[j + 1] += dptr[j];
    std::vector<int32_t> pos(dptr.begin(), dptr.end() - 1);
    for (int32_t i = 0; i < N; i++)
        for (int32_t j = rp[i]; j < rp[i + 1] - 1; j++) drow[pos[c[j]]++] = i;
    hipStream_t st = a->ctx->stream;
    bool ok = hipMalloc(&a->d_dep_ptr, 4 * ((size_t)N + 1)) == hipSuccess &&
              hipMalloc(&a->d_dep_row, 4 * drow.size()) == hipSuccess &&
              hipMalloc(&a->d_indeg, 4 * (size_t)N) == hipSuccess && hipMalloc(&a->d_cnt, 4 * (size_t)N) == hipSuccess;
    if (ok)
        ok = hipMemcpyAsync(a->d_dep_ptr, dptr.data(), 4 * dptr.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
             hipMemcpyAsync(a->d_dep_row, drow.data(), 4 * drow.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
             hipMemcpyAsync(a->d_indeg, indeg.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        int32_t** ptrs[] = {&a->d_dep_ptr, &a->d_dep_row, &a->d_indeg, &a->d_cnt};
        for (int32_t** p : ptrs) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
        return fail(SABLE_ENOMEM, "in-degree tables: allocation or upload failed");
    }
    return SABLE_OK;
}

template <typename T>
static void launch_indeg(hipStream_t st, int num_cu, sable_spts_s* a, const T* b, T* x) {
    const int32_t N = a->N;
    const int G = a->lanes;
    hipLaunchKernelGGL(k_indeg_reset, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_cnt, a->d_ticket);
    const int64_t waves_needed = ((int64_t)N * G + WAVE - 1) / WAVE;
    int64_t blocks = (waves_needed + (TB / WAVE) - 1) / (TB / WAVE);
    const int64_t cap = (int64_t)num_cu * 8;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
#define ID_LAUNCH(GG)                                                                                             \
    hipLaunchKernelGGL((k_syncfree_indeg<T, GG>), dim3((unsigned)blocks), dim3(TB), 0, st, N, a->rp, a->col,     \
                       (const T*)a->val, a->d_dep_ptr, a->d_dep_row, a->d_indeg, a->d_cnt, b, x, a->d_ticket,      \
                       a->d_err, DEFAULT_SPIN)
    switch (G) {
        case 1: ID_LAUNCH(1); break;
        case 4: ID_LAUNCH(4); break;
        case 8: ID_LAUNCH(8); break;
        case 16: ID_LAUNCH(16); break;
        default: ID_LAUNCH(64); break;
    }
#undef ID_LAUNCH
}

static void free_analysis(sable_spts_s* a) {
    if (!a) return;
    free_blocked(a);
    {
        void* ptrs[] = {a->d_dep_ptr, a->d_dep_row, a->d_indeg, a->d_cnt};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    for (int k = 0; k < 3; k++) {
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

// Captures body(stream) into graph slot k of the analysis, once per lane
// count.  The capture stream lives until the graph is instantiated, the
// capture is always ended (also when a launch inside it failed), and a
// failed capture leaves no graph behind.
template <typename F>
static int capture_graph(sable_spts_s* a, int k, F&& body) {
    if (a->exec[k] && a->graph_lanes[k] == a->lanes) return SABLE_OK;
    if (a->exec[k]) {
        // a replay of the old graph may still be queued on the stream (the
        // lanes were changed without a wait): let it finish first
        (void)hipStreamSynchronize(a->ctx->stream);
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
    const int lrc = body(cs);
    e = hipStreamEndCapture(cs, &a->graph[k]);
    if (e == hipSuccess && lrc == SABLE_OK) e = hipGraphInstantiate(&a->exec[k], a->graph[k], nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess || lrc != SABLE_OK) {
        if (a->exec[k]) (void)hipGraphExecDestroy(a->exec[k]);
        if (a->graph[k]) (void)hipGraphDestroy(a->graph[k]);
        a->exec[k] = nullptr;
        a->graph[k] = nullptr;
        return lrc != SABLE_OK ? lrc : fail(SABLE_EHIP, "level graph capture: %s", hipGetErrorString(e));
    }
    a->graph_lanes[k] = a->lanes;
    return SABLE_OK;
}

template <typename T>
static int capture_levels(sable_spts_s* a, bool use32) {
    const T* val = use32 ? (const T*)a->d_lval32 : (const T*)a->d_val;
    return capture_graph(a, use32 ? 1 : 0, [&](hipStream_t cs) {
        return levels_on_stream<T>(cs, a->lanes, a->h_level_ptr, a->d_rp, a->d_col, val, (const T*)a->d_bl,
                                   (T*)a->d_yl);
    });
}

// Level solve on level-ordered buffers bl -> yl: graph replay, or the
// per-level launches themselves.
template <typename T>
static int solve_level_ordered(sable_spts_s* a, bool use32) {
    if (a->schedule == SABLE_SPTS_LEVEL_LAUNCHES) {
        const T* val = use32 ? (const T*)a->d_lval32 : (const T*)a->d_val;
        return levels_on_stream<T>(a->ctx->stream, a->lanes, a->h_level_ptr, a->d_rp, a->d_col, val,
                                   (const T*)a->d_bl, (T*)a->d_yl);
    }
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
    a->h_rp = std::move(rp);
    a->h_col = std::move(c);
    a->h_level = std::move(level);
    rc = build_blocked(a, default_block_rows(N));
    if (rc) {
        free_analysis(a);
        return rc;
    }
    *out = a;
    return SABLE_OK;
}

SABLE_API int sable_spts_set_schedule(sable_spts_t* a, int schedule, int32_t block_rows) {
    SABLE_CHECK(a, SABLE_EINVAL, "NULL analysis");
    SABLE_CHECK(schedule == SABLE_SPTS_BLOCKS || schedule == SABLE_SPTS_LEVEL_GRAPH ||
                    schedule == SABLE_SPTS_LEVEL_LAUNCHES,
                SABLE_EINVAL, "unknown schedule %d", schedule);
    if (schedule == SABLE_SPTS_BLOCKS && block_rows > 0 && block_rows != a->bl_rows) {
        (void)hipStreamSynchronize(a->ctx->stream);
        const int rc = build_blocked(a, block_rows);
        if (rc) return rc;
    }
    a->schedule = schedule;
    return SABLE_OK;
}

SABLE_API int sable_spts_schedule_info(const sable_spts_t* a, int* schedule, int32_t* block_rows,
                                       int32_t* nblocks, int32_t* nsteps, int64_t* slots) {
    SABLE_CHECK(a, SABLE_EINVAL, "NULL analysis");
    if (schedule) *schedule = a->schedule;
    if (block_rows) *block_rows = a->bl_rows;
    if (nblocks) *nblocks = a->bl_nblk;
    if (nsteps) *nsteps = a->bl_nsteps;
    if (slots) *slots = a->bl_slots;
    return SABLE_OK;
}

// Times solve_level under each schedule (row blocks of 1024 to 16384 rows,
// the level graph at 1, 4, 8, 16 and 64 lanes per row) on b = L*1 and keeps the
// fastest. Every schedule and lane count gives the same x bit for bit, so
// the choice changes only the time.
SABLE_API int sable_spts_tune(sable_ctx_t* ctx, sable_spts_t* a, int* schedule, int32_t* block_rows, int* lanes,
                              float* best_ms) {
    SABLE_CHECK(ctx && a, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == a->ctx, SABLE_EINVAL, "analysis belongs to another context");
    hipStream_t st = ctx->stream;
    const int32_t N = a->N;
    const size_t vs = a->dtype == SABLE_F64 ? 8 : 4;
    void* b = nullptr;
    void* x = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SABLE_OK;
    if (hipMalloc(&b, vs * (size_t)N) != hipSuccess || hipMalloc(&x, vs * (size_t)N) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(SABLE_ENOMEM, "tune: allocation failed");
    }
    if (!rc) {
        if (a->dtype == SABLE_F64)
            hipLaunchKernelGGL(k_rowsum<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->rp, (const double*)a->val,
                               (double*)b);
        else
            hipLaunchKernelGGL(k_rowsum<float>, dim3(gsz(N)), dim3(TB), 0, st, N, a->rp, (const float*)a->val,
                               (float*)b);
        if (hipGetLastError() != hipSuccess) rc = fail(SABLE_EHIP, "tune: launch failed");
    }
    // one untimed solve (it also captures the graph), then the best of 5
    auto time_one = [&](float* out) -> int {
        int r = sable_spts_solve_level(ctx, a, b, x);
        float best = 1e30f;
        for (int it = 0; it < 5 && !r; it++) {
            if (hipEventRecord(e0, st) != hipSuccess) return fail(SABLE_EHIP, "tune: event record failed");
            r = sable_spts_solve_level(ctx, a, b, x);
            if (r) break;
            float ms = 0.f;
            if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                return fail(SABLE_EHIP, "tune: event timing failed");
            best = std::min(best, ms);
        }
        if (!r) r = sable_spts_status(ctx, a);
        *out = best;
        return r;
    };
    int best_s = SABLE_SPTS_BLOCKS;
    int32_t best_b = 0;
    float best = 1e30f;
    const int saved = a->schedule, saved_lanes = a->lanes;
    int graph_lanes = saved_lanes;
    for (int32_t B = 1024; B <= 16384 && !rc; B *= 2) {
        if (B > 1024 && B >= 2 * (int64_t)N) break;
        rc = sable_spts_set_schedule(a, SABLE_SPTS_BLOCKS, B);
        float ms = 0.f;
        if (!rc) rc = time_one(&ms);
        if (!rc && ms < best) {
            best = ms;
            best_b = B;
            best_s = SABLE_SPTS_BLOCKS;
        }
    }
    for (int G : {1, 4, 8, 16, 64}) {
        if (rc) break;
        a->schedule = SABLE_SPTS_LEVEL_GRAPH;
        a->lanes = G;
        float ms = 0.f;
        rc = time_one(&ms);
        if (!rc && ms < best) {
            best = ms;
            best_s = SABLE_SPTS_LEVEL_GRAPH;
            graph_lanes = G;
        }
    }
    if (!rc) {
        // the row-block schedule stays built at its best size, also when the
        // graph wins (the sync-free solve runs on it)
        rc = sable_spts_set_schedule(a, SABLE_SPTS_BLOCKS, best_b);
        if (!rc) a->schedule = best_s;
        a->lanes = best_s == SABLE_SPTS_LEVEL_GRAPH ? graph_lanes : saved_lanes;
    } else {
        a->schedule = saved;
        a->lanes = saved_lanes;
    }
    (void)hipStreamSynchronize(st);
    if (b) (void)hipFree(b);
    if (x) (void)hipFree(x);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (schedule) *schedule = a->schedule;
    if (block_rows) *block_rows = a->bl_rows;
    if (lanes) *lanes = a->lanes;
    if (best_ms) *best_ms = best;
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
    if (a->schedule == SABLE_SPTS_BLOCKS) {
        if (a->dtype == SABLE_F64)
            return launch_blocked<double, false>(st, ctx->num_cu, a, (const double*)a->d_bl_diag,
                                                 (const double*)a->d_bl_val, (const double*)b, (double*)x);
        return launch_blocked<float, false>(st, ctx->num_cu, a, (const float*)a->d_bl_diag, (const float*)a->d_bl_val,
                                            (const float*)b, (float*)x);
    }
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
    if (a->schedule == SABLE_SPTS_BLOCKS) {
        if (a->dtype == SABLE_F64)
            return launch_blocked<double, true>(ctx->stream, ctx->num_cu, a, (const double*)a->d_bl_diag,
                                                (const double*)a->d_bl_val, (const double*)b, (double*)x);
        return launch_blocked<float, true>(ctx->stream, ctx->num_cu, a, (const float*)a->d_bl_diag,
                                           (const float*)a->d_bl_val, (const float*)b, (float*)x);
    }
    if (a->dtype == SABLE_F64)
        launch_syncfree<double>(ctx->stream, a->lanes, ctx->num_cu, a->N, a->rp, a->col, (const double*)a->val,
                                (const double*)b, (double*)x, a->d_ticket, a->d_err, DEFAULT_SPIN);
    else
        launch_syncfree<float>(ctx->stream, a->lanes, ctx->num_cu, a->N, a->rp, a->col, (const float*)a->val,
                               (const float*)b, (float*)x, a->d_ticket, a->d_err, DEFAULT_SPIN);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_spts_solve_indegree(sable_ctx_t* ctx, sable_spts_t* a, const void* b, void* x) {
    SABLE_CHECK(ctx && a && b && x, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == a->ctx, SABLE_EINVAL, "analysis belongs to another context");
    SABLE_CHECK(b != x, SABLE_EINVAL, "b and x must be distinct");
    const int rc = ensure_indeg(a);
    if (rc) return rc;
    if (a->dtype == SABLE_F64)
        launch_indeg<double>(ctx->stream, ctx->num_cu, a, (const double*)b, (double*)x);
    else
        launch_indeg<float>(ctx->stream, ctx->num_cu, a, (const float*)b, (float*)x);
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
    SABLE_CHECK(iters >= 0 && (inner == 0 || inner == 1 || inner == 2), SABLE_EINVAL, "bad iters/inner");
    SABLE_CHECK(inner != 2 || iters == 1, SABLE_EINVAL, "the fused refined solve takes one refinement step");
    int rc = ensure_f32(a);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const int32_t N = a->N;
    if (inner == 2) {
        // FP32 solve, FP64 residual, FP32 correction and FP64 update in one
        // sweep of the analysis's schedule
        if (a->schedule == SABLE_SPTS_BLOCKS) return launch_blocked_ref(st, ctx->num_cu, a, b, x);
        float2* XD = (float2*)a->d_work;  // N (x0, d) pairs in the 2N floats of work space
        hipLaunchKernelGGL(k_gather_perm<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm, b, (double*)a->d_bl);
        auto body = [&](hipStream_t s2) {
            return levels_ref_on_stream(s2, a->lanes, a->h_level_ptr, a->d_rp, a->d_col, (const double*)a->d_val,
                                        (const double*)a->d_bl, XD, (double*)a->d_yl);
        };
        if (a->schedule == SABLE_SPTS_LEVEL_LAUNCHES) {
            rc = body(st);
        } else {
            rc = capture_graph(a, 2, body);
            if (!rc) SABLE_HIP(hipGraphLaunch(a->exec[2], st));
        }
        if (rc) return rc;
        hipLaunchKernelGGL(k_scatter_perm<double>, dim3(gsz(N)), dim3(TB), 0, st, N, a->d_perm,
                           (const double*)a->d_yl, x);
        SABLE_LAUNCH_CHECK();
        return SABLE_OK;
    }
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
        if (a->schedule == SABLE_SPTS_BLOCKS) {
            rc = ensure_bl32(a);
            if (rc) return rc;
            rc = inner == 1 ? launch_blocked<float, true>(st, ctx->num_cu, a, a->d_bl_diag32, a->d_bl_val32, r32, d32)
                            : launch_blocked<float, false>(st, ctx->num_cu, a, a->d_bl_diag32, a->d_bl_val32, r32,
                                                           d32);
            if (rc) return rc;
        } else if (inner == 1) {
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
