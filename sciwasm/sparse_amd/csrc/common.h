// common.h -- shared internals of libsablesparse_hip (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>
#include <mutex>

#include "../../../include/sable_sparse.h"

namespace sable {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define SABLE_HIP(call)                                                          \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess)                                                    \
            return ::sable::fail(SABLE_EHIP, "%s failed: %s (%s:%d)", #call,    \
                                 hipGetErrorString(e_), __FILE__, __LINE__);     \
    } while (0)

#define SABLE_CHECK(cond, code, ...)                                             \
    do {                                                                         \
        if (!(cond)) return ::sable::fail((code), __VA_ARGS__);                  \
    } while (0)

#define SABLE_LAUNCH_CHECK() SABLE_HIP(hipGetLastError())

}  // namespace sable

// ---------------------------------------------------------------- context
struct sable_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cu = 256;
    // grow-only device scratch for transient plans and flags
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    // small pinned host buffer for reading device flags / counters
    void* pinned = nullptr;
    // stream-ordered pool for per-call temporaries (created on first use;
    // keeps its memory until the context is destroyed)
    hipMemPool_t pool = nullptr;
    bool pool_tried = false;
    // pinned staging halves for copies from/to pageable host memory
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_used[2] = {false, false};
    int stage_next = 0;
    // device address of the allocation-free work space on this device
    void* noalloc = nullptr;
    std::mutex mu;
};

namespace sable {
// Stream-explicit launchers shared by the batch server (spmv.hip, spts.hip).
int csr_launch(hipStream_t st, sable_dtype_t dt, const int32_t* blk, int64_t nb, const int32_t* rp,
               const int32_t* col, const void* val, const void* x, void* y, double beta, int lds_cap);
int csr_plan_view(const sable_csr_plan_t* p, const int32_t** blk, int64_t* nb, int* lds_cap);
int64_t plan_nblocks(int64_t nnz, int32_t N, int T, int C);
int plan_lds_cap(int T);
int plan_fill_stream(hipStream_t st, int32_t N, const int32_t* rp, int T, int C, int64_t B, int32_t* blk, int* err);
// y = A.x + beta*y with a column-sorted tile plan, on stream st.
int tiles_launch(hipStream_t st, const sable_csr_tiles_t* p, const void* x, void* y, double beta);
int batch_tiles_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int nrhs, const sable_batch_tiles_item_t* d_items,
                       int nitems, int64_t total_units, int32_t max_tile_rows, unsigned long long* counter, int* err,
                       int workgroups);
int syncfree_launch(hipStream_t st, int num_cu, sable_dtype_t dt, int lanes, int32_t N, const int32_t* rp,
                    const int32_t* col, const void* val, const void* b, void* x, unsigned* ticket, int* err,
                    long long spin);
int ctx_scratch(sable_ctx_t* ctx, size_t bytes, void** out);
// Address on `device` (the current device) of the allocation-free CSR
// product's work space, looked up once per device.
int spmv_noalloc_setup(int device, void** out);
// Temporary device memory ordered on ctx->stream: usable by work enqueued on
// that stream after the call, released by ctx_pool_free once the work that
// uses it has been enqueued.  Without a memory pool it is plain hipMalloc.
int ctx_pool_alloc(sable_ctx_t* ctx, size_t bytes, void** out);
void ctx_pool_free(sable_ctx_t* ctx, void* p);
int read_device_i64(sable_ctx_t* ctx, const void* dsrc, size_t count, int64_t* host_out);

constexpr int WAVE = 64;

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 0x7fffffff) g = 0x7fffffff;
    return (unsigned)g;
}

// Rounded multiply and add without contraction: the reference's WebAssembly
// arithmetic is a separately rounded mul followed by add.  HIP's __fmul_rn
// family are plain operators that the default -ffp-contract=fast would fuse
// into an FMA, so contraction is switched off inside each helper.
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double sub_rn(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ float div_rn(float a, float b) { return __fdiv_rn(a, b); }
__device__ __forceinline__ double div_rn(double a, double b) { return a / b; }

// Sum over aligned groups of L lanes (L a power of two <= 64).
template <typename T>
__device__ __forceinline__ T group_sum(T v, int L) {
    for (int off = L >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// s += p (SUB: s -= p) for the values p of lanes 0 .. n-1 of an aligned
// group of G lanes, in lane order: the serial sum's bits.  The values are
// read across the group eight at a time before they are added, so the reads
// overlap; every lane of the group gets the same s.
template <bool SUB, int G, typename T>
__device__ __forceinline__ T group_ordered_sum(T s, T p, int n) {
    constexpr int W = G < 8 ? G : 8;
#pragma unroll
    for (int q0 = 0; q0 < G; q0 += W) {
        T v[W];
#pragma unroll
        for (int q = 0; q < W; q++) v[q] = __shfl(p, q0 + q, G);
#pragma unroll
        for (int q = 0; q < W; q++)
            if (q0 + q < n) s = SUB ? sub_rn(s, v[q]) : add_rn(s, v[q]);
    }
    return s;
}

// The same ordered sum for groups of 16 or more lanes through LDS: each
// lane stages its value in stage[threadIdx.x] (one T per thread of the
// workgroup), then every lane of the group reads the group's n values in
// lane order (the group's lanes read one address at a time, a broadcast)
// and adds them itself: the serial sum's bits, with one LDS read per value
// instead of one cross-lane read per value and 32 bits.  Smaller groups
// take the cross-lane reads, which measured faster there.
template <bool SUB, int G, typename T>
__device__ __forceinline__ T group_ordered_sum_lds(T s, T p, int n, T* stage) {
    if constexpr (G <= 8) return group_ordered_sum<SUB, G>(s, p, n);
    T* grp = stage + (threadIdx.x - (threadIdx.x & (G - 1)));
    stage[threadIdx.x] = p;
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < n; q++) {
        const T v = grp[q];
        s = SUB ? sub_rn(s, v) : add_rn(s, v);
    }
    __builtin_amdgcn_wave_barrier();
    return s;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

}  // namespace sable
