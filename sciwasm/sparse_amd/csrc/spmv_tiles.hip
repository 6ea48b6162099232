// spmv_tiles.hip -- CSR SpMV over column-sorted tiles, for matrices whose
// columns are scattered over an x much larger than one XCD's L2.
//
// With random columns every x[col] gather is its own L2 request: the row-
// block kernel spends its time in the L2 request rate, not in HBM bytes.
// This schedule regroups the matrix so that consecutive lanes of a wave read
// neighbouring x values:
//  * the rows are cut into tiles of `tile_rows` rows and the columns into
//    panels of `panel_cols` columns (x panel = panel_cols values, sized for
//    one XCD's L2);
//  * within each (panel, row tile) the entries are sorted by column (stable:
//    equal columns keep ascending rows) and stored as one 32-bit word,
//    column-in-panel << row_bits | row-in-tile, plus the value;
//  * one workgroup per row tile streams its entries panel by panel (two
//    panels per step in the one-launch schedule), gathers x (consecutive
//    lanes hit the same or adjacent x lines, so the gathers of a wave
//    coalesce into fewer L2 requests), and accumulates the products into
//    the tile's rows held in LDS with LDS atomic adds; y is written once at
//    the end (or, in the per-panel schedule, one launch per panel adds the
//    panel's sums to y);
//  * in the one-launch schedule the workgroups of each XCD also load the x
//    lines of the next step's panels into their L2 ahead of the gathers.
// Each row's products are summed in an order that depends on the LDS atomic
// order, so the result is not bit-reproducible from run to run; it meets the
// same per-row bound as every other CSR schedule.
#include "common.h"
#include "prims.h"

#include <algorithm>
#include <cstring>

struct sable_csr_tiles_s {
    sable_ctx_t* ctx = nullptr;
    sable_dtype_t dtype = SABLE_F64;
    int32_t N = 0;
    int64_t nnz = 0;
    int32_t tile_rows = 0;
    int32_t panel_cols = 0;
    int row_bits = 0;
    int col_bits = 0;
    int32_t ntr = 0;        // row tiles
    int32_t npanels = 0;
    int32_t* tile_ptr = nullptr;  // npanels * ntr + 1 offsets, panel-major
    uint32_t* packed = nullptr;   // column-in-panel << row_bits | row-in-tile
    void* val = nullptr;
    size_t bytes = 0;
};

namespace sable {

constexpr int TILE_BS = 1024;
constexpr int TILE_U = 4;

static int log2_exact(int64_t v) {
    int b = 0;
    while (((int64_t)1 << b) < v) b++;
    return ((int64_t)1 << b) == v ? b : -1;
}

__global__ void k_tile_rows(int32_t N, const int32_t* __restrict__ rp, int32_t* __restrict__ rows) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
        for (int64_t j = rp[i]; j < rp[i + 1]; j++) rows[j] = (int32_t)i;
}

// key = (panel * ntr + row tile) << col_bits | column-in-panel: a panel's
// entries are contiguous, row tile after row tile, each tile by column, so
// that the workgroups working on one panel at the same time read one
// compact region of memory
// Sort key of entry k: (panel * ntr + row tile) << col_bits | column-in-
// panel in the low sort_bits bits (the sorted part), the row-in-tile from
// bit rin_shift = sort_bits on (carried through the sort untouched); the value's bits are the
// sort's value (P: an unsigned integer of the value's size), so the packing
// pass reads everything in sorted order and gathers nothing.
template <typename T, typename P>
__global__ void k_tile_keys(int64_t nnz, int32_t ncols, const int32_t* __restrict__ rows,
                            const int32_t* __restrict__ col, const T* __restrict__ val, int32_t tile_rows,
                            int col_bits, int rin_shift, int32_t ntr, uint64_t* __restrict__ keys,
                            P* __restrict__ pay, int* __restrict__ err) {
    static_assert(sizeof(T) == sizeof(P), "value bits of the value's size");
    const uint32_t cmask = (1u << col_bits) - 1u;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = col[k];
        if (c < 0 || c >= ncols) {
            atomicOr(err, 1);
            keys[k] = 0;
        } else {
            const int32_t r = rows[k];
            const int64_t t = (int64_t)(c >> col_bits) * ntr + (r / tile_rows);
            keys[k] = ((uint64_t)(uint32_t)(r % tile_rows) << rin_shift) | ((uint64_t)t << col_bits) |
                      ((uint32_t)c & cmask);
        }
        pay[k] = __builtin_bit_cast(P, val[k]);
    }
}
template <typename T, typename P>
__global__ void k_tile_pack(int64_t nnz, const uint64_t* __restrict__ keys, const P* __restrict__ pay,
                            int row_bits, int col_bits, int rin_shift, uint32_t* __restrict__ packed,
                            T* __restrict__ vout) {
    const uint32_t cmask = (1u << col_bits) - 1u;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nnz; s += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = keys[s];
        packed[s] = (((uint32_t)key & cmask) << row_bits) | (uint32_t)(key >> rin_shift);
        vout[s] = __builtin_bit_cast(T, pay[s]);
    }
}
__global__ void k_tile_heads(int64_t nnz, int64_t ntiles, int col_bits, int sort_bits,
                             const uint64_t* __restrict__ keys, int32_t* __restrict__ tile_ptr) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nnz) return;
    const uint64_t m = ((uint64_t)1 << sort_bits) - 1;
    const int64_t prev = s > 0 ? (int64_t)((keys[s - 1] & m) >> col_bits) : -1;
    const int64_t cur = s < nnz ? (int64_t)((keys[s] & m) >> col_bits) : ntiles;
    for (int64_t t = prev + 1; t <= cur; t++) tile_ptr[t] = (int32_t)s;
}
template <typename V>
__device__ __forceinline__ V ld_nt(const V* p) {
    return __builtin_nontemporal_load(p);
}

// Row sums are held in LDS in FP64 for either value type: each product is
// formed in the value type and added to its row with an FP64 LDS atomic,
// and the sum is rounded to the value type once.  On gfx950 the FP32 LDS
// atomic add has measured about 2.5x slower than the FP64 one in this loop,
// and the FP64 sum only tightens the FP32 result.

// One panel: workgroup = one row tile.  The tile's row sums live in LDS.
template <typename T>
__global__ __launch_bounds__(TILE_BS) void k_tiles(int32_t N, int32_t tile_rows, int row_bits, int32_t ntr,
                                                   int32_t panel, int32_t panel_cols,
                                                   const int32_t* __restrict__ tile_ptr,
                                                   const uint32_t* __restrict__ packed, const T* __restrict__ val,
                                                   const T* __restrict__ x, T* __restrict__ y, T beta, int first) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* acc = reinterpret_cast<double*>(smem_raw);
    const int tr = blockIdx.x;
    for (int r = threadIdx.x; r < tile_rows; r += TILE_BS) acc[r] = 0.0;
    __syncthreads();
    const int64_t t = (int64_t)panel * ntr + tr;
    const int64_t k0 = tile_ptr[t], k1 = tile_ptr[t + 1];
    const T* __restrict__ xp = x + (int64_t)panel * panel_cols;
    const uint32_t rmask = (1u << row_bits) - 1u;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += (int64_t)TILE_U * TILE_BS) {
        uint32_t pk[TILE_U];
        T v[TILE_U];
#pragma unroll
        for (int u = 0; u < TILE_U; u++) {
            const int64_t kk = k + (int64_t)u * TILE_BS;
            pk[u] = kk < k1 ? ld_nt(packed + kk) : 0u;
            v[u] = kk < k1 ? ld_nt(val + kk) : T(0);
        }
        T xv[TILE_U];
#pragma unroll
        for (int u = 0; u < TILE_U; u++) xv[u] = (k + (int64_t)u * TILE_BS < k1) ? xp[pk[u] >> row_bits] : T(0);
#pragma unroll
        for (int u = 0; u < TILE_U; u++)
            if (k + (int64_t)u * TILE_BS < k1) atomicAdd(&acc[pk[u] & rmask], (double)(v[u] * xv[u]));
    }
    __syncthreads();
    const int64_t r0 = (int64_t)tr * tile_rows;
    for (int r = threadIdx.x; r < tile_rows; r += TILE_BS) {
        const int64_t i = r0 + r;
        if (i >= N) break;
        const T a = (T)acc[r];
        if (first) {
            if (beta == T(0)) y[i] = a;
            else if (beta == T(1)) { if (a != T(0)) y[i] = y[i] + a; }
            else y[i] = beta * y[i] + a;
        } else if (a != T(0)) {
            y[i] = y[i] + a;
        }
    }
}

// All panels in one launch: the workgroup of a row tile walks the panels in
// order with its row sums kept in LDS throughout, and writes y once.
//  * A step takes TILE_G panels: each thread loads up to TILE_GU entries of
//    each (the loads of both panels in flight together), gathers their x,
//    and adds the products; entries past TILE_GU * TILE_BS in a panel are
//    taken TILE_GU at a time after that.  There is no barrier between
//    steps: each wave walks on its own.
//  * The workgroups start together and go through the panels at about the
//    same pace, since every step costs one load -> gather -> add chain, so
//    the x panels in use at any time stay few and L2-resident.
//  * x prefetch: workgroup w runs on XCD w % 8, so the w / 8-th workgroup
//    of each XCD, with its 16 waves, is one of that XCD's (ntr + 7) / 8
//    slots.  Each wave loads one 128-byte x line of each panel of the next
//    step (line = slot * 16 + wave), so that the lines of those panels are
//    in the XCD's L2 before its workgroups gather from them; otherwise the
//    first gather of a line misses and the gathers of the other workgroups
//    on that XCD wait behind its fill.  The loaded values are kept in
//    registers only until the step ends (an empty asm statement consumes
//    them); a panel wider than the XCD's waves can cover is prefetched in
//    part.  The mapping only decides which L2 a line warms: results do not
//    depend on it.
constexpr int TILE_G = 2;
constexpr int TILE_GU = 2;

template <typename T>
__global__ __launch_bounds__(TILE_BS) void k_tiles_fused(int32_t N, int32_t tile_rows, int row_bits, int32_t ntr,
                                                         int32_t npanels, int32_t panel_cols,
                                                         const int32_t* __restrict__ tile_ptr,
                                                         const uint32_t* __restrict__ packed,
                                                         const T* __restrict__ val, const T* __restrict__ x,
                                                         T* __restrict__ y, T beta) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* acc = reinterpret_cast<double*>(smem_raw);
    const int tr = blockIdx.x;
    constexpr int XL = 128 / (int)sizeof(T);  // x values per 128-byte line
    const int pf_line = (int)(blockIdx.x >> 3) * (TILE_BS / 64) + (int)(threadIdx.x >> 6);
    const bool pf_on = (int64_t)pf_line * XL < panel_cols;
    T pfs = T(0);
    for (int r = threadIdx.x; r < tile_rows; r += TILE_BS) acc[r] = 0.0;
    __syncthreads();
    const uint32_t rmask = (1u << row_bits) - 1u;
    // entry positions fit 32 unsigned bits: nnz < 2^31 (int32 row_ptr), and
    // a position runs at most TILE_GU * TILE_BS past its range's end
    for (int32_t p = 0; p < npanels; p += TILE_G) {
        uint32_t k0[TILE_G], k1[TILE_G];
#pragma unroll
        for (int g = 0; g < TILE_G; g++) {
            if (p + g < npanels) {
                const int64_t t = (int64_t)(p + g) * ntr + tr;
                k0[g] = (uint32_t)tile_ptr[t];
                k1[g] = (uint32_t)tile_ptr[t + 1];
            } else {
                k0[g] = k1[g] = 0;
            }
        }
        uint32_t pk[TILE_G][TILE_GU];
        T v[TILE_G][TILE_GU];
#pragma unroll
        for (int g = 0; g < TILE_G; g++)
#pragma unroll
            for (int u = 0; u < TILE_GU; u++) {
                const uint32_t kk = k0[g] + threadIdx.x + (uint32_t)(u * TILE_BS);
                pk[g][u] = kk < k1[g] ? ld_nt(packed + kk) : 0xffffffffu;
                v[g][u] = kk < k1[g] ? ld_nt(val + kk) : T(0);
            }
        // 0xffffffff marks an empty slot: a packed word never has it, since
        // the column-in-panel is below 2^col_bits and col_bits + row_bits < 32
        T pf = T(0);
        if (pf_on) {
#pragma unroll
            for (int g = 0; g < TILE_G; g++) {
                const int64_t i = (int64_t)(p + TILE_G + g) * panel_cols + (int64_t)pf_line * XL;
                if (p + TILE_G + g < npanels && i < N) pf += x[i];
            }
        }
        T xv[TILE_G][TILE_GU];
#pragma unroll
        for (int g = 0; g < TILE_G; g++)
#pragma unroll
            for (int u = 0; u < TILE_GU; u++)
                xv[g][u] = pk[g][u] != 0xffffffffu ? x[(int64_t)(p + g) * panel_cols + (pk[g][u] >> row_bits)] : T(0);
#pragma unroll
        for (int g = 0; g < TILE_G; g++)
#pragma unroll
            for (int u = 0; u < TILE_GU; u++)
                if (pk[g][u] != 0xffffffffu) atomicAdd(&acc[pk[g][u] & rmask], (double)(v[g][u] * xv[g][u]));
        // the rest of a panel holding more than TILE_GU * TILE_BS entries
#pragma unroll
        for (int g = 0; g < TILE_G; g++) {
            const T* __restrict__ xp = x + (int64_t)(p + g) * panel_cols;
            for (uint32_t k = k0[g] + threadIdx.x + (uint32_t)(TILE_GU * TILE_BS); k < k1[g]; k += TILE_BS) {
                const uint32_t q = packed[k];
                atomicAdd(&acc[q & rmask], (double)(val[k] * xp[q >> row_bits]));
            }
        }
        pfs += pf;
    }
    // the prefetched values are consumed once, after the walk, so that the
    // loads are kept without tying the walk's instruction order to them
    asm volatile("" ::"v"(pfs));
    __syncthreads();
    const int64_t r0 = (int64_t)tr * tile_rows;
    for (int r = threadIdx.x; r < tile_rows; r += TILE_BS) {
        const int64_t i = r0 + r;
        if (i >= N) break;
        const T a = (T)acc[r];
        if (beta == T(0)) y[i] = a;
        else if (beta == T(1)) { if (a != T(0)) y[i] = y[i] + a; }
        else y[i] = beta * y[i] + a;
    }
}

// Many tile plans in one persistent launch: a workgroup claims one unit
// (item, row tile) at a time, finds its item by binary search over the
// items' unit_begin, and runs the row tile as k_tiles_fused does.  The
// claiming thread searches while the others wait at the barrier; a unit is
// a whole row tile, so the search is paid once per tile_rows rows.
// NR right-hand sides: x and y are row-major [N][NR], so one gather brings
// an entry's NR x values; the LDS row sums are RHS-major (sum r of row q at
// r * tile_rows + q), so that the 64 lanes of one atomic instruction, on
// random rows of one right-hand side, spread over the LDS banks.
// The item's arrays are read through global-address-space pointers:
// pointers loaded from memory are otherwise generic, and generic loads
// share the LDS counter, so every wait for a gather would also wait for the
// LDS atomics in flight.
constexpr int BT_BS = 256;
constexpr int BT_XCDS = 8;  // claim ranges (MI355X has 8 XCDs); counters used: BT_XCDS
#define SABLE_GLOBAL __attribute__((address_space(1)))

// The NR values of row c of a row-major [N][NR] x: 16-byte loads when a row
// is a multiple of 16 bytes (x is then 16-byte aligned), else one load.
typedef unsigned int tile_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int tile_u32x2 __attribute__((ext_vector_type(2)));

template <typename T, int NR>
__device__ __forceinline__ void tile_xrow(const SABLE_GLOBAL T* x, uint32_t c, T (&out)[NR]) {
    constexpr int BYTES = NR * (int)sizeof(T);
    const SABLE_GLOBAL char* p = (const SABLE_GLOBAL char*)(x + (int64_t)c * NR);
    if constexpr (BYTES % 16 == 0) {
#pragma unroll
        for (int q = 0; q < BYTES / 16; q++) {
            const tile_u32x4 w = *(const SABLE_GLOBAL tile_u32x4*)(p + 16 * q);
            __builtin_memcpy((char*)out + 16 * q, &w, 16);
        }
    } else if constexpr (BYTES == 8) {
        const tile_u32x2 w = *(const SABLE_GLOBAL tile_u32x2*)p;
        __builtin_memcpy(out, &w, 8);
    } else {
        out[0] = *(const SABLE_GLOBAL T*)p;
    }
}

template <typename T, int NR>
__global__ __launch_bounds__(BT_BS) void k_batch_tiles(const sable_batch_tiles_item_t* __restrict__ items,
                                                       int nitems, int64_t total, int32_t max_rows,
                                                       unsigned long long* __restrict__ counter,
                                                       int* __restrict__ err) {
    constexpr int U = NR * (int)sizeof(T) >= 32 ? 2 : TILE_U;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* acc = reinterpret_cast<double*>(smem_raw);
    __shared__ long long s_g;
    __shared__ int s_k;
    // claim ranges: the units are cut into BT_XCDS contiguous ranges with a
    // counter each; a workgroup claims from range blockIdx % BT_XCDS (the XCD
    // it is dispatched to) and, once that range is used up, from the others
    // in turn.  Consecutive units mostly belong to one matrix, so its x is
    // fetched into few XCDs' L2s.  Every unit is handed out exactly once
    // whatever the mapping; the mapping only decides whose L2 it warms.
    int exhausted = 0;  // ranges this workgroup found used up (thread 0)
    const int q0 = (int)(blockIdx.x % BT_XCDS);
    while (true) {
        if (threadIdx.x == 0) {
            long long g = total;
            for (int i = 0; i < BT_XCDS; i++) {
                const int q = (q0 + i) % BT_XCDS;
                if (exhausted & (1 << q)) continue;
                const long long e = total * (q + 1) / BT_XCDS;
                const long long c = total * q / BT_XCDS + (long long)atomicAdd(&counter[q], 1ull);
                if (c < e) {
                    g = c;
                    break;
                }
                exhausted |= 1 << q;
            }
            int lo = 0;
            if (g < total) {
                int hi = nitems - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (items[mid].unit_begin <= g) lo = mid;
                    else hi = mid - 1;
                }
            }
            s_g = g;
            s_k = lo;
        }
        __syncthreads();
        const long long g = s_g;
        if (g >= total) break;
        const sable_batch_tiles_item_t it = items[s_k];
        if (it.tile_rows > max_rows) {
            if (threadIdx.x == 0) atomicOr(err, 1);
            __syncthreads();  // s_g and s_k are read before the next claim rewrites them
            continue;
        }
        const int32_t tr = (int32_t)(g - it.unit_begin);
        const int64_t r0 = (int64_t)tr * it.tile_rows;
        const int32_t nrows = (int32_t)min((int64_t)it.tile_rows, (int64_t)it.N - r0);
        const int32_t stride = it.tile_rows;
        for (int q = threadIdx.x; q < stride * NR; q += BT_BS) acc[q] = 0.0;
        __syncthreads();
        const uint32_t rmask = (1u << it.row_bits) - 1u;
        const int row_bits = it.row_bits;
        const SABLE_GLOBAL int32_t* tile_ptr = (const SABLE_GLOBAL int32_t*)it.tile_ptr;
        const SABLE_GLOBAL uint32_t* packed = (const SABLE_GLOBAL uint32_t*)it.packed;
        const SABLE_GLOBAL T* val = (const SABLE_GLOBAL T*)it.val;
        for (int32_t panel = 0; panel < it.npanels; panel++) {
            const int64_t t = (int64_t)panel * it.ntr + tr;
            const int64_t k0 = tile_ptr[t], k1 = tile_ptr[t + 1];
            const SABLE_GLOBAL T* xp = (const SABLE_GLOBAL T*)it.x + (int64_t)panel * it.panel_cols * NR;
            for (int64_t k = k0 + threadIdx.x; k < k1; k += (int64_t)U * BT_BS) {
                uint32_t pk[U];
                T v[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int64_t kk = k + (int64_t)u * BT_BS;
                    pk[u] = kk < k1 ? __builtin_nontemporal_load(packed + kk) : 0u;
                    v[u] = kk < k1 ? __builtin_nontemporal_load(val + kk) : T(0);
                }
                T xv[U][NR];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (k + (int64_t)u * BT_BS < k1) tile_xrow<T, NR>(xp, pk[u] >> row_bits, xv[u]);
                    else {
#pragma unroll
                        for (int r = 0; r < NR; r++) xv[u][r] = T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (k + (int64_t)u * BT_BS < k1) {
                        const uint32_t row = pk[u] & rmask;
#pragma unroll
                        for (int r = 0; r < NR; r++) atomicAdd(&acc[r * stride + row], (double)(v[u] * xv[u][r]));
                    }
            }
        }
        __syncthreads();
        SABLE_GLOBAL T* y = (SABLE_GLOBAL T*)it.y + r0 * NR;
        const T beta = (T)it.beta;
        for (int q = threadIdx.x; q < nrows * NR; q += BT_BS) {
            const int rr = q / NR, r = q - rr * NR;
            const T a = (T)acc[r * stride + rr];
            if (beta == T(0)) y[q] = a;
            else if (beta == T(1)) { if (a != T(0)) y[q] = y[q] + a; }
            else y[q] = beta * y[q] + a;
        }
        __syncthreads();  // acc, s_g and s_k are rewritten for the next unit
    }
}
#undef SABLE_GLOBAL

__global__ void k_batch_tiles_reset(unsigned long long* counter, int* err) {
    for (int q = 0; q < BT_XCDS; q++) counter[q] = 0ull;
    *err = 0;
}

// Launch only (no wait, no check of the items): the BT_XCDS counters at
// `counter` and err are reset on the stream first.  max_tile_rows * nrhs * 8
// must fit the LDS.
int batch_tiles_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int nrhs, const sable_batch_tiles_item_t* d_items,
                       int nitems, int64_t total_units, int32_t max_tile_rows, unsigned long long* counter, int* err,
                       int workgroups) {
    if (nitems <= 0 || total_units <= 0) return SABLE_OK;
    const size_t shm = (size_t)max_tile_rows * sizeof(double) * (size_t)nrhs;
    int64_t grid = workgroups > 0 ? workgroups : (int64_t)num_cu * 4;
    if (grid > total_units) grid = total_units;
    hipLaunchKernelGGL(k_batch_tiles_reset, dim3(1), dim3(1), 0, st, counter, err);
    hipError_t ea = hipSuccess;
#define BT_LAUNCH(T, NR)                                                                                       \
    do {                                                                                                        \
        if (shm > 65536 - 256) /* with the kernel's static LDS, above the default limit */                     \
            ea = hipFuncSetAttribute((const void*)k_batch_tiles<T, NR>,                                         \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);                     \
        if (ea == hipSuccess)                                                                                   \
            hipLaunchKernelGGL((k_batch_tiles<T, NR>), dim3((unsigned)grid), dim3(BT_BS), shm, st, d_items,     \
                               nitems, total_units, max_tile_rows, counter, err);                               \
    } while (0)
#define BT_SWITCH(T)                     \
    switch (nrhs) {                      \
        case 1: BT_LAUNCH(T, 1); break;  \
        case 2: BT_LAUNCH(T, 2); break;  \
        case 4: BT_LAUNCH(T, 4); break;  \
        default: BT_LAUNCH(T, 8); break; \
    }
    if (dtype == SABLE_F64) {
        BT_SWITCH(double)
    } else {
        BT_SWITCH(float)
    }
#undef BT_SWITCH
#undef BT_LAUNCH
    if (ea != hipSuccess) {
        (void)hipGetLastError();
        return fail(SABLE_EINVAL, "tile of %zu LDS bytes not supported: %s", shm, hipGetErrorString(ea));
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// 0: one launch per panel; 1: one launch over all panels (process-wide).
int g_tiles_mode = 1;

static unsigned gs(int64_t n) {
    unsigned g = grid_for(n, 256);
    return g > 65536 ? 65536 : g;
}

template <typename T>
static int tiles_run_t(hipStream_t st, const sable_csr_tiles_s* p, const void* x, void* y, double beta) {
    const size_t shm = (size_t)p->tile_rows * sizeof(double);
    if (g_tiles_mode == 1) {
        auto kern = k_tiles_fused<T>;
        hipLaunchKernelGGL(kern, dim3((unsigned)p->ntr), dim3(TILE_BS), shm, st, p->N,
                           p->tile_rows, p->row_bits, p->ntr, p->npanels, p->panel_cols, p->tile_ptr, p->packed,
                           (const T*)p->val, (const T*)x, (T*)y, (T)beta);
        SABLE_LAUNCH_CHECK();
        return SABLE_OK;
    }
    for (int32_t q = 0; q < p->npanels; q++) {
        hipLaunchKernelGGL(k_tiles<T>, dim3((unsigned)p->ntr), dim3(TILE_BS), shm, st, p->N, p->tile_rows,
                           p->row_bits, p->ntr, q, p->panel_cols, p->tile_ptr, p->packed, (const T*)p->val,
                           (const T*)x, (T*)y, (T)beta, q == 0 ? 1 : 0);
    }
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

int tiles_launch(hipStream_t st, const sable_csr_tiles_t* p, const void* x, void* y, double beta) {
    return p->dtype == SABLE_F64 ? tiles_run_t<double>(st, p, x, y, beta) : tiles_run_t<float>(st, p, x, y, beta);
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_csr_tiles_create(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, const int32_t* row_ptr,
                                     const int32_t* col, const void* val, int32_t tile_rows, int32_t panel_cols,
                                     sable_csr_tiles_t** out) {
    return sable_csr_tiles_create_rect(ctx, dtype, N, N, row_ptr, col, val, tile_rows, panel_cols, out);
}

SABLE_API int sable_csr_tiles_create_rect(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t ncols,
                                          const int32_t* row_ptr, const int32_t* col, const void* val,
                                          int32_t tile_rows, int32_t panel_cols, sable_csr_tiles_t** out) {
    SABLE_CHECK(ctx && out && row_ptr && col && val, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N > 0 && ncols > 0, SABLE_EINVAL, "N and ncols must be positive");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    *out = nullptr;
    const size_t vs = dtype == SABLE_F64 ? 8 : 4;
    if (tile_rows <= 0) tile_rows = 16384;
    if (panel_cols <= 0) panel_cols = 8192;
    // tile_rows: a multiple of 64 (row-in-tile takes bits_for(tile_rows - 1)
    // bits of the packed word); panel_cols: a power of two
    SABLE_CHECK(tile_rows >= 64 && tile_rows % 64 == 0, SABLE_EINVAL, "tile_rows must be a multiple of 64");
    const int rb = bits_for(tile_rows - 1), cb = log2_exact(panel_cols);
    SABLE_CHECK(cb >= 6, SABLE_EINVAL, "panel_cols must be a power of two >= 64");
    SABLE_CHECK(rb + cb <= 32, SABLE_EINVAL, "tile_rows * panel_cols must not exceed 2^32");
    SABLE_CHECK((size_t)tile_rows * sizeof(double) <= 160 * 1024, SABLE_EINVAL,
                "a tile's FP64 row sums must fit the LDS (160 KiB): tile_rows <= 20480");
    hipStream_t st = ctx->stream;
    int32_t ends[2];
    SABLE_HIP(hipMemcpyAsync(ends, row_ptr, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipMemcpyAsync(ends + 1, row_ptr + N, 4, hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    SABLE_CHECK(ends[0] == 0 && ends[1] >= 0, SABLE_EMALFORMED, "row_ptr must start at 0 and end at nnz >= 0");
    const int64_t nnz = ends[1];
    const int32_t ntr = (int32_t)(((int64_t)N + tile_rows - 1) / tile_rows);
    const int32_t npanels = (int32_t)(((int64_t)ncols + panel_cols - 1) / panel_cols);
    const int64_t ntiles = (int64_t)ntr * npanels;
    SABLE_CHECK(ntiles < (int64_t)1 << 30, SABLE_EINVAL, "too many tiles");
    const int sort_bits = bits_for(ntiles) + cb;
    const int rin_shift = sort_bits;
    SABLE_CHECK(rin_shift + rb <= 64, SABLE_EINVAL, "tile plan keys exceed 64 bits");
    auto* p = new sable_csr_tiles_s();
    p->ctx = ctx;
    p->dtype = dtype;
    p->N = N;
    p->nnz = nnz;
    p->tile_rows = tile_rows;
    p->panel_cols = panel_cols;
    p->row_bits = rb;
    p->col_bits = cb;
    p->ntr = ntr;
    p->npanels = npanels;
    // build buffers: rows, keys, the value bits and the sort's work space
    const size_t tmp_sort = sort_tmp_bytes(nnz, (int)vs);
    const size_t need = Carver::round(4 * (size_t)nnz) + Carver::round(8 * (size_t)nnz) +
                        Carver::round(vs * (size_t)nnz) + Carver::round(tmp_sort) + 256;
    void* work = nullptr;
    bool ok = hipMalloc(&work, need) == hipSuccess &&
              hipMalloc(&p->tile_ptr, 4 * (size_t)(ntiles + 1)) == hipSuccess &&
              hipMalloc(&p->packed, 4 * (size_t)(nnz ? nnz : 1)) == hipSuccess &&
              hipMalloc(&p->val, vs * (size_t)(nnz ? nnz : 1)) == hipSuccess;
    auto cleanup = [&](int code) {
        if (work) (void)hipFree(work);
        if (code) sable_csr_tiles_destroy(p);
        else *out = p;
        return code;
    };
    if (!ok) {
        (void)hipGetLastError();
        return cleanup(fail(SABLE_ENOMEM, "tile plan allocation failed (%zu bytes of work space)", need));
    }
    p->bytes = 4 * (size_t)(ntiles + 1) + (4 + vs) * (size_t)nnz;
    Carver cv(work);
    int32_t* rows = cv.take<int32_t>((size_t)nnz);
    uint64_t* keys = cv.take<uint64_t>((size_t)nnz);
    void* pay = cv.take<char>(vs * (size_t)nnz);
    void* stmp = cv.take<char>(tmp_sort);
    int* err = cv.take<int>(64);
    SABLE_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_tile_rows, dim3(gs(N)), dim3(256), 0, st, N, row_ptr, rows);
    if (nnz > 0) {
        if (dtype == SABLE_F64)
            hipLaunchKernelGGL((k_tile_keys<double, uint64_t>), dim3(gs(nnz)), dim3(256), 0, st, nnz, ncols, rows, col,
                               (const double*)val, tile_rows, cb, rin_shift, p->ntr, keys, (uint64_t*)pay, err);
        else
            hipLaunchKernelGGL((k_tile_keys<float, uint32_t>), dim3(gs(nnz)), dim3(256), 0, st, nnz, ncols, rows, col,
                               (const float*)val, tile_rows, cb, rin_shift, p->ntr, keys, (uint32_t*)pay, err);
    }
    if (hipGetLastError() != hipSuccess) return cleanup(fail(SABLE_EHIP, "tile key launch failed"));
    if (hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return cleanup(fail(SABLE_EHIP, "tile key check failed"));
    if (*(int*)ctx->pinned) return cleanup(fail(SABLE_EMALFORMED, "column index outside [0, ncols)"));
    if (nnz > 0) {
        // stable on the sorted bits: equal (tile, column) keep ascending rows
        if (dtype == SABLE_F64) {
            const int rc = sort_pairs_u64_u64(st, keys, (uint64_t*)pay, nnz, sort_bits, stmp);
            if (rc) return cleanup(rc);
            hipLaunchKernelGGL((k_tile_pack<double, uint64_t>), dim3(gs(nnz)), dim3(256), 0, st, nnz, keys,
                               (const uint64_t*)pay, rb, cb, rin_shift, p->packed, (double*)p->val);
        } else {
            const int rc = sort_pairs_u64(st, keys, (uint32_t*)pay, nnz, sort_bits, stmp);
            if (rc) return cleanup(rc);
            hipLaunchKernelGGL((k_tile_pack<float, uint32_t>), dim3(gs(nnz)), dim3(256), 0, st, nnz, keys,
                               (const uint32_t*)pay, rb, cb, rin_shift, p->packed, (float*)p->val);
        }
    }
    hipLaunchKernelGGL(k_tile_heads, dim3((unsigned)grid_for(nnz + 1, 256)), dim3(256), 0, st, nnz, ntiles, cb,
                       sort_bits, keys, p->tile_ptr);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return cleanup(fail(SABLE_EHIP, "tile plan build failed"));
    // LDS above the default dynamic limit must be requested per kernel
    const size_t shm = (size_t)tile_rows * sizeof(double);
    hipError_t ea = dtype == SABLE_F64
                        ? hipFuncSetAttribute((const void*)k_tiles<double>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm)
                        : hipFuncSetAttribute((const void*)k_tiles<float>,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (ea == hipSuccess)
        ea = dtype == SABLE_F64 ? hipFuncSetAttribute((const void*)k_tiles_fused<double>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm)
                                : hipFuncSetAttribute((const void*)k_tiles_fused<float>,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (ea != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(fail(SABLE_EINVAL, "tile of %zu LDS bytes not supported: %s", shm, hipGetErrorString(ea)));
    }
    return cleanup(SABLE_OK);
}

SABLE_API void sable_csr_tiles_destroy(sable_csr_tiles_t* p) {
    if (!p) return;
    if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
    if (p->tile_ptr) (void)hipFree(p->tile_ptr);
    if (p->packed) (void)hipFree(p->packed);
    if (p->val) (void)hipFree(p->val);
    delete p;
}

SABLE_API int sable_csr_tiles_info(const sable_csr_tiles_t* p, int32_t* row_tiles, int32_t* npanels, int64_t* bytes) {
    SABLE_CHECK(p, SABLE_EINVAL, "NULL plan");
    if (row_tiles) *row_tiles = p->ntr;
    if (npanels) *npanels = p->npanels;
    if (bytes) *bytes = (int64_t)p->bytes;
    return SABLE_OK;
}

SABLE_API int sable_csr_tiles_select(int mode) {
    SABLE_CHECK(mode == 0 || mode == 1, SABLE_EINVAL, "mode must be 0 (per-panel launches) or 1 (one launch)");
    g_tiles_mode = mode;
    return SABLE_OK;
}

SABLE_API int sable_spmv_csr_tiles(sable_ctx_t* ctx, const sable_csr_tiles_t* p, const void* x, void* y, double beta,
                                   int inner_max) {
    SABLE_CHECK(ctx && p && x && y, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(inner_max >= 0, SABLE_EINVAL, "negative inner_max");
    for (int it = 0; it < inner_max; it++) {
        const int rc = tiles_launch(ctx->stream, p, x, y, beta);
        if (rc) return rc;
    }
    return SABLE_OK;
}

SABLE_API int sable_csr_tiles_batch_item(const sable_csr_tiles_t* p, const void* x, void* y, double beta,
                                         int64_t unit_begin, sable_batch_tiles_item_t* out) {
    SABLE_CHECK(p && x && y && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(unit_begin >= 0, SABLE_EINVAL, "negative unit_begin");
    out->tile_ptr = p->tile_ptr;
    out->packed = p->packed;
    out->val = p->val;
    out->x = x;
    out->y = y;
    out->unit_begin = unit_begin;
    out->beta = beta;
    out->N = p->N;
    out->tile_rows = p->tile_rows;
    out->ntr = p->ntr;
    out->npanels = p->npanels;
    out->panel_cols = p->panel_cols;
    out->row_bits = p->row_bits;
    return SABLE_OK;
}

SABLE_API int sable_batch_spmv_tiles(sable_ctx_t* ctx, sable_dtype_t dtype, int nrhs,
                                     const sable_batch_tiles_item_t* d_items, int nitems, int64_t total_units,
                                     int32_t max_tile_rows, int workgroups) {
    SABLE_CHECK(ctx && (d_items || nitems == 0), SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    SABLE_CHECK(nrhs == 1 || nrhs == 2 || nrhs == 4 || nrhs == 8, SABLE_EINVAL, "nrhs must be 1, 2, 4 or 8");
    SABLE_CHECK(nitems >= 0 && total_units >= 0 && workgroups >= 0, SABLE_EINVAL, "negative size");
    if (nitems == 0 || total_units == 0) return SABLE_OK;
    // row sums are held in FP64 for either dtype, nrhs per row
    const size_t vs = sizeof(double) * (size_t)nrhs;
    SABLE_CHECK(max_tile_rows > 0 && (size_t)max_tile_rows * vs <= 160 * 1024, SABLE_EINVAL,
                "max_tile_rows * nrhs FP64 row sums must fit the LDS (160 KiB)");
    void* s;
    int rc = ctx_scratch(ctx, 256, &s);
    if (rc) return rc;
    unsigned long long* counter = (unsigned long long*)s;  // BT_XCDS counters
    int* err = (int*)((char*)s + 128);
    hipStream_t st = ctx->stream;
    rc = batch_tiles_launch(st, ctx->num_cu, dtype, nrhs, d_items, nitems, total_units, max_tile_rows, counter, err,
                            workgroups);
    if (rc) return rc;
    SABLE_HIP(hipMemcpyAsync(ctx->pinned, err, sizeof(int), hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    SABLE_CHECK(*(int*)ctx->pinned == 0, SABLE_EINVAL, "an item's tile_rows exceeds max_tile_rows");
    return SABLE_OK;
}

}  // extern "C"
