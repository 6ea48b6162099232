// prims.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Scan: tiles of 4096 int32 (256 threads x 16 items, blocked), a per-tile
// sum pass, a recursive scan of the tile sums, then a per-tile scan that
// adds the tile offset.  In place is allowed.
//
// Radix sort: 8-bit digits, tiles of 4096 keys (4 waves x 1024 contiguous
// keys).  Per pass: digit histogram per tile (digit-major layout), exclusive
// scan of the histogram, then a scatter that ranks the tile's keys by digit
// with wave ballots, reorders them in LDS and writes each digit's run
// contiguously.  Equal digits keep their input order, so the sort is stable.
#include "common.h"
#include "prims.h"

namespace sable {

constexpr int SCAN_BS = 256;
constexpr int SCAN_IT = 16;
constexpr int SCAN_TILE = SCAN_BS * SCAN_IT;

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        int t = __shfl_up(v, off, WAVE);
        if (lane >= off) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the 256-thread block.
__device__ __forceinline__ int block_excl_scan(int v, int* total, int* lds /* >= 4 */) {
    const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
    int inc = wave_incl_scan(v);
    if (lane == WAVE - 1) lds[w] = inc;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < SCAN_BS / WAVE; k++) {
        int s = lds[k];
        if (k < w) before += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + inc - v;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_sums(const int32_t* __restrict__ in, int64_t n,
                                                       int32_t* __restrict__ sums) {
    __shared__ int lds[SCAN_BS / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IT;
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IT; i++)
        if (base + i < n) s += in[base + i];
    int tot;
    block_excl_scan(s, &tot, lds);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BS) void k_scan_tiles(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                        int64_t n, const int32_t* __restrict__ tile_off,
                                                        int32_t* __restrict__ total_out) {
    __shared__ int lds[SCAN_BS / WAVE];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IT;
    int v[SCAN_IT];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IT; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int tot;
    int pre = block_excl_scan(s, &tot, lds) + (tile_off ? tile_off[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < SCAN_IT; i++) {
        if (base + i < n) out[base + i] = pre;
        pre += v[i];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BS - 1) *total_out = pre;
}

size_t scan_tmp_bytes(int64_t n) {
    size_t b = 0;
    while (n > SCAN_TILE) {
        int64_t t = (n + SCAN_TILE - 1) / SCAN_TILE;
        b += ((size_t)t * 4 + 255) & ~(size_t)255;
        n = t;
    }
    return b + 256;
}

// out[0..n) = exclusive scan of in[0..n); *total_dev (if given) = sum.
int scan_exclusive_i32(hipStream_t st, const int32_t* in, int32_t* out, int64_t n, int32_t* total_dev, void* tmp) {
    if (n <= 0) return SABLE_OK;
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_BS), 0, st, in, out, n, (const int32_t*)nullptr, total_dev);
        SABLE_LAUNCH_CHECK();
        return SABLE_OK;
    }
    SABLE_CHECK(tiles < 0x7fffffffLL, SABLE_EINVAL, "scan too large");
    int32_t* sums = (int32_t*)tmp;
    void* rest = (char*)tmp + (((size_t)tiles * 4 + 255) & ~(size_t)255);
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)tiles), dim3(SCAN_BS), 0, st, in, n, sums);
    SABLE_LAUNCH_CHECK();
    int rc = scan_exclusive_i32(st, sums, sums, tiles, nullptr, rest);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_BS), 0, st, in, out, n, sums, total_dev);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

// ---------------------------------------------------------------- radix sort
constexpr int RS_BS = 256;                 // 4 waves
constexpr int RS_WAVES = RS_BS / WAVE;
constexpr int RS_ROUNDS = 16;              // 64-key rounds per wave
constexpr int RS_WKEYS = WAVE * RS_ROUNDS; // keys per wave (contiguous)
constexpr int RS_TILE = RS_BS * RS_ROUNDS; // 4096 keys per workgroup
constexpr int RADIX = 256;

// Per-tile digit histogram, digit-major: counts[d * ntiles + tile].  Each
// wave counts into its own LDS row so that skewed digits do not serialise
// on one counter.
__global__ __launch_bounds__(RS_BS) void k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                   int32_t* __restrict__ counts, int64_t ntiles) {
    __shared__ int hist[RS_WAVES][RADIX];
    const int tid = threadIdx.x, w = tid / WAVE;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) hist[k][tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = base + r * RS_BS + tid;
        if (i < n) atomicAdd(&hist[w][(int)((keys[i] >> shift) & (RADIX - 1))], 1);
    }
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) t += hist[k][tid];
    counts[(int64_t)tid * ntiles + blockIdx.x] = t;
}

// One stable scatter pass over a 4096-key tile.
//  1. Wave w ranks its 1024 contiguous keys, 64 per round: the 8 digit-bit
//     ballots give each lane the set of lanes with its digit; a lane's rank
//     is the number of such lanes below it plus the wave's running count of
//     that digit (kept in LDS by the group's lowest lane).
//  2. Tile-local start of each digit (scan over digits) and each wave's
//     offset inside it give every key its position in the tile sorted by
//     digit; keys and values are placed there in LDS.
//  3. The tile is written out in sorted order: consecutive threads write
//     consecutive positions of one bucket, so the stores coalesce into runs
//     instead of one scattered 8-byte store per key.
__global__ __launch_bounds__(RS_BS) void k_rs_scatter(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                      uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      int64_t n, int shift, const int32_t* __restrict__ offsets,
                                                      int64_t ntiles) {
    __shared__ uint64_t sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ int wcnt[RS_WAVES][RADIX];  // running, then exclusive-over-waves counts
    __shared__ int dstart[RADIX];           // tile-local start of each digit
    __shared__ int gbase[RADIX];            // global start of this tile's run of each digit
    __shared__ int scan_lds[RS_WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) wcnt[k][tid] = 0;
    gbase[tid] = offsets[(int64_t)tid * ntiles + blockIdx.x];
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t tbase = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wbase = tbase + (int64_t)w * RS_WKEYS;
    uint64_t key[RS_ROUNDS];
    uint32_t val[RS_ROUNDS];
    int rank[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = wbase + r * WAVE + lane;
        const bool valid = i < n;
        key[r] = valid ? kin[i] : 0;
        val[r] = valid ? vin[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = wbase + r * WAVE + lane;
        const bool valid = i < n;
        const int d = (int)((key[r] >> shift) & (RADIX - 1));
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int below = __popcll(peers & lt);
        int before = 0;
        if (valid) before = wcnt[w][d];
        rank[r] = valid ? before + below : -1;
        // all lanes have read the running count before the leader bumps it
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wcnt[w][d] = before + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive over waves, and the tile total
    int tot = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) {
        const int c = wcnt[k][tid];
        wcnt[k][tid] = tot;
        tot += c;
    }
    int all;
    const int st = block_excl_scan(tot, &all, scan_lds);
    dstart[tid] = st;
    gbase[tid] -= st;  // global position = gbase[d] + tile-local position
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        if (rank[r] >= 0) {
            const int d = (int)((key[r] >> shift) & (RADIX - 1));
            const int pos = dstart[d] + wcnt[w][d] + rank[r];
            sk[pos] = key[r];
            sv[pos] = val[r];
        }
    }
    __syncthreads();
    const int64_t tn = (n - tbase) < RS_TILE ? (n - tbase) : RS_TILE;
#pragma unroll 4
    for (int i = tid; i < tn; i += RS_BS) {
        const uint64_t k = sk[i];
        const int d = (int)((k >> shift) & (RADIX - 1));
        const int64_t g = (int64_t)gbase[d] + i;
        kout[g] = k;
        vout[g] = sv[i];
    }
}

size_t sort_tmp_bytes(int64_t n) {
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    size_t cnt = ((size_t)ntiles * RADIX * 4 + 255) & ~(size_t)255;
    size_t kv = (((size_t)n * 8 + 255) & ~(size_t)255) + (((size_t)n * 4 + 255) & ~(size_t)255);
    return cnt + kv + scan_tmp_bytes(ntiles * RADIX) + 256;
}

// Sorts (keys, vals) in place by bits [0, end_bit) of the keys, stably.
int sort_pairs_u64(hipStream_t st, uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, void* tmp) {
    if (n <= 1 || end_bit <= 0) return SABLE_OK;
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    SABLE_CHECK(ntiles * RADIX < 0x7fffffffLL, SABLE_EINVAL, "sort too large");
    char* p = (char*)tmp;
    int32_t* counts = (int32_t*)p;
    p += ((size_t)ntiles * RADIX * 4 + 255) & ~(size_t)255;
    uint64_t* k2 = (uint64_t*)p;
    p += ((size_t)n * 8 + 255) & ~(size_t)255;
    uint32_t* v2 = (uint32_t*)p;
    p += ((size_t)n * 4 + 255) & ~(size_t)255;
    void* stmp = p;
    uint64_t *ka = keys, *kb = k2;
    uint32_t *va = vals, *vb = v2;
    for (int shift = 0; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_BS), 0, st, ka, n, shift, counts, ntiles);
        SABLE_LAUNCH_CHECK();
        int rc = scan_exclusive_i32(st, counts, counts, ntiles * RADIX, nullptr, stmp);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)ntiles), dim3(RS_BS), 0, st, ka, va, kb, vb, n, shift,
                           counts, ntiles);
        SABLE_LAUNCH_CHECK();
        uint64_t* tk = ka;
        ka = kb;
        kb = tk;
        uint32_t* tv = va;
        va = vb;
        vb = tv;
    }
    if (ka != keys) {
        SABLE_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        SABLE_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
    return SABLE_OK;
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_scan_exclusive_i32(sable_ctx_t* ctx, const int32_t* in, int32_t* out, int64_t n) {
    SABLE_CHECK(ctx && n >= 0, SABLE_EINVAL, "bad argument");
    if (n == 0) return SABLE_OK;
    void* tmp;
    int rc = ctx_scratch(ctx, scan_tmp_bytes(n), &tmp);
    if (rc) return rc;
    return scan_exclusive_i32(ctx->stream, in, out, n, out + n, tmp);
}

SABLE_API int sable_sort_pairs_u64(sable_ctx_t* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int end_bit) {
    SABLE_CHECK(ctx && n >= 0 && end_bit >= 0 && end_bit <= 64, SABLE_EINVAL, "bad argument");
    if (n <= 1) return SABLE_OK;
    void* tmp;
    int rc = ctx_scratch(ctx, sort_tmp_bytes(n), &tmp);
    if (rc) return rc;
    return sort_pairs_u64(ctx->stream, keys, vals, n, end_bit, tmp);
}

}  // extern "C"
