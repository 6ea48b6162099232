// prims.hip -- device-wide exclusive scan and stable LSD radix sort.
//
// Scan: tiles of 4096 int32 (256 threads x 16 items, blocked), a per-tile
// sum pass, a recursive scan of the tile sums, then a per-tile scan that
// adds the tile offset.  In place is allowed.
//
// Radix sort: 8-bit digits, tiles of 4096 keys (4 waves x 1024 contiguous
// keys), 32- or 64-bit values.  Per pass: digit histogram per tile
// (tile-major), the start of every (tile, digit) run from a chunked scan down
// the digit columns, then a scatter that ranks the tile's keys by digit
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

// Per-tile digit histogram, tile-major: counts[tile * RADIX + d], so each
// workgroup writes one contiguous 1 KiB row.  Each wave counts into its own
// LDS row so that skewed digits do not serialise on one counter.
__global__ __launch_bounds__(RS_BS) void k_rs_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                   uint32_t dmask, int32_t* __restrict__ counts, int64_t ntiles) {
    __shared__ int hist[RS_WAVES][RADIX];
    const int tid = threadIdx.x, w = tid / WAVE;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) hist[k][tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    // every load of the tile in flight before the first count
    uint64_t k[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = base + r * RS_BS + tid;
        k[r] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++)
        if (base + r * RS_BS + tid < n) atomicAdd(&hist[w][(int)((k[r] >> shift) & dmask)], 1);
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) t += hist[k][tid];
    counts[(int64_t)blockIdx.x * RADIX + tid] = t;
}

// The scatter needs, per tile and digit, the global start of the tile's run
// of that digit: all keys of smaller digits plus the digit's keys in earlier
// tiles.  Over the tile-major counts that is a scan down each column, done
// in chunks of RS_CHUNK tiles: column sums per chunk (written digit-major),
// a flat exclusive scan of those (digit-major order gives exactly the
// smaller-digit totals plus the earlier chunks), then a walk down each
// chunk's column that replaces the counts by the starts.
constexpr int RS_CHUNK = 64;

__global__ __launch_bounds__(RADIX) void k_rs_chunk_sums(const int32_t* __restrict__ counts, int64_t ntiles,
                                                         int64_t nchunks, int32_t* __restrict__ csum) {
    const int d = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * RS_CHUNK;
    const int m = (int)((ntiles - t0) < RS_CHUNK ? (ntiles - t0) : RS_CHUNK);
    int s = 0;
#pragma unroll 16
    for (int j = 0; j < m; j++) s += counts[(t0 + j) * RADIX + d];
    csum[(int64_t)d * nchunks + blockIdx.x] = s;
}

__global__ __launch_bounds__(RADIX) void k_rs_chunk_starts(int32_t* __restrict__ counts, int64_t ntiles,
                                                           int64_t nchunks, const int32_t* __restrict__ cstart) {
    const int d = threadIdx.x;
    const int64_t t0 = (int64_t)blockIdx.x * RS_CHUNK;
    const int m = (int)((ntiles - t0) < RS_CHUNK ? (ntiles - t0) : RS_CHUNK);
    int run = cstart[(int64_t)d * nchunks + blockIdx.x];
    for (int j0 = 0; j0 < m; j0 += 16) {
        int c[16];
#pragma unroll
        for (int j = 0; j < 16; j++) c[j] = j0 + j < m ? counts[(t0 + j0 + j) * RADIX + d] : 0;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j0 + j < m) {
                counts[(t0 + j0 + j) * RADIX + d] = run;
                run += c[j];
            }
    }
}

// Workgroup b runs on XCD b % 8.  Each XCD is given a contiguous range of
// tiles (in launch order), so tiles t and t + 1, whose runs of one digit
// meet inside a 128-byte line, write that line through the same L2 at
// about the same time; with tile = blockIdx.x the two halves of most lines
// came from different XCDs and reached HBM as partial writes.
__device__ __forceinline__ int64_t rs_tile_of_block(int64_t b, int64_t ntiles) {
    const int64_t q = ntiles >> 3, rem = ntiles & 7, x = b & 7;
    return x * q + (x < rem ? x : rem) + (b >> 3);
}

// One stable scatter pass over a 4096-key tile.
//  1. Wave w ranks its 1024 contiguous keys, 64 per round: the 8 digit-bit
//     ballots give each lane the set of lanes with its digit; a lane's rank
//     is the number of such lanes below it plus the wave's running count of
//     that digit (kept in LDS by the group's lowest lane).
//  2. Tile-local start of each digit (scan over digits) and each wave's
//     offset inside it give every key its position in the tile sorted by
//     digit; the keys are placed there in LDS.
//  3. The tile is written out in sorted order: consecutive threads write
//     consecutive positions of one bucket, so the stores coalesce into runs
//     instead of one scattered 8-byte store per key.  The values (32 or
//     64 bits) then take the same LDS buffer and are written the same way,
//     so one 32 KiB buffer serves both and four workgroups fit in a CU's
//     LDS.
template <typename V>
__global__ __launch_bounds__(RS_BS, sizeof(V) > 4 ? 3 : 4) void k_rs_scatter(const uint64_t* __restrict__ kin, const V* __restrict__ vin,
                                                         uint64_t* __restrict__ kout, V* __restrict__ vout,
                                                         int64_t n, int shift, uint32_t dmask,
                                                         const int32_t* __restrict__ offsets, int64_t ntiles) {
    static_assert(sizeof(V) <= sizeof(uint64_t), "values share the key buffer");
    __shared__ uint64_t sk[RS_TILE];        // keys, then values
    __shared__ int wcnt[RS_WAVES][RADIX];  // running, then exclusive-over-waves counts
    __shared__ int dstart[RADIX];           // tile-local start of each digit
    __shared__ int gbase[RADIX];            // global start of this tile's run of each digit
    __shared__ int scan_lds[RS_WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), w = tid / WAVE;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) wcnt[k][tid] = 0;
    const int64_t tile = rs_tile_of_block(blockIdx.x, ntiles);
    gbase[tid] = offsets[tile * RADIX + tid];
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t tbase = tile * RS_TILE;
    const int64_t wbase = tbase + (int64_t)w * RS_WKEYS;
    uint64_t key[RS_ROUNDS];
    V val[RS_ROUNDS];
    int pos[RS_ROUNDS];  // rank in the wave's run of the digit, then position in the tile
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = wbase + r * WAVE + lane;
        const bool valid = i < n;
        key[r] = valid ? kin[i] : 0;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = wbase + r * WAVE + lane;
        const bool valid = i < n;
        const int d = (int)((key[r] >> shift) & dmask);
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
        pos[r] = valid ? before + below : -1;
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
        if (pos[r] >= 0) {
            const int d = (int)((key[r] >> shift) & dmask);
            pos[r] += dstart[d] + wcnt[w][d];
            sk[pos[r]] = key[r];
        }
    }
    // the values are loaded only once the keys are placed (fewer live
    // registers); they arrive while the keys are written
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++) {
        const int64_t i = wbase + r * WAVE + lane;
        val[r] = i < n ? vin[i] : V(0);
    }
    __syncthreads();
    const int tn = (n - tbase) < RS_TILE ? (int)(n - tbase) : RS_TILE;
    int dst[RS_ROUNDS];  // global position - tile-local position of slot tid + j * RS_BS
#pragma unroll
    for (int j = 0; j < RS_ROUNDS; j++) {
        const int i = tid + j * RS_BS;
        if (i < tn) {
            const uint64_t k = sk[i];
            dst[j] = gbase[(int)((k >> shift) & dmask)];
            kout[(int64_t)dst[j] + i] = k;
        }
    }
    __syncthreads();
    V* sv = (V*)sk;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; r++)
        if (pos[r] >= 0) sv[pos[r]] = val[r];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ROUNDS; j++) {
        const int i = tid + j * RS_BS;
        if (i < tn) vout[(int64_t)dst[j] + i] = sv[i];
    }
}

size_t sort_tmp_bytes(int64_t n, int value_bytes) {
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    int64_t nchunks = (ntiles + RS_CHUNK - 1) / RS_CHUNK;
    size_t cnt = ((size_t)ntiles * RADIX * 4 + 255) & ~(size_t)255;
    size_t cs = ((size_t)nchunks * RADIX * 4 + 255) & ~(size_t)255;
    // the second key and value buffers
    size_t kv = (((size_t)n * 8 + 255) & ~(size_t)255) + (((size_t)n * value_bytes + 255) & ~(size_t)255);
    return cnt + cs + kv + scan_tmp_bytes(nchunks * RADIX) + 256;
}

template <typename V>
static int sort_pairs_t(hipStream_t st, uint64_t* keys, V* vals, int64_t n, int end_bit, void* tmp) {
    if (n <= 1 || end_bit <= 0) return SABLE_OK;
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    // positions and digit counts are 32-bit
    SABLE_CHECK(n <= 0x7fffffffLL && ntiles * RADIX < 0x7fffffffLL, SABLE_EINVAL, "sort too large");
    const int64_t nchunks = (ntiles + RS_CHUNK - 1) / RS_CHUNK;
    char* p = (char*)tmp;
    int32_t* counts = (int32_t*)p;
    p += ((size_t)ntiles * RADIX * 4 + 255) & ~(size_t)255;
    int32_t* csum = (int32_t*)p;
    p += ((size_t)nchunks * RADIX * 4 + 255) & ~(size_t)255;
    uint64_t* k2 = (uint64_t*)p;
    p += ((size_t)n * 8 + 255) & ~(size_t)255;
    V* v2 = (V*)p;
    p += ((size_t)n * sizeof(V) + 255) & ~(size_t)255;
    void* stmp = p;
    uint64_t *ka = keys, *kb = k2;
    V *va = vals, *vb = v2;
    for (int shift = 0; shift < end_bit; shift += 8) {
        // the last digit stops at end_bit: key bits above it are carried, not sorted
        const uint32_t dmask = end_bit - shift >= 8 ? RADIX - 1 : (1u << (end_bit - shift)) - 1u;
        hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)ntiles), dim3(RS_BS), 0, st, ka, n, shift, dmask, counts,
                           ntiles);
        SABLE_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_rs_chunk_sums, dim3((unsigned)nchunks), dim3(RADIX), 0, st, counts, ntiles, nchunks,
                           csum);
        SABLE_LAUNCH_CHECK();
        int rc = scan_exclusive_i32(st, csum, csum, nchunks * RADIX, nullptr, stmp);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_chunk_starts, dim3((unsigned)nchunks), dim3(RADIX), 0, st, counts, ntiles, nchunks,
                           csum);
        SABLE_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_rs_scatter<V>, dim3((unsigned)ntiles), dim3(RS_BS), 0, st, ka, va, kb, vb, n, shift,
                           dmask, counts, ntiles);
        SABLE_LAUNCH_CHECK();
        uint64_t* tk = ka;
        ka = kb;
        kb = tk;
        V* tv = va;
        va = vb;
        vb = tv;
    }
    if (ka != keys) {
        SABLE_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
        SABLE_HIP(hipMemcpyAsync(vals, va, (size_t)n * sizeof(V), hipMemcpyDeviceToDevice, st));
    }
    return SABLE_OK;
}

// Sorts (keys, vals) in place by bits [0, end_bit) of the keys, stably.
int sort_pairs_u64(hipStream_t st, uint64_t* keys, uint32_t* vals, int64_t n, int end_bit, void* tmp) {
    return sort_pairs_t<uint32_t>(st, keys, vals, n, end_bit, tmp);
}

int sort_pairs_u64_u64(hipStream_t st, uint64_t* keys, uint64_t* vals, int64_t n, int end_bit, void* tmp) {
    return sort_pairs_t<uint64_t>(st, keys, vals, n, end_bit, tmp);
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
