// auto.hip -- a CSR product whose schedule is chosen by measurement.
//
// sable_csr_auto_create builds the candidate schedules of one resident
// matrix, times each with HIP events on the context stream and keeps the
// fastest:
//  * row blocks (k_csr_block) at three block/row weights;
//  * the row-vector kernel with lanes around the mean row length, for
//    short, regular rows;
//  * column-sorted tiles (16384-row tiles over 8192-column panels when x
//    outgrows one XCD's 4 MiB L2, and over 16384-column panels too when a
//    tile holds few entries per panel; single-panel tiles of up to 4096
//    rows otherwise), unless the caller asks for bit-reproducible results.
// The winner is cached per matrix class, keyed on row-length statistics
// (log2 buckets of N, of the mean and of the longest row, a coefficient-of-
// variation bucket), the value type and whether x fits one XCD's L2, so
// later matrices of a measured class are built straight away.
//
// This is the schedule the Node-API addon keeps for each resident matrix
// behind the reference's spmv_csr / spmv_csr_wrapper and spmv_coo exports.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

struct Cand {
    int sched = SABLE_SCHED_BLOCKS;
    int p0 = 0, p1 = 0;  // blocks: block/row weight; vector: lanes/rows; tiles: tile rows/panel cols
};

std::mutex g_auto_mu;
std::map<std::string, Cand> g_auto_cache;

}  // namespace

struct sable_csr_auto_s {
    sable_ctx_t* ctx = nullptr;
    sable_dtype_t dtype = SABLE_F64;
    int32_t N = 0;
    int32_t ncols = 0;
    int64_t nnz = 0;
    const int32_t* row_ptr = nullptr;
    const int32_t* col = nullptr;
    const void* val = nullptr;
    Cand chosen;
    float ms = 0.f;
    bool from_cache = false;
    sable_csr_plan_t* plan = nullptr;
    sable_csr_tiles_t* tiles = nullptr;
};

namespace sable {

__global__ void k_auto_colmax(int64_t n, const int32_t* __restrict__ col, int* __restrict__ out) {
    int m = -1;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        m = max(m, col[k]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0 && m >= 0) atomicMax(out, m);
}

__global__ void k_auto_rebase(int32_t N, const int32_t* __restrict__ rp, int32_t* __restrict__ out) {
    const int32_t b = rp[0];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= N; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = rp[i] - b;
}

__global__ void k_auto_set(int* p, int v) {
    if (threadIdx.x == 0) *p = v;
}

template <typename T>
__global__ void k_auto_ones(int64_t n, T* __restrict__ p) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = T(1);
}

static int bucket(double v) { return (int)std::floor(std::log2(std::max(v, 1.0))); }

static void release_sched(sable_csr_auto_s* p) {
    if (p->plan) sable_csr_plan_destroy(p->plan);
    if (p->tiles) sable_csr_tiles_destroy(p->tiles);
    p->plan = nullptr;
    p->tiles = nullptr;
}

// Builds candidate c for p (replacing any previous schedule).
static int build_sched(sable_csr_auto_s* p, const Cand& c, int32_t base) {
    release_sched(p);
    sable_ctx_t* ctx = p->ctx;
    if (c.sched == SABLE_SCHED_TILES) {
        const size_t vs = p->dtype == SABLE_F64 ? 8 : 4;
        const int32_t* rp = p->row_ptr;
        int32_t* rebased = nullptr;
        if (base != 0) {
            // a row slice: the tile plan takes a row_ptr starting at 0 over
            // col/val shifted to the slice's first entry
            SABLE_HIP(hipMalloc(&rebased, 4 * ((size_t)p->N + 1)));
            unsigned g = grid_for((int64_t)p->N + 1, 256);
            if (g > 4096) g = 4096;
            hipLaunchKernelGGL(k_auto_rebase, dim3(g), dim3(256), 0, ctx->stream, p->N, p->row_ptr, rebased);
            rp = rebased;
        }
        const int rc = sable_csr_tiles_create_rect(ctx, p->dtype, p->N, p->ncols, rp, p->col + base,
                                                   (const char*)p->val + vs * (size_t)base, c.p0, c.p1, &p->tiles);
        if (rebased) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(rebased);
        }
        return rc;
    }
    sable_csr_tune_t t;
    memset(&t, 0, sizeof t);
    if (c.sched == SABLE_SCHED_VECTOR) {
        t.vector_lanes = c.p0;
        t.vector_rows = c.p1;
    } else {
        t.block_weight = c.p0;
        t.row_weight = c.p1;
    }
    return sable_csr_plan_create(ctx, p->N, p->row_ptr, &t, &p->plan);
}

static int run_sched(sable_ctx_t* ctx, const sable_csr_auto_s* p, const void* x, void* y, double beta,
                     int inner_max) {
    if (p->tiles) return sable_spmv_csr_tiles(ctx, p->tiles, x, y, beta, inner_max);
    return sable_spmv_csr_plan(ctx, p->plan, p->dtype, p->col, p->val, x, y, beta, inner_max);
}

// Fewer than 1024 entries per (16384-row tile, 8192-column panel) on
// average: a panel step then does too little work for its fixed cost, and
// panels twice as wide are the better walk (Zipf rows of mean 16 over 4 M
// columns: 0.326 against 0.424 ms; mean 49: 0.857 against 0.642 ms).
static bool sparse_panels(const sable_csr_auto_s* p) {
    const double tiles = std::ceil((double)p->N / 16384.0);
    return (double)p->nnz / tiles * (8192.0 / (double)p->ncols) < 1024.0;
}

static std::vector<Cand> candidates(const sable_csr_auto_s* p, double mean, int32_t mx, int flags) {
    std::vector<Cand> out;
    out.push_back({SABLE_SCHED_BLOCKS, 1024, 2});
    out.push_back({SABLE_SCHED_BLOCKS, 512, 1});
    out.push_back({SABLE_SCHED_BLOCKS, 2048, 4});
    int g0 = 2;
    while (g0 < mean && g0 < 64) g0 *= 2;
    for (int g : {g0 / 2, g0, 2 * g0})
        if (g >= 2 && g <= 64 && mx <= 8 * g) out.push_back({SABLE_SCHED_VECTOR, g, 4});
    const size_t vs = p->dtype == SABLE_F64 ? 8 : 4;
    if (!(flags & SABLE_AUTO_DETERMINISTIC) && p->nnz >= 4096) {
        if ((size_t)p->ncols * vs > ((size_t)4 << 20)) {
            out.push_back({SABLE_SCHED_TILES, 16384, 8192});
            if (sparse_panels(p)) out.push_back({SABLE_SCHED_TILES, 16384, 16384});
        } else {
            int pc = 64;
            while (pc < p->ncols) pc *= 2;
            const int tr = (int)std::min<int64_t>(4096, ((int64_t)p->N + 63) / 64 * 64);
            if ((int64_t)tr * pc <= ((int64_t)1 << 32)) out.push_back({SABLE_SCHED_TILES, tr, pc});
        }
    }
    return out;
}

// Untimed choice: tiles when x outgrows one XCD's L2 and rows are not tiny,
// the row-vector kernel for short regular rows, row blocks otherwise.
static Cand heuristic(const sable_csr_auto_s* p, double mean, int32_t mx, double cv, int flags) {
    const size_t vs = p->dtype == SABLE_F64 ? 8 : 4;
    if (!(flags & SABLE_AUTO_DETERMINISTIC) && (size_t)p->ncols * vs > ((size_t)4 << 20) && mean >= 8.0)
        return {SABLE_SCHED_TILES, 16384, sparse_panels(p) ? 16384 : 8192};
    if (mean <= 16.0 && cv < 0.5 && mx <= 64) {
        int g = 2;
        while (g < mean && g < 64) g *= 2;
        if (mx <= 8 * g) return {SABLE_SCHED_VECTOR, g, 4};
    }
    return {SABLE_SCHED_BLOCKS, 1024, 2};
}

}  // namespace sable

using namespace sable;

extern "C" {

SABLE_API int sable_csr_auto_create(sable_ctx_t* ctx, sable_dtype_t dtype, int32_t N, int32_t ncols,
                                    const int32_t* row_ptr, const int32_t* col, const void* val, int flags,
                                    sable_csr_auto_t** out) {
    SABLE_CHECK(ctx && out && row_ptr && col && val, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(N > 0, SABLE_EINVAL, "N must be positive");
    SABLE_CHECK(dtype == SABLE_F32 || dtype == SABLE_F64, SABLE_EINVAL, "bad dtype");
    *out = nullptr;
    hipStream_t st = ctx->stream;
    std::vector<int32_t> rp((size_t)N + 1);
    SABLE_HIP(hipMemcpyAsync(rp.data(), row_ptr, 4 * ((size_t)N + 1), hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipStreamSynchronize(st));
    const int32_t base = rp[0];
    SABLE_CHECK(base >= 0 && rp[N] >= base, SABLE_EMALFORMED, "row_ptr[N]=%d < row_ptr[0]=%d", rp[N], base);
    int32_t mx = 0;
    double sum = 0, sq = 0;
    for (int32_t i = 0; i < N; i++) {
        const int32_t l = rp[i + 1] - rp[i];
        SABLE_CHECK(l >= 0, SABLE_EMALFORMED, "row_ptr decreases at row %d", i);
        mx = std::max(mx, l);
        sum += l;
        sq += (double)l * l;
    }
    const int64_t nnz = (int64_t)rp[N] - base;
    const double mean = sum / N;
    const double var = std::max(0.0, sq / N - mean * mean);
    const double cv = mean > 0 ? std::sqrt(var) / mean : 0.0;
    if (ncols <= 0) {
        // x length = largest column + 1 (synchronises)
        int* d = (int*)ctx->pinned;
        void* s;
        int rc = ctx_scratch(ctx, 256, &s);
        if (rc) return rc;
        int* dm = (int*)s;
        hipLaunchKernelGGL(k_auto_set, dim3(1), dim3(64), 0, st, dm, -1);
        if (nnz > 0) {
            unsigned g = grid_for(nnz, 256);
            if (g > 2048) g = 2048;
            hipLaunchKernelGGL(k_auto_colmax, dim3(g), dim3(256), 0, st, nnz, col + base, dm);
        }
        SABLE_LAUNCH_CHECK();
        SABLE_HIP(hipMemcpyAsync(d, dm, 4, hipMemcpyDeviceToHost, st));
        SABLE_HIP(hipStreamSynchronize(st));
        ncols = *d + 1;
        if (ncols < 1) ncols = 1;
    }
    auto* p = new sable_csr_auto_s();
    p->ctx = ctx;
    p->dtype = dtype;
    p->N = N;
    p->ncols = ncols;
    p->nnz = nnz;
    p->row_ptr = row_ptr;
    p->col = col;
    p->val = val;
    const size_t vs = dtype == SABLE_F64 ? 8 : 4;
    char key[160];
    snprintf(key, sizeof key, "N%d_mean%d_max%d_cv%d_%s_x%s_f%d", bucket(N), bucket(mean), bucket(mx),
             std::min((int)(cv * 2), 8), dtype == SABLE_F64 ? "f64" : "f32",
             (size_t)ncols * vs > ((size_t)4 << 20) ? "big" : "l2", flags & SABLE_AUTO_DETERMINISTIC);
    Cand chosen;
    bool cached = false;
    {
        std::lock_guard<std::mutex> hold(g_auto_mu);
        auto it = g_auto_cache.find(key);
        if (it != g_auto_cache.end()) {
            chosen = it->second;
            cached = true;
        }
    }
    int rc = SABLE_OK;
    if (!cached && (flags & SABLE_AUTO_NO_TIMING)) {
        chosen = heuristic(p, mean, mx, cv, flags);
    } else if (!cached) {
        // time every candidate on y = A.x with x = 1
        void *dx = nullptr, *dy = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipMalloc(&dx, vs * (size_t)ncols) != hipSuccess || hipMalloc(&dy, vs * (size_t)N) != hipSuccess ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            (void)hipGetLastError();
            if (dx) (void)hipFree(dx);
            if (dy) (void)hipFree(dy);
            if (e0) (void)hipEventDestroy(e0);
            delete p;
            return fail(SABLE_ENOMEM, "autotune buffers");
        }
        unsigned g = grid_for(ncols, 256);
        if (g > 4096) g = 4096;
        if (dtype == SABLE_F64)
            hipLaunchKernelGGL(k_auto_ones<double>, dim3(g), dim3(256), 0, st, (int64_t)ncols, (double*)dx);
        else
            hipLaunchKernelGGL(k_auto_ones<float>, dim3(g), dim3(256), 0, st, (int64_t)ncols, (float*)dx);
        float best = 1e30f;
        const int reps = 3;
        for (const Cand& c : candidates(p, mean, mx, flags)) {
            if (build_sched(p, c, base) != SABLE_OK) continue;  // e.g. a tile plan that does not fit
            if (run_sched(ctx, p, dx, dy, 0.0, 1) != SABLE_OK) continue;
            (void)hipEventRecord(e0, st);
            int r2 = run_sched(ctx, p, dx, dy, 0.0, reps);
            (void)hipEventRecord(e1, st);
            float ms = 1e30f;
            if (r2 == SABLE_OK && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                ms /= reps;
            if (ms < best) {
                best = ms;
                chosen = c;
            }
        }
        (void)hipGetLastError();
        release_sched(p);
        (void)hipStreamSynchronize(st);
        (void)hipFree(dx);
        (void)hipFree(dy);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (best >= 1e30f) {
            delete p;
            return fail(SABLE_EHIP, "no CSR schedule could be built for this matrix");
        }
        p->ms = best;
        std::lock_guard<std::mutex> hold(g_auto_mu);
        g_auto_cache[key] = chosen;
    }
    rc = build_sched(p, chosen, base);
    if (rc) {
        release_sched(p);
        delete p;
        return rc;
    }
    p->chosen = chosen;
    p->from_cache = cached;
    *out = p;
    return SABLE_OK;
}

SABLE_API void sable_csr_auto_destroy(sable_csr_auto_t* p) {
    if (!p) return;
    release_sched(p);
    delete p;
}

SABLE_API int sable_csr_auto_info(const sable_csr_auto_t* p, sable_csr_auto_info_t* info) {
    SABLE_CHECK(p && info, SABLE_EINVAL, "NULL argument");
    memset(info, 0, sizeof *info);
    info->schedule = p->chosen.sched;
    info->param0 = p->chosen.p0;
    info->param1 = p->chosen.p1;
    info->ncols = p->ncols;
    info->nnz = p->nnz;
    info->ms = p->ms;
    info->from_cache = p->from_cache ? 1 : 0;
    if (p->tiles) {
        int32_t a, b;
        int64_t bytes = 0;
        (void)sable_csr_tiles_info(p->tiles, &a, &b, &bytes);
        info->extra_bytes = bytes;
    }
    return SABLE_OK;
}

SABLE_API int sable_spmv_csr_auto(sable_ctx_t* ctx, const sable_csr_auto_t* p, const void* x, void* y, double beta,
                                  int inner_max) {
    SABLE_CHECK(ctx && p && x && y, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(ctx == p->ctx, SABLE_EINVAL, "plan belongs to another context");
    SABLE_CHECK(inner_max >= 0, SABLE_EINVAL, "negative inner_max");
    if (inner_max == 0) return SABLE_OK;
    return run_sched(ctx, p, x, y, beta, inner_max);
}

SABLE_API void sable_csr_auto_cache_clear(void) {
    std::lock_guard<std::mutex> hold(g_auto_mu);
    g_auto_cache.clear();
}

}  // extern "C"
