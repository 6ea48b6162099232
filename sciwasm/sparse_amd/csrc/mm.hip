// mm.hip -- MatrixMarket ingest on the device: the text of a coordinate
// file goes to HBM once, and the entry lines are found, parsed, expanded and
// sorted there, ending in a COO ready for sable_coo_csr.
//
// Reference: seq/src/sparse32.js (FP64: seq/src/sparse64.js:1536-1760)
//   read_MM_header           :1889-1911   banner tokens 3 and 4, '%' lines, size line
//   calculate_actual_nnz     :1914-1951   nnz: symmetric off-diagonals twice, explicit zeros dropped
//   read_matrix_MM_files     :1953-1981   nnz == 0 -> nnz = nentries
//   create_LCOO_from_MM      :1983-2097   strictly lower part + unit diagonal
//   create_COO_from_MM       :2100-2172   expansion in file order, then quick_sort_COO by (row, col)
//   mmread                   :2498-2510
//
// Pipeline (one stream, the context's):
//   1. k_mm_count / k_mm_starts: every byte position that begins an entry
//      line (after '\n', not '\n', '\r' or '%') is compacted into starts[]
//      through a scan of per-4 KiB-tile counts.
//   2. k_mm_parse: one lane per line, 256 lines per workgroup staged into
//      LDS; tokens, indices and values (mm_parse.h).  Values the fast
//      conversion cannot decide are queued for k_mm_exact.
//   3. sable_mm_coo: scan of per-line output counts, expansion into
//      (row << b | col) keys, stable radix sort, gather.
#include "common.h"
#include "prims.h"

#define SABLE_HD __host__ __device__
#include "mm_parse.h"

#include <string.h>
#include <algorithm>
#include <string>

using namespace sable;
using namespace sable::mm;

__device__ const Pow5 kPow5Dev[] = {
#include "pow5_table.inc"
};

namespace {

constexpr int MB = 256;                 // threads per workgroup
constexpr int TILE_BYTES = MB * 16;     // bytes per line-start tile
constexpr int STAGE_BYTES = 32 * 1024;  // LDS staging per 256 lines

// counters[] slots
enum { C_ENTRIES = 0, C_NNZ, C_LOWER, C_AMB, C_ERR, C_COUNT };

__device__ __forceinline__ bool line_head(int prev, int c) {
    return prev == '\n' && c != '\n' && c != '\r' && c != '%';
}

// Bit k set: byte p + k (p 16-byte aligned in memory) begins a line at or
// after b0.
__device__ __forceinline__ uint32_t head_mask(const uint8_t* __restrict__ t, int64_t len, int64_t b0, int64_t p) {
    uint8_t by[16];
    if (p >= 0 && p + 16 <= len) {
        const uint4 v = *(const uint4*)(t + p);
        memcpy(by, &v, 16);
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) by[k] = (p + k >= 0 && p + k < len) ? t[p + k] : 0;
    }
    int prev = (p >= 1 && p - 1 < len) ? t[p - 1] : '\n';
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int64_t pos = p + k;
        const int c = by[k];
        if (pos >= b0 && pos < len && line_head(prev, c)) m |= 1u << k;
        prev = c;
    }
    return m;
}

__global__ __launch_bounds__(MB) void k_mm_count(const uint8_t* __restrict__ t, int64_t len, int64_t b0, int64_t a0,
                                                 int32_t* __restrict__ tile_counts) {
    const int64_t p = a0 + (int64_t)blockIdx.x * TILE_BYTES + (int64_t)threadIdx.x * 16;
    int n = __popc(head_mask(t, len, b0, p));
    n = wave_sum(n);
    __shared__ int ws[MB / WAVE];
    if ((threadIdx.x & (WAVE - 1)) == 0) ws[threadIdx.x / WAVE] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < MB / WAVE; ++w) s += ws[w];
        tile_counts[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(MB) void k_mm_starts(const uint8_t* __restrict__ t, int64_t len, int64_t b0, int64_t a0,
                                                  const int32_t* __restrict__ tile_offs,
                                                  int64_t* __restrict__ starts) {
    const int64_t p = a0 + (int64_t)blockIdx.x * TILE_BYTES + (int64_t)threadIdx.x * 16;
    uint32_t m = head_mask(t, len, b0, p);
    const int n = __popc(m);
    // block exclusive scan of n
    const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    int inc = n;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const int o = __shfl_up(inc, off, WAVE);
        if (lane >= off) inc += o;
    }
    __shared__ int ws[MB / WAVE];
    if (lane == WAVE - 1) ws[wv] = inc;
    __syncthreads();
    int base = tile_offs[blockIdx.x];
    for (int w = 0; w < wv; ++w) base += ws[w];
    int64_t o = (int64_t)base + inc - n;
    while (m) {
        const int k = __ffs(m) - 1;
        starts[o++] = p + k;
        m &= m - 1;
    }
}

struct GSrc {
    const uint8_t* t;
    int64_t end;
    __device__ int operator()(int64_t p) const { return p < end ? (int)t[p] : -1; }
};

struct LSrc {
    const uint8_t* lds;
    int64_t base, end;
    __device__ int operator()(int64_t p) const { return p < end ? (int)lds[p - base] : -1; }
};

__device__ __forceinline__ void entry_counts(int64_t r, int64_t c, bool kept, bool symmetric, int* full, int* lower) {
    *full = kept ? ((symmetric && r != c) ? 2 : 1) : 0;
    *lower = (kept && (symmetric ? r != c : r > c)) ? 1 : 0;
}

__device__ __forceinline__ void report_error(unsigned long long* counters, int64_t pos, int code) {
    // smallest byte offset wins; the code rides in the low byte
    atomicMin(&counters[C_ERR], ((unsigned long long)pos << 8) | (unsigned long long)(-code));
}

template <class Src>
__device__ void parse_one(const Src& s, int64_t l, int64_t p, bool pattern, bool symmetric, int64_t nrows,
                          int64_t ncols, int32_t* rows, int32_t* cols, double* vals, int32_t* cnt,
                          int32_t* amb_list, unsigned long long* counters, int* n_entry, int* n_full, int* n_lower) {
    Entry e;
    parse_line(s, p, pattern, &e);
    int32_t r32 = 0, c32 = 0, cn = 0;
    double v = 0.0;
    if (e.status == LINE_ENTRY) {
        if (e.row < 1 || e.row > nrows || e.col < 1 || e.col > ncols) {
            report_error(counters, p, E_RANGE);
        } else {
            r32 = (int32_t)e.row;
            c32 = (int32_t)e.col;
            *n_entry += 1;
            bool amb = false;
            if (pattern) {
                v = 1.0;
            } else {
                v = __longlong_as_double((long long)fast_bits(e.v, kPow5Dev, &amb));
            }
            if (amb) {
                const unsigned long long k = atomicAdd(&counters[C_AMB], 1ull);
                amb_list[k] = (int32_t)l;
            } else {
                int f, lo;
                entry_counts(e.row, e.col, pattern || v != 0.0, symmetric, &f, &lo);
                cn = f;
                *n_full += f;
                *n_lower += lo;
            }
        }
    } else if (e.status < 0) {
        report_error(counters, p, e.status);
    }
    rows[l] = r32;
    cols[l] = c32;
    vals[l] = v;
    cnt[l] = cn;
}

__global__ __launch_bounds__(MB) void k_mm_parse(const uint8_t* __restrict__ t, int64_t len,
                                                 const int64_t* __restrict__ starts, int64_t nlines, int pattern,
                                                 int symmetric, int64_t nrows, int64_t ncols,
                                                 int32_t* __restrict__ rows, int32_t* __restrict__ cols,
                                                 double* __restrict__ vals, int32_t* __restrict__ cnt,
                                                 int32_t* __restrict__ amb_list,
                                                 unsigned long long* __restrict__ counters,
                                                 int32_t* __restrict__ partials) {
    __shared__ __align__(16) uint8_t stage[STAGE_BYTES + 32];
    __shared__ int wsum[3][MB / WAVE];
    const int64_t l0 = (int64_t)blockIdx.x * MB;
    const int64_t l1 = l0 + MB < nlines ? l0 + MB : nlines;
    const int64_t s0 = starts[l0];
    const int64_t s1 = l1 < nlines ? starts[l1] : len;
    // 16-byte aligned window [a0, a1) covering [s0, s1)
    const int64_t mis = (int64_t)((uintptr_t)t & 15);
    const int64_t a0 = s0 - ((s0 + mis) & 15);
    const int64_t a1 = s1 + ((16 - ((s1 + mis) & 15)) & 15);
    const bool staged = (a1 - a0) <= STAGE_BYTES;
    if (staged) {
        for (int64_t q = a0 + 16 * (int64_t)threadIdx.x; q < a1; q += 16 * MB) {
            uint8_t* d = stage + (q - a0);
            if (q + 16 <= len) {
                *(uint4*)d = *(const uint4*)(t + q);
            } else {
                for (int k = 0; k < 16; ++k) d[k] = (q + k < len) ? t[q + k] : 0;
            }
        }
    }
    __syncthreads();
    const int64_t l = l0 + threadIdx.x;
    int n_entry = 0, n_full = 0, n_lower = 0;
    if (l < l1) {
        const int64_t p = starts[l];
        const int64_t e = (l + 1 < nlines) ? starts[l + 1] : len;
        if (staged)
            parse_one(LSrc{stage, a0, e}, l, p, pattern, symmetric, nrows, ncols, rows, cols, vals, cnt, amb_list,
                      counters, &n_entry, &n_full, &n_lower);
        else
            parse_one(GSrc{t, e}, l, p, pattern, symmetric, nrows, ncols, rows, cols, vals, cnt, amb_list, counters,
                      &n_entry, &n_full, &n_lower);
    }
    // per-workgroup sums; k_mm_reduce adds them (one device-wide atomic per
    // wave on three addresses serialises across the XCDs)
    n_entry = wave_sum(n_entry);
    n_full = wave_sum(n_full);
    n_lower = wave_sum(n_lower);
    const int wv = threadIdx.x / WAVE;
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        wsum[0][wv] = n_entry;
        wsum[1][wv] = n_full;
        wsum[2][wv] = n_lower;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int t = 0;
        for (int w = 0; w < MB / WAVE; ++w) t += wsum[threadIdx.x][w];
        partials[3 * (int64_t)blockIdx.x + threadIdx.x] = t;
    }
}

// counters[C_ENTRIES, C_NNZ, C_LOWER] += sums of the per-workgroup partials.
__global__ __launch_bounds__(1024) void k_mm_reduce(const int32_t* __restrict__ partials, int64_t nblocks,
                                                    unsigned long long* __restrict__ counters) {
    unsigned long long s[3] = {0, 0, 0};
    for (int64_t b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < 3; ++k) s[k] += (unsigned long long)partials[3 * b + k];
    __shared__ unsigned long long red[3][1024 / WAVE];
    for (int k = 0; k < 3; ++k) {
        unsigned long long v = s[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
        if ((threadIdx.x & (WAVE - 1)) == 0) red[k][threadIdx.x / WAVE] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) t += red[threadIdx.x][w];
        counters[C_ENTRIES + threadIdx.x] += t;
    }
}

// Exact conversion of the queued values (rare: ties, exact binary fractions
// outside the fast path, long digit strings near a rounding boundary).
__global__ __launch_bounds__(64) void k_mm_exact(const uint8_t* __restrict__ t, int64_t len,
                                                 const int64_t* __restrict__ starts, int64_t nlines, int symmetric,
                                                 const int32_t* __restrict__ amb_list,
                                                 const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                                                 double* __restrict__ vals, int32_t* __restrict__ cnt,
                                                 unsigned long long* __restrict__ counters) {
    Big big[3];
    const int64_t namb = (int64_t)counters[C_AMB];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < namb;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = amb_list[k];
        const int64_t p = starts[l];
        const int64_t e = (l + 1 < nlines) ? starts[l + 1] : len;
        GSrc s{t, e};
        Entry en;
        parse_line(s, p, false, &en);
        bool amb = false;
        const uint64_t cand = fast_bits(en.v, kPow5Dev, &amb);
        bool fail = false;
        const uint64_t b = exact_bits(s, en.vpos, cand, big, &fail);
        if (fail) {
            report_error(counters, p, E_EXACT);
            continue;
        }
        const double v = __longlong_as_double((long long)b);
        vals[l] = v;
        int f, lo;
        entry_counts(rows[l], cols[l], v != 0.0, symmetric != 0, &f, &lo);
        cnt[l] = f;
        if (f) atomicAdd(&counters[C_NNZ], (unsigned long long)f);
        if (lo) atomicAdd(&counters[C_LOWER], (unsigned long long)lo);
    }
}

__global__ void k_mm_lower_count(int64_t nlines, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                                 const double* __restrict__ vals, int pattern, int symmetric,
                                 int32_t* __restrict__ out) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const int32_t r = rows[l];
    int f = 0, lo = 0;
    if (r) entry_counts(r, cols[l], pattern || vals[l] != 0.0, symmetric != 0, &f, &lo);
    out[l] = lo;
}

// create_COO_from_MM order: entry (r, c) then, symmetric off-diagonal, (c, r).
template <typename T>
__global__ void k_mm_expand(int64_t nlines, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                            const double* __restrict__ vals, const int32_t* __restrict__ cnt,
                            const int32_t* __restrict__ pos, int cbits, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ idx, T* __restrict__ xval) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const int n = cnt[l];
    if (!n) return;
    const uint64_t r = (uint64_t)(rows[l] - 1), c = (uint64_t)(cols[l] - 1);
    const T v = (T)vals[l];
    const int64_t o = pos[l];
    keys[o] = (r << cbits) | c;
    idx[o] = (uint32_t)o;
    xval[o] = v;
    if (n == 2) {
        keys[o + 1] = (c << cbits) | r;
        idx[o + 1] = (uint32_t)(o + 1);
        xval[o + 1] = v;
    }
}

// create_LCOO_from_MM: strictly lower entries (upper ones mirrored when
// symmetric), values 1 for pattern files, then the unit diagonal.
template <typename T>
__global__ void k_mm_lower_expand(int64_t nlines, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                                  const double* __restrict__ vals, const int32_t* __restrict__ lcnt,
                                  const int32_t* __restrict__ pos, int pattern, int cbits,
                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ idx, T* __restrict__ xval) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines || !lcnt[l]) return;
    uint64_t r = (uint64_t)(rows[l] - 1), c = (uint64_t)(cols[l] - 1);
    if (r < c) {
        const uint64_t tmp = r;
        r = c;
        c = tmp;
    }
    const int64_t o = pos[l];
    keys[o] = (r << cbits) | c;
    idx[o] = (uint32_t)o;
    xval[o] = pattern ? (T)1 : (T)vals[l];
}

template <typename T>
__global__ void k_mm_diag(int64_t n, int64_t base, int cbits, uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                          T* __restrict__ xval) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[base + i] = ((uint64_t)i << cbits) | (uint64_t)i;
    idx[base + i] = (uint32_t)(base + i);
    xval[base + i] = (T)1;
}

template <typename T>
__global__ void k_mm_gather(int64_t n, int cbits, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                            const T* __restrict__ xval, int32_t* __restrict__ row, int32_t* __restrict__ col,
                            T* __restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    row[i] = (int32_t)(k >> cbits);
    col[i] = (int32_t)(k & ((1ull << cbits) - 1));
    val[i] = xval[idx[i]];
}

template <typename T>
__global__ void k_mm_zero(int64_t n, int32_t* __restrict__ row, int32_t* __restrict__ col, T* __restrict__ val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    row[i] = 0;
    col[i] = 0;
    val[i] = (T)0;
}

unsigned gs(int64_t n) { return grid_for(n, MB); }

// ---------------------------------------------------------------- header
std::vector<std::string> split_blanks(const char* s, size_t n) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < n) {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r')) ++i;
        size_t j = i;
        while (j < n && !(s[j] == ' ' || s[j] == '\t' || s[j] == '\r')) ++j;
        if (j > i) out.emplace_back(s + i, j - i);
        i = j;
    }
    return out;
}

bool parse_i64(const std::string& s, int64_t* v) {
    if (s.empty() || s.size() > 18) return false;
    int64_t x = 0;
    for (char ch : s) {
        if (ch < '0' || ch > '9') return false;
        x = x * 10 + (ch - '0');
    }
    *v = x;
    return true;
}

// 1 = complete, 0 = need more bytes, <0 = error code (message set)
int parse_header(const char* h, size_t n, bool at_end, sable_mm_info_t* info) {
    size_t p = 0;
    auto line_end = [&](size_t from) -> size_t {
        const void* nl = memchr(h + from, '\n', n - from);
        return nl ? (size_t)((const char*)nl - h) : n;
    };
    size_t e = line_end(0);
    if (e == n && !at_end) return 0;
    std::vector<std::string> tok = split_blanks(h, e);
    if (tok.size() < 5 || tok[0].compare(0, 14, "%%MatrixMarket") != 0)
        return fail(SABLE_EMALFORMED, "MatrixMarket banner expected: %%%%MatrixMarket matrix coordinate <field> <symmetry>");
    if (tok[2] != "coordinate")
        return fail(SABLE_EINVAL, "only the coordinate format is read (banner says '%s')", tok[2].c_str());
    const std::string& f = tok[3];
    info->field = f == "real" ? SABLE_MM_REAL
                : f == "integer" ? SABLE_MM_INTEGER
                : f == "complex" ? SABLE_MM_COMPLEX
                : f == "pattern" ? SABLE_MM_PATTERN
                : SABLE_MM_OTHER;
    info->symmetric = tok[4] == "symmetric";
    // comment lines, then the size line
    p = e + 1;
    for (;;) {
        if (p >= n) {
            if (!at_end) return 0;
            return fail(SABLE_EMALFORMED, "MatrixMarket size line missing");
        }
        e = line_end(p);
        if (e == n && !at_end) return 0;
        std::vector<std::string> t = split_blanks(h + p, e - p);
        if (h[p] == '%' || t.empty()) {
            p = e + 1;
            continue;
        }
        int64_t a, b, c;
        if (t.size() < 3 || !parse_i64(t[0], &a) || !parse_i64(t[1], &b) || !parse_i64(t[2], &c))
            return fail(SABLE_EMALFORMED, "MatrixMarket size line must hold three integers");
        info->nrows = a;
        info->ncols = b;
        info->nentries = c;
        info->body_offset = (int64_t)(e < n ? e + 1 : n);
        return 1;
    }
}

}  // namespace

struct sable_mm_s {
    sable_mm_info_t info;
    int64_t nlines = 0;
    void* block = nullptr;  // one allocation holding the four arrays below
    int32_t* rows = nullptr;
    int32_t* cols = nullptr;
    double* vals = nullptr;
    int32_t* cnt = nullptr;
    int64_t full_sum = 0;  // sum of cnt (before the nnz == 0 rule)
    int64_t exact_values = 0;
};

static void mm_free(sable_mm_t* m) {
    if (!m) return;
    if (m->block) (void)hipFree(m->block);
    delete m;
}

static const char* err_name(int code) {
    switch (code) {
        case E_ROW_TOKEN: return "row index is not a positive decimal integer";
        case E_COL_TOKEN: return "column index missing or not a positive decimal integer";
        case E_VAL_TOKEN: return "value missing or not a decimal number";
        case E_RANGE: return "index outside the size line's rows x columns";
        case E_EXACT: return "value has more significant digits than the exact conversion holds";
        default: return "malformed entry";
    }
}

static int mm_parse_device(sable_ctx_t* ctx, const uint8_t* t, int64_t len, sable_mm_info_t info, sable_mm_t** out) {
    hipStream_t st = ctx->stream;
    const int64_t b0 = info.body_offset;
    SABLE_CHECK(info.nrows >= 0 && info.nrows <= 0x7fffffffLL && info.ncols >= 0 && info.ncols <= 0x7fffffffLL,
                SABLE_EINVAL, "matrix dimensions exceed the int32 index range");
    SABLE_CHECK(info.nentries <= 0x7fffffffLL, SABLE_EINVAL, "more than 2^31-1 entries");
    SABLE_CHECK(!info.symmetric || info.nrows == info.ncols, SABLE_EMALFORMED, "a symmetric matrix must be square");
    sable_mm_t* m = new sable_mm_t();
    m->info = info;
    // 1. line heads.  Temporaries come from the context's stream-ordered
    // pool, so repeated parses do not pay for device allocation.
    const int64_t mis = (int64_t)((uintptr_t)t & 15);
    const int64_t a0 = b0 - ((b0 + mis) & 15);
    const int64_t ntiles = len > a0 ? (len - a0 + TILE_BYTES - 1) / TILE_BYTES : 0;
    SABLE_CHECK(ntiles < 0x7fffffffLL, SABLE_EINVAL, "text too large");
    int64_t nlines = 0;
    void* tiles = nullptr;  // [counts | offsets | total | scan tmp]
    void* work = nullptr;   // [starts | amb | counters | partials]
    auto cleanup = [&](int rc) {
        ctx_pool_free(ctx, tiles);
        ctx_pool_free(ctx, work);
        tiles = work = nullptr;
        if (rc) mm_free(m);
        return rc;
    };
    int64_t counters[C_COUNT] = {0, 0, 0, 0, -1};
    if (ntiles > 0) {
        const size_t tb = Carver::round(4 * (size_t)ntiles);
        int rc = ctx_pool_alloc(ctx, 2 * tb + 256 + scan_tmp_bytes(ntiles), &tiles);
        if (rc) return cleanup(rc);
        Carver tv(tiles);
        int32_t* tc = tv.take<int32_t>(ntiles);
        int32_t* to = tv.take<int32_t>(ntiles);
        int32_t* tot = tv.take<int32_t>(1);
        void* stmp = tv.p + tv.used;
        hipLaunchKernelGGL(k_mm_count, dim3((unsigned)ntiles), dim3(MB), 0, st, t, len, b0, a0, tc);
        SABLE_LAUNCH_CHECK();
        rc = scan_exclusive_i32(st, tc, to, ntiles, tot, stmp);
        if (rc) return cleanup(rc);
        SABLE_HIP(hipMemcpyAsync(ctx->pinned, tot, 4, hipMemcpyDeviceToHost, st));
        SABLE_HIP(hipStreamSynchronize(st));
        int32_t total = 0;
        memcpy(&total, ctx->pinned, 4);
        nlines = total;
    }
    m->nlines = nlines;
    const size_t L = (size_t)(nlines ? nlines : 1);
    // rows | cols | cnt | vals in one allocation
    if (hipMalloc(&m->block, 20 * L + 64) != hipSuccess) {
        (void)hipGetLastError();
        m->block = nullptr;
        return cleanup(fail(SABLE_ENOMEM, "allocation for %lld parsed lines failed", (long long)nlines));
    }
    m->vals = (double*)m->block;
    m->rows = (int32_t*)((char*)m->block + 8 * L);
    m->cols = m->rows + L;
    m->cnt = m->cols + L;
    if (nlines > 0) {
        // 2. starts, then the parse
        const int64_t nblk = (nlines + MB - 1) / MB;
        const size_t need = Carver::round(8 * (size_t)nlines) + Carver::round(4 * (size_t)nlines) + 256 +
                            Carver::round(12 * (size_t)nblk);
        int rc = ctx_pool_alloc(ctx, need, &work);
        if (rc) return cleanup(rc);
        Carver cv(work);
        int64_t* starts = cv.take<int64_t>(nlines);
        int32_t* amb = cv.take<int32_t>(nlines);
        unsigned long long* ctr = cv.take<unsigned long long>(C_COUNT);
        int32_t* partials = cv.take<int32_t>(3 * nblk);
        Carver tv(tiles);
        (void)tv.take<int32_t>(ntiles);
        const int32_t* to = tv.take<int32_t>(ntiles);
        hipLaunchKernelGGL(k_mm_starts, dim3((unsigned)ntiles), dim3(MB), 0, st, t, len, b0, a0, to, starts);
        SABLE_LAUNCH_CHECK();
        const unsigned long long init[C_COUNT] = {0, 0, 0, 0, ~0ull};
        memcpy(ctx->pinned, init, sizeof(init));
        SABLE_HIP(hipMemcpyAsync(ctr, ctx->pinned, sizeof(init), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_mm_parse, dim3(gs(nlines)), dim3(MB), 0, st, t, len, starts, nlines,
                           (int)(info.field == SABLE_MM_PATTERN), (int)info.symmetric, info.nrows, info.ncols, m->rows,
                           m->cols, m->vals, m->cnt, amb, ctr, partials);
        SABLE_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_mm_reduce, dim3(1), dim3(1024), 0, st, partials, nblk, ctr);
        SABLE_LAUNCH_CHECK();
        rc = read_device_i64(ctx, ctr, C_COUNT, counters);
        if (rc) return cleanup(rc);
        if (counters[C_AMB] > 0 && counters[C_ERR] == -1) {
            // 3. values the fast conversion left open
            const unsigned g = (unsigned)std::min<int64_t>((counters[C_AMB] + 63) / 64, 1024);
            hipLaunchKernelGGL(k_mm_exact, dim3(g), dim3(64), 0, st, t, len, starts, nlines, (int)info.symmetric, amb,
                               m->rows, m->cols, m->vals, m->cnt, ctr);
            SABLE_LAUNCH_CHECK();
            rc = read_device_i64(ctx, ctr, C_COUNT, counters);
            if (rc) return cleanup(rc);
        }
    }
    cleanup(SABLE_OK);
    const uint64_t err = (uint64_t)counters[C_ERR];
    if (err != ~0ull) {
        const int64_t pos = (int64_t)(err >> 8);
        const int code = -(int)(err & 0xff);
        char line[81] = {0};
        const int64_t n = std::min<int64_t>(80, len - pos);
        if (n > 0) {
            SABLE_HIP(hipMemcpyAsync(ctx->pinned, t + pos, (size_t)n, hipMemcpyDeviceToHost, st));
            SABLE_HIP(hipStreamSynchronize(st));
            memcpy(line, ctx->pinned, (size_t)n);
            for (int i = 0; i < n; ++i)
                if (line[i] == '\n' || line[i] == '\r') line[i] = 0;
        }
        mm_free(m);
        return fail(SABLE_EMALFORMED, "MatrixMarket entry at byte %lld: %s: '%s'", (long long)pos, err_name(code), line);
    }
    if (counters[C_ENTRIES] != info.nentries) {
        mm_free(m);
        return fail(SABLE_EMALFORMED, "the file holds %lld entry lines but its size line says %lld",
                    (long long)counters[C_ENTRIES], (long long)info.nentries);
    }
    m->full_sum = counters[C_NNZ];
    m->exact_values = counters[C_AMB];
    m->info.nnz = counters[C_NNZ] ? counters[C_NNZ] : info.nentries;
    m->info.lower_nnz = info.nrows + counters[C_LOWER];
    *out = m;
    return SABLE_OK;
}

SABLE_API int sable_mm_header(const char* text, size_t len, sable_mm_info_t* info) {
    SABLE_CHECK(text && info, SABLE_EINVAL, "NULL argument");
    memset(info, 0, sizeof(*info));
    const int r = parse_header(text, len, true, info);
    return r < 0 ? r : SABLE_OK;
}

SABLE_API int sable_mm_parse(sable_ctx_t* ctx, const char* text, size_t len, int text_on_device, sable_mm_t** out) {
    SABLE_CHECK(ctx && text && out, SABLE_EINVAL, "NULL argument");
    *out = nullptr;
    sable_mm_info_t info;
    memset(&info, 0, sizeof(info));
    if (!text_on_device) {
        const int r = parse_header(text, len, true, &info);
        if (r < 0) return r;
    } else {
        // read the header a chunk at a time
        std::vector<char> h;
        size_t got = 0, want = 65536;
        for (;;) {
            want = std::min(want, len);
            h.resize(want);
            SABLE_HIP(hipMemcpy(h.data() + got, text + got, want - got, hipMemcpyDeviceToHost));
            got = want;
            const int r = parse_header(h.data(), got, got == len, &info);
            if (r < 0) return r;
            if (r == 1) break;
            want *= 4;
        }
    }
    const uint8_t* dt = (const uint8_t*)text;
    void* own = nullptr;
    if (!text_on_device) {
        if (hipMalloc(&own, len + 16) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SABLE_ENOMEM, "device copy of the %zu-byte text failed", len);
        }
        SABLE_HIP(hipMemcpyAsync(own, text, len, hipMemcpyHostToDevice, ctx->stream));
        dt = (const uint8_t*)own;
    }
    const int rc = mm_parse_device(ctx, dt, (int64_t)len, info, out);
    if (own) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(own);
    }
    return rc;
}

SABLE_API int sable_mm_get_info(const sable_mm_t* mm, sable_mm_info_t* info) {
    SABLE_CHECK(mm && info, SABLE_EINVAL, "NULL argument");
    *info = mm->info;
    return SABLE_OK;
}

SABLE_API int sable_mm_stats(const sable_mm_t* mm, int64_t* lines, int64_t* exact_values) {
    SABLE_CHECK(mm, SABLE_EINVAL, "NULL argument");
    if (lines) *lines = mm->nlines;
    if (exact_values) *exact_values = mm->exact_values;
    return SABLE_OK;
}

template <typename T>
static int mm_coo_t(sable_ctx_t* ctx, const sable_mm_t* m, int lower, int32_t* row, int32_t* col, T* val) {
    hipStream_t st = ctx->stream;
    const sable_mm_info_t& I = m->info;
    const int64_t n = lower ? I.lower_nnz : I.nnz;
    SABLE_CHECK(n <= 0x7fffffffLL, SABLE_EINVAL, "%lld entries exceed the int32 range", (long long)n);
    if (n == 0) return SABLE_OK;
    SABLE_CHECK(row && col && val, SABLE_EINVAL, "NULL output array");
    if (!lower && m->full_sum == 0) {
        // every entry was an explicit zero: nnz falls back to nentries and
        // the entries are left zero
        hipLaunchKernelGGL(k_mm_zero<T>, dim3(gs(n)), dim3(MB), 0, st, n, row, col, val);
        SABLE_LAUNCH_CHECK();
        return SABLE_OK;
    }
    const int64_t L = m->nlines;
    const int cbits = bits_for(std::max<int64_t>(std::max(I.nrows, I.ncols) - 1, 0));
    const size_t need = 2 * Carver::round(4 * (size_t)(L + 1)) + Carver::round(8 * (size_t)n) +
                        Carver::round(4 * (size_t)n) + Carver::round(sizeof(T) * (size_t)n) + 256 +
                        std::max(sort_tmp_bytes(n), scan_tmp_bytes(L));
    void* s;
    int rc = ctx_scratch(ctx, need, &s);
    if (rc) return rc;
    Carver cv(s);
    int32_t* pos = cv.take<int32_t>(L + 1);
    int32_t* lcnt = cv.take<int32_t>(L + 1);
    uint64_t* keys = cv.take<uint64_t>(n);
    uint32_t* idx = cv.take<uint32_t>(n);
    T* xval = cv.take<T>(n);
    void* tmp = cv.p + cv.used;
    const int pattern = I.field == SABLE_MM_PATTERN;
    if (L > 0) {
        if (!lower) {
            rc = scan_exclusive_i32(st, m->cnt, pos, L, nullptr, tmp);
            if (rc) return rc;
            hipLaunchKernelGGL(k_mm_expand<T>, dim3(gs(L)), dim3(MB), 0, st, L, m->rows, m->cols, m->vals, m->cnt,
                               pos, cbits, keys, idx, xval);
        } else {
            hipLaunchKernelGGL(k_mm_lower_count, dim3(gs(L)), dim3(MB), 0, st, L, m->rows, m->cols, m->vals, pattern,
                               (int)I.symmetric, lcnt);
            SABLE_LAUNCH_CHECK();
            rc = scan_exclusive_i32(st, lcnt, pos, L, nullptr, tmp);
            if (rc) return rc;
            hipLaunchKernelGGL(k_mm_lower_expand<T>, dim3(gs(L)), dim3(MB), 0, st, L, m->rows, m->cols, m->vals, lcnt,
                               pos, pattern, cbits, keys, idx, xval);
        }
        SABLE_LAUNCH_CHECK();
    }
    if (lower) {
        hipLaunchKernelGGL(k_mm_diag<T>, dim3(gs(I.nrows)), dim3(MB), 0, st, I.nrows, n - I.nrows, cbits, keys, idx,
                           xval);
        SABLE_LAUNCH_CHECK();
    }
    rc = sort_pairs_u64(st, keys, idx, n, 2 * cbits, tmp);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mm_gather<T>, dim3(gs(n)), dim3(MB), 0, st, n, cbits, keys, idx, xval, row, col, val);
    SABLE_LAUNCH_CHECK();
    return SABLE_OK;
}

SABLE_API int sable_mm_coo(sable_ctx_t* ctx, const sable_mm_t* mm, sable_dtype_t dtype, int lower, int32_t* row,
                           int32_t* col, void* val) {
    SABLE_CHECK(ctx && mm, SABLE_EINVAL, "NULL argument");
    if (dtype == SABLE_F32) return mm_coo_t<float>(ctx, mm, lower, row, col, (float*)val);
    if (dtype == SABLE_F64) return mm_coo_t<double>(ctx, mm, lower, row, col, (double*)val);
    return fail(SABLE_EINVAL, "bad dtype");
}

SABLE_API void sable_mm_destroy(sable_mm_t* mm) { mm_free(mm); }
