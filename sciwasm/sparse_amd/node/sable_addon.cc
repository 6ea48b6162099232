// sable_addon.cc -- Node-API addon that stands in for the reference's
// WebAssembly modules over libsablesparse_hip.
//
// The reference's JS layer (seq/src/sparse32.js, sparse64.js) talks to
//   * a linear memory (Module['wasmMemory'].buffer) that it views through
//     typed arrays at byte offsets,
//   * malloc_instance.exports._malloc/_free returning such offsets, and
//   * sparse_instance.exports.<kernel>(offsets..., sizes...).
// This addon provides the same three things: `wasmMemory.buffer` is an
// ArrayBuffer over host memory owned by the addon, `malloc_exports`
// allocates inside it, and `sparse_opt_32` / `sparse_opt_64` (and the
// parallel tree's modules) carry every hot-path export with the reference's
// names and argument orders.
//
// Matrices stay resident on the device between calls:
//  * every array an export only reads is mirrored to HBM once, keyed by
//    (offset, length); the mirror stays valid while its bytes are unchanged;
//  * changes are found by write tracking, not by hashing: the whole pages of
//    a mirrored array are made read-only, and the first write to such a page
//    (from JS through a typed array, or from any thread) takes a fault that
//    the addon's SIGSEGV handler turns into "page written": it counts the
//    write on the page and makes the page writable again.  A call re-checks a
//    mirror in O(1) when no page anywhere was written since its last check,
//    else by summing the write counts of its pages; the partial pages at the
//    two ends of an array are hashed (at most 8 KiB);
//  * a CSR or COO matrix keeps an autotuned product schedule
//    (sable_csr_auto_*) built on its first product and rebuilt only when one
//    of its arrays changes;
//  * x and y move per call through two grow-only device vectors and the
//    context's pinned staging buffers; nothing is allocated per product.
// SABLE_WRITE_TRACKING=0 replaces the tracking by a full hash of each
// mirrored array on every call; SABLE_LINEAR_PINNED=1 places the linear
// memory in pinned memory (which is then never write-protected).
#include <node_api.h>

#include <signal.h>
#include <sys/mman.h>
#include <unistd.h>

#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../../include/sable_sparse.h"
#include "write_track.h"

namespace {

// ------------------------------------------------------------ linear memory
struct Linear {
    uint8_t* base = nullptr;
    size_t size = 0;
    bool pinned = false;
    bool mapped = false;
    std::map<size_t, size_t> used;   // offset -> bytes
    std::map<size_t, size_t> freeb;  // offset -> bytes
};
Linear g_mem;
sable_ctx_t* g_ctx = nullptr;

// ------------------------------------------------------------ write tracking
// Page-granular write tracking of the linear memory (write_track.h).
using sable_track::full_pages;
using sable_track::protect_pages;
using sable_track::unprotect_range;
using sable_track::wsum;

bool init_linear(size_t bytes) {
    bytes = (bytes + 65535) & ~(size_t)65535;
    const char* pin = getenv("SABLE_LINEAR_PINNED");
    void* p = nullptr;
    if (pin && atoi(pin) == 1 && sable_host_alloc(bytes, &p) == SABLE_OK) {
        g_mem.pinned = true;
        memset(p, 0, bytes);
    } else {
        p = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (p == MAP_FAILED) return false;
        g_mem.mapped = true;  // anonymous memory reads as zero
    }
    g_mem.base = (uint8_t*)p;
    g_mem.size = bytes;
    // offset 0 stays reserved so that 0 can mean "no allocation"
    g_mem.freeb[64] = bytes - 64;
    const char* wt = getenv("SABLE_WRITE_TRACKING");
    if (g_mem.mapped && !(wt && atoi(wt) == 0)) sable_track::init(g_mem.base, g_mem.size);
    return true;
}

size_t lm_malloc(size_t bytes) {
    bytes = (bytes + 63) & ~(size_t)63;
    if (!bytes) bytes = 64;
    for (auto it = g_mem.freeb.begin(); it != g_mem.freeb.end(); ++it) {
        if (it->second >= bytes) {
            size_t off = it->first, rest = it->second - bytes;
            g_mem.freeb.erase(it);
            if (rest) g_mem.freeb[off + bytes] = rest;
            g_mem.used[off] = bytes;
            return off;
        }
    }
    return 0;
}

void lm_free(size_t off) {
    auto it = g_mem.used.find(off);
    if (it == g_mem.used.end()) return;
    size_t bytes = it->second;
    g_mem.used.erase(it);
    auto nx = g_mem.freeb.lower_bound(off);
    if (nx != g_mem.freeb.end() && off + bytes == nx->first) {
        bytes += nx->second;
        g_mem.freeb.erase(nx);
    }
    auto pv = g_mem.freeb.lower_bound(off);
    if (pv != g_mem.freeb.begin()) {
        --pv;
        if (pv->first + pv->second == off) {
            pv->second += bytes;
            return;
        }
    }
    g_mem.freeb[off] = bytes;
}

// ------------------------------------------------------------ device mirrors
uint64_t hash_bytes(const uint8_t* p, size_t n) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ n;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, p + i, 8);
        h = (h ^ w) * 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
    }
    for (; i < n; i++) h = (h ^ p[i]) * 0x100000001B3ull;
    return h;
}

struct Err {
    std::string msg;
};

void check(int rc, const char* what) {
    if (rc != SABLE_OK) throw Err{std::string(what) + ": " + sable_last_error()};
}

void ensure_ctx() {
    if (!g_ctx) check(sable_ctx_create(0, nullptr, &g_ctx), "sable_ctx_create");
}

void check_range(size_t off, size_t bytes) {
    if (off > g_mem.size || bytes > g_mem.size - off) throw Err{"offset outside the linear memory"};
}

void upload(void* d, size_t off, size_t bytes) {
    if (!bytes) return;
    if (g_mem.pinned)
        check(sable_memcpy_h2d(g_ctx, d, g_mem.base + off, bytes), "upload");
    else
        check(sable_memcpy_h2d_staged(g_ctx, d, g_mem.base + off, bytes), "upload");
}

void download_raw(size_t off, const void* d, size_t bytes, bool count = true) {
    check_range(off, bytes);
    if (!bytes) return;
    unprotect_range(off, bytes, count);
    if (g_mem.pinned)
        check(sable_memcpy_d2h(g_ctx, g_mem.base + off, d, bytes), "download");
    else
        check(sable_memcpy_d2h_staged(g_ctx, g_mem.base + off, d, bytes), "download");
}

// A device copy of a region that exports only read.  `version` changes on
// every (re)upload, so that plans built over a mirror know when to rebuild.
struct Mirror {
    void* d = nullptr;
    size_t off = 0, bytes = 0;
    uint64_t version = 0;
    bool tracked = false;
    size_t p0 = 0, p1 = 0;   // whole pages, write-protected
    uint64_t sum = 0;        // their write counts at upload
    uint64_t seen = 0;       // g_faults when last found valid
    uint64_t hash = 0;       // partial pages (tracked) or the whole region
};
std::map<std::pair<size_t, size_t>, std::unique_ptr<Mirror>> g_cache;
uint64_t g_version = 0;

uint64_t edge_hash(const Mirror& m) {
    const size_t a = m.p0 << sable_track::g_page_shift, b = m.p1 << sable_track::g_page_shift;
    if (a >= b) return hash_bytes(g_mem.base + m.off, m.bytes);
    return hash_bytes(g_mem.base + m.off, a - m.off) * 31u + hash_bytes(g_mem.base + b, m.off + m.bytes - b);
}

bool mirror_valid(Mirror& m) {
    if (!m.tracked) return hash_bytes(g_mem.base + m.off, m.bytes) == m.hash;
    if (edge_hash(m) != m.hash) return false;
    const uint64_t f = sable_track::faults();
    if (f == m.seen) return true;
    if (wsum(m.p0, m.p1) != m.sum) return false;
    m.seen = f;
    return true;
}

uint64_t g_uploads = 0, g_upload_bytes = 0;  // mirror uploads of matrix arrays (statistics)
uint64_t g_vec_uploads = 0;                   // ... of product vectors x and y
int g_last_sched = -1;                         // SABLE_SCHED_* of the last plan built
double g_last_sched_ms = 0;

void fill_mirror(Mirror& m, bool vector = false) {
    // order: counts first, then protection, then the copy, so that a write
    // racing with the copy is always seen by the next check
    if (vector) {
        g_vec_uploads++;
    } else {
        g_uploads++;
        g_upload_bytes += m.bytes;
    }
    m.version = ++g_version;
    m.seen = sable_track::faults();
    if (m.tracked) {
        m.sum = wsum(m.p0, m.p1);
        protect_pages(m.p0, m.p1);
    }
    m.hash = m.tracked ? edge_hash(m) : hash_bytes(g_mem.base + m.off, m.bytes);
    upload(m.d, m.off, m.bytes);
}

Mirror& mirror_rec(size_t off, size_t bytes, bool vector = false) {
    check_range(off, bytes);
    auto& slot = g_cache[{off, bytes}];
    if (slot && slot->d && mirror_valid(*slot)) return *slot;
    if (!slot) {
        slot.reset(new Mirror());
        slot->off = off;
        slot->bytes = bytes;
        // arrays of 16 KiB or more are write-tracked; smaller ones hashed
        slot->tracked = sable_track::g_on && bytes >= 4 * sable_track::g_page;
        if (slot->tracked) full_pages(off, bytes, &slot->p0, &slot->p1);
        check(sable_device_alloc(g_ctx, bytes ? bytes : 64, &slot->d), "device alloc");
    }
    fill_mirror(*slot, vector);
    return *slot;
}

void* mirror(size_t off, size_t bytes) { return mirror_rec(off, bytes).d; }

// The mirror's device bytes were just made equal to the host's (written on
// both sides by the export): it is valid again, as a new version.
void revalidate(Mirror& m) {
    m.version = ++g_version;
    m.seen = sable_track::faults();
    if (m.tracked) {
        m.sum = wsum(m.p0, m.p1);
        protect_pages(m.p0, m.p1);
    }
    m.hash = m.tracked ? edge_hash(m) : hash_bytes(g_mem.base + m.off, m.bytes);
}

// After an export wrote `bytes` at `off` from device memory d (the bytes
// are now the same on both sides), d becomes that region's mirror, so
// that a later export reading the result does not upload it again.
void adopt(size_t off, size_t bytes, void* d) {
    auto& slot = g_cache[{off, bytes}];
    if (slot && slot->d) check(sable_device_free(g_ctx, slot->d), "device free");
    slot.reset(new Mirror());
    Mirror& m = *slot;
    m.off = off;
    m.bytes = bytes;
    m.d = d;
    m.tracked = sable_track::g_on && bytes >= 4 * sable_track::g_page;
    if (m.tracked) full_pages(off, bytes, &m.p0, &m.p1);
    revalidate(m);
}

// Fresh device buffers for one call's outputs and per-call inputs.
std::vector<std::pair<void*, size_t>> g_scratch;

void* scratch(size_t off, size_t bytes, bool up) {
    check_range(off, bytes);
    void* d = nullptr;
    check(sable_device_alloc(g_ctx, bytes ? bytes : 64, &d), "device alloc");
    g_scratch.push_back({d, bytes});
    if (up) upload(d, off, bytes);
    return d;
}

// Device bytes to the linear memory.  A whole per-call output buffer
// (from scratch) is kept as the mirror of the region it was written to.
void download(size_t off, const void* d, size_t bytes) {
    download_raw(off, d, bytes);
    for (auto it = g_scratch.begin(); it != g_scratch.end(); ++it)
        if (it->first == d && it->second == bytes) {
            void* keep = it->first;
            g_scratch.erase(it);
            adopt(off, bytes, keep);
            break;
        }
}

void release_scratch() {
    for (auto& s : g_scratch) sable_device_free(g_ctx, s.first);
    g_scratch.clear();
}

// Two grow-only device vectors for x and y of the products.
struct VecPool {
    void* d[2] = {nullptr, nullptr};
    size_t bytes[2] = {0, 0};
} g_vec;

void* vec(int k, size_t bytes) {
    if (bytes > g_vec.bytes[k]) {
        if (g_vec.d[k]) check(sable_device_free(g_ctx, g_vec.d[k]), "device free");
        g_vec.d[k] = nullptr;
        g_vec.bytes[k] = 0;
        const size_t want = bytes + bytes / 4 + 256;
        check(sable_device_alloc(g_ctx, want, &g_vec.d[k]), "device alloc");
        g_vec.bytes[k] = want;
    }
    return g_vec.d[k];
}

void* vec_in(int k, size_t off, size_t bytes) {
    check_range(off, bytes);
    void* d = vec(k, bytes);
    upload(d, off, bytes);
    return d;
}

// ------------------------------------------------------------ product vectors
// x and y of the whole-matrix products.  A vector the host left unchanged
// since the last call need not cross PCIe again, so x and y are kept as
// mirrors (write-tracked like the matrix arrays); y is updated on the
// device in place, downloaded, and stays the valid mirror of its region.
// A vector found written by the host since the call that mirrored it goes
// to the pool vectors for the next kStreamCalls calls (uploaded on every
// call, its pages left writable), so that a caller that rewrites x or y
// before each call pays one round of page faults, not one per call.  When y
// overlaps x or one of the matrix arrays, both take the pool vectors.
// SABLE_RESIDENT_VECTORS=0 keeps x and y out of the mirrors altogether
// (their pages are then never write-protected, for callers that fill them
// with system calls such as fs.readSync).
constexpr int kStreamCalls = 64;

bool resident_vectors() {
    static const bool on = [] {
        const char* e = getenv("SABLE_RESIDENT_VECTORS");
        return !(e && atoi(e) == 0);
    }();
    return on;
}
std::map<std::pair<size_t, size_t>, int> g_stream;  // calls left in pool mode

bool overlaps(size_t a, size_t an, size_t b, size_t bn) { return an && bn && a < b + bn && b < a + an; }

Mirror* product_mirror(size_t off, size_t bytes) {
    check_range(off, bytes);
    const auto key = std::make_pair(off, bytes);
    auto st = g_stream.find(key);
    if (st != g_stream.end()) {
        if (--st->second > 0) return nullptr;
        g_stream.erase(st);
    }
    auto it = g_cache.find(key);
    if (it != g_cache.end() && it->second->d && !mirror_valid(*it->second)) {
        g_stream[key] = kStreamCalls;
        return nullptr;
    }
    return &mirror_rec(off, bytes, true);
}

// A solve's right-hand side: mirrored like x of the products (uploaded
// only when the host wrote it since the last call), unless it overlaps the
// region the solution is written to.
void* rhs_in(size_t b, size_t bb, size_t y, size_t yb) {
    Mirror* m = (!resident_vectors() || overlaps(b, bb, y, yb)) ? nullptr : product_mirror(b, bb);
    return m ? m->d : vec_in(0, b, bb);
}

struct ProductXY {
    void* dx = nullptr;
    void* dy = nullptr;
    Mirror* my = nullptr;  // y's mirror (null: pool vector)
    size_t y = 0, yb = 0;
};

ProductXY xy_in(size_t x, size_t xb, size_t y, size_t yb, std::initializer_list<std::pair<size_t, size_t>> mats) {
    bool alias = !resident_vectors() || overlaps(x, xb, y, yb);
    for (const auto& m : mats) alias = alias || overlaps(m.first, m.second, y, yb);
    ProductXY r;
    r.y = y;
    r.yb = yb;
    Mirror* mx = alias ? nullptr : product_mirror(x, xb);
    r.dx = mx ? mx->d : vec_in(0, x, xb);
    r.my = alias ? nullptr : product_mirror(y, yb);
    r.dy = r.my ? r.my->d : vec_in(1, y, yb);
    return r;
}

// Runs the product `f` (which writes y on the device) and brings y back.
template <typename F>
void xy_run(ProductXY& r, F&& f) {
    try {
        f();
    } catch (...) {
        // y's device copy may be part-written: no longer a mirror of the host
        if (r.my) r.my->hash = ~r.my->hash;
        throw;
    }
    if (!r.my) {
        download(r.y, r.dy, r.yb);
        return;
    }
    bool shared = false;  // another mirror covers some of y's pages
    for (const auto& e : g_cache)
        if (e.second.get() != r.my && overlaps(e.first.first, e.first.second, r.y, r.yb)) shared = true;
    download_raw(r.y, r.dy, r.yb, shared);
    revalidate(*r.my);
}

const int32_t* host_i32(size_t off, size_t count) {
    check_range(off, count * 4);
    return (const int32_t*)(g_mem.base + off);
}

struct ResidentBase {
    virtual ~ResidentBase() {}
};
std::map<std::tuple<int, size_t, size_t, size_t, int64_t>, std::unique_ptr<ResidentBase>> g_resident;

void drop_mirrors_in(size_t off, size_t bytes) {
    // outputs written into the linear memory invalidate overlapping mirrors
    for (auto it = g_cache.begin(); it != g_cache.end();) {
        const size_t a = it->first.first, b = a + it->first.second;
        if (a < off + bytes && off < b) {
            sable_device_free(g_ctx, it->second->d);
            it = g_cache.erase(it);
        } else {
            ++it;
        }
    }
}

// Resident matrices whose arrays lie in [off, off + bytes) (freed memory).
void drop_resident_in(size_t off, size_t bytes) {
    for (auto it = g_resident.begin(); it != g_resident.end();) {
        const size_t a = std::get<1>(it->first), b = std::get<2>(it->first), c = std::get<3>(it->first);
        auto in = [&](size_t v) { return v >= off && v < off + bytes; };
        if (in(a) || in(b) || in(c)) it = g_resident.erase(it);
        else ++it;
    }
}

// ------------------------------------------------------------ resident matrices
// A CSR matrix (or a CSR built from a COO) with its autotuned schedule.
struct ResidentCsr : ResidentBase {
    sable_csr_auto_t* plan = nullptr;
    uint64_t v[3] = {0, 0, 0};  // mirror versions the plan was built on
    int32_t ncols = 0;
    // COO input: the sorted CSR owned here, or the COO plan's row heads
    sable_coo_plan_t* coo = nullptr;
    void* own[3] = {nullptr, nullptr, nullptr};
    int32_t N = 0;
    ~ResidentCsr() override { clear(); }
    void clear() {
        if (plan) sable_csr_auto_destroy(plan);
        if (coo) sable_coo_plan_destroy(coo);
        for (void*& p : own)
            if (p) {
                sable_device_free(g_ctx, p);
                p = nullptr;
            }
        plan = nullptr;
        coo = nullptr;
    }
};

int auto_flags() {
    const char* e = getenv("SABLE_DETERMINISTIC");
    return (e && atoi(e) == 1) ? SABLE_AUTO_DETERMINISTIC : 0;
}

template <typename R>
R& resident(int kind, size_t a, size_t b, size_t c, int64_t n) {
    auto& slot = g_resident[std::make_tuple(kind, a, b, c, n)];
    if (!slot) slot.reset(new R());
    return *static_cast<R*>(slot.get());
}

// ------------------------------------------------------------ N-API helpers
// Arguments of one call; `base` skips the parallel tree's leading worker id
// so that one implementation serves both signatures.
struct Args {
    napi_env env;
    std::vector<napi_value> v;
    size_t base;
    int64_t i(size_t k) const {
        int64_t r = 0;
        k += base;
        if (k >= v.size() || napi_get_value_int64(env, v[k], &r) != napi_ok) throw Err{"expected a number argument"};
        return r;
    }
    int64_t id() const { return base ? Args{env, v, 0}.i(0) : 0; }
    double d(size_t k) const {
        double r = 0;
        k += base;
        if (k >= v.size() || napi_get_value_double(env, v[k], &r) != napi_ok) throw Err{"expected a number argument"};
        return r;
    }
};

// Module flavour, carried as the callback data of every export.
struct Prec {
    sable_dtype_t dt;
    size_t vs;
    bool stride_n_minus_min;  // seq/src/sparse64.js:1066-1092 num_diags
    size_t lead;              // 1 for the parallel tree (leading worker id)
};

Prec* prec_of(napi_env env, napi_callback_info info) {
    void* data = nullptr;
    napi_get_cb_info(env, info, nullptr, nullptr, nullptr, &data);
    return (Prec*)data;
}

Args get_args(napi_env env, napi_callback_info info, size_t n) {
    Prec* P = prec_of(env, info);
    const size_t lead = P ? P->lead : 0;
    Args a{env, std::vector<napi_value>(n + lead), lead};
    size_t argc = n + lead;
    napi_get_cb_info(env, info, &argc, a.v.data(), nullptr, nullptr);
    if (argc < n + lead) throw Err{"too few arguments"};
    a.v.resize(argc);
    return a;
}

napi_value undef(napi_env env) {
    napi_value u;
    napi_get_undefined(env, &u);
    return u;
}

napi_value num(napi_env env, double d) {
    napi_value r;
    napi_create_double(env, d, &r);
    return r;
}

// Every export runs under one process-wide lock: the parallel tree loads
// the addon in several worker threads, which share the linear memory, the
// device mirrors and the HIP context.
std::mutex g_lock;

template <typename F>
napi_value guarded(napi_env env, F&& f) {
    std::lock_guard<std::mutex> hold(g_lock);
    try {
        return f();
    } catch (const Err& e) {
        release_scratch();
        napi_throw_error(env, nullptr, e.msg.c_str());
        return nullptr;
    }
}

// ------------------------------------------------------------ kernel exports
Prec g_p32{SABLE_F32, 4, false, 0};
Prec g_p64{SABLE_F64, 8, true, 0};
Prec g_par32{SABLE_F32, 4, false, 1};
Prec g_par64{SABLE_F64, 8, false, 1};

// The resident CSR (row_ptr, col, val) of N rows with its schedule; row_ptr
// may be a slice: col/val are mirrored from the first entry it references
// and the device base pointers shifted back, so that the absolute row_ptr
// values index them.
ResidentCsr& csr_resident(Prec* P, size_t rp, size_t col, size_t val, int32_t N) {
    const int32_t* hrp = host_i32(rp, (size_t)N + 1);
    const int64_t j0 = hrp[0], j1 = hrp[N];
    if (j0 < 0 || j1 < j0) throw Err{"row_ptr is not monotone"};
    Mirror& mr = mirror_rec(rp, 4 * ((size_t)N + 1));
    Mirror& mc = mirror_rec(col + 4 * j0, 4 * (j1 - j0));
    Mirror& mv = mirror_rec(val + P->vs * j0, P->vs * (j1 - j0));
    ResidentCsr& R = resident<ResidentCsr>((int)P->dt, rp, col, val, N);
    if (!R.plan || R.v[0] != mr.version || R.v[1] != mc.version || R.v[2] != mv.version) {
        R.clear();
        const int32_t* dc = (const int32_t*)((char*)mc.d - 4 * j0);
        const void* dv = (const char*)mv.d - P->vs * j0;
        check(sable_csr_auto_create(g_ctx, P->dt, N, 0, (const int32_t*)mr.d, dc, dv, auto_flags(), &R.plan),
              "CSR schedule");
        sable_csr_auto_info_t inf;
        check(sable_csr_auto_info(R.plan, &inf), "CSR schedule");
        R.ncols = inf.ncols;
        g_last_sched = inf.schedule;
        g_last_sched_ms = inf.ms;
        R.N = N;
        R.v[0] = mr.version;
        R.v[1] = mc.version;
        R.v[2] = mv.version;
    }
    return R;
}

// The resident COO (row, col, val) of nnz entries: row-sorted input keeps
// its row heads (the COO's own col/val form the CSR), other input is sorted
// once on the device into a CSR owned here (duplicates kept, so every
// entry's product is still added to its row).  N = largest row + 1.
ResidentCsr& coo_resident(Prec* P, size_t row, size_t col, size_t val, int64_t nnz) {
    Mirror& mr = mirror_rec(row, 4 * (size_t)nnz);
    Mirror& mc = mirror_rec(col, 4 * (size_t)nnz);
    Mirror& mv = mirror_rec(val, P->vs * (size_t)nnz);
    ResidentCsr& R = resident<ResidentCsr>(16 + (int)P->dt, row, col, val, nnz);
    if (!R.plan || R.v[0] != mr.version || R.v[1] != mc.version || R.v[2] != mv.version) {
        R.clear();
        int32_t mx = -1;
        check(sable_max_i32(g_ctx, (const int32_t*)mr.d, nnz, &mx), "spmv_coo");
        const int32_t N = mx + 1;
        if (N <= 0) throw Err{"spmv_coo: negative row index"};
        check(sable_coo_plan_create(g_ctx, N, nnz, (const int32_t*)mr.d, &R.coo), "spmv_coo");
        int sorted = 0;
        int64_t nb = 0;
        check(sable_coo_plan_info(R.coo, &sorted, &nb), "spmv_coo");
        const int32_t* rp;
        const int32_t* c = (const int32_t*)mc.d;
        const void* v = mv.d;
        if (sorted) {
            check(sable_coo_plan_row_ptr(R.coo, &rp), "spmv_coo");
        } else {
            // sorted as a square matrix of max(rows, columns) (columns may lie
            // beyond the last row); rows past the last stored one are empty,
            // so the first N + 1 row heads are the CSR of the N rows
            int32_t mc = -1;
            check(sable_max_i32(g_ctx, c, nnz, &mc), "spmv_coo");
            const int32_t Ns = mc + 1 > N ? mc + 1 : N;
            check(sable_device_alloc(g_ctx, 4 * ((size_t)Ns + 1), &R.own[0]), "device alloc");
            check(sable_device_alloc(g_ctx, 4 * (size_t)nnz, &R.own[1]), "device alloc");
            check(sable_device_alloc(g_ctx, P->vs * (size_t)nnz, &R.own[2]), "device alloc");
            check(sable_coo_csr(g_ctx, P->dt, Ns, nnz, (const int32_t*)mr.d, c, v, (int32_t*)R.own[0],
                                (int32_t*)R.own[1], R.own[2]),
                  "spmv_coo");
            rp = (const int32_t*)R.own[0];
            c = (const int32_t*)R.own[1];
            v = R.own[2];
        }
        check(sable_csr_auto_create(g_ctx, P->dt, N, 0, rp, c, v, auto_flags(), &R.plan), "COO schedule");
        sable_csr_auto_info_t inf;
        check(sable_csr_auto_info(R.plan, &inf), "COO schedule");
        R.ncols = inf.ncols;
        g_last_sched = inf.schedule;
        g_last_sched_ms = inf.ms;
        R.N = N;
        R.v[0] = mr.version;
        R.v[1] = mc.version;
        R.v[2] = mv.version;
    }
    return R;
}

// One product call: x (ncols values) and y (N values) in, y += A.x
// inner_max times on the resident schedule, y out.  `mats`: the matrix's
// regions of the linear memory.
void run_resident(Prec* P, ResidentCsr& R, size_t x, size_t y, int im,
                  std::initializer_list<std::pair<size_t, size_t>> mats) {
    ProductXY r = xy_in(x, P->vs * (size_t)R.ncols, y, P->vs * (size_t)R.N, mats);
    xy_run(r, [&] { check(sable_spmv_csr_auto(g_ctx, R.plan, r.dx, r.dy, 1.0, im), "spmv"); });
}

// spmv_coo(row, col, val, x, y, len[, inner_max])
napi_value spmv_coo_impl(napi_env env, napi_callback_info info, bool wrapper) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, wrapper ? 7 : 6);
        ensure_ctx();
        const size_t row = a.i(0), col = a.i(1), val = a.i(2), x = a.i(3), y = a.i(4);
        const int64_t nnz = a.i(5);
        const int im = wrapper ? (int)a.i(6) : 1;
        if (nnz <= 0 || im <= 0) return undef(env);
        run_resident(P, coo_resident(P, row, col, val, nnz), x, y, im,
                     {{row, 4 * (size_t)nnz}, {col, 4 * (size_t)nnz}, {val, P->vs * (size_t)nnz}});
        return undef(env);
    });
}

// spmv_csr(rowptr, col, val, x, y, N[, inner_max]); rowptr may be a slice
napi_value spmv_csr_impl(napi_env env, napi_callback_info info, bool wrapper) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, wrapper ? 7 : 6);
        ensure_ctx();
        const size_t rp = a.i(0), col = a.i(1), val = a.i(2), x = a.i(3), y = a.i(4);
        const int32_t N = (int32_t)a.i(5);
        const int im = wrapper ? (int)a.i(6) : 1;
        if (N <= 0 || im <= 0) return undef(env);
        ResidentCsr& R = csr_resident(P, rp, col, val, N);
        const size_t j1 = (size_t)host_i32(rp, (size_t)N + 1)[N];
        run_resident(P, R, x, y, im, {{rp, 4 * ((size_t)N + 1)}, {col, 4 * j1}, {val, P->vs * j1}});
        return undef(env);
    });
}

// Parallel-tree CSR variants with short-row groups
// (par/src/sparse_opt_32.wat:1791-1989 and the unroll/gs forms):
// (id, rowptr, col, val, x, y, len, n1, n2, n3, inner_max) covers n1 rows of
// one entry, n2 of two, n3 of three, then len rows of any length, all from
// rowptr on; each row adds its entries to y in order, so the product is
// the plain CSR product over the n1 + n2 + n3 + len rows.  The group sizes
// are checked against rowptr (the reference would read them as given).
napi_value spmv_csr_short_impl(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 10);
        ensure_ctx();
        const size_t rp = a.i(0), col = a.i(1), val = a.i(2), x = a.i(3), y = a.i(4);
        const int64_t len = a.i(5), n1 = a.i(6), n2 = a.i(7), n3 = a.i(8);
        const int im = (int)a.i(9);
        if (len < 0 || n1 < 0 || n2 < 0 || n3 < 0) throw Err{"negative row count"};
        const int64_t N = n1 + n2 + n3 + len;
        if (N <= 0 || im <= 0) return undef(env);
        if (N > 0x7fffffff) throw Err{"too many rows"};
        const int32_t* h = host_i32(rp, (size_t)N + 1);
        if (h[n1] - h[0] != n1 || h[n1 + n2] - h[n1] != 2 * n2 || h[n1 + n2 + n3] - h[n1 + n2] != 3 * n3)
            throw Err{"short-row group sizes do not match rowptr"};
        // the workers' calls share y: the pool vectors, as for the other slices
        ResidentCsr& R = csr_resident(P, rp, col, val, (int32_t)N);
        void* dx = vec_in(0, x, P->vs * (size_t)R.ncols);
        void* dy = vec_in(1, y, P->vs * (size_t)R.N);
        check(sable_spmv_csr_auto(g_ctx, R.plan, dx, dy, 1.0, im), "spmv");
        download(y, dy, P->vs * (size_t)R.N);
        return undef(env);
    });
}

// spmv_dia(offset, data, N, nd, stride, x, y[, inner_max])
napi_value spmv_dia_impl(napi_env env, napi_callback_info info, bool wrapper) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, wrapper ? 8 : 7);
        ensure_ctx();
        const size_t off = a.i(0), data = a.i(1), x = a.i(5), y = a.i(6);
        const int32_t N = (int32_t)a.i(2), nd = (int32_t)a.i(3), stride = (int32_t)a.i(4);
        const int im = wrapper ? (int)a.i(7) : 1;
        if (nd <= 0 || N <= 0 || im <= 0) return undef(env);
        void* doff = mirror(off, 4 * (size_t)nd);
        void* dd = mirror(data, P->vs * (size_t)nd * stride);
        ProductXY r = xy_in(x, P->vs * N, y, P->vs * N, {{off, 4 * (size_t)nd}, {data, P->vs * (size_t)nd * stride}});
        xy_run(r, [&] {
            check(sable_spmv_dia(g_ctx, P->dt, N, nd, stride, (const int32_t*)doff, dd, r.dx, r.dy, 1.0, im),
                  "spmv_dia");
        });
        return undef(env);
    });
}

// Largest index of an int32 mirror, found once per upload.
std::map<const Mirror*, std::pair<uint64_t, int32_t>> g_maxidx;
int32_t max_index(const Mirror& m, int64_t n) {
    auto& e = g_maxidx[&m];
    if (e.first != m.version) {
        int32_t mx = -1;
        check(sable_max_i32(g_ctx, (const int32_t*)m.d, n, &mx), "index scan");
        e = {m.version, mx};
    }
    return e.second;
}

// spmv_ell(indices, data, N, nc, x, y[, inner_max])
napi_value spmv_ell_impl(napi_env env, napi_callback_info info, bool wrapper) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, wrapper ? 7 : 6);
        ensure_ctx();
        const size_t idx = a.i(0), data = a.i(1), x = a.i(4), y = a.i(5);
        const int32_t N = (int32_t)a.i(2), nc = (int32_t)a.i(3);
        const int im = wrapper ? (int)a.i(6) : 1;
        if (nc <= 0 || N <= 0 || im <= 0) return undef(env);
        const size_t slots = (size_t)nc * N;
        Mirror& mi = mirror_rec(idx, 4 * slots);
        void* dd = mirror(data, P->vs * slots);
        const int32_t ncol = max_index(mi, (int64_t)slots) + 1;
        ProductXY r = xy_in(x, P->vs * (ncol > N ? ncol : N), y, P->vs * N, {{idx, 4 * slots}, {data, P->vs * slots}});
        xy_run(r, [&] {
            check(sable_spmv_ell(g_ctx, P->dt, N, nc, (const int32_t*)mi.d, dd, r.dx, r.dy, 1.0, im), "spmv_ell");
        });
        return undef(env);
    });
}

#define WRAP2(NAME, IMPL)                                                                         \
    napi_value NAME(napi_env env, napi_callback_info info) { return IMPL(env, info, false); }    \
    napi_value NAME##_wrapper(napi_env env, napi_callback_info info) { return IMPL(env, info, true); }
WRAP2(spmv_coo, spmv_coo_impl)
WRAP2(spmv_csr, spmv_csr_impl)
WRAP2(spmv_dia, spmv_dia_impl)
WRAP2(spmv_ell, spmv_ell_impl)

// transpose_csr(in_row_ptr, in_col, in_val, out_row_ptr, out_col, out_val, N, nnz)
napi_value transpose_csr(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 8);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(6);
        const int64_t nnz = a.i(7);
        if (N <= 0) return undef(env);
        void* drp = mirror(a.i(0), 4 * (size_t)(N + 1));
        void* dc = mirror(a.i(1), 4 * nnz);
        void* dv = mirror(a.i(2), P->vs * nnz);
        void* orp = scratch(a.i(3), 4 * (size_t)(N + 1), false);
        void* oc = scratch(a.i(4), 4 * nnz, false);
        void* ov = scratch(a.i(5), P->vs * nnz, false);
        check(sable_transpose_csr(g_ctx, P->dt, N, (const int32_t*)drp, (const int32_t*)dc, dv, (int32_t*)orp,
                                  (int32_t*)oc, ov),
              "transpose_csr");
        drop_mirrors_in(a.i(3), 4 * (size_t)(N + 1));
        drop_mirrors_in(a.i(4), 4 * nnz);
        drop_mirrors_in(a.i(5), P->vs * nnz);
        download(a.i(3), orp, 4 * (size_t)(N + 1));
        download(a.i(4), oc, 4 * nnz);
        download(a.i(5), ov, P->vs * nnz);
        release_scratch();
        return undef(env);
    });
}

// transpose_dia(in_offset, in_data, out_offset, out_data, ndiags, stride, N)
napi_value transpose_dia(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int32_t nd = (int32_t)a.i(4), stride = (int32_t)a.i(5), N = (int32_t)a.i(6);
        if (nd <= 0 || N <= 0) return undef(env);
        const size_t db = P->vs * (size_t)nd * stride;
        void* doff = mirror(a.i(0), 4 * (size_t)nd);
        void* dd = mirror(a.i(1), db);
        void* oo = scratch(a.i(2), 4 * (size_t)nd, false);
        void* od = scratch(a.i(3), db, false);
        check(sable_transpose_dia(g_ctx, P->dt, N, nd, stride, (const int32_t*)doff, dd, (int32_t*)oo, od),
              "transpose_dia");
        drop_mirrors_in(a.i(2), 4 * (size_t)nd);
        drop_mirrors_in(a.i(3), db);
        download(a.i(2), oo, 4 * (size_t)nd);
        download(a.i(3), od, db);
        release_scratch();
        return undef(env);
    });
}

// count_transpose_cols_ell(data, indices, nnz_row, ncols, N) -> i32
napi_value count_transpose_cols_ell(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 5);
        ensure_ctx();
        const int32_t nc = (int32_t)a.i(3), N = (int32_t)a.i(4);
        if (!(nc > 0 && N > 0)) return num(env, 0);
        const size_t slots = (size_t)nc * N;
        void* dd = mirror(a.i(0), P->vs * slots);
        void* di = mirror(a.i(1), 4 * slots);
        int32_t nco = 0;
        check(sable_count_transpose_cols_ell(g_ctx, P->dt, N, nc, dd, (const int32_t*)di, &nco),
              "count_transpose_cols_ell");
        return num(env, nco);
    });
}

// transpose_ell(in_data, in_indices, out_data, out_indices, nnz_row, ncols, N);
// the output width is the count of the preceding count_transpose_cols_ell call
napi_value transpose_ell(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int32_t nc = (int32_t)a.i(5), N = (int32_t)a.i(6);
        if (!(nc > 0 && N > 0)) return undef(env);
        const size_t slots = (size_t)nc * N;
        void* dd = mirror(a.i(0), P->vs * slots);
        void* di = mirror(a.i(1), 4 * slots);
        int32_t nco = 0;
        check(sable_count_transpose_cols_ell(g_ctx, P->dt, N, nc, dd, (const int32_t*)di, &nco), "count");
        const size_t oslots = (size_t)nco * N;
        void* od = scratch(a.i(2), P->vs * oslots, false);
        void* oi = scratch(a.i(3), 4 * oslots, false);
        check(sable_transpose_ell(g_ctx, P->dt, N, nc, nco, dd, (const int32_t*)di, od, (int32_t*)oi),
              "transpose_ell");
        drop_mirrors_in(a.i(2), P->vs * oslots);
        drop_mirrors_in(a.i(3), 4 * oslots);
        download(a.i(2), od, P->vs * oslots);
        download(a.i(3), oi, 4 * oslots);
        release_scratch();
        return undef(env);
    });
}

// memcpy(src, dst, bytes) inside the linear memory
napi_value lm_memcpy(napi_env env, napi_callback_info info) {
    return guarded(env, [&] {
        Args a = get_args(env, info, 3);
        const size_t src = a.i(0), dst = a.i(1), n = a.i(2);
        check_range(src, n);
        check_range(dst, n);
        unprotect_range(dst, n);
        memmove(g_mem.base + dst, g_mem.base + src, n);
        return undef(env);
    });
}

// ---- the three conversions the reference does in JS (now exports)
// coo_csr(row, col, val, nnz, N, out_row_ptr, out_col, out_val)
napi_value coo_csr(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 8);
        ensure_ctx();
        const int64_t nnz = a.i(3);
        const int32_t N = (int32_t)a.i(4);
        void* dr = mirror(a.i(0), 4 * nnz);
        void* dc = mirror(a.i(1), 4 * nnz);
        void* dv = mirror(a.i(2), P->vs * nnz);
        void* orp = scratch(a.i(5), 4 * (size_t)(N + 1), false);
        void* oc = scratch(a.i(6), 4 * nnz, false);
        void* ov = scratch(a.i(7), P->vs * nnz, false);
        check(sable_coo_csr(g_ctx, P->dt, N, nnz, (const int32_t*)dr, (const int32_t*)dc, dv, (int32_t*)orp,
                            (int32_t*)oc, ov),
              "coo_csr");
        drop_mirrors_in(a.i(5), 4 * (size_t)(N + 1));
        drop_mirrors_in(a.i(6), 4 * nnz);
        drop_mirrors_in(a.i(7), P->vs * nnz);
        download(a.i(5), orp, 4 * (size_t)(N + 1));
        download(a.i(6), oc, 4 * nnz);
        download(a.i(7), ov, P->vs * nnz);
        release_scratch();
        return undef(env);
    });
}

// coo_csc(row, col, val, nnz, N, out_col_ptr, out_row, out_val)
napi_value coo_csc(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 8);
        ensure_ctx();
        const int64_t nnz = a.i(3);
        const int32_t N = (int32_t)a.i(4);
        void* dr = mirror(a.i(0), 4 * nnz);
        void* dc = mirror(a.i(1), 4 * nnz);
        void* dv = mirror(a.i(2), P->vs * nnz);
        void* ocp = scratch(a.i(5), 4 * (size_t)(N + 1), false);
        void* orow = scratch(a.i(6), 4 * nnz, false);
        void* ov = scratch(a.i(7), P->vs * nnz, false);
        check(sable_coo_csc(g_ctx, P->dt, N, nnz, (const int32_t*)dr, (const int32_t*)dc, dv, (int32_t*)ocp,
                            (int32_t*)orow, ov),
              "coo_csc");
        drop_mirrors_in(a.i(5), 4 * (size_t)(N + 1));
        drop_mirrors_in(a.i(6), 4 * nnz);
        drop_mirrors_in(a.i(7), P->vs * nnz);
        download(a.i(5), ocp, 4 * (size_t)(N + 1));
        download(a.i(6), orow, 4 * nnz);
        download(a.i(7), ov, P->vs * nnz);
        release_scratch();
        return undef(env);
    });
}

// A lower-triangular matrix with its solve analysis (level schedule, the
// level-ordered copy), kept while its arrays are unchanged.  Every schedule
// (row blocks, lanes per row) adds a row's products in the reference's
// per-row order, so the solves keep the reference's bits.
struct ResidentSpts : ResidentBase {
    sable_spts_t* a = nullptr;
    uint64_t v[3] = {0, 0, 0};
    bool sf_tuned = false;  // the sync-free schedule has been timed
    ~ResidentSpts() override {
        if (a) sable_spts_destroy(a);
    }
};

// KIND_SPTS_LEVEL: the level-set exports' analysis, whose schedule is tuned
// once (sable_spts_tune: the fastest of the row-block and level-graph
// schedules, all of which give the level solve's bits); the sync-free
// export keeps its own analysis, its schedule timed once (tune_syncfree).
enum { KIND_SPTS_CSR = 32, KIND_SPTS_CSC = 48, KIND_SPTS_LEVEL = 64 };

ResidentSpts& spts_resident(Prec* P, int kind, size_t p, size_t i, size_t v, int32_t N) {
    const int32_t* h = host_i32(p, (size_t)N + 1);
    const int64_t nnz = h[N];
    if (h[0] != 0 || nnz < N) throw Err{"the pointer array must start at 0 and hold the N diagonals"};
    Mirror& mp = mirror_rec(p, 4 * ((size_t)N + 1));
    Mirror& mi = mirror_rec(i, 4 * (size_t)nnz);
    Mirror& mv = mirror_rec(v, P->vs * (size_t)nnz);
    ResidentSpts& R = resident<ResidentSpts>(kind + (int)P->dt, p, i, v, N);
    if (!R.a || R.v[0] != mp.version || R.v[1] != mi.version || R.v[2] != mv.version) {
        if (R.a) sable_spts_destroy(R.a);
        R.a = nullptr;
        R.sf_tuned = false;
        if (kind == KIND_SPTS_CSC)
            check(sable_spts_csc_analyze(g_ctx, P->dt, N, (const int32_t*)mp.d, (const int32_t*)mi.d, mv.d, 1, &R.a),
                  "spts_csc");
        else
            check(sable_spts_analyze(g_ctx, P->dt, N, (const int32_t*)mp.d, (const int32_t*)mi.d, mv.d, 1, &R.a),
                  "spts analysis");
        if (kind == KIND_SPTS_LEVEL) check(sable_spts_tune(g_ctx, R.a, nullptr, nullptr, nullptr, nullptr), "spts tune");
        R.v[0] = mp.version;
        R.v[1] = mi.version;
        R.v[2] = mv.version;
    }
    return R;
}

// The sync-free exports' schedule (spts_csr_sync_free_wrapper, spts_csc),
// timed once per resident matrix on its first call: row blocks of 2048,
// 8192 and 16384 rows (at most twice N; each size costs a rebuild of the
// block layout, so the set is kept short) and the natural-order sweep;
// every one gives the reference sync-free solve's bits.
void tune_syncfree(ResidentSpts& R, const void* db, void* dy, int32_t N) {
    struct C {
        int s;
        int32_t br;
    };
    std::vector<C> cands;
    for (int64_t br : {2048, 8192, 16384})
        if (br <= 2 * (int64_t)N || br == 2048) cands.push_back({SABLE_SPTS_BLOCKS, (int32_t)br});
    cands.push_back({SABLE_SPTS_LEVEL_GRAPH, 0});
    void *e0 = nullptr, *e1 = nullptr;
    check(sable_event_create(&e0), "event");
    if (sable_event_create(&e1) != SABLE_OK) {
        sable_event_destroy(e0);
        throw Err{"event"};
    }
    C pick = cands[0];
    float best = 0.f;
    try {
        for (const C& c : cands) {
            check(sable_spts_set_schedule(R.a, c.s, c.br), "spts schedule");
            check(sable_spts_solve_syncfree(g_ctx, R.a, db, dy), "spts schedule timing");
            check(sable_event_record(g_ctx, e0), "event");
            for (int r = 0; r < 3; r++) check(sable_spts_solve_syncfree(g_ctx, R.a, db, dy), "spts schedule timing");
            check(sable_event_record(g_ctx, e1), "event");
            check(sable_ctx_sync(g_ctx), "spts schedule timing");
            float ms = 0.f;
            check(sable_event_elapsed_ms(e0, e1, &ms), "event");
            if (&c == &cands[0] || ms < best) {
                best = ms;
                pick = c;
            }
        }
        check(sable_spts_status(g_ctx, R.a), "spts schedule timing");
        check(sable_spts_set_schedule(R.a, pick.s, pick.br), "spts schedule");
    } catch (...) {
        sable_event_destroy(e0);
        sable_event_destroy(e1);
        throw;
    }
    sable_event_destroy(e0);
    sable_event_destroy(e1);
    R.sf_tuned = true;
}

// spts_csc(col_ptr, row, val, x, y, N[, inner_max]): y = L^-1 x, L lower in
// CSC with the diagonal first in each column
napi_value spts_csc_impl(napi_env env, napi_callback_info info, bool wrapper) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, wrapper ? 7 : 6);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(5);
        const int im = wrapper ? (int)a.i(6) : 1;
        if (N <= 0 || im <= 0) return undef(env);
        ResidentSpts& R = spts_resident(P, KIND_SPTS_CSC, a.i(0), a.i(1), a.i(2), N);
        void* dx = rhs_in(a.i(3), P->vs * N, a.i(4), P->vs * N);
        void* dy = vec(1, P->vs * N);
        if (!R.sf_tuned) tune_syncfree(R, dx, dy, N);
        for (int it = 0; it < im; it++) check(sable_spts_solve_syncfree(g_ctx, R.a, dx, dy), "spts_csc");
        check(sable_spts_status(g_ctx, R.a), "spts_csc");
        download(a.i(4), dy, P->vs * N);
        return undef(env);
    });
}
napi_value spts_csc(napi_env env, napi_callback_info info) { return spts_csc_impl(env, info, false); }
napi_value spts_csc_wrapper(napi_env env, napi_callback_info info) { return spts_csc_impl(env, info, true); }

// mmread(text[, lower]) -> {row_index, col_index, val_index, N, nnz, nrows,
// ncols, nentries, field, symmetry}: the whole MatrixMarket text (Buffer,
// TypedArray, ArrayBuffer or string) parsed on the device; the COO arrays are
// allocated in the linear memory with _malloc.  lower = 1 builds the lower
// COO with a unit diagonal.  Replaces mmread / read_matrix_MM_files +
// create_COO_from_MM (seq/src/sparse32.js:1953-2172, 2498-2510) and
// create_LCOO_from_MM (:1983-2097).
napi_value mmread(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        size_t argc = 2;
        napi_value argv[2];
        napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
        if (argc < 1) throw Err{"mmread(text[, lower])"};
        int64_t lower = 0;
        if (argc > 1) napi_get_value_int64(env, argv[1], &lower);
        const char* text = nullptr;
        size_t len = 0;
        std::string copy;
        bool is = false;
        if (napi_is_buffer(env, argv[0], &is) == napi_ok && is) {
            void* d = nullptr;
            napi_get_buffer_info(env, argv[0], &d, &len);
            text = (const char*)d;
        } else if (napi_is_typedarray(env, argv[0], &is) == napi_ok && is) {
            napi_typedarray_type t;
            size_t n = 0, off = 0;
            void* d = nullptr;
            napi_value ab;
            napi_get_typedarray_info(env, argv[0], &t, &n, &d, &ab, &off);
            size_t es = (t == napi_int8_array || t == napi_uint8_array || t == napi_uint8_clamped_array) ? 1
                        : (t == napi_int16_array || t == napi_uint16_array) ? 2
                        : (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
                        : 4;
            text = (const char*)d;
            len = n * es;
        } else if (napi_is_arraybuffer(env, argv[0], &is) == napi_ok && is) {
            void* d = nullptr;
            napi_get_arraybuffer_info(env, argv[0], &d, &len);
            text = (const char*)d;
        } else {
            size_t n = 0;
            if (napi_get_value_string_utf8(env, argv[0], nullptr, 0, &n) != napi_ok)
                throw Err{"mmread: text must be a Buffer, TypedArray, ArrayBuffer or string"};
            copy.resize(n + 1);
            napi_get_value_string_utf8(env, argv[0], &copy[0], n + 1, &n);
            copy.resize(n);
            text = copy.data();
            len = n;
        }
        if (!len) throw Err{"mmread: empty text"};
        ensure_ctx();
        sable_mm_t* mm = nullptr;
        check(sable_mm_parse(g_ctx, text, len, 0, &mm), "mmread");
        sable_mm_info_t mi;
        sable_mm_get_info(mm, &mi);
        const int64_t n = lower ? mi.lower_nnz : mi.nnz;
        const size_t ro = lm_malloc(4 * (size_t)n), co = lm_malloc(4 * (size_t)n), vo = lm_malloc(P->vs * (size_t)n);
        if (n > 0 && (!ro || !co || !vo)) {
            if (ro) lm_free(ro);
            if (co) lm_free(co);
            if (vo) lm_free(vo);
            sable_mm_destroy(mm);
            throw Err{"mmread: the linear memory cannot hold the COO"};
        }
        if (n > 0) {
            void* dr = scratch(ro, 4 * (size_t)n, false);
            void* dc = scratch(co, 4 * (size_t)n, false);
            void* dv = scratch(vo, P->vs * (size_t)n, false);
            const int rc = sable_mm_coo(g_ctx, mm, P->dt, (int)lower, (int32_t*)dr, (int32_t*)dc, dv);
            if (rc != SABLE_OK) {
                sable_mm_destroy(mm);
                check(rc, "mmread");
            }
            download(ro, dr, 4 * (size_t)n);
            download(co, dc, 4 * (size_t)n);
            download(vo, dv, P->vs * (size_t)n);
            release_scratch();
        }
        sable_mm_destroy(mm);
        static const char* fields[] = {"real", "integer", "complex", "pattern", "other"};
        napi_value o, s1, s2;
        napi_create_object(env, &o);
        napi_set_named_property(env, o, "row_index", num(env, (double)ro));
        napi_set_named_property(env, o, "col_index", num(env, (double)co));
        napi_set_named_property(env, o, "val_index", num(env, (double)vo));
        napi_set_named_property(env, o, "N", num(env, (double)mi.nrows));
        napi_set_named_property(env, o, "nnz", num(env, (double)n));
        napi_set_named_property(env, o, "nrows", num(env, (double)mi.nrows));
        napi_set_named_property(env, o, "ncols", num(env, (double)mi.ncols));
        napi_set_named_property(env, o, "nentries", num(env, (double)mi.nentries));
        napi_create_string_utf8(env, fields[mi.field >= 0 && mi.field <= 4 ? mi.field : 4], NAPI_AUTO_LENGTH, &s1);
        napi_create_string_utf8(env, mi.symmetric ? "symmetric" : "general", NAPI_AUTO_LENGTH, &s2);
        napi_set_named_property(env, o, "field", s1);
        napi_set_named_property(env, o, "symmetry", s2);
        return o;
    });
}

// num_cols(row_ptr, N) -> i32
napi_value num_cols(napi_env env, napi_callback_info info) {
    return guarded(env, [&] {
        Args a = get_args(env, info, 2);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(1);
        void* drp = mirror(a.i(0), 4 * (size_t)(N + 1));
        int32_t nc = 0;
        check(sable_csr_num_cols(g_ctx, N, (const int32_t*)drp, &nc), "num_cols");
        return num(env, nc);
    });
}

// csr_ell(row_ptr, col, val, N, nc, out_indices, out_data); csr_ell_row: the
// same into the parallel tree's row-major layout
napi_value csr_ell_any(napi_env env, napi_callback_info info, bool row_major);
napi_value csr_ell(napi_env env, napi_callback_info info) { return csr_ell_any(env, info, false); }
napi_value csr_ell_row(napi_env env, napi_callback_info info) { return csr_ell_any(env, info, true); }
napi_value csr_ell_any(napi_env env, napi_callback_info info, bool row_major) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(3), nc = (int32_t)a.i(4);
        const int32_t* hrp = host_i32(a.i(0), N + 1);
        const int64_t nnz = hrp[N];
        void* drp = mirror(a.i(0), 4 * (size_t)(N + 1));
        void* dc = mirror(a.i(1), 4 * nnz);
        void* dv = mirror(a.i(2), P->vs * nnz);
        const size_t slots = (size_t)nc * N;
        void* oi = scratch(a.i(5), 4 * slots, false);
        void* od = scratch(a.i(6), P->vs * slots, false);
        check((row_major ? sable_csr_ell_rowmajor : sable_csr_ell)(g_ctx, P->dt, N, nc, (const int32_t*)drp,
                                                                  (const int32_t*)dc, dv, (int32_t*)oi, od),
              "csr_ell");
        drop_mirrors_in(a.i(5), 4 * slots);
        drop_mirrors_in(a.i(6), P->vs * slots);
        download(a.i(5), oi, 4 * slots);
        download(a.i(6), od, P->vs * slots);
        release_scratch();
        return undef(env);
    });
}

// num_diags(row_ptr, col, N) -> [nd, stride] (stride N, or N - min|offset|
// for the FP64 sequential module as seq/src/sparse64.js:1066-1092)
napi_value num_diags(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 3);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(2);
        const int32_t* hrp = host_i32(a.i(0), N + 1);
        const int64_t nnz = hrp[N];
        void* drp = mirror(a.i(0), 4 * (size_t)(N + 1));
        void* dc = mirror(a.i(1), 4 * nnz);
        int32_t nd = 0, mn = 0;
        check(sable_csr_num_diags(g_ctx, N, (const int32_t*)drp, (const int32_t*)dc, &nd, &mn), "num_diags");
        napi_value arr;
        napi_create_array_with_length(env, 2, &arr);
        napi_set_element(env, arr, 0, num(env, nd));
        napi_set_element(env, arr, 1, num(env, P->stride_n_minus_min ? N - mn : N));
        return arr;
    });
}

// csr_dia(row_ptr, col, val, N, nd, stride, out_offset, out_data); csr_dia_row:
// the parallel tree's row-major layout (stride N)
napi_value csr_dia_any(napi_env env, napi_callback_info info, bool row_major);
napi_value csr_dia(napi_env env, napi_callback_info info) { return csr_dia_any(env, info, false); }
napi_value csr_dia_row(napi_env env, napi_callback_info info) { return csr_dia_any(env, info, true); }
napi_value csr_dia_any(napi_env env, napi_callback_info info, bool row_major) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 8);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(3), nd = (int32_t)a.i(4), stride = (int32_t)a.i(5);
        const int32_t* hrp = host_i32(a.i(0), N + 1);
        const int64_t nnz = hrp[N];
        void* drp = mirror(a.i(0), 4 * (size_t)(N + 1));
        void* dc = mirror(a.i(1), 4 * nnz);
        void* dv = mirror(a.i(2), P->vs * nnz);
        const size_t db = P->vs * (size_t)nd * stride;
        void* oo = scratch(a.i(6), 4 * (size_t)(nd > 0 ? nd : 1), false);
        void* od = scratch(a.i(7), db, false);
        if (row_major && stride != N) throw Err{"row-major DIA storage has stride N"};
        check(row_major ? sable_csr_dia_rowmajor(g_ctx, P->dt, N, nd, (const int32_t*)drp, (const int32_t*)dc, dv,
                                                 (int32_t*)oo, od)
                        : sable_csr_dia(g_ctx, P->dt, N, nd, stride, (const int32_t*)drp, (const int32_t*)dc, dv,
                                        (int32_t*)oo, od),
              "csr_dia");
        drop_mirrors_in(a.i(6), 4 * (size_t)nd);
        drop_mirrors_in(a.i(7), db);
        download(a.i(6), oo, 4 * (size_t)nd);
        download(a.i(7), od, db);
        release_scratch();
        return undef(env);
    });
}

// ---- triangular solves (parallel tree; the whole solve runs on the GPU in
// the call of worker 0, the other workers return at once)

// Level-set solve on the level-ordered matrix: b and y in level order.  The
// level-ordered matrix is its own level order (rows sorted by level), so its
// analysis reproduces the caller's level_ptr and the solve is the
// reference's level-by-level arithmetic, t = 0 + v0*y0 + ...; y = (b - t)/d.
napi_value level_solve(napi_env env, Prec* P, const Args& a, size_t lp_at, size_t row_at, size_t b_at, size_t y_at,
                       int32_t nl, int32_t N, int im) {
    if (a.id() != 0 || nl <= 0 || im <= 0 || N <= 0) return undef(env);
    ensure_ctx();
    (void)host_i32(a.i(lp_at), (size_t)nl + 1);
    ResidentSpts& R = spts_resident(P, KIND_SPTS_LEVEL, a.i(row_at), a.i(row_at + 1), a.i(row_at + 2), N);
    void* db = rhs_in(a.i(b_at), P->vs * N, a.i(y_at), P->vs * N);
    void* dy = vec(1, P->vs * N);
    for (int it = 0; it < im; it++) check(sable_spts_solve_level(g_ctx, R.a, db, dy), "spts_level_csr");
    download(a.i(y_at), dy, P->vs * N);
    return undef(env);
}

// spts_level_csr(id, level_ptr, rowptr, col, val, b, y, perm, nlevels,
//                barrier, flag, nthreads, N, inner_max)
napi_value spts_level_csr(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 13);
        return level_solve(env, P, a, 0, 1, 4, 5, (int32_t)a.i(7), (int32_t)a.i(11), (int)a.i(12));
    });
}

// spts_csr_level_sync_free_wrapper(id, level_ptr, rowptr, col, val, b, y,
//     perm, nlevels, barrier, flag, array_flag, nthreads, N, inner_max):
// the same level-ordered arithmetic as spts_level_csr
napi_value spts_csr_level_sync_free_wrapper(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 14);
        return level_solve(env, P, a, 0, 1, 4, 5, (int32_t)a.i(7), (int32_t)a.i(12), (int)a.i(13));
    });
}

// spts_csr_opt_level_sync_free_wrapper(id, level_ptr, rowptr, col, val, b,
//     y, perm, nlevels, barrier, flag, global_level, global_rows,
//     array_level, row_level, row_worker, worker_level, nthreads, N,
//     inner_max): combined level + sync-free solve, level-ordered arithmetic
napi_value spts_csr_opt_level_sync_free_wrapper(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 19);
        return level_solve(env, P, a, 0, 1, 4, 5, (int32_t)a.i(7), (int32_t)a.i(17), (int)a.i(18));
    });
}

// metadata(id, level_ptr, nlevels, nthreads, N, row_level, row_worker): the
// per-worker ownership tables of the CPU combined solve; the GPU solve needs
// none, so the call does nothing
napi_value metadata(napi_env env, napi_callback_info info) {
    return guarded(env, [&] {
        (void)get_args(env, info, 6);
        return undef(env);
    });
}

// spts_csr_sync_free_wrapper(id, rowptr, col, val, b, y, N, barrier,
//                            global_flag, flag, nthreads, inner_max)
napi_value spts_csr_sync_free_wrapper(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 11);
        if (a.id() != 0) return undef(env);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(5);
        const int im = (int)a.i(10);
        if (im <= 0 || N <= 0) return undef(env);
        ResidentSpts& R = spts_resident(P, KIND_SPTS_CSR, a.i(0), a.i(1), a.i(2), N);
        void* db = rhs_in(a.i(3), P->vs * N, a.i(4), P->vs * N);
        void* dy = vec(1, P->vs * N);
        if (!R.sf_tuned) tune_syncfree(R, db, dy, N);
        for (int it = 0; it < im; it++) check(sable_spts_solve_syncfree(g_ctx, R.a, db, dy), "spts_csr_sync_free");
        check(sable_spts_status(g_ctx, R.a), "spts_csr_sync_free");
        download(a.i(4), dy, P->vs * N);
        return undef(env);
    });
}

// ---- parallel-tree SpMV slices whose kernels take a row range

// The worker's rows [r0, r1) of y on the device: the pooled vector holds
// only those rows, and the pointer handed to a row-range kernel is shifted
// back by r0 (the kernel touches rows r0..r1-1 only).
char* y_rows(Prec* P, size_t y, int64_t r0, int64_t r1) {
    return (char*)vec_in(1, y + P->vs * r0, P->vs * (size_t)(r1 - r0)) - P->vs * r0;
}

void y_rows_out(Prec* P, size_t y, const char* dy, int64_t r0, int64_t r1) {
    download(y + P->vs * r0, dy + P->vs * r0, P->vs * (size_t)(r1 - r0));
}

// Column-major DIA slices: spmv_dia_col_wrapper and spmv_dia_col_basic_wrapper
// (id, offset, data, start, end_incl, nd, N, stride, x, y, inner_max) and
// spmv_bdia_col_wrapper (id, offset, data, istart, iend, start, end_incl, nd,
// N, stride, x, y, inner_max): rows [start, end_incl] of y += A.x, each row
// adding its diagonals in stored order (bit-identical).  istart/iend are the
// blocked kernel's work arrays; the device sweep needs none and leaves them
// untouched.
napi_value dia_col_slice(napi_env env, napi_callback_info info, bool blocked) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, blocked ? 12 : 10);
        ensure_ctx();
        const size_t k = blocked ? 2 : 0;
        const int64_t r0 = a.i(2 + k), r1 = a.i(3 + k) + 1;
        const int32_t nd = (int32_t)a.i(4 + k), N = (int32_t)a.i(5 + k), stride = (int32_t)a.i(6 + k);
        const size_t x = a.i(7 + k), y = a.i(8 + k);
        const int im = (int)a.i(9 + k);
        if (nd <= 0 || N <= 0 || im <= 0 || r0 >= r1) return undef(env);
        if (r0 < 0 || r1 > N) throw Err{"row range outside the matrix"};
        void* doff = mirror(a.i(0), 4 * (size_t)nd);
        void* dd = mirror(a.i(1), P->vs * (size_t)nd * stride);
        void* dx = vec_in(0, x, P->vs * N);
        char* dy = y_rows(P, y, r0, r1);
        check(sable_spmv_dia_range(g_ctx, P->dt, N, nd, stride, (const int32_t*)doff, dd, (int32_t)r0, (int32_t)r1,
                                   dx, dy, 1.0, im),
              "spmv_dia");
        y_rows_out(P, y, dy, r0, r1);
        return undef(env);
    });
}
napi_value par_spmv_dia_col_wrapper(napi_env env, napi_callback_info info) { return dia_col_slice(env, info, false); }
napi_value par_spmv_bdia_col_wrapper(napi_env env, napi_callback_info info) { return dia_col_slice(env, info, true); }

// Column-major ELL slices: spmv_ell_col_wrapper, spmv_ell_col_gs_wrapper and
// spmv_bell_col_gs_wrapper (id, indices, data, start, end, nc, N, x, y,
// inner_max): rows [start, end) of y += A.x, slots in order per row.
napi_value par_spmv_ell_col_wrapper(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 9);
        ensure_ctx();
        const int64_t r0 = a.i(2), r1 = a.i(3);
        const int32_t nc = (int32_t)a.i(4), N = (int32_t)a.i(5);
        const size_t x = a.i(6), y = a.i(7);
        const int im = (int)a.i(8);
        if (nc <= 0 || N <= 0 || im <= 0 || r0 >= r1) return undef(env);
        if (r0 < 0 || r1 > N) throw Err{"row range outside the matrix"};
        const size_t slots = (size_t)nc * N;
        Mirror& mi = mirror_rec(a.i(0), 4 * slots);
        void* dd = mirror(a.i(1), P->vs * slots);
        const int32_t ncol = max_index(mi, (int64_t)slots) + 1;
        void* dx = vec_in(0, x, P->vs * (ncol > N ? ncol : N));
        char* dy = y_rows(P, y, r0, r1);
        check(sable_spmv_ell_range(g_ctx, P->dt, N, nc, (const int32_t*)mi.d, dd, (int32_t)r0, (int32_t)r1, dx, dy,
                                   1.0, im),
              "spmv_ell");
        y_rows_out(P, y, dy, r0, r1);
        return undef(env);
    });
}

// Row-major DIA slice: spmv_dia_wrapper(id, offset, data, start, end, nd, N,
// x, y, inner_max), rows [start, end), data[i*nd + k].
napi_value par_spmv_dia_row_wrapper(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 9);
        ensure_ctx();
        const int64_t r0 = a.i(2), r1 = a.i(3);
        const int32_t nd = (int32_t)a.i(4), N = (int32_t)a.i(5);
        const size_t x = a.i(6), y = a.i(7);
        const int im = (int)a.i(8);
        if (nd <= 0 || N <= 0 || im <= 0 || r0 >= r1) return undef(env);
        if (r0 < 0 || r1 > N) throw Err{"row range outside the matrix"};
        void* doff = mirror(a.i(0), 4 * (size_t)nd);
        void* dd = mirror(a.i(1), P->vs * (size_t)nd * N);
        void* dx = vec_in(0, x, P->vs * N);
        char* dy = y_rows(P, y, r0, r1);
        check(sable_spmv_dia_rowmajor(g_ctx, P->dt, N, nd, (const int32_t*)doff, dd, (int32_t)r0, (int32_t)r1, dx,
                                      dy, im),
              "spmv_dia");
        y_rows_out(P, y, dy, r0, r1);
        return undef(env);
    });
}

// Row-major ELL slices: spmv_ell_wrapper and spmv_ell_row_gs_wrapper (id,
// indices, data, start, end, nc, N, x, y, inner_max), rows [start, end),
// slots at i*nc + k, in the plain or the gs order.
napi_value ell_row_slice(napi_env env, napi_callback_info info, int order) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 9);
        ensure_ctx();
        const int64_t r0 = a.i(2), r1 = a.i(3);
        const int32_t nc = (int32_t)a.i(4), N = (int32_t)a.i(5);
        const size_t x = a.i(6), y = a.i(7);
        const int im = (int)a.i(8);
        if (nc <= 0 || N <= 0 || im <= 0 || r0 >= r1) return undef(env);
        if (r0 < 0 || r1 > N) throw Err{"row range outside the matrix"};
        const size_t slots = (size_t)nc * N;
        Mirror& mi = mirror_rec(a.i(0), 4 * slots);
        void* dd = mirror(a.i(1), P->vs * slots);
        const int32_t ncol = max_index(mi, (int64_t)slots) + 1;
        void* dx = vec_in(0, x, P->vs * (ncol > N ? ncol : N));
        char* dy = y_rows(P, y, r0, r1);
        check(sable_spmv_ell_rowmajor(g_ctx, P->dt, N, nc, (const int32_t*)mi.d, dd, (int32_t)r0, (int32_t)r1, dx,
                                      dy, order, im),
              "spmv_ell");
        y_rows_out(P, y, dy, r0, r1);
        return undef(env);
    });
}
napi_value par_spmv_ell_row_wrapper(napi_env env, napi_callback_info info) { return ell_row_slice(env, info, 0); }
napi_value par_spmv_ell_row_gs_wrapper(napi_env env, napi_callback_info info) { return ell_row_slice(env, info, 1); }

// sum(y, w, start, end[, N]): y[i] = y[i] + w[i] for i in [start, end)
napi_value par_sum(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 4);
        a.base = 0;  // sum has no worker id
        ensure_ctx();
        const int64_t r0 = a.i(2), r1 = a.i(3);
        if (r0 >= r1) return undef(env);
        const size_t y = a.i(0) + P->vs * r0, w = a.i(1) + P->vs * r0, n = r1 - r0;
        void* dy = vec_in(1, y, P->vs * n);
        void* dw = vec_in(0, w, P->vs * n);
        check(sable_vec_add(g_ctx, P->dt, dy, dw, n), "sum");
        download(y, dy, P->vs * n);
        return undef(env);
    });
}

// ------------------------------------------------------------ element-wise exports
// <op>_coo / self_<op>_coo, <op>_dia / self_<op>_dia, <op>_ell /
// self_<op>_ell (seq/src/sparse_opt_32.wat:597-6347; the parallel tree's
// forms take a leading worker id and, for DIA and ELL, a row range:
// par/src/sparse_opt_32.wat).  Signatures:
//   COO  [scalar,] in_val, out_val, len            self: [scalar,] val, len
//   DIA  [scalar,] offset, in, out, [r0, r1,] ndiags, stride, N
//   ELL  [scalar,] in, out, [r0, r1,] ncols, N
// where scalar is p (pow), pi (deg2rad, rad2deg) or the multiplier.
enum EwFmt { EW_COO = 0, EW_DIA = 1, EW_ELL = 2 };
struct EwDesc {
    Prec* P;
    EwFmt fmt;
    int op;
    bool self;
    bool scalar;
};

napi_value ew_export(napi_env env, napi_callback_info info) {
    void* data = nullptr;
    napi_get_cb_info(env, info, nullptr, nullptr, nullptr, &data);
    const EwDesc* D = (const EwDesc*)data;
    Prec* P = D->P;
    return guarded(env, [&] {
        const bool par = P->lead != 0;
        size_t n = D->scalar ? 1 : 0;
        if (D->fmt == EW_COO) n += D->self ? 2 : 3;
        else if (D->fmt == EW_DIA) n += (D->self ? 5 : 6) + (par ? 2 : 0);
        else n += (D->self ? 3 : 4) + (par ? 2 : 0);
        size_t argc = n + P->lead;
        std::vector<napi_value> v(argc);
        napi_get_cb_info(env, info, &argc, v.data(), nullptr, nullptr);
        if (argc < n + P->lead) throw Err{"too few arguments"};
        Args a{env, v, P->lead};
        size_t k = 0;
        const double param = D->scalar ? a.d(k++) : 0.0;
        ensure_ctx();
        if (D->fmt == EW_COO) {
            const size_t in = a.i(k++);
            const size_t out = D->self ? in : (size_t)a.i(k++);
            const int64_t len = a.i(k++);
            if (len <= 0) return undef(env);
            const size_t bytes = P->vs * (size_t)len;
            void* din = scratch(in, bytes, true);
            void* dout = D->self ? din : scratch(out, bytes, false);
            check(sable_ew_values(g_ctx, P->dt, D->op, param, len, din, dout), "element-wise");
            drop_mirrors_in(out, bytes);
            download(out, dout, bytes);
        } else if (D->fmt == EW_DIA) {
            const size_t off = a.i(k++), in = a.i(k++);
            const size_t out = D->self ? in : (size_t)a.i(k++);
            int64_t r0 = 0, r1 = 0;
            if (par) {
                r0 = a.i(k++);
                r1 = a.i(k++);
            }
            const int32_t nd = (int32_t)a.i(k++), stride = (int32_t)a.i(k++), N = (int32_t)a.i(k++);
            if (!par) r1 = N;
            if (nd <= 0 || N <= 0 || r0 >= r1) return undef(env);
            const size_t bytes = P->vs * (size_t)nd * stride;
            void* doff = mirror(off, 4 * (size_t)nd);
            void* din = scratch(in, bytes, true);
            // out is uploaded too: slots outside the diagonals' rows keep their bytes
            void* dout = D->self ? din : scratch(out, bytes, true);
            check(sable_ew_dia(g_ctx, P->dt, D->op, param, N, nd, stride, (const int32_t*)doff, (int32_t)r0,
                               (int32_t)r1, din, dout),
                  "element-wise");
            drop_mirrors_in(out, bytes);
            download(out, dout, bytes);
        } else {
            const size_t in = a.i(k++);
            const size_t out = D->self ? in : (size_t)a.i(k++);
            int64_t r0 = 0, r1 = 0;
            if (par) {
                r0 = a.i(k++);
                r1 = a.i(k++);
            }
            const int32_t nc = (int32_t)a.i(k++), N = (int32_t)a.i(k++);
            if (!par) r1 = N;
            if (nc <= 0 || N <= 0 || r0 >= r1) return undef(env);
            const size_t bytes = P->vs * (size_t)nc * N;
            void* din = scratch(in, bytes, true);
            void* dout = D->self ? din : scratch(out, bytes, true);
            check(sable_ew_ell(g_ctx, P->dt, D->op, param, N, nc, (int32_t)r0, (int32_t)r1, din, dout),
                  "element-wise");
            drop_mirrors_in(out, bytes);
            download(out, dout, bytes);
        }
        release_scratch();
        return undef(env);
    });
}

// count_nonzeros_coo(val, nnz) -> i32 (seq/src/sparse_opt_32.wat:500-521)
napi_value count_nonzeros_coo(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 2);
        ensure_ctx();
        const int64_t n = a.i(1);
        if (n <= 0) return num(env, 0);
        void* dv = mirror(a.i(0), P->vs * n);
        int64_t c = 0;
        check(sable_count_nonzeros(g_ctx, P->dt, n, dv, &c), "count_nonzeros_coo");
        return num(env, (double)c);
    });
}

// copy_nonzeros_coo(in_row, in_col, in_val, out_row, out_col, out_val, nnz) (:523-546)
napi_value copy_nonzeros_coo(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int64_t n = a.i(6);
        if (n <= 0) return undef(env);
        void* dr = mirror(a.i(0), 4 * n);
        void* dc = mirror(a.i(1), 4 * n);
        void* dv = mirror(a.i(2), P->vs * n);
        int64_t c = 0;
        check(sable_count_nonzeros(g_ctx, P->dt, n, dv, &c), "copy_nonzeros_coo");
        void* orow = scratch(a.i(3), 4 * c, false);
        void* ocol = scratch(a.i(4), 4 * c, false);
        void* oval = scratch(a.i(5), P->vs * c, false);
        check(sable_eliminate_zeros_coo(g_ctx, P->dt, n, (const int32_t*)dr, (const int32_t*)dc, dv, (int32_t*)orow,
                                        (int32_t*)ocol, oval, &c),
              "copy_nonzeros_coo");
        for (int j = 3; j < 6; ++j) drop_mirrors_in(a.i(j), (j == 5 ? P->vs : 4) * c);
        download(a.i(3), orow, 4 * c);
        download(a.i(4), ocol, 4 * c);
        download(a.i(5), oval, P->vs * c);
        release_scratch();
        return undef(env);
    });
}

// copy_nonzeros_csr(in_row_ptr, in_col, in_val, out_row_ptr, out_col, out_val, N) (:548-595)
napi_value copy_nonzeros_csr(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int32_t N = (int32_t)a.i(6);
        if (N <= 0) return undef(env);
        const int32_t* hrp = host_i32(a.i(0), N + 1);
        const int64_t nnz = (int64_t)hrp[N] - hrp[0];
        if (nnz < 0) throw Err{"row_ptr[N] < row_ptr[0]"};
        void* drp = mirror(a.i(0), 4 * (size_t)(N + 1));
        void* dc = mirror(a.i(1), 4 * (size_t)(hrp[N]));
        void* dv = mirror(a.i(2), P->vs * (size_t)(hrp[N]));
        int64_t c = 0;
        check(sable_count_nonzeros(g_ctx, P->dt, nnz, (const char*)dv + P->vs * hrp[0], &c), "copy_nonzeros_csr");
        void* orp = scratch(a.i(3), 4 * (size_t)(N + 1), false);
        void* ocol = scratch(a.i(4), 4 * c, false);
        void* oval = scratch(a.i(5), P->vs * c, false);
        check(sable_eliminate_zeros_csr(g_ctx, P->dt, N, (const int32_t*)drp, (const int32_t*)dc, dv, (int32_t*)orp,
                                        (int32_t*)ocol, oval, &c),
              "copy_nonzeros_csr");
        drop_mirrors_in(a.i(3), 4 * (size_t)(N + 1));
        drop_mirrors_in(a.i(4), 4 * c);
        drop_mirrors_in(a.i(5), P->vs * c);
        download(a.i(3), orp, 4 * (size_t)(N + 1));
        download(a.i(4), ocol, 4 * c);
        download(a.i(5), oval, P->vs * c);
        release_scratch();
        return undef(env);
    });
}

napi_value pair(napi_env env, double x, double y) {
    napi_value arr;
    napi_create_array_with_length(env, 2, &arr);
    napi_set_element(env, arr, 0, num(env, x));
    napi_set_element(env, arr, 1, num(env, y));
    return arr;
}

// count_nonzeros_dia(offset, data, ndiags, stride, N) -> [count_diags, count] (:1948-2039)
napi_value count_nonzeros_dia(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 5);
        ensure_ctx();
        const int32_t nd = (int32_t)a.i(2), stride = (int32_t)a.i(3), N = (int32_t)a.i(4);
        if (nd <= 0 || N <= 0) return pair(env, 0, 0);
        void* doff = mirror(a.i(0), 4 * (size_t)nd);
        void* dd = mirror(a.i(1), P->vs * (size_t)nd * stride);
        int32_t cd = 0;
        int64_t c = 0;
        check(sable_count_nonzeros_dia(g_ctx, P->dt, N, nd, stride, (const int32_t*)doff, dd, &cd, &c),
              "count_nonzeros_dia");
        return pair(env, cd, (double)c);
    });
}

// copy_nonzeros_dia(in_offset, in_data, out_offset, out_data, ndiags, stride, N) (:2041-2135)
napi_value copy_nonzeros_dia(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int32_t nd = (int32_t)a.i(4), stride = (int32_t)a.i(5), N = (int32_t)a.i(6);
        if (nd <= 0 || N <= 0) return undef(env);
        void* doff = mirror(a.i(0), 4 * (size_t)nd);
        void* dd = mirror(a.i(1), P->vs * (size_t)nd * stride);
        int32_t cd = 0;
        int64_t c = 0;
        check(sable_count_nonzeros_dia(g_ctx, P->dt, N, nd, stride, (const int32_t*)doff, dd, &cd, &c),
              "copy_nonzeros_dia");
        if (cd == 0) return undef(env);
        void* ooff = scratch(a.i(2), 4 * (size_t)cd, false);
        void* odata = scratch(a.i(3), P->vs * (size_t)cd * stride, false);
        check(sable_eliminate_zeros_dia(g_ctx, P->dt, N, nd, stride, (const int32_t*)doff, dd, (int32_t*)ooff, odata,
                                        &cd),
              "copy_nonzeros_dia");
        drop_mirrors_in(a.i(2), 4 * (size_t)cd);
        drop_mirrors_in(a.i(3), P->vs * (size_t)cd * stride);
        download(a.i(2), ooff, 4 * (size_t)cd);
        download(a.i(3), odata, P->vs * (size_t)cd * stride);
        release_scratch();
        return undef(env);
    });
}

// count_nonzeros_ell(data, nnz_row, ncols, N) -> [count_cols, count] (:5821-5893)
napi_value count_nonzeros_ell(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 4);
        ensure_ctx();
        const int32_t nc = (int32_t)a.i(2), N = (int32_t)a.i(3);
        if (nc <= 0 || N <= 0) return pair(env, 0, 0);
        void* dd = mirror(a.i(0), P->vs * (size_t)nc * N);
        int32_t cc = 0;
        int64_t c = 0;
        check(sable_count_nonzeros_ell(g_ctx, P->dt, N, nc, dd, &cc, &c), "count_nonzeros_ell");
        return pair(env, cc, (double)c);
    });
}

// copy_nonzeros_ell(in_data, in_indices, out_data, out_indices, nnz_row, ncols, N) (:5895-5949):
// the output width is the largest per-row nonzero count
napi_value copy_nonzeros_ell(napi_env env, napi_callback_info info) {
    Prec* P = prec_of(env, info);
    return guarded(env, [&] {
        Args a = get_args(env, info, 7);
        ensure_ctx();
        const int32_t nc = (int32_t)a.i(5), N = (int32_t)a.i(6);
        if (nc <= 0 || N <= 0) return undef(env);
        const size_t slots = (size_t)nc * N;
        void* dd = mirror(a.i(0), P->vs * slots);
        void* di = mirror(a.i(1), 4 * slots);
        int32_t cc = 0;
        int64_t c = 0;
        check(sable_count_nonzeros_ell(g_ctx, P->dt, N, nc, dd, &cc, &c), "copy_nonzeros_ell");
        if (cc == 0) return undef(env);
        const size_t oslots = (size_t)cc * N;
        void* od = scratch(a.i(2), P->vs * oslots, false);
        void* oi = scratch(a.i(3), 4 * oslots, false);
        check(sable_eliminate_zeros_ell(g_ctx, P->dt, N, nc, cc, dd, (const int32_t*)di, od, (int32_t*)oi),
              "copy_nonzeros_ell");
        drop_mirrors_in(a.i(2), P->vs * oslots);
        drop_mirrors_in(a.i(3), 4 * oslots);
        download(a.i(2), od, P->vs * oslots);
        download(a.i(3), oi, 4 * oslots);
        release_scratch();
        return undef(env);
    });
}

// Registers every element-wise export of one module.  The descriptors live
// for the process: each JS environment that loads the addon gets functions
// over the same ones.
const char* const kEwNames[] = {"expm1", "log1p", "sin", "tan", "pow", "deg2rad", "rad2deg", "sign", "sign",
                                "abs", "neg", "sqrt", "ceil", "floor", "trunc", "nearest", "multiply_scalar"};

void add_fn(napi_env env, napi_value obj, const char* name, napi_callback cb, void* data);

void add_ew(napi_env env, napi_value m, Prec* P) {
    static std::mutex mu;
    static std::map<Prec*, std::vector<EwDesc*>> made;
    std::vector<EwDesc*>* list;
    {
        std::lock_guard<std::mutex> hold(mu);
        list = &made[P];
        if (list->empty()) {
            const bool par = P->lead != 0;
            for (int fmt = 0; fmt < 3; ++fmt)
                for (int op = SABLE_EW_EXPM1; op <= SABLE_EW_MULTIPLY; ++op) {
                    const bool coo = fmt == EW_COO;
                    if (op == SABLE_EW_SIGN && !coo) continue;   // DIA/ELL sign keeps zero
                    if (op == SABLE_EW_SIGN0 && coo) continue;
                    const bool scalar = op == SABLE_EW_POW || op == SABLE_EW_DEG2RAD || op == SABLE_EW_RAD2DEG ||
                                        op == SABLE_EW_MULTIPLY;
                    for (int self = 0; self < 2; ++self) {
                        // serial tree: multiply_scalar only as self_multiply_scalar_coo
                        if (op == SABLE_EW_MULTIPLY && !par && (!coo || !self)) continue;
                        list->push_back(new EwDesc{P, (EwFmt)fmt, op, self != 0, scalar});
                    }
                }
        }
    }
    static const char* const fmts[] = {"coo", "dia", "ell"};
    for (EwDesc* D : *list) {
        std::string name = std::string(D->self ? "self_" : "") + kEwNames[D->op] + "_" + fmts[D->fmt];
        add_fn(env, m, name.c_str(), ew_export, D);
    }
}

// ------------------------------------------------------------ module objects
napi_value malloc_fn(napi_env env, napi_callback_info info) {
    return guarded(env, [&] {
        Args a = get_args(env, info, 1);
        return num(env, (double)lm_malloc((size_t)a.i(0)));
    });
}

napi_value free_fn(napi_env env, napi_callback_info info) {
    return guarded(env, [&] {
        Args a = get_args(env, info, 1);
        const size_t off = a.i(0);
        auto it = g_mem.used.find(off);
        if (it != g_mem.used.end()) {
            // the region's mirrors and resident matrices go; its pages become
            // writable so that the next owner's writes do not fault
            drop_resident_in(off, it->second);
            drop_mirrors_in(off, it->second);
            unprotect_range(off, it->second);
            for (auto st = g_stream.begin(); st != g_stream.end();)
                st = overlaps(st->first.first, st->first.second, off, it->second) ? g_stream.erase(st) : std::next(st);
        }
        lm_free(off);
        return undef(env);
    });
}

// One ArrayBuffer per JS environment (main thread or worker), all over the
// same host memory, so that every `memory.buffer` sees the same bytes.
struct EnvData {
    napi_ref buffer = nullptr;
};

void env_cleanup(napi_env env, void* data, void*) {
    EnvData* d = (EnvData*)data;
    if (d->buffer) napi_delete_reference(env, d->buffer);
    delete d;
}

napi_value buffer_getter(napi_env env, napi_callback_info) {
    EnvData* d = nullptr;
    napi_get_instance_data(env, (void**)&d);
    napi_value ab;
    if (d && d->buffer) {
        napi_get_reference_value(env, d->buffer, &ab);
        return ab;
    }
    napi_create_external_arraybuffer(env, g_mem.base, g_mem.size, nullptr, nullptr, &ab);
    if (d) napi_create_reference(env, ab, 1, &d->buffer);
    return ab;
}

napi_value stats_fn(napi_env env, napi_callback_info) {
    std::lock_guard<std::mutex> hold(g_lock);
    napi_value o;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "mirrors", num(env, (double)g_cache.size()));
    napi_set_named_property(env, o, "linear_bytes", num(env, (double)g_mem.size));
    napi_set_named_property(env, o, "pinned", num(env, g_mem.pinned ? 1 : 0));
    napi_set_named_property(env, o, "allocations", num(env, (double)g_mem.used.size()));
    napi_set_named_property(env, o, "write_tracking", num(env, sable_track::g_on ? 1 : 0));
    napi_set_named_property(env, o, "writes_caught", num(env, (double)sable_track::faults()));
    napi_set_named_property(env, o, "uploads", num(env, (double)g_uploads));
    napi_set_named_property(env, o, "upload_bytes", num(env, (double)g_upload_bytes));
    napi_set_named_property(env, o, "vector_uploads", num(env, (double)g_vec_uploads));
    napi_set_named_property(env, o, "resident", num(env, (double)g_resident.size()));
    napi_set_named_property(env, o, "schedule", num(env, g_last_sched));
    napi_set_named_property(env, o, "schedule_ms", num(env, g_last_sched_ms));
    return o;
}

void add_fn(napi_env env, napi_value obj, const char* name, napi_callback cb, void* data) {
    napi_value f;
    napi_create_function(env, name, NAPI_AUTO_LENGTH, cb, data, &f);
    napi_set_named_property(env, obj, name, f);
}

struct Fn {
    const char* name;
    napi_callback cb;
};

// Serial tree: the exports of seq/src/sparse_opt_32.wat / sparse_opt_64.wat
// on the hot path, plus the three conversions the reference does in JS.
napi_value make_serial(napi_env env, Prec* P) {
    napi_value m;
    napi_create_object(env, &m);
    const Fn fns[] = {
        {"spmv_coo", spmv_coo}, {"spmv_coo_wrapper", spmv_coo_wrapper},
        {"spmv_csr", spmv_csr}, {"spmv_csr_wrapper", spmv_csr_wrapper},
        {"spmv_dia", spmv_dia}, {"spmv_dia_wrapper", spmv_dia_wrapper},
        {"spmv_ell", spmv_ell}, {"spmv_ell_wrapper", spmv_ell_wrapper},
        {"transpose_csr", transpose_csr}, {"transpose_dia", transpose_dia},
        {"count_transpose_cols_ell", count_transpose_cols_ell}, {"transpose_ell", transpose_ell},
        {"memcpy", lm_memcpy}, {"coo_csr", coo_csr}, {"num_cols", num_cols}, {"csr_ell", csr_ell},
        {"num_diags", num_diags}, {"csr_dia", csr_dia}, {"coo_csc", coo_csc},
        {"csr_ell_row", csr_ell_row}, {"csr_dia_row", csr_dia_row},
        {"spts_csc", spts_csc}, {"spts_csc_wrapper", spts_csc_wrapper}, {"mmread", mmread},
        {"count_nonzeros_coo", count_nonzeros_coo}, {"copy_nonzeros_coo", copy_nonzeros_coo},
        {"copy_nonzeros_csr", copy_nonzeros_csr}, {"count_nonzeros_dia", count_nonzeros_dia},
        {"copy_nonzeros_dia", copy_nonzeros_dia}, {"count_nonzeros_ell", count_nonzeros_ell},
        {"copy_nonzeros_ell", copy_nonzeros_ell},
    };
    for (auto& f : fns) add_fn(env, m, f.name, f.cb, P);
    add_ew(env, m, P);
    return m;
}

// Parallel tree: the exports that par/src/worker32.js calls on the hot path,
// each with its leading worker id.
napi_value make_parallel(napi_env env, Prec* P, Prec* serial) {
    napi_value m;
    napi_create_object(env, &m);
    const Fn fns[] = {
        {"spmv_coo_wrapper", spmv_coo_wrapper}, {"spmv_csr_wrapper", spmv_csr_wrapper},
        // the SIMD variants compute the same products (CSR/COO within the
        // summation bound, the column-major DIA/ELL ones bit for bit)
        {"spmv_coo_gs_wrapper", spmv_coo_wrapper}, {"spmv_csr_gs_wrapper", spmv_csr_wrapper},
        {"spmv_csr_short_rows_wrapper", spmv_csr_short_impl}, {"spmv_csr_gs_short_wrapper", spmv_csr_short_impl},
        {"spmv_csr_unroll2_wrapper", spmv_csr_short_impl}, {"spmv_csr_unroll3_wrapper", spmv_csr_short_impl},
        {"spmv_csr_unroll4_wrapper", spmv_csr_short_impl}, {"spmv_csr_unroll6_wrapper", spmv_csr_short_impl},
        {"spmv_dia_col_wrapper", par_spmv_dia_col_wrapper},
        {"spmv_dia_col_basic_wrapper", par_spmv_dia_col_wrapper},
        {"spmv_bdia_col_wrapper", par_spmv_bdia_col_wrapper},
        {"spmv_dia_wrapper", par_spmv_dia_row_wrapper},
        {"spmv_ell_col_wrapper", par_spmv_ell_col_wrapper}, {"spmv_ell_col_gs_wrapper", par_spmv_ell_col_wrapper},
        {"spmv_bell_col_gs_wrapper", par_spmv_ell_col_wrapper},
        {"spmv_ell_wrapper", par_spmv_ell_row_wrapper}, {"spmv_ell_row_gs_wrapper", par_spmv_ell_row_gs_wrapper},
        {"memcpy", lm_memcpy}, {"spts_level_csr", spts_level_csr},
        {"spts_csr_sync_free_wrapper", spts_csr_sync_free_wrapper},
        {"spts_csr_level_sync_free_wrapper", spts_csr_level_sync_free_wrapper},
        {"spts_csr_opt_level_sync_free_wrapper", spts_csr_opt_level_sync_free_wrapper},
        {"metadata", metadata},
    };
    for (auto& f : fns) add_fn(env, m, f.name, f.cb, P);
    add_fn(env, m, "sum", par_sum, serial);  // sum takes no worker id
    // the row-major conversions are JS functions of the parallel tree (no worker id)
    add_fn(env, m, "csr_ell_row", csr_ell_row, serial);
    add_fn(env, m, "csr_dia_row", csr_dia_row, serial);
    add_ew(env, m, P);
    return m;
}

napi_value Init(napi_env env, napi_value exports) {
    {
        std::lock_guard<std::mutex> hold(g_lock);
        size_t mb = 1024;
        if (const char* e = getenv("SABLE_LINEAR_MEMORY_MB")) mb = (size_t)atol(e);
        if (!g_mem.base && !init_linear(mb << 20)) {
            napi_throw_error(env, nullptr, "cannot allocate the linear memory");
            return nullptr;
        }
    }
    napi_set_instance_data(env, new EnvData(), env_cleanup, nullptr);
    napi_value mem;
    napi_create_object(env, &mem);
    napi_property_descriptor d = {"buffer", nullptr, nullptr, buffer_getter, nullptr, nullptr, napi_default, nullptr};
    napi_define_properties(env, mem, 1, &d);
    napi_set_named_property(env, exports, "wasmMemory", mem);
    napi_value mal;
    napi_create_object(env, &mal);
    add_fn(env, mal, "_malloc", malloc_fn, nullptr);
    add_fn(env, mal, "_free", free_fn, nullptr);
    napi_set_named_property(env, exports, "malloc_exports", mal);
    napi_set_named_property(env, exports, "sparse_opt_32", make_serial(env, &g_p32));
    napi_set_named_property(env, exports, "sparse_opt_64", make_serial(env, &g_p64));
    napi_set_named_property(env, exports, "par_sparse_opt_32", make_parallel(env, &g_par32, &g_p32));
    napi_set_named_property(env, exports, "par_sparse_opt_64", make_parallel(env, &g_par64, &g_p64));
    add_fn(env, exports, "stats", stats_fn, nullptr);
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
