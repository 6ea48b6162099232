// serve.hip -- a request server for SpMV and lower SpTS over resident matrices.
//
// Matrices (and their SpMV row-block plans and FP32 copies) are uploaded
// once and stay in HBM; they are charged against the server's HBM budget.
// Requests (SpMV y = A.x, or lower solve L.x = b) queue on the host.  A
// scheduler thread forms batches continuously: whenever one of `slots`
// in-flight batch slots is free it admits queued requests in arrival order
// until the batch is full or the slot's device arena (inputs, outputs and
// refinement work space) is exhausted, stages the inputs, and launches
//   mode 0 (plain):   one request per batch, FP64, one SpMV launch or one
//                     sync-free solve launch per request;
//   mode 1 (batched): FP64, one persistent SpMV launch per right-hand-side
//                     width and one sync-free solve launch for the whole
//                     batch; SpMV requests on the same matrix share one
//                     multi-RHS product (up to `group_rhs` of them, x and y
//                     row-major [N][w] with w = 1, 2, 4 or 8), so the
//                     matrix is read once for all of them;
//   mode 2 (mixed):   batched, SpMV with FP32 values/vectors and FP32
//                     accumulation; solves in FP32 with FP64 residual and
//                     update for `refine_iters` refinement steps (one
//                     step, the default, is fused into the solve's sweep).
// Batched solves (modes 1 and 2) run level by level (spts_sweep 0): every
// lower matrix carries its level schedule from registration, and launch l
// of a batch solves level l of all of the batch's solves at once (a batch
// needs as many launches as its deepest matrix has levels); or (spts_sweep
// 1) in one sync-free launch over all of the batch's rows.
// A completion thread waits on each slot's event in order, copies results
// out and records latencies.  Several slots in flight let a new batch be
// staged and launched while earlier ones run on the GPU.
#include "common.h"
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <random>
#include <thread>
#include <unordered_map>

namespace sable {

static double now_us() {
    using namespace std::chrono;
    return duration<double, std::micro>(steady_clock::now().time_since_epoch()).count();
}

struct SMatrix {
    int32_t N = 0;
    int64_t nnz = 0;
    int kind = 0;  // 0 general, 1 lower triangular (diagonal last)
    int32_t* rp = nullptr;
    int32_t* col = nullptr;
    double* val64 = nullptr;
    float* val32 = nullptr;
    sable_csr_plan_t* plan = nullptr;
    const int32_t* blk = nullptr;
    int64_t nblocks = 0;
    int lds_cap = 0;
    size_t bytes = 0;
    // lower matrices: level schedule (level_ptr on the host, rows listed by
    // level on the device) for the batched level-set solves
    std::vector<int32_t> lp;
    int32_t* perm = nullptr;
    // large general matrices: column-sorted tile plans (FP64 and FP32) for
    // their products, which run one request per launch
    sable_csr_tiles_t* tiles64 = nullptr;
    sable_csr_tiles_t* tiles32 = nullptr;
    // smaller general matrices in the batched and mixed modes: a column-
    // sorted tile plan in the mode's precision for the batched tile launch
    sable_csr_tiles_t* btiles = nullptr;
};

// Matrices at least this large get tile plans: their x is far larger than
// one XCD's L2, where the row-block kernel is bound by x gathers.
constexpr int32_t TILE_MIN_ROWS = 1 << 20;

struct SRequest {
    int64_t id = 0;
    int32_t mat = 0;
    int op = 0;  // 0 SpMV, 1 lower solve
    const double* in = nullptr;
    double* out = nullptr;
    uint64_t seed = 0;
    bool seeded = false;
    double t_arrival = 0, t_done = 0;
    int status = 0;
    bool done = false;
    size_t in_off = 0, out_off = 0, work_off = 0;
    int32_t col = 0, ld = 1;  // SpMV: column of the request in its group's [N][ld] x and y
};

struct FillItem {
    void* ptr;
    int64_t begin;  // first global element
    int32_t n;
    int32_t is_f32;
    uint64_t seed;
    int64_t stride;  // element i is written at ptr[i * stride]
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    char* arena = nullptr;
    char* stage_in = nullptr;
    char* stage_out = nullptr;
    char* h_items = nullptr;  // pinned staging for item arrays
    char* d_items = nullptr;  // device item arrays
    char* h_chunks = nullptr;  // pinned staging for the batch's work tables
    char* d_chunks = nullptr;
    unsigned long long* d_counter = nullptr;
    unsigned* d_ticket = nullptr;
    int* d_err = nullptr;
    int* h_err = nullptr;
    size_t out_begin = 0, out_end = 0;
    std::vector<SRequest*> reqs;
    bool busy = false;
    int launch_rc = 0;  // a failed launch fails the batch's requests
};

constexpr size_t ITEM_BYTES = 1 << 20;  // per slot, per item kind


__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// A chunk of one fill item: elements first .. first + count - 1.
struct FillChunk {
    int32_t fill;
    int32_t first;
    int32_t count;
    int32_t pad;
};
constexpr int FILL_CHUNK = 1024;

// U(-1,1) inputs from (seed, index) for seeded requests, one workgroup per
// chunk of FILL_CHUNK elements.
__global__ __launch_bounds__(256) void k_fill_uniform(const FillItem* __restrict__ items,
                                                      const FillChunk* __restrict__ chunks) {
    const FillChunk c = chunks[blockIdx.x];
    const FillItem& it = items[c.fill];
    for (int t = threadIdx.x; t < c.count; t += 256) {
        const int64_t i = (int64_t)c.first + t;
        const double u = (double)(mix64(it.seed * 0x100000001B3ull + (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
        const double v = 2.0 * u - 1.0;
        if (it.is_f32) ((float*)it.ptr)[i * it.stride] = (float)v;
        else ((double*)it.ptr)[i * it.stride] = v;
    }
}

__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ o) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        o[i] = (float)a[i];
}

static size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

// The server's matrices: append-only, at addresses that never change.  The
// scheduler and completion threads and the submitting threads index it
// without the server lock while add_matrix appends under it: a matrix is
// complete before the count that covers it is published (release), and
// every id is checked against the count (acquire) before it is used.  A
// fixed directory of chunks holds up to MAX matrices, so no append moves
// the directory a reader may be walking.
class MatrixTable {
  public:
    static constexpr int32_t CHUNK = 1024, DIRS = 1024, MAX = CHUNK * DIRS;
    MatrixTable() {
        for (auto& d : dir_) d.store(nullptr, std::memory_order_relaxed);
    }
    ~MatrixTable() {
        for (auto& d : dir_) delete[] d.load(std::memory_order_relaxed);
    }
    MatrixTable(const MatrixTable&) = delete;
    MatrixTable& operator=(const MatrixTable&) = delete;
    int32_t size() const { return n_.load(std::memory_order_acquire); }
    SMatrix& operator[](int32_t i) { return dir_[i / CHUNK].load(std::memory_order_acquire)[i % CHUNK]; }
    const SMatrix& operator[](int32_t i) const {
        return dir_[i / CHUNK].load(std::memory_order_acquire)[i % CHUNK];
    }
    // One writer at a time (the caller holds the server lock).  Returns the
    // new id, or -1 when the table is full.
    int32_t push_back(SMatrix&& m) {
        const int32_t i = n_.load(std::memory_order_relaxed);
        if (i >= MAX) return -1;
        SMatrix* c = dir_[i / CHUNK].load(std::memory_order_relaxed);
        if (!c) {
            c = new SMatrix[CHUNK];
            dir_[i / CHUNK].store(c, std::memory_order_release);
        }
        c[i % CHUNK] = std::move(m);
        n_.store(i + 1, std::memory_order_release);
        return i;
    }

  private:
    std::atomic<SMatrix*> dir_[DIRS];
    std::atomic<int32_t> n_{0};
};

}  // namespace sable

using namespace sable;

struct sable_server_s {
    sable_ctx_t* ctx = nullptr;
    sable_server_config_t cfg{};
    int64_t used = 0;
    MatrixTable mats;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_pending, cv_slot, cv_done, cv_complete;
    std::deque<SRequest*> pending;
    std::deque<int> inflight;  // slot indices in launch order
    std::unordered_map<int64_t, std::unique_ptr<SRequest>> reqs;
    int64_t next_id = 1;
    bool stop = false;
    bool sched_done = false;
    std::thread sched, compl_;
    // statistics
    std::vector<double> latencies;
    int64_t completed = 0, rejected = 0, batches = 0, batched_reqs = 0;
    double t_last_done = 0;
    int error = 0;
    char errmsg[256] = "";

    size_t vbytes_spmv() const { return cfg.mode == 2 ? 4 : 8; }
    // Work-table buffer per slot (fill, level and row chunks of a batch):
    // an eighth of the arena, at least 8 MiB.  Admission keeps each batch's
    // tables within it (table_need).
    size_t chunk_bytes() const { return std::max((size_t)8 << 20, (size_t)cfg.arena_bytes / 8); }
    int chunk_rows() const { return 256 / (cfg.spts_lanes > 0 ? cfg.spts_lanes : 4); }
    // Upper bound of the table bytes request r adds: its fill chunks, and
    // for a solve one level chunk per CH rows plus one per level, and one
    // row chunk per CH rows.
    size_t table_need(const SRequest* r) const {
        const SMatrix& m = mats[r->mat];
        size_t b = r->seeded ? ((size_t)m.N / 1024 + 1) * 16 : 0;
        if (r->op == 1) {
            const size_t ch = (size_t)chunk_rows();
            b += ((size_t)m.N / ch + m.lp.size() + (size_t)m.N / ch + 2) * sizeof(RowChunk);
        }
        return b + 64;
    }
    void scheduler();
    void completer();
    int launch_batch(Slot& s, std::vector<SRequest*>& batch);
    int group_max() const { return cfg.mode == 0 ? 1 : std::max(1, std::min(8, cfg.group_rhs)); }
    // Arena bytes that admitting r adds to a batch in which `before` SpMV
    // requests on the same matrix precede it: SpMV requests on one matrix
    // form groups of up to group_max() in arrival order, each holding x and
    // y of width 1, 2, 4 or 8.
    size_t need_bytes(const SRequest* r, int before) const;
};

static int width_for(size_t k) { return k == 0 ? 0 : k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : 8; }

size_t sable_server_s::need_bytes(const SRequest* r, int before) const {
    const size_t n = (size_t)mats[r->mat].N;
    if (r->op == 1) return 2 * al(n * 8) + (cfg.mode == 2 ? 2 * al(n * 4) : 0);
    const int g = mats[r->mat].tiles64 ? 1 : group_max();
    const int k = before % g;  // requests already in the open group
    auto cost = [&](int w) { return w ? 2 * al(n * (size_t)w * vbytes_spmv()) : (size_t)0; };
    return cost(width_for((size_t)k + 1)) - cost(width_for((size_t)k));
}

void sable_server_s::scheduler() {
    (void)hipSetDevice(ctx->device);
    size_t next_slot = 0;
    while (true) {
        Slot* sl;
        std::vector<SRequest*> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_pending.wait(lk, [&] { return stop || !pending.empty(); });
            if (!stop) cv_slot.wait(lk, [&] { return stop || !slots[next_slot].busy; });
            if (stop) {
                // requests still queued at shutdown fail; in-flight batches finish
                for (SRequest* r : pending) {
                    r->status = SABLE_EINVAL;
                    r->done = true;
                    r->t_done = now_us();
                }
                pending.clear();
                sched_done = true;
                cv_done.notify_all();
                cv_complete.notify_all();
                return;
            }
            sl = &slots[next_slot];
            const size_t cap = (size_t)cfg.arena_bytes;
            size_t used_b = 0;
            const int maxb = cfg.mode == 0 ? 1 : cfg.max_batch;
            std::unordered_map<int32_t, int> per_mat;  // SpMV requests per matrix in this batch
            size_t tables = 768;  // alignment slack of the three tables
            while (!pending.empty() && (int)batch.size() < maxb) {
                SRequest* r = pending.front();
                int& before = per_mat[r->op == 0 ? r->mat : -1 - r->mat];
                const size_t nb = need_bytes(r, r->op == 0 ? before : 0);
                const size_t tb = table_need(r);
                if (need_bytes(r, 0) > cap || table_need(r) + 768 > chunk_bytes()) {  // can never fit: reject
                    pending.pop_front();
                    r->status = SABLE_EBUDGET;
                    r->done = true;
                    r->t_done = now_us();
                    rejected++;
                    cv_done.notify_all();
                    continue;
                }
                if (used_b + nb > cap || tables + tb > chunk_bytes()) break;
                used_b += nb;
                tables += tb;
                if (r->op == 0) before++;
                batch.push_back(r);
                pending.pop_front();
            }
            if (batch.empty()) continue;
            sl->busy = true;
        }
        int rc = launch_batch(*sl, batch);
        if (rc) (void)hipEventRecord(sl->ev, sl->stream);  // keep completion order
        {
            std::lock_guard<std::mutex> lk(mu);
            sl->launch_rc = rc;
            if (rc) {
                error = rc;
                snprintf(errmsg, sizeof errmsg, "%s", sable_last_error());
            }
            inflight.push_back((int)next_slot);
            batches++;
            batched_reqs += (int64_t)batch.size();
        }
        cv_complete.notify_all();
        next_slot = (next_slot + 1) % slots.size();
    }
}

namespace {
// SpMV requests of one batch that share a matrix and one [N][w] x and y.
struct SpmvGroup {
    int32_t mat;
    int w;  // 1, 2, 4 or 8 columns (requests, padded to a power of two)
    std::vector<SRequest*> rs;
    size_t x_off = 0, y_off = 0;
};

}  // namespace

int sable_server_s::launch_batch(Slot& s, std::vector<SRequest*>& batch) {
    s.reqs = batch;
    const bool mixed = cfg.mode == 2;
    const size_t vs = vbytes_spmv();
    // group SpMV requests by matrix (arrival order within a group); plain
    // mode keeps one request per product
    const int gmax = group_max();
    std::vector<SpmvGroup> groups;
    {
        std::unordered_map<int32_t, size_t> open;
        for (auto* r : batch) {
            if (r->op != 0) continue;
            auto it = open.find(r->mat);
            const size_t cap = mats[r->mat].tiles64 ? 1 : (size_t)gmax;
            if (it == open.end() || groups[it->second].rs.size() == cap) {
                open[r->mat] = groups.size();
                groups.push_back(SpmvGroup{r->mat, 1, {}, 0, 0});
                it = open.find(r->mat);
            }
            groups[it->second].rs.push_back(r);
        }
        for (auto& g : groups) g.w = width_for(g.rs.size());
    }
    // arena layout: inputs, then outputs, then refinement work
    size_t off = 0;
    for (auto& g : groups) {
        g.x_off = off;
        off += al((size_t)mats[g.mat].N * g.w * vs);
        for (size_t c = 0; c < g.rs.size(); c++) {
            g.rs[c]->in_off = g.x_off;
            g.rs[c]->col = (int32_t)c;
            g.rs[c]->ld = g.w;
        }
    }
    for (auto* r : batch) {
        if (r->op == 0) continue;
        r->in_off = off;
        off += al((size_t)mats[r->mat].N * 8);
    }
    const size_t in_end = off;
    s.out_begin = off;
    for (auto& g : groups) {
        g.y_off = off;
        off += al((size_t)mats[g.mat].N * g.w * vs);
        for (auto* r : g.rs) r->out_off = g.y_off;
    }
    for (auto* r : batch) {
        if (r->op == 0) continue;
        r->out_off = off;
        off += al((size_t)mats[r->mat].N * 8);
    }
    s.out_end = off;
    for (auto* r : batch) {
        if (r->op == 1 && mixed) {
            r->work_off = off;
            off += 2 * al((size_t)mats[r->mat].N * 4);
        }
    }
    SABLE_CHECK(off <= (size_t)cfg.arena_bytes, SABLE_EINVAL, "batch exceeds the slot arena");
    hipStream_t st = s.stream;
    // host inputs -> pinned staging -> device; seeded inputs are generated
    // on the device
    bool any_host = false;
    std::vector<FillItem> fills;
    int64_t fill_total = 0;
    for (auto* r : batch) {
        const int32_t n = mats[r->mat].N;
        const bool f32 = r->op == 0 && mixed;
        const size_t es = r->op == 0 ? vs : 8;
        char* dst = s.arena + r->in_off + (size_t)r->col * es;
        if (r->seeded) {
            fills.push_back(FillItem{dst, fill_total, n, f32 ? 1 : 0, r->seed, r->ld});
            fill_total += n;
        } else {
            any_host = true;
            char* h = s.stage_in + r->in_off + (size_t)r->col * es;
            if (f32) {
                float* d = (float*)h;
                for (int32_t i = 0; i < n; i++) d[(size_t)i * r->ld] = (float)r->in[i];
            } else if (r->ld == 1) {
                memcpy(h, r->in, 8 * (size_t)n);
            } else {
                double* d = (double*)h;
                for (int32_t i = 0; i < n; i++) d[(size_t)i * r->ld] = r->in[i];
            }
        }
    }
    if (any_host) SABLE_HIP(hipMemcpyAsync(s.arena, s.stage_in, in_end, hipMemcpyHostToDevice, st));
    // item arrays: [fill | spmv (by width) | spts or refine | spts view]
    char* hi = s.h_items;
    const size_t fill_off = 0, spmv_off = ITEM_BYTES, spts_off = 2 * ITEM_BYTES, view_off = 3 * ITEM_BYTES;
    SABLE_CHECK(fills.size() * sizeof(FillItem) <= ITEM_BYTES, SABLE_EINVAL, "too many requests in a batch");
    memcpy(hi + fill_off, fills.data(), fills.size() * sizeof(FillItem));
    std::vector<sable_batch_spmv_item_t> sv[4];  // widths 1, 2, 4, 8
    std::vector<sable_batch_tiles_item_t> bt[4];  // the same, matrices with a batch tile plan
    int64_t blk_total[4] = {0, 0, 0, 0}, unit_total[4] = {0, 0, 0, 0};
    int32_t bt_rows[4] = {0, 0, 0, 0};
    for (auto& g : groups) {
        const SMatrix& m = mats[g.mat];
        if (m.tiles64) continue;  // launched on its own below
        const int wi = g.w == 1 ? 0 : g.w == 2 ? 1 : g.w == 4 ? 2 : 3;
        if (m.btiles) {
            sable_batch_tiles_item_t ti;
            const int trc = sable_csr_tiles_batch_item(m.btiles, s.arena + g.x_off, s.arena + g.y_off, 0.0,
                                                       unit_total[wi], &ti);
            if (trc) return trc;
            unit_total[wi] += ti.ntr;
            bt_rows[wi] = std::max(bt_rows[wi], ti.tile_rows);
            bt[wi].push_back(ti);
            continue;
        }
        sable_batch_spmv_item_t it;
        it.row_ptr = m.rp;
        it.col = m.col;
        it.val = mixed ? (const void*)m.val32 : (const void*)m.val64;
        it.blk = m.blk;
        it.x = s.arena + g.x_off;
        it.y = s.arena + g.y_off;
        it.blk_begin = blk_total[wi];
        it.beta = 0.0;
        blk_total[wi] += m.nblocks;
        sv[wi].push_back(it);
    }
    std::vector<sable_batch_spts_item_t> tv;
    std::vector<sable_batch_refine_item_t> rv;
    int64_t row_total = 0;
    for (auto* r : batch) {
        if (r->op == 0) continue;
        const SMatrix& m = mats[r->mat];
        if (!mixed) {
            sable_batch_spts_item_t it;
            it.row_ptr = m.rp;
            it.col = m.col;
            it.val = m.val64;
            it.b = s.arena + r->in_off;
            it.x = s.arena + r->out_off;
            it.row_begin = row_total;
            it.N = m.N;
            it.pad = 0;
            row_total += m.N;
            tv.push_back(it);
        } else {
            sable_batch_refine_item_t it;
            it.row_ptr = m.rp;
            it.col = m.col;
            it.val64 = m.val64;
            it.val32 = m.val32;
            it.b64 = (const double*)(s.arena + r->in_off);
            it.x64 = (double*)(s.arena + r->out_off);
            it.r32 = (float*)(s.arena + r->work_off);
            it.d32 = (float*)(s.arena + r->work_off + al((size_t)m.N * 4));
            it.row_begin = row_total;
            it.N = m.N;
            it.pad = 0;
            row_total += m.N;
            rv.push_back(it);
        }
    }
    size_t sbytes[4], sofs[4], tbytes[4], tofs[4], sb = 0;
    for (int wi = 0; wi < 4; wi++) {
        sofs[wi] = spmv_off + sb;
        sbytes[wi] = sv[wi].size() * sizeof(sable_batch_spmv_item_t);
        sb += sbytes[wi];
    }
    for (int wi = 0; wi < 4; wi++) {
        tofs[wi] = spmv_off + sb;
        tbytes[wi] = bt[wi].size() * sizeof(sable_batch_tiles_item_t);
        sb += tbytes[wi];
    }
    SABLE_CHECK(sb <= ITEM_BYTES && tv.size() * sizeof(sable_batch_spts_item_t) <= ITEM_BYTES &&
                    rv.size() * sizeof(sable_batch_refine_item_t) <= ITEM_BYTES,
                SABLE_EINVAL, "too many requests in a batch");
    for (int wi = 0; wi < 4; wi++) {
        if (sbytes[wi]) memcpy(hi + sofs[wi], sv[wi].data(), sbytes[wi]);
        if (tbytes[wi]) memcpy(hi + tofs[wi], bt[wi].data(), tbytes[wi]);
    }
    if (mixed) memcpy(hi + spts_off, rv.data(), rv.size() * sizeof(rv[0]));
    else memcpy(hi + spts_off, tv.data(), tv.size() * sizeof(tv[0]));
    const size_t fb = fills.size() * sizeof(FillItem);
    const size_t tb = mixed ? rv.size() * sizeof(rv[0]) : tv.size() * sizeof(tv[0]);
    if (fb) SABLE_HIP(hipMemcpyAsync(s.d_items + fill_off, hi + fill_off, fb, hipMemcpyHostToDevice, st));
    if (sb) SABLE_HIP(hipMemcpyAsync(s.d_items + spmv_off, hi + spmv_off, sb, hipMemcpyHostToDevice, st));
    if (tb) SABLE_HIP(hipMemcpyAsync(s.d_items + spts_off, hi + spts_off, tb, hipMemcpyHostToDevice, st));
    const int lanes = cfg.spts_lanes > 0 ? cfg.spts_lanes : 4;
    const long long spin = 1ll << 24;
    // Work tables of the batch, cut into chunks on the host (layout in the
    // slot's chunk buffer: fill chunks | level chunks | row chunks):
    //  * fill chunks of FILL_CHUNK elements of each seeded input;
    //  * level chunks: level l lists, item by item, the rows of level l of
    //    every solve that has one, CH rows (one workgroup) per chunk;
    //  * row chunks: every solve's rows in natural order, CH per chunk.
    const int CH = 256 / lanes;
    std::vector<FillChunk> fch;
    for (size_t f = 0; f < fills.size(); f++)
        for (int32_t a = 0; a < fills[f].n; a += FILL_CHUNK)
            fch.push_back(FillChunk{(int32_t)f, a, std::min(FILL_CHUNK, fills[f].n - a), 0});
    LevelPlan plan;
    std::vector<RowChunk> lch, rch;
    if (cfg.mode != 0 && cfg.spts_sweep != 1) {
        std::vector<const SRequest*> solves;
        for (auto* r : batch)
            if (r->op == 1) solves.push_back(r);
        int top = 0;
        for (auto* r : solves) top = std::max(top, (int)mats[r->mat].lp.size() - 1);
        plan.chunk_lo.assign((size_t)top + 1, 0);
        for (int lv = 0; lv < top; lv++) {
            plan.chunk_lo[lv] = (int32_t)lch.size();
            for (size_t k = 0; k < solves.size(); k++) {
                const SMatrix& m = mats[solves[k]->mat];
                if (lv + 1 >= (int)m.lp.size()) continue;
                for (int32_t a = m.lp[lv]; a < m.lp[lv + 1]; a += CH)
                    lch.push_back(RowChunk{m.perm + a, (int32_t)k, 0, std::min(CH, m.lp[lv + 1] - a), 0});
            }
        }
        plan.chunk_lo[top] = (int32_t)lch.size();
        for (size_t k = 0; k < solves.size(); k++) {
            const int32_t n = mats[solves[k]->mat].N;
            for (int32_t a = 0; a < n; a += CH) rch.push_back(RowChunk{nullptr, (int32_t)k, a, std::min(CH, n - a), 0});
        }
    }
    const size_t fcb = fch.size() * sizeof(FillChunk), lcb = lch.size() * sizeof(RowChunk),
                 rcb = rch.size() * sizeof(RowChunk);
    const size_t lc_off = (fcb + 255) & ~(size_t)255, rc_off = lc_off + ((lcb + 255) & ~(size_t)255);
    SABLE_CHECK(rc_off + rcb <= chunk_bytes(), SABLE_EINVAL, "batch work tables exceed the slot's chunk buffer");
    if (fcb) memcpy(s.h_chunks, fch.data(), fcb);
    if (lcb) memcpy(s.h_chunks + lc_off, lch.data(), lcb);
    if (rcb) memcpy(s.h_chunks + rc_off, rch.data(), rcb);
    if (rc_off + rcb) SABLE_HIP(hipMemcpyAsync(s.d_chunks, s.h_chunks, rc_off + rcb, hipMemcpyHostToDevice, st));
    if (!fch.empty()) {
        hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)fch.size()), dim3(256), 0, st,
                           (const FillItem*)(s.d_items + fill_off), (const FillChunk*)s.d_chunks);
        SABLE_LAUNCH_CHECK();
    }
    const RowChunk* d_lch = (const RowChunk*)(s.d_chunks + lc_off);
    const RowChunk* d_rch = (const RowChunk*)(s.d_chunks + rc_off);
    SABLE_HIP(hipMemsetAsync(s.d_err, 0, sizeof(int), st));
    int rc = SABLE_OK;
    if (cfg.mode == 0) {
        // plain: exactly one request, its own single-matrix launch
        SRequest* r = batch[0];
        const SMatrix& m = mats[r->mat];
        if (r->op == 0 && m.tiles64)
            rc = tiles_launch(st, m.tiles64, s.arena + r->in_off, s.arena + r->out_off, 0.0);
        else if (r->op == 0)
            rc = csr_launch(st, SABLE_F64, m.blk, m.nblocks, m.rp, m.col, m.val64, s.arena + r->in_off,
                            s.arena + r->out_off, 0.0, m.lds_cap);
        else
            rc = syncfree_launch(st, ctx->num_cu, SABLE_F64, lanes, m.N, m.rp, m.col, m.val64, s.arena + r->in_off,
                                 s.arena + r->out_off, s.d_ticket, s.d_err, spin);
    } else {
        static const int widths[4] = {1, 2, 4, 8};
        for (auto& g : groups) {
            const SMatrix& m = mats[g.mat];
            if (!m.tiles64 || rc) continue;
            rc = tiles_launch(st, mixed ? m.tiles32 : m.tiles64, s.arena + g.x_off, s.arena + g.y_off, 0.0);
        }
        for (int wi = 0; wi < 4 && !rc; wi++)
            if (!sv[wi].empty())
                rc = batch_spmv_launch(st, ctx->num_cu, mixed ? SABLE_F32 : SABLE_F64, widths[wi],
                                       (const sable_batch_spmv_item_t*)(s.d_items + sofs[wi]), (int)sv[wi].size(),
                                       blk_total[wi], s.d_counter, 0);
        // the items come from the matrices' own plans, so no tile exceeds
        // bt_rows; the launch's flag word is the batch's error word, which
        // the solves below set only on failure
        for (int wi = 0; wi < 4 && !rc; wi++)
            if (!bt[wi].empty())
                rc = batch_tiles_launch(st, ctx->num_cu, mixed ? SABLE_F32 : SABLE_F64, widths[wi],
                                        (const sable_batch_tiles_item_t*)(s.d_items + tofs[wi]), (int)bt[wi].size(),
                                        unit_total[wi], bt_rows[wi], s.d_counter, s.d_err, 0);
        const auto* titems = (const sable_batch_spts_item_t*)(s.d_items + spts_off);
        const auto* ritems = (const sable_batch_refine_item_t*)(s.d_items + spts_off);
        auto* view = (sable_batch_spts_item_t*)(s.d_items + view_off);
        if (!rc && !tv.empty()) {
            if (cfg.spts_sweep == 1)
                rc = batch_spts_launch(st, ctx->num_cu, SABLE_F64, lanes, titems, (int)tv.size(), row_total,
                                       s.d_ticket, s.d_err, spin);
            else
                rc = batch_level_launch(st, SABLE_F64, lanes, titems, d_lch, plan);
        }
        if (!rc && !rv.empty()) {
            if (cfg.spts_sweep == 1)
                rc = batch_refine_launch(st, ctx->num_cu, lanes, ritems, view, (int)rv.size(), row_total,
                                         cfg.refine_iters, s.d_ticket, s.d_err, spin);
            else
                rc = batch_refine_level_launch(st, lanes, ritems, view, (int)rv.size(), cfg.refine_iters, d_lch,
                                               plan, d_rch, (int)rch.size());
        }
    }
    // results go back to the host only for requests that brought a host
    // output buffer (requests generated on the device keep theirs in HBM)
    bool any_out = false;
    for (auto* r : batch) any_out = any_out || r->out != nullptr;
    if (any_out)
        SABLE_HIP(hipMemcpyAsync(s.stage_out + s.out_begin, s.arena + s.out_begin, s.out_end - s.out_begin,
                                 hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipMemcpyAsync(s.h_err, s.d_err, sizeof(int), hipMemcpyDeviceToHost, st));
    SABLE_HIP(hipEventRecord(s.ev, st));
    return rc;
}

void sable_server_s::completer() {
    (void)hipSetDevice(ctx->device);
    while (true) {
        int si;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_complete.wait(lk, [&] { return !inflight.empty() || sched_done; });
            if (inflight.empty()) return;
            si = inflight.front();
        }
        Slot& s = slots[si];
        hipError_t e = hipEventSynchronize(s.ev);
        const double t = now_us();
        const bool mixed = cfg.mode == 2;
        const int errflag = *s.h_err;
        for (auto* r : s.reqs) {
            if (r->out) {
                const int32_t n = mats[r->mat].N;
                const size_t ld = (size_t)r->ld;
                if (r->op == 0 && mixed) {
                    const float* src = (const float*)(s.stage_out + r->out_off) + r->col;
                    for (int32_t i = 0; i < n; i++) r->out[i] = (double)src[i * ld];
                } else if (ld == 1) {
                    memcpy(r->out, s.stage_out + r->out_off, 8 * (size_t)n);
                } else {
                    const double* src = (const double*)(s.stage_out + r->out_off) + r->col;
                    for (int32_t i = 0; i < n; i++) r->out[i] = src[i * ld];
                }
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto* r : s.reqs) {
                r->t_done = t;
                r->status = s.launch_rc ? s.launch_rc
                                        : e != hipSuccess ? SABLE_EHIP : (errflag ? SABLE_ETIMEOUT : SABLE_OK);
                r->done = true;
                latencies.push_back(t - r->t_arrival);
                completed++;
            }
            t_last_done = t;
            s.reqs.clear();
            s.busy = false;
            inflight.pop_front();
        }
        cv_slot.notify_all();
        cv_done.notify_all();
    }
}

static void free_matrix(SMatrix& m) {
    if (m.plan) sable_csr_plan_destroy(m.plan);
    if (m.tiles64) sable_csr_tiles_destroy(m.tiles64);
    if (m.tiles32) sable_csr_tiles_destroy(m.tiles32);
    if (m.btiles) sable_csr_tiles_destroy(m.btiles);
    void* ps[] = {m.rp, m.col, m.val64, m.val32, m.perm};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m = SMatrix();
}

static void free_server(sable_server_s* s) {
    for (int32_t k = 0; k < s->mats.size(); k++) free_matrix(s->mats[k]);
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.arena) (void)hipFree(sl.arena);
        if (sl.d_items) (void)hipFree(sl.d_items);
        if (sl.d_chunks) (void)hipFree(sl.d_chunks);
        if (sl.h_chunks) (void)hipHostFree(sl.h_chunks);
        if (sl.d_counter) (void)hipFree(sl.d_counter);
        if (sl.stage_in) (void)hipHostFree(sl.stage_in);
        if (sl.stage_out) (void)hipHostFree(sl.stage_out);
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        if (sl.h_err) (void)hipHostFree(sl.h_err);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    delete s;
}

extern "C" {

SABLE_API int sable_server_create(sable_ctx_t* ctx, const sable_server_config_t* cfg, sable_server_t** out) {
    SABLE_CHECK(ctx && cfg && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(cfg->mode >= 0 && cfg->mode <= 2, SABLE_EINVAL, "mode must be 0, 1 or 2");
    SABLE_CHECK(cfg->spts_sweep == 0 || cfg->spts_sweep == 1, SABLE_EINVAL, "spts_sweep must be 0 or 1");
    *out = nullptr;
    auto* s = new sable_server_s();
    s->ctx = ctx;
    s->cfg = *cfg;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 256;
    if (s->cfg.slots <= 0) s->cfg.slots = 3;
    if (s->cfg.arena_bytes <= 0) s->cfg.arena_bytes = 256ll << 20;
    if (s->cfg.refine_iters < 0) s->cfg.refine_iters = 1;
    if (s->cfg.max_batch > 4096) s->cfg.max_batch = 4096;
    if (s->cfg.group_rhs <= 0) s->cfg.group_rhs = 8;
    if (s->cfg.group_rhs > 8) s->cfg.group_rhs = 8;
    const int64_t slot_bytes = s->cfg.arena_bytes + 4 * (int64_t)ITEM_BYTES + (int64_t)s->chunk_bytes() + 4096;
    if (cfg->hbm_budget > 0 && slot_bytes * s->cfg.slots > cfg->hbm_budget) {
        delete s;
        return fail(SABLE_EBUDGET, "batch arenas (%lld bytes) exceed the HBM budget", (long long)(slot_bytes * cfg->slots));
    }
    (void)hipSetDevice(ctx->device);
    s->slots.resize(s->cfg.slots);
    for (auto& sl : s->slots) {
        bool ok = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) == hipSuccess &&
                  hipMalloc(&sl.arena, (size_t)s->cfg.arena_bytes) == hipSuccess &&
                  hipMalloc(&sl.d_items, 4 * ITEM_BYTES) == hipSuccess &&
                  hipMalloc(&sl.d_counter, 4096) == hipSuccess &&
                  hipHostMalloc(&sl.stage_in, (size_t)s->cfg.arena_bytes, hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc(&sl.stage_out, (size_t)s->cfg.arena_bytes, hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc(&sl.h_items, 4 * ITEM_BYTES, hipHostMallocDefault) == hipSuccess &&
                  hipMalloc(&sl.d_chunks, s->chunk_bytes()) == hipSuccess &&
                  hipHostMalloc(&sl.h_chunks, s->chunk_bytes(), hipHostMallocDefault) == hipSuccess &&
                  hipHostMalloc(&sl.h_err, 256, hipHostMallocDefault) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            free_server(s);
            return fail(SABLE_ENOMEM, "server slot allocation failed");
        }
        sl.d_ticket = (unsigned*)((char*)sl.d_counter + 256);
        sl.d_err = (int*)((char*)sl.d_counter + 512);
        s->used += slot_bytes;
    }
    s->sched = std::thread([s] { s->scheduler(); });
    s->compl_ = std::thread([s] { s->completer(); });
    *out = s;
    return SABLE_OK;
}

SABLE_API void sable_server_destroy(sable_server_t* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_pending.notify_all();
    s->cv_slot.notify_all();
    s->cv_complete.notify_all();
    if (s->sched.joinable()) s->sched.join();
    s->cv_complete.notify_all();
    if (s->compl_.joinable()) s->compl_.join();
    free_server(s);
}

SABLE_API int sable_server_add_matrix(sable_server_t* s, int32_t N, const int32_t* row_ptr, const int32_t* col,
                                      const double* val, int kind, int32_t* matrix_id) {
    SABLE_CHECK(s && row_ptr && matrix_id && N > 0, SABLE_EINVAL, "bad argument");
    SABLE_CHECK(kind == 0 || kind == 1, SABLE_EINVAL, "kind must be 0 (general) or 1 (lower)");
    const int64_t nnz = (int64_t)row_ptr[N] - row_ptr[0];
    SABLE_CHECK(row_ptr[0] == 0 && nnz >= 0, SABLE_EMALFORMED, "row_ptr must start at 0 and be monotone");
    if (kind == 1) {
        for (int32_t i = 0; i < N; i++) {
            SABLE_CHECK(row_ptr[i + 1] > row_ptr[i] && col[row_ptr[i + 1] - 1] == i && val[row_ptr[i + 1] - 1] != 0.0,
                        SABLE_EMALFORMED, "row %d: the last entry must be a nonzero diagonal", i);
            for (int32_t j = row_ptr[i]; j < row_ptr[i + 1] - 1; j++)
                SABLE_CHECK(col[j] >= 0 && col[j] < i, SABLE_EMALFORMED, "row %d is not lower triangular", i);
        }
    }
    SMatrix m;
    m.N = N;
    m.nnz = nnz;
    m.kind = kind;
    // level schedule of a lower matrix (par/src/sparse32.js:2952-3007: level
    // 0 for a one-entry row, else 1 + the largest level among its other
    // columns; rows listed level by level, ascending within a level)
    std::vector<int32_t> h_lp, h_perm;
    if (kind == 1) {
        std::vector<int32_t> level(N);
        int32_t top = 0;
        for (int32_t i = 0; i < N; i++) {
            int32_t m = -1;
            for (int32_t j = row_ptr[i]; j < row_ptr[i + 1] - 1; j++) m = std::max(m, level[col[j]]);
            level[i] = m + 1;
            top = std::max(top, level[i]);
        }
        h_lp.assign((size_t)top + 2, 0);
        for (int32_t i = 0; i < N; i++) h_lp[(size_t)level[i] + 1]++;
        for (int32_t l = 0; l <= top; l++) h_lp[(size_t)l + 1] += h_lp[l];
        std::vector<int32_t> next(h_lp.begin(), h_lp.end() - 1);
        h_perm.resize(N);
        for (int32_t i = 0; i < N; i++) h_perm[next[level[i]]++] = i;
    }
    const size_t bytes = al(4 * (size_t)(N + 1)) + al(4 * (size_t)nnz) + al(8 * (size_t)nnz) + al(4 * (size_t)nnz) +
                         al(4 * ((size_t)(nnz + N) / 512 + 16)) + (kind == 1 ? al(4 * (size_t)N) : 0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->cfg.hbm_budget > 0 && s->used + (int64_t)bytes > s->cfg.hbm_budget)
            return fail(SABLE_EBUDGET, "matrix of %zu bytes exceeds the HBM budget (%lld of %lld used)", bytes,
                        (long long)s->used, (long long)s->cfg.hbm_budget);
        s->used += (int64_t)bytes;
    }
    m.bytes = bytes;
    hipStream_t st = s->ctx->stream;
    bool ok = hipMalloc(&m.rp, 4 * (size_t)(N + 1)) == hipSuccess &&
              hipMalloc(&m.col, 4 * (size_t)(nnz ? nnz : 1)) == hipSuccess &&
              hipMalloc(&m.val64, 8 * (size_t)(nnz ? nnz : 1)) == hipSuccess &&
              hipMalloc(&m.val32, 4 * (size_t)(nnz ? nnz : 1)) == hipSuccess &&
              (kind == 0 || hipMalloc(&m.perm, 4 * (size_t)N) == hipSuccess);
    int rc = SABLE_OK;
    if (!ok) {
        (void)hipGetLastError();
        rc = fail(SABLE_ENOMEM, "matrix allocation failed");
    } else if (hipMemcpyAsync(m.rp, row_ptr, 4 * (size_t)(N + 1), hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(m.col, col, 4 * (size_t)nnz, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpyAsync(m.val64, val, 8 * (size_t)nnz, hipMemcpyHostToDevice, st) != hipSuccess ||
               (kind == 1 && hipMemcpyAsync(m.perm, h_perm.data(), 4 * (size_t)N, hipMemcpyHostToDevice, st) !=
                                 hipSuccess)) {
        rc = fail(SABLE_EHIP, "matrix upload failed");
    } else {
        unsigned g = grid_for(nnz, 256);
        if (g > 16384) g = 16384;
        if (nnz) hipLaunchKernelGGL(k_f64_to_f32, dim3(g), dim3(256), 0, st, nnz, m.val64, m.val32);
        rc = sable_csr_plan_create(s->ctx, N, m.rp, nullptr, &m.plan);
        if (!rc) csr_plan_view(m.plan, &m.blk, &m.nblocks, &m.lds_cap);
        if (!rc && kind == 0 && N >= TILE_MIN_ROWS) {
            rc = sable_csr_tiles_create(s->ctx, SABLE_F64, N, m.rp, m.col, m.val64, 0, 0, &m.tiles64);
            if (!rc) rc = sable_csr_tiles_create(s->ctx, SABLE_F32, N, m.rp, m.col, m.val32, 0, 0, &m.tiles32);
            if (!rc) {
                int64_t b64 = 0, b32 = 0;
                sable_csr_tiles_info(m.tiles64, nullptr, nullptr, &b64);
                sable_csr_tiles_info(m.tiles32, nullptr, nullptr, &b32);
                const size_t extra = (size_t)(b64 + b32);
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->cfg.hbm_budget > 0 && s->used + (int64_t)extra > s->cfg.hbm_budget)
                    rc = fail(SABLE_EBUDGET, "tile plans of %zu bytes exceed the HBM budget", extra);
                else {
                    s->used += (int64_t)extra;
                    m.bytes += extra;
                }
            }
        }
        if (!rc && kind == 0 && N > 0 && N < TILE_MIN_ROWS && s->cfg.mode != 0 && s->cfg.batch_tiles) {
            // rows per tile: min(4096, 8192 / widest group), so that a
            // group's FP64 row sums take at most 64 KiB of LDS; one panel
            const int32_t tr = std::min(std::min(4096, 8192 / s->group_max()) / 64 * 64, (N + 63) / 64 * 64);
            int32_t pc = 64;
            while (pc < N) pc *= 2;
            const bool mx = s->cfg.mode == 2;
            rc = sable_csr_tiles_create(s->ctx, mx ? SABLE_F32 : SABLE_F64, N, m.rp, m.col,
                                        mx ? (const void*)m.val32 : (const void*)m.val64, tr, pc, &m.btiles);
            if (!rc) {
                int64_t bb = 0;
                sable_csr_tiles_info(m.btiles, nullptr, nullptr, &bb);
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->cfg.hbm_budget > 0 && s->used + bb > s->cfg.hbm_budget)
                    rc = fail(SABLE_EBUDGET, "batch tile plan of %lld bytes exceeds the HBM budget", (long long)bb);
                else {
                    s->used += bb;
                    m.bytes += (size_t)bb;
                }
            }
        }
    }
    if (rc) {
        // m.bytes counts the matrix and every plan already charged
        const int64_t charged = (int64_t)m.bytes;
        free_matrix(m);
        std::lock_guard<std::mutex> lk(s->mu);
        s->used -= charged;
        return rc;
    }
    SABLE_HIP(hipStreamSynchronize(st));
    m.lp = std::move(h_lp);
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t charged = (int64_t)m.bytes;
    const int32_t id = s->mats.push_back(std::move(m));
    if (id < 0) {
        free_matrix(m);
        s->used -= charged;
        return fail(SABLE_EINVAL, "the server holds at most %d matrices", MatrixTable::MAX);
    }
    *matrix_id = id;
    return SABLE_OK;
}

static int submit_req(sable_server_s* s, int32_t mid, int op, const double* in, double* out, bool seeded,
                      uint64_t seed, double t_arrival, int64_t* rid) {
    SABLE_CHECK(mid >= 0 && mid < (int32_t)s->mats.size(), SABLE_EINVAL, "unknown matrix %d", mid);
    SABLE_CHECK(op == 0 || op == 1, SABLE_EINVAL, "op must be 0 (SpMV) or 1 (lower solve)");
    SABLE_CHECK(op == 0 || s->mats[mid].kind == 1, SABLE_EINVAL, "matrix %d is not lower triangular", mid);
    SABLE_CHECK(seeded || in, SABLE_EINVAL, "NULL input");
    auto r = std::make_unique<SRequest>();
    r->mat = mid;
    r->op = op;
    r->in = in;
    r->out = out;
    r->seeded = seeded;
    r->seed = seed;
    r->t_arrival = t_arrival > 0 ? t_arrival : now_us();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        SABLE_CHECK(!s->stop, SABLE_EINVAL, "server is stopping");
        r->id = s->next_id++;
        if (rid) *rid = r->id;
        s->pending.push_back(r.get());
        s->reqs.emplace(r->id, std::move(r));
    }
    s->cv_pending.notify_one();
    return SABLE_OK;
}

SABLE_API int sable_server_submit(sable_server_t* s, int32_t matrix_id, int op, const double* in, double* out,
                                  int64_t* request_id) {
    SABLE_CHECK(s, SABLE_EINVAL, "NULL server");
    return submit_req(s, matrix_id, op, in, out, false, 0, 0, request_id);
}

SABLE_API int sable_server_wait(sable_server_t* s, int64_t request_id, double* latency_us) {
    SABLE_CHECK(s, SABLE_EINVAL, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto itr = s->reqs.find(request_id);
    SABLE_CHECK(itr != s->reqs.end(), SABLE_EINVAL, "unknown request %lld", (long long)request_id);
    SRequest* r = itr->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (latency_us) *latency_us = r->t_done - r->t_arrival;
    const int st = r->status;
    s->reqs.erase(itr);
    if (st) return fail(st, "request %lld failed", (long long)request_id);
    return SABLE_OK;
}

SABLE_API int sable_server_run_load(sable_server_t* s, const sable_load_t* load, sable_load_stats_t* out) {
    SABLE_CHECK(s && load && out, SABLE_EINVAL, "NULL argument");
    SABLE_CHECK(load->rate > 0 && load->duration_s > 0, SABLE_EINVAL, "rate and duration must be positive");
    std::vector<int32_t> general, lower;
    for (int32_t k = 0; k < (int32_t)s->mats.size(); k++) {
        general.push_back(k);
        if (s->mats[k].kind == 1) lower.push_back(k);
    }
    SABLE_CHECK(!general.empty(), SABLE_EINVAL, "no matrices");
    SABLE_CHECK(load->spts_fraction <= 0 || !lower.empty(), SABLE_EINVAL, "no lower-triangular matrices");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->latencies.clear();
        s->completed = s->rejected = s->batches = s->batched_reqs = 0;
    }
    std::mt19937_64 rng(load->seed ? load->seed : 1);
    std::exponential_distribution<double> gap(load->rate);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    const double t0 = now_us();
    double t = t0;
    int64_t submitted = 0;
    std::vector<int64_t> ids;
    while (true) {
        t += gap(rng) * 1e6;
        if (t - t0 > load->duration_s * 1e6) break;
        if (load->max_requests > 0 && submitted >= load->max_requests) break;
        while (now_us() < t) {
            const double d = t - now_us();
            if (d > 200) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(d - 100)));
        }
        const bool spts = u01(rng) < load->spts_fraction;
        const int32_t m = spts ? lower[rng() % lower.size()] : general[rng() % general.size()];
        int64_t id;
        int rc = submit_req(s, m, spts ? 1 : 0, nullptr, nullptr, true, rng(), t, &id);
        if (rc) return rc;
        ids.push_back(id);
        submitted++;
    }
    int64_t fails = 0;
    for (int64_t id : ids) {
        std::unique_lock<std::mutex> lk(s->mu);
        auto itr = s->reqs.find(id);
        if (itr == s->reqs.end()) continue;
        SRequest* r = itr->second.get();
        s->cv_done.wait(lk, [&] { return r->done; });
        if (r->status) fails++;
        s->reqs.erase(itr);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> lat = s->latencies;
    std::sort(lat.begin(), lat.end());
    memset(out, 0, sizeof *out);
    out->submitted = submitted;
    out->completed = s->completed;
    out->failed = fails;
    out->rejected = s->rejected;
    out->elapsed_s = (s->t_last_done - t0) * 1e-6;
    out->solves_per_s = out->elapsed_s > 0 ? (double)s->completed / out->elapsed_s : 0;
    if (!lat.empty()) {
        out->p50_us = lat[lat.size() / 2];
        out->p99_us = lat[std::min(lat.size() - 1, (size_t)(0.99 * (double)lat.size()))];
        out->max_us = lat.back();
        double sm = 0;
        for (double v : lat) sm += v;
        out->mean_us = sm / (double)lat.size();
    }
    out->batches = s->batches;
    out->mean_batch = s->batches ? (double)s->batched_reqs / (double)s->batches : 0;
    if (s->error) return fail(s->error, "server error: %s", s->errmsg);
    return SABLE_OK;
}

SABLE_API int sable_server_info(sable_server_t* s, int64_t* hbm_used, int32_t* nmatrices) {
    SABLE_CHECK(s, SABLE_EINVAL, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (hbm_used) *hbm_used = s->used;
    if (nmatrices) *nmatrices = (int32_t)s->mats.size();
    return SABLE_OK;
}

}  // extern "C"
