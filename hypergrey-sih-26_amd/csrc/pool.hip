// Instance pool and serving loop: many box QPs resident in one HBM budget,
// advanced together by graph-replayed check periods (pool_kernels.hip).
#include "device_util.hpp"
#include "hgqp.h"
#include "internal.hpp"
#include "pool_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace hgqp {
namespace {

// First-fit allocator of [0, capacity) ranges with coalescing on release.
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;  // start -> length
    void init(int64_t capacity) {
        free_.clear();
        if (capacity > 0) free_[0] = capacity;
    }
    int64_t alloc(int64_t len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            const int64_t s = it->first, rest = it->second - len;
            free_.erase(it);
            if (rest) free_[s + len] = rest;
            return s;
        }
        return -1;
    }
    void release(int64_t s, int64_t len) {
        auto it = free_.emplace(s, len).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct SlotHost {
    bool used = false;
    int64_t tag = 0;
    int64_t n = 0, nnz = 0;
    int64_t row0 = -1, ent0 = -1, blk0 = -1, nblk = 0;
    double* x_out = nullptr;
    std::vector<int32_t> start;  // local row starts (kept for re-blocking)
};

constexpr int kLoadsChoices[3] = {4, 8, 16};

}  // namespace
}  // namespace hgqp

using namespace hgqp;

struct hgqp_pool {
    hgqp_pool_options opt{};
    int loads = 8, rows_per_block = 256;
    bool tuned = false;
    int64_t max_blocks = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    PoolView v{};
    int* blk_r0 = nullptr;
    int* blk_r1 = nullptr;
    int* blk_inst = nullptr;
    PoolInst* table_host = nullptr;  // pinned mirror of the slot table
    hipEvent_t table_event = nullptr;
    bool table_pending = false;
    RangeAlloc rows, ents, blks;
    std::vector<SlotHost> slots;
    std::vector<int> free_slots;
    hipGraphExec_t graph = nullptr;
    int resident = 0;
    std::vector<hgqp_pool_result> ready;  // collected results not yet returned

    ~hgqp_pool() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        if (table_event) (void)hipEventDestroy(table_event);
        if (table_host) (void)hipHostFree(table_host);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void check_pool(const hgqp_pool* p) {
    if (!p) throw Error(HGQP_INVALID_ARGUMENT, "null hgqp_pool handle");
}

// Upper bound of row blocks for `rows` rows and `nnz` entries at the smallest
// block shape (cap 256 entries, 64 rows), plus one per instance.
int64_t block_budget(int64_t rows, int64_t nnz, int64_t inst) { return nnz / 256 + rows / 64 + 2 * inst + 16; }

size_t pool_bytes(const hgqp_pool_options& o) {
    const size_t R = size_t(o.max_rows), E = size_t(std::max<int64_t>(o.max_nnz, 1));
    const size_t B = size_t(block_budget(o.max_rows, o.max_nnz, o.max_instances)), S = size_t(o.max_instances);
    return Arena::span(R + 1, 4) + Arena::span(E, 4) + Arena::span(E, 8) + Arena::span(E, 4) +  // start col val val32
           3 * Arena::span(B, 4) +                                                            // blocks
           5 * Arena::span(R, 8) +                                                            // c lo up x g
           7 * Arena::span(R, 4) +                                                            // c32 lo32 up32 d[2] w[2]
           Arena::span(S, sizeof(PoolInst)) + 2 * Arena::span(B, 4) + Arena::span(4, 4) + Arena::span(S, 4) +
           Arena::span(B * kPartial, 8) + Arena::span(B, 8);
}

void validate_options(const hgqp_pool_options& o) {
    if (o.max_rows < 2 || o.max_rows > 2000000000ll || o.max_nnz < 1 || o.max_nnz > 2000000000ll ||
        o.max_instances < 1 || o.max_instances > 65536)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: need 2 <= max_rows <= 2e9, 1 <= max_nnz <= 2e9, "
                                           "1 <= max_instances <= 65536");
    if (o.check_every < 2 || o.check_every % 2)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: check_every must be even and >= 2");
    if (o.block_loads != 0 && o.block_loads != 4 && o.block_loads != 8 && o.block_loads != 16)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: block_loads must be 0, 4, 8 or 16");
    if (o.block_rows != 0 && o.block_rows != 64 && o.block_rows != 256)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: block_rows must be 0, 64 or 256");
    if (o.momentum != HGQP_MOMENTUM_RESTART && o.momentum != HGQP_MOMENTUM_REFERENCE)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: unknown momentum schedule");
    if (!(o.fp32_switch > 0) || !std::isfinite(o.fp32_switch))
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: fp32_switch must be positive");
}

// Writes one slot's blocks for the pool's block shape; returns false when the
// block budget has no room.
bool place_blocks(hgqp_pool* p, int s) {
    SlotHost& h = p->slots[s];
    std::vector<int> blk = row_blocks(h.n, h.start.data(), kWave * p->loads, p->rows_per_block);
    const int64_t nb = int64_t(blk.size()) - 1;
    const int64_t b0 = p->blks.alloc(nb);
    if (b0 < 0) return false;
    h.blk0 = b0;
    h.nblk = nb;
    std::vector<int> r0(nb), r1(nb), inst(nb, s);
    for (int64_t k = 0; k < nb; ++k) {
        r0[k] = int(h.row0 + blk[k]);
        r1[k] = int(h.row0 + blk[k + 1]);
    }
    HIPCHECK(hipMemcpyAsync(p->blk_r0 + b0, r0.data(), nb * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipMemcpyAsync(p->blk_r1 + b0, r1.data(), nb * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipMemcpyAsync(p->blk_inst + b0, inst.data(), nb * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));  // the host vectors go out of scope
    return true;
}

// Re-blocks every resident slot for the current shape and writes b0/nblk into
// the device table (used by the autotuner before the period graph exists).
void reblock_all(hgqp_pool* p) {
    p->blks.init(p->max_blocks);
    std::vector<PoolInst> table(p->opt.max_instances);
    HIPCHECK(hipMemcpyAsync(table.data(), p->v.inst, table.size() * sizeof(PoolInst), hipMemcpyDeviceToHost,
                            p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
    for (int s = 0; s < p->opt.max_instances; ++s) {
        if (!p->slots[s].used) continue;
        if (!place_blocks(p, s)) throw Error(HGQP_OUT_OF_MEMORY, "hgqp_pool: block budget exceeded while re-blocking");
        table[s].b0 = int(p->slots[s].blk0);
        table[s].nblk = int(p->slots[s].nblk);
    }
    HIPCHECK(hipMemcpyAsync(p->v.inst, table.data(), table.size() * sizeof(PoolInst), hipMemcpyHostToDevice,
                            p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
}

void enqueue_periods(hgqp_pool* p, int periods) {
    for (int k = 0; k < periods; ++k) {
        if (p->opt.use_graph) {
            if (!p->graph) {
                hipGraph_t g;
                HIPCHECK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
                launch_pool_period(p->v, p->loads, int(std::min<int64_t>(p->max_blocks, 1 << 30)), p->opt.check_every,
                                   p->stream);
                HIPCHECK(hipStreamEndCapture(p->stream, &g));
                hipError_t e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHECK(e);
            }
            HIPCHECK(hipGraphLaunch(p->graph, p->stream));
        } else {
            launch_pool_period(p->v, p->loads, int(std::min<int64_t>(p->max_blocks, 1 << 30)), p->opt.check_every,
                               p->stream);
            HIPCHECK(hipGetLastError());
        }
    }
    HIPCHECK(hipMemcpyAsync(p->table_host, p->v.inst, sizeof(PoolInst) * p->opt.max_instances, hipMemcpyDeviceToHost,
                            p->stream));
    HIPCHECK(hipEventRecord(p->table_event, p->stream));
    p->table_pending = true;
}

// Tunes block shape on the resident slots: times FP32 steps for each shape.
void autotune(hgqp_pool* p) {
    if (p->tuned) return;
    p->tuned = true;
    const bool tune_loads = p->opt.block_loads == 0, tune_rows = p->opt.block_rows == 0;
    if (!tune_loads && !tune_rows) return;
    if (!p->resident) return;  // nothing to time: keep the defaults
    // a step needs c32/lo32/up32 and d/w: they hold whatever the slots left; steps
    // are timed only, and the slots are re-initialized (phase INIT) afterwards
    std::vector<int> loads_c, rows_c;
    for (int l : kLoadsChoices)
        if (tune_loads || l == p->opt.block_loads) loads_c.push_back(l);
    for (int r : {64, 256})
        if (tune_rows || r == p->opt.block_rows) rows_c.push_back(r);
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    float best = std::numeric_limits<float>::infinity();
    int best_l = p->loads, best_r = p->rows_per_block;
    // make every resident slot "active" for the timing, then restore INIT
    std::vector<PoolInst> table(p->opt.max_instances), saved;
    HIPCHECK(hipMemcpyAsync(table.data(), p->v.inst, table.size() * sizeof(PoolInst), hipMemcpyDeviceToHost, p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
    saved = table;
    for (auto& t : table)
        if (t.phase != kSlotFree) { t.phase = kSlotActive; t.t[0] = t.t[1] = 1; t.use_x[0] = t.use_x[1] = 0; }
    for (int l : loads_c)
        for (int r : rows_c) {
            p->loads = l;
            p->rows_per_block = r;
            reblock_all(p);
            std::vector<PoolInst> t2(p->opt.max_instances);
            HIPCHECK(hipMemcpyAsync(t2.data(), p->v.inst, t2.size() * sizeof(PoolInst), hipMemcpyDeviceToHost, p->stream));
            HIPCHECK(hipStreamSynchronize(p->stream));
            for (int s = 0; s < p->opt.max_instances; ++s) { table[s].b0 = t2[s].b0; table[s].nblk = t2[s].nblk; }
            HIPCHECK(hipMemcpyAsync(p->v.inst, table.data(), table.size() * sizeof(PoolInst), hipMemcpyHostToDevice,
                                    p->stream));
            launch_pool_lists(p->v, p->stream);
            const int bmax = int(std::min<int64_t>(p->max_blocks, 1 << 30));
            launch_pool_steps(p->v, l, bmax, 2, p->stream);  // warm
            HIPCHECK(hipEventRecord(a, p->stream));
            launch_pool_steps(p->v, l, bmax, 6, p->stream);
            HIPCHECK(hipEventRecord(b, p->stream));
            HIPCHECK(hipEventSynchronize(b));
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, a, b));
            if (ms < best) { best = ms; best_l = l; best_r = r; }
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    p->loads = best_l;
    p->rows_per_block = best_r;
    reblock_all(p);
    std::vector<PoolInst> t2(p->opt.max_instances);
    HIPCHECK(hipMemcpyAsync(t2.data(), p->v.inst, t2.size() * sizeof(PoolInst), hipMemcpyDeviceToHost, p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
    for (int s = 0; s < p->opt.max_instances; ++s) { saved[s].b0 = t2[s].b0; saved[s].nblk = t2[s].nblk; }
    HIPCHECK(hipMemcpyAsync(p->v.inst, saved.data(), saved.size() * sizeof(PoolInst), hipMemcpyHostToDevice, p->stream));
    launch_pool_lists(p->v, p->stream);
    HIPCHECK(hipStreamSynchronize(p->stream));
}

hgqp_pool* create_pool(const hgqp_pool_options& o) {
    validate_options(o);
    require_device();
    auto p = std::make_unique<hgqp_pool>();
    p->opt = o;
    p->loads = o.block_loads ? o.block_loads : 8;
    p->rows_per_block = o.block_rows ? o.block_rows : 256;
    p->tuned = o.block_loads && o.block_rows;
    p->max_blocks = block_budget(o.max_rows, o.max_nnz, o.max_instances);
    HIPCHECK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    p->arena.reserve(pool_bytes(o));
    const size_t R = size_t(o.max_rows), E = size_t(o.max_nnz), B = size_t(p->max_blocks), S = size_t(o.max_instances);
    PoolView& v = p->v;
    int* start = p->arena.take<int>(R + 1);
    int* col = p->arena.take<int>(E);
    double* val = p->arena.take<double>(E);
    float* val32 = p->arena.take<float>(E);
    p->blk_r0 = p->arena.take<int>(B);
    p->blk_r1 = p->arena.take<int>(B);
    p->blk_inst = p->arena.take<int>(B);
    v.start = start;
    v.col = col;
    v.val = val;
    v.val32 = val32;
    v.blk_r0 = p->blk_r0;
    v.blk_r1 = p->blk_r1;
    v.blk_inst = p->blk_inst;
    double* c = p->arena.take<double>(R);
    double* lo = p->arena.take<double>(R);
    double* up = p->arena.take<double>(R);
    v.c = c;
    v.lo = lo;
    v.up = up;
    v.x = p->arena.take<double>(R);
    v.g = p->arena.take<double>(R);
    v.c32 = p->arena.take<float>(R);
    v.lo32 = p->arena.take<float>(R);
    v.up32 = p->arena.take<float>(R);
    for (int k = 0; k < 2; ++k) {
        v.d[k] = p->arena.take<float>(R);
        v.w[k] = p->arena.take<float>(R);
    }
    v.inst = p->arena.take<PoolInst>(S);
    v.max_inst = int(S);
    v.list_active = p->arena.take<int>(B);
    v.list_verify = p->arena.take<int>(B);
    v.counts = p->arena.take<int>(4);
    v.offsets = p->arena.take<int>(S);
    v.partials = p->arena.take<double>(B * kPartial);
    v.resid = p->arena.take<double>(B);
    v.switch_stat = o.fp32_switch;
    v.restart = o.momentum == HGQP_MOMENTUM_RESTART ? 1 : 0;
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&p->table_host), sizeof(PoolInst) * S, hipHostMallocDefault));
    std::memset(p->table_host, 0, sizeof(PoolInst) * S);  // all kSlotFree
    HIPCHECK(hipMemcpyAsync(v.inst, p->table_host, sizeof(PoolInst) * S, hipMemcpyHostToDevice, p->stream));
    int zero[4] = {0, 0, 0, 0};
    HIPCHECK(hipMemcpyAsync(v.counts, zero, sizeof zero, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipEventCreateWithFlags(&p->table_event, hipEventDisableTiming));
    HIPCHECK(hipStreamSynchronize(p->stream));
    p->rows.init(o.max_rows);
    p->ents.init(o.max_nnz);
    p->blks.init(p->max_blocks);
    p->slots.resize(S);
    for (int s = int(S) - 1; s >= 0; --s) p->free_slots.push_back(s);
    return p.release();
}

// Admission (host side, between periods). Returns HGQP_BUSY through Error.
int admit(hgqp_pool* p, int64_t n, int64_t nnz, const int32_t* start, const int32_t* index, const double* value,
          const double* cost, const double* lower, const double* upper, double lipschitz, int32_t maximize,
          double offset, double tol, uint64_t limit, int64_t tag, double* x_out) {
    if (n <= 0 || nnz < 0 || !start || (nnz && (!index || !value)) || !cost || !lower || !upper || !x_out)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_admit: null array or bad size");
    if (!(lipschitz > 0) || !std::isfinite(lipschitz))
        throw Error(HGQP_UNSUPPORTED, "hgqp_pool_admit: needs a positive finite Lipschitz bound");
    if (!(tol > 0) || !std::isfinite(tol)) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_admit: bad tolerance");
    if (start[0] != 0 || start[n] != nnz) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_admit: malformed CSR start");
    bool diag = true;
    for (int64_t i = 0; i < n; ++i) {
        if (start[i + 1] < start[i]) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_admit: CSR start not monotone");
        bool pos = false;
        for (int32_t k = start[i]; k < start[i + 1]; ++k) {
            if (index[k] < 0 || index[k] >= n) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_admit: column out of range");
            if (index[k] == i) pos = value[k] > 0;
            else diag = false;
        }
        diag = diag && pos;
        if (std::isnan(lower[i]) || std::isnan(upper[i]) || lower[i] > upper[i])
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_admit: bounds must satisfy lower <= upper");
    }
    if (p->free_slots.empty()) return HGQP_BUSY;
    const int64_t r0 = p->rows.alloc(n + 1);
    if (r0 < 0) return HGQP_BUSY;
    const int64_t e0 = p->ents.alloc(std::max<int64_t>(nnz, 1));
    if (e0 < 0) { p->rows.release(r0, n + 1); return HGQP_BUSY; }
    const int s = p->free_slots.back();
    SlotHost& h = p->slots[s];
    h.n = n;
    h.nnz = nnz;
    h.row0 = r0;
    h.ent0 = e0;
    h.start.assign(start, start + n + 1);
    if (!place_blocks(p, s)) {
        p->rows.release(r0, n + 1);
        p->ents.release(e0, std::max<int64_t>(nnz, 1));
        return HGQP_BUSY;
    }
    p->free_slots.pop_back();
    h.used = true;
    h.tag = tag;
    h.x_out = x_out;
    // matrix in global numbering: starts offset by e0, columns by r0
    std::vector<int32_t> gs(n + 1), gc(nnz);
    for (int64_t i = 0; i <= n; ++i) gs[i] = int32_t(e0 + start[i]);
    for (int64_t k = 0; k < nnz; ++k) gc[k] = int32_t(r0 + index[k]);
    hipStream_t st = p->stream;
    HIPCHECK(hipMemcpyAsync(const_cast<int*>(p->v.start) + r0, gs.data(), (n + 1) * 4, hipMemcpyHostToDevice, st));
    if (nnz) {
        HIPCHECK(hipMemcpyAsync(const_cast<int*>(p->v.col) + e0, gc.data(), nnz * 4, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.val) + e0, value, nnz * 8, hipMemcpyHostToDevice, st));
        launch_convert<float, double>(nnz, p->v.val + e0, const_cast<float*>(p->v.val32) + e0, st);
    }
    HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.c) + r0, cost, n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.lo) + r0, lower, n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.up) + r0, upper, n * 8, hipMemcpyHostToDevice, st));
    PoolInst t;
    std::memset(&t, 0, sizeof t);
    t.lipschitz = lipschitz;
    t.tol = tol;
    t.t[0] = t.t[1] = 1.0;
    t.phase = kSlotInit;
    t.status = HGQP_STATUS_RUNNING;
    t.limit = int64_t(std::min<uint64_t>(limit, uint64_t(std::numeric_limits<int64_t>::max())));
    t.best_stat = std::numeric_limits<double>::infinity();
    t.sign = maximize ? -1.0 : 1.0;
    t.offset = offset;
    t.diagonal = diag ? 1 : 0;
    t.b0 = int(h.blk0);
    t.nblk = int(h.nblk);
    t.r0 = int(r0);
    t.n = int(n);
    HIPCHECK(hipMemcpyAsync(p->v.inst + s, &t, sizeof t, hipMemcpyHostToDevice, st));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));  // gs / gc / t are host temporaries
    ++p->resident;
    return HGQP_OK;
}

// Reads the table after the last enqueued periods; finished slots are copied
// out, freed and queued in p->ready.
void harvest(hgqp_pool* p) {
    if (!p->table_pending) {
        HIPCHECK(hipMemcpyAsync(p->table_host, p->v.inst, sizeof(PoolInst) * p->opt.max_instances,
                                hipMemcpyDeviceToHost, p->stream));
        HIPCHECK(hipEventRecord(p->table_event, p->stream));
    }
    HIPCHECK(hipEventSynchronize(p->table_event));
    p->table_pending = false;
    std::vector<int> done;
    for (int s = 0; s < p->opt.max_instances; ++s)
        if (p->slots[s].used && p->table_host[s].phase == kSlotDone) done.push_back(s);
    if (done.empty()) return;
    const int free_phase = kSlotFree;
    for (int s : done) {
        SlotHost& h = p->slots[s];
        HIPCHECK(hipMemcpyAsync(h.x_out, p->v.x + h.row0, h.n * 8, hipMemcpyDeviceToHost, p->stream));
        HIPCHECK(hipMemcpyAsync(&p->v.inst[s].phase, &free_phase, sizeof(int), hipMemcpyHostToDevice, p->stream));
    }
    HIPCHECK(hipStreamSynchronize(p->stream));
    for (int s : done) {
        SlotHost& h = p->slots[s];
        const PoolInst& t = p->table_host[s];
        hgqp_pool_result r;
        std::memset(&r, 0, sizeof r);
        r.tag = h.tag;
        r.slot = s;
        r.status = t.status;
        r.iterations = uint64_t(t.iter);
        r.restarts = t.restarts;
        r.checks = t.checks;
        r.cycles = t.cycle + 1;
        r.check.objective = t.objective;
        r.check.primal_residual = t.primal_residual;
        r.check.stationarity_residual = t.stationarity;
        r.check.complementarity_residual = std::isfinite(t.objective) ? 0.0 : std::numeric_limits<double>::infinity();
        r.check.dual_bound = t.dual_bound;
        r.check.relative_gap = t.relative_gap;
        r.check.optimality_verified = t.verified;
        r.check.finite = t.finite;
        p->ready.push_back(r);
        p->rows.release(h.row0, h.n + 1);
        p->ents.release(h.ent0, std::max<int64_t>(h.nnz, 1));
        p->blks.release(h.blk0, h.nblk);
        h = SlotHost();
        p->free_slots.push_back(s);
        --p->resident;
    }
}

}  // namespace

// ---- serving loop -------------------------------------------------------------------
struct hgqp_server {
    hgqp_pool* pool = nullptr;
    std::thread thread;
    std::mutex mu;
    std::condition_variable cv;
    bool stopping = false;
    std::string error;
    struct Request {
        int64_t n, nnz;
        std::vector<int32_t> start, index;
        std::vector<double> value, cost, lower, upper;
        double lipschitz, offset, tol;
        int32_t maximize;
        uint64_t limit;
        int64_t tag;
        double* x_out;
        std::chrono::steady_clock::time_point submitted, admitted;
    };
    std::deque<std::unique_ptr<Request>> queue;              // submitted, not admitted
    std::map<int64_t, std::unique_ptr<Request>> resident;     // by slot
    std::deque<hgqp_served> done;
    int64_t outstanding = 0;                                 // submitted, not polled

    void loop() {
        try {
            for (;;) {
                std::vector<std::unique_ptr<Request>> admit_now;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stopping || !queue.empty() || !resident.empty(); });
                    if (stopping && queue.empty() && resident.empty()) return;
                    while (!queue.empty()) {
                        admit_now.push_back(std::move(queue.front()));
                        queue.pop_front();
                    }
                }
                // admit in arrival order while slots and budget allow; the rest wait
                std::vector<std::unique_ptr<Request>> back;
                bool full = false;
                for (auto& r : admit_now) {
                    if (full) { back.push_back(std::move(r)); continue; }
                    const int s = pool->free_slots.empty() ? -1 : pool->free_slots.back();
                    int rc = admit(pool, r->n, r->nnz, r->start.data(), r->index.data(), r->value.data(),
                                   r->cost.data(), r->lower.data(), r->upper.data(), r->lipschitz, r->maximize,
                                   r->offset, r->tol, r->limit, r->tag, r->x_out);
                    if (rc == HGQP_BUSY) { full = true; back.push_back(std::move(r)); continue; }
                    r->admitted = std::chrono::steady_clock::now();
                    r->start.clear(); r->start.shrink_to_fit();
                    r->index.clear(); r->index.shrink_to_fit();
                    r->value.clear(); r->value.shrink_to_fit();
                    std::lock_guard<std::mutex> lk(mu);
                    resident[s] = std::move(r);
                }
                if (!back.empty()) {
                    std::lock_guard<std::mutex> lk(mu);
                    for (auto it = back.rbegin(); it != back.rend(); ++it) queue.push_front(std::move(*it));
                }
                if (pool->resident == 0) continue;
                if (!pool->tuned) autotune(pool);
                enqueue_periods(pool, 1);
                harvest(pool);
                if (!pool->ready.empty()) {
                    const auto now = std::chrono::steady_clock::now();
                    std::lock_guard<std::mutex> lk(mu);
                    for (const auto& res : pool->ready) {
                        auto it = resident.find(res.slot);
                        hgqp_served sv;
                        sv.result = res;
                        sv.queued_seconds = std::chrono::duration<double>(it->second->admitted - it->second->submitted).count();
                        sv.latency_seconds = std::chrono::duration<double>(now - it->second->submitted).count();
                        done.push_back(sv);
                        resident.erase(it);
                    }
                    pool->ready.clear();
                }
            }
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(mu);
            error = e.what();
            stopping = true;
        }
    }

    ~hgqp_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = true;
        }
        cv.notify_all();
        if (thread.joinable()) thread.join();
        delete pool;
    }
};

extern "C" {

void hgqp_pool_options_default(hgqp_pool_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_rows = 1 << 22;
    o->max_nnz = 1 << 26;
    o->max_instances = 1024;
    o->check_every = 10;
    o->block_loads = 8;
    o->block_rows = 256;
    o->use_graph = 1;
    o->momentum = HGQP_MOMENTUM_RESTART;
    o->fp32_switch = 10.0;
}

int hgqp_pool_bytes(const hgqp_pool_options* options, int64_t* bytes) {
    return guarded([&] {
        if (!options || !bytes) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_bytes: null argument");
        validate_options(*options);
        *bytes = int64_t(pool_bytes(*options));
    });
}

int hgqp_pool_create(const hgqp_pool_options* options, hgqp_pool** out) {
    return guarded([&] {
        if (!out) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_create: null output");
        *out = nullptr;
        hgqp_pool_options o;
        if (options) o = *options;
        else hgqp_pool_options_default(&o);
        *out = create_pool(o);
    });
}

int hgqp_pool_admit(hgqp_pool* pool, int64_t n, int64_t nnz, const int32_t* start, const int32_t* index,
                    const double* value, const double* cost, const double* lower, const double* upper,
                    double lipschitz, int32_t maximize, double offset, double tolerance, uint64_t iteration_limit,
                    int64_t tag, double* x_out, int32_t* slot) {
    int rc = HGQP_OK;
    int g = guarded([&] {
        check_pool(pool);
        const int s = pool->free_slots.empty() ? -1 : pool->free_slots.back();
        rc = admit(pool, n, nnz, start, index, value, cost, lower, upper, lipschitz, maximize, offset, tolerance,
                   iteration_limit, tag, x_out);
        if (rc == HGQP_BUSY) set_last_error("hgqp_pool_admit: no free slot or no room in the HBM budget");
        if (slot) *slot = rc == HGQP_OK ? s : -1;
    });
    return g != HGQP_OK ? g : rc;
}

int hgqp_pool_run(hgqp_pool* pool, int32_t periods) {
    return guarded([&] {
        check_pool(pool);
        if (periods < 1) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_run: periods must be >= 1");
        if (!pool->tuned) autotune(pool);
        enqueue_periods(pool, periods);
    });
}

int hgqp_pool_collect(hgqp_pool* pool, hgqp_pool_result* results, int32_t capacity, int32_t* count,
                      int32_t* active) {
    return guarded([&] {
        check_pool(pool);
        if (!count || capacity < 0 || (capacity && !results))
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_collect: bad output");
        harvest(pool);
        const int32_t k = int32_t(std::min<size_t>(size_t(capacity), pool->ready.size()));
        for (int32_t i = 0; i < k; ++i) results[i] = pool->ready[i];
        pool->ready.erase(pool->ready.begin(), pool->ready.begin() + k);
        *count = k;
        if (active) *active = pool->resident;
    });
}

int hgqp_pool_autotune(hgqp_pool* pool, int32_t* block_loads, int32_t* block_rows) {
    return guarded([&] {
        check_pool(pool);
        autotune(pool);
        if (block_loads) *block_loads = pool->loads;
        if (block_rows) *block_rows = pool->rows_per_block;
    });
}

int hgqp_pool_destroy(hgqp_pool* pool) {
    return guarded([&] { delete pool; });
}

int hgqp_server_start(const hgqp_pool_options* options, hgqp_server** out) {
    return guarded([&] {
        if (!out) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_start: null output");
        *out = nullptr;
        hgqp_pool_options o;
        if (options) o = *options;
        else hgqp_pool_options_default(&o);
        auto s = std::make_unique<hgqp_server>();
        s->pool = create_pool(o);
        hgqp_server* raw = s.get();
        s->thread = std::thread([raw] { raw->loop(); });
        *out = s.release();
    });
}

int hgqp_server_submit(hgqp_server* server, int64_t n, int64_t nnz, const int32_t* start, const int32_t* index,
                       const double* value, const double* cost, const double* lower, const double* upper,
                       double lipschitz, int32_t maximize, double offset, double tolerance,
                       uint64_t iteration_limit, int64_t tag, double* x_out) {
    return guarded([&] {
        if (!server) throw Error(HGQP_INVALID_ARGUMENT, "null hgqp_server handle");
        if (n <= 0 || nnz < 0 || !start || (nnz && (!index || !value)) || !cost || !lower || !upper || !x_out ||
            start[0] != 0 || start[n] != nnz)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit: null array or malformed CSR");
        if (n + 1 > server->pool->opt.max_rows || nnz > server->pool->opt.max_nnz)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit: instance larger than the pool budget");
        auto r = std::make_unique<hgqp_server::Request>();
        r->n = n;
        r->nnz = nnz;
        r->start.assign(start, start + n + 1);
        r->index.assign(index, index + nnz);
        r->value.assign(value, value + nnz);
        r->cost.assign(cost, cost + n);
        r->lower.assign(lower, lower + n);
        r->upper.assign(upper, upper + n);
        r->lipschitz = lipschitz;
        r->maximize = maximize;
        r->offset = offset;
        r->tol = tolerance;
        r->limit = iteration_limit;
        r->tag = tag;
        r->x_out = x_out;
        r->submitted = std::chrono::steady_clock::now();
        {
            std::lock_guard<std::mutex> lk(server->mu);
            if (!server->error.empty()) throw Error(HGQP_DEVICE_ERROR, "hgqp_server: " + server->error);
            if (server->stopping) throw Error(HGQP_BUSY, "hgqp_server: stopping");
            server->queue.push_back(std::move(r));
            ++server->outstanding;
        }
        server->cv.notify_all();
    });
}

int hgqp_server_poll(hgqp_server* server, hgqp_served* results, int32_t capacity, int32_t* count) {
    return guarded([&] {
        if (!server || !count || capacity < 0 || (capacity && !results))
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_poll: bad argument");
        std::lock_guard<std::mutex> lk(server->mu);
        int32_t k = 0;
        while (k < capacity && !server->done.empty()) {
            results[k++] = server->done.front();
            server->done.pop_front();
        }
        server->outstanding -= k;
        *count = k;
        if (k == 0 && !server->error.empty()) throw Error(HGQP_DEVICE_ERROR, "hgqp_server: " + server->error);
    });
}

int hgqp_server_pending(hgqp_server* server, int64_t* pending) {
    return guarded([&] {
        if (!server || !pending) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_pending: bad argument");
        std::lock_guard<std::mutex> lk(server->mu);
        *pending = server->outstanding;
    });
}

int hgqp_server_stop(hgqp_server* server) {
    return guarded([&] { delete server; });
}

}  // extern "C"
