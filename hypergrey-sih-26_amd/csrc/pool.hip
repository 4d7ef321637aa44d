// Instance pool and serving loop: many box QPs resident in one HBM budget,
// advanced together by graph-replayed check periods (pool_kernels.hip).
#include "device_util.hpp"
#include "hgqp.h"
#include "ingest.hpp"
#include "internal.hpp"
#include "pool_kernels.hpp"
#include "slabs.hpp"

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
    int64_t rows_len = 0;              // rows held in the row range
    int64_t sblk0 = -1, nsblk = 0;     // slab blocks (pools with slabs)
    std::vector<int> srb;              // their first rows (local), then n
    std::vector<int32_t> srb_stage;    // host staging of their slot and local indices
    std::shared_ptr<SlabStore> slab;   // the instance's slab layout
    PoolSlab slab_desc{};              // host copy of its device descriptor
    double* x_out = nullptr;
    int32_t gen = 0;             // admission number, mirrored in the device slot entry
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
    int* sblk_inst = nullptr;        // slab blocks: owning slot and local block (pools with slabs)
    int* sblk_local = nullptr;
    PoolSlab* slab_dev = nullptr;    // per slot
    PoolInst* table_host = nullptr;  // pinned mirror of the slot table
    double* x_stage = nullptr;       // pinned staging of collected iterates (grown on demand)
    size_t x_stage_bytes = 0;
    int32_t next_gen = 1;
    hipEvent_t table_event = nullptr;
    bool table_pending = false;
    RangeAlloc rows, ents, blks, sblks;
    // slab layouts of released slots: freed at the next admission (which waits
    // for the stream anyway), so a release never stalls the queued periods
    std::vector<std::shared_ptr<SlabStore>> retired;
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
        if (x_stage) (void)hipHostFree(x_stage);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void check_pool(const hgqp_pool* p) {
    if (!p) throw Error(HGQP_INVALID_ARGUMENT, "null hgqp_pool handle");
}

// Upper bound of row blocks for `rows` rows and `nnz` entries at the smallest
// block shape (cap 256 entries, 64 rows), plus one per instance. It also bounds
// the slab blocks of a pool with slabs (at most 4000 rows and about 64K
// entries each, plus one per instance).
int64_t block_budget(int64_t rows, int64_t nnz, int64_t inst) { return nnz / 256 + rows / 64 + 2 * inst + 16; }

// Rows the row allocator hands out: with slabs every instance starts at a
// multiple of 4 rows (16-byte aligned vectors for the slab copies).
int64_t row_capacity(const hgqp_pool_options& o) {
    return o.max_rows + (o.slabs ? 3 * int64_t(o.max_instances) : 0);
}

// Length of the row vectors: with slabs a slab of padding follows the last
// row, so every slab copy of an instance stays inside the vector.
int64_t vector_rows(const hgqp_pool_options& o) { return row_capacity(o) + (o.slabs ? kSlabWidth : 0); }

size_t pool_bytes(const hgqp_pool_options& o) {
    const size_t Ra = size_t(row_capacity(o)), R = size_t(vector_rows(o)), E = size_t(std::max<int64_t>(o.max_nnz, 1));
    const size_t B = size_t(block_budget(o.max_rows, o.max_nnz, o.max_instances)), S = size_t(o.max_instances);
    size_t bytes = Arena::span(Ra + 1, 4) + Arena::span(E, 4) + Arena::span(E, 8) + Arena::span(E, 4) +  // start col val val32
                   3 * Arena::span(B, 4) +                                                             // blocks
                   5 * Arena::span(R, 8) +                                                             // c lo up x g
                   8 * Arena::span(R, 4) +                                             // c32 lo32 up32 d[2] w[2] dstep
                   Arena::span(S, sizeof(PoolInst)) + 2 * Arena::span(B, 4) + Arena::span(4, 4) + Arena::span(S, 4) +
                   Arena::span(B * kPartial, 8) + 2 * Arena::span(B, 8);
    if (o.slabs) bytes += 2 * Arena::span(B, 4) + Arena::span(S, sizeof(PoolSlab));
    return bytes;
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
    if (o.step_metric != HGQP_STEP_SCALAR && o.step_metric != HGQP_STEP_JACOBI)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: unknown step metric");
    if (o.slab_block_entries != 0 && (o.slab_block_entries < 4096 || o.slab_block_entries > (1 << 22)))
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: slab_block_entries must be 0 or in [4096, 4194304]");
    if (o.slabs != 0 && o.slabs != 1) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool: slabs must be 0 or 1");
}

// Computes one slot's row blocks for the pool's block shape into `stage`
// (3 x nblk ints: first rows, end rows, slot; room for 3 (n + 1)) and enqueues
// their upload. `stage` must stay valid until the stream passes this point.
// Returns false when the block budget has no room.
bool place_blocks(hgqp_pool* p, int s, const int32_t* start, int32_t* stage) {
    SlotHost& h = p->slots[s];
    std::vector<int> blk = row_blocks(h.n, start, kWave * p->loads, p->rows_per_block);
    const int64_t nb = int64_t(blk.size()) - 1;
    const int64_t b0 = p->blks.alloc(nb);
    if (b0 < 0) return false;
    h.blk0 = b0;
    h.nblk = nb;
    int32_t *r0 = stage, *r1 = stage + nb, *inst = stage + 2 * nb;
    for (int64_t k = 0; k < nb; ++k) {
        r0[k] = int32_t(h.row0 + blk[k]);
        r1[k] = int32_t(h.row0 + blk[k + 1]);
        inst[k] = s;
    }
    HIPCHECK(hipMemcpyAsync(p->blk_r0 + b0, r0, nb * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipMemcpyAsync(p->blk_r1 + b0, r1, nb * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipMemcpyAsync(p->blk_inst + b0, inst, nb * 4, hipMemcpyHostToDevice, p->stream));
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
    std::vector<std::vector<int32_t>> stages(p->opt.max_instances);
    for (int s = 0; s < p->opt.max_instances; ++s) {
        if (!p->slots[s].used) continue;
        stages[s].resize(3 * (p->slots[s].n + 1));
        if (!place_blocks(p, s, p->slots[s].start.data(), stages[s].data()))
            throw Error(HGQP_OUT_OF_MEMORY, "hgqp_pool: block budget exceeded while re-blocking");
        table[s].b0 = int(p->slots[s].blk0);
        table[s].nblk = int(p->slots[s].nblk);
    }
    HIPCHECK(hipMemcpyAsync(p->v.inst, table.data(), table.size() * sizeof(PoolInst), hipMemcpyHostToDevice,
                            p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
}

// Enqueues `periods` check periods (graph replays after the first capture).
void launch_periods(hgqp_pool* p, int periods) {
    for (int k = 0; k < periods; ++k) {
        if (p->opt.use_graph) {
            if (!p->graph) {
                hipGraph_t g;
                // capture on a drained stream: the admissions' uploads and first
                // checks queued before the first period complete first
                HIPCHECK(hipStreamSynchronize(p->stream));
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
}

// Periods, then a snapshot of the slot table for hgqp_pool_collect.
void enqueue_periods(hgqp_pool* p, int periods) {
    launch_periods(p, periods);
    HIPCHECK(hipMemcpyAsync(p->table_host, p->v.inst, sizeof(PoolInst) * p->opt.max_instances, hipMemcpyDeviceToHost,
                            p->stream));
    HIPCHECK(hipEventRecord(p->table_event, p->stream));
    p->table_pending = true;
}

// Tunes block shape on the resident slots: times FP32 steps for each shape.
// Times `reps` FP32 steps with every resident slot active at the current block
// shape (2 untimed steps first), then restores the slot table and work lists.
// Returns the average step time in microseconds.
double time_all_active(hgqp_pool* p, int reps) {
    std::vector<PoolInst> table(p->opt.max_instances);
    HIPCHECK(hipMemcpyAsync(table.data(), p->v.inst, table.size() * sizeof(PoolInst), hipMemcpyDeviceToHost, p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
    std::vector<PoolInst> active = table;
    for (auto& t : active)
        if (t.phase != kSlotFree) { t.phase = kSlotActive; t.t[0] = t.t[1] = 1; t.use_x[0] = t.use_x[1] = 0; }
    HIPCHECK(hipMemcpyAsync(p->v.inst, active.data(), active.size() * sizeof(PoolInst), hipMemcpyHostToDevice,
                            p->stream));
    launch_pool_lists(p->v, p->stream);
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    const int bmax = int(std::min<int64_t>(p->max_blocks, 1 << 30));
    launch_pool_steps(p->v, p->loads, bmax, 2, p->stream);  // warm
    HIPCHECK(hipEventRecord(a, p->stream));
    launch_pool_steps(p->v, p->loads, bmax, reps, p->stream);
    HIPCHECK(hipEventRecord(b, p->stream));
    hipError_t e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIPCHECK(e);
    HIPCHECK(hipMemcpyAsync(p->v.inst, table.data(), table.size() * sizeof(PoolInst), hipMemcpyHostToDevice, p->stream));
    launch_pool_lists(p->v, p->stream);
    HIPCHECK(hipStreamSynchronize(p->stream));
    return double(ms) * 1e3 / reps;
}

void autotune(hgqp_pool* p) {
    if (p->tuned) return;
    p->tuned = true;
    const bool tune_loads = p->opt.block_loads == 0, tune_rows = p->opt.block_rows == 0;
    if (!tune_loads && !tune_rows) return;
    if (!p->resident) return;  // nothing to time: keep the defaults
    if (p->opt.slabs) {
        // the steps run over slabs; the row blocks serve the FP64 checks, whose
        // best shape on the batched workload was 16 loads and 256 rows
        p->loads = tune_loads ? 16 : p->opt.block_loads;
        p->rows_per_block = tune_rows ? 256 : p->opt.block_rows;
        reblock_all(p);
        launch_pool_lists(p->v, p->stream);
        HIPCHECK(hipStreamSynchronize(p->stream));
        return;
    }
    // Steps are timed only: they run on whatever c32/lo32/up32 and d/w the slots
    // hold, and time_all_active restores the slot table afterwards.
    std::vector<int> loads_c, rows_c;
    for (int l : kLoadsChoices)
        if (tune_loads || l == p->opt.block_loads) loads_c.push_back(l);
    for (int r : {64, 256})
        if (tune_rows || r == p->opt.block_rows) rows_c.push_back(r);
    double best = std::numeric_limits<double>::infinity();
    int best_l = p->loads, best_r = p->rows_per_block;
    for (int l : loads_c)
        for (int r : rows_c) {
            p->loads = l;
            p->rows_per_block = r;
            reblock_all(p);
            const double us = time_all_active(p, 6);
            if (us < best) { best = us; best_l = l; best_r = r; }
        }
    p->loads = best_l;
    p->rows_per_block = best_r;
    reblock_all(p);
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
    const size_t Ra = size_t(row_capacity(o)), R = size_t(vector_rows(o)), E = size_t(o.max_nnz),
                 B = size_t(p->max_blocks), S = size_t(o.max_instances);
    PoolView& v = p->v;
    int* start = p->arena.take<int>(Ra + 1);
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
    float* dstep = p->arena.take<float>(R);
    v.dstep = o.step_metric == HGQP_STEP_JACOBI ? dstep : nullptr;
    v.inst = p->arena.take<PoolInst>(S);
    v.max_inst = int(S);
    v.list_active = p->arena.take<int>(B);
    v.list_verify = p->arena.take<int>(B);
    v.counts = p->arena.take<int>(4);
    v.offsets = p->arena.take<int>(S);
    v.partials = p->arena.take<double>(B * kPartial);
    v.dots = p->arena.take<double>(B);  // apart from the check partials, whose index space is 10 b
    v.resid = p->arena.take<double>(B);
    v.switch_stat = o.fp32_switch;
    v.restart = o.momentum == HGQP_MOMENTUM_RESTART ? 1 : 0;
    v.slabs = o.slabs;
    if (o.slabs) {
        p->sblk_inst = p->arena.take<int>(B);
        p->sblk_local = p->arena.take<int>(B);
        p->slab_dev = p->arena.take<PoolSlab>(S);
        v.sblk_inst = p->sblk_inst;
        v.sblk_local = p->sblk_local;
        v.slab = p->slab_dev;
        int dev = 0, cus = 0;
        HIPCHECK(hipGetDevice(&dev));
        HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        v.slab_groups = std::max(cus, 1);  // 160 KiB of LDS per workgroup: one per CU
    }
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&p->table_host), sizeof(PoolInst) * S, hipHostMallocDefault));
    std::memset(p->table_host, 0, sizeof(PoolInst) * S);  // all kSlotFree
    HIPCHECK(hipMemcpyAsync(v.inst, p->table_host, sizeof(PoolInst) * S, hipMemcpyHostToDevice, p->stream));
    int zero[4] = {0, 0, 0, 0};
    HIPCHECK(hipMemcpyAsync(v.counts, zero, sizeof zero, hipMemcpyHostToDevice, p->stream));
    HIPCHECK(hipEventCreateWithFlags(&p->table_event, hipEventDisableTiming));
    HIPCHECK(hipStreamSynchronize(p->stream));
    p->rows.init(row_capacity(o));
    p->ents.init(o.max_nnz);
    p->blks.init(p->max_blocks);
    p->sblks.init(o.slabs ? p->max_blocks : 0);
    p->slots.resize(S);
    for (int s = int(S) - 1; s >= 0; --s) p->free_slots.push_back(s);
    return p.release();
}

// Checks an instance in the CSR form of hgqp_box_create: sizes, monotone
// starts, columns in range, lower <= upper, positive finite L and tolerance.
// Returns whether H is positive diagonal (the verdict then adds the gap test,
// qp_admm.cpp:115-116).
bool validate_instance(int64_t n, int64_t nnz, const int32_t* start, const int32_t* index, const double* value,
                       const double* cost, const double* lower, const double* upper, const double* x_out,
                       double lipschitz, double tol, const char* who, bool* zero_start) {
    const std::string w(who);
    if (n <= 0 || nnz < 0 || !start || (nnz && (!index || !value)) || !cost || !lower || !upper || !x_out)
        throw Error(HGQP_INVALID_ARGUMENT, w + ": null array or bad size");
    if (n >= (int64_t(1) << 31) - 1 || nnz >= (int64_t(1) << 31) - 1)
        throw Error(HGQP_INVALID_ARGUMENT, w + ": instance too large for 32-bit indices");
    if (!(lipschitz > 0) || !std::isfinite(lipschitz))
        throw Error(HGQP_UNSUPPORTED, w + ": needs a positive finite Lipschitz bound");
    if (!(tol > 0) || !std::isfinite(tol)) throw Error(HGQP_INVALID_ARGUMENT, w + ": bad tolerance");
    if (start[0] != 0 || start[n] != nnz) throw Error(HGQP_INVALID_ARGUMENT, w + ": malformed CSR start");
    bool diag = true;
    *zero_start = true;
    for (int64_t i = 0; i < n; ++i) {
        if (start[i + 1] < start[i]) throw Error(HGQP_INVALID_ARGUMENT, w + ": CSR start not monotone");
        *zero_start = *zero_start && lower[i] <= 0 && 0 <= upper[i];
        bool pos = false;
        for (int32_t k = start[i]; k < start[i + 1]; ++k) {
            if (index[k] < 0 || index[k] >= n) throw Error(HGQP_INVALID_ARGUMENT, w + ": column out of range");
            if (index[k] == i) pos = value[k] > 0;
            else diag = false;
        }
        diag = diag && pos;
        if (std::isnan(lower[i]) || std::isnan(upper[i]) || lower[i] > upper[i])
            throw Error(HGQP_INVALID_ARGUMENT, w + ": bounds must satisfy lower <= upper");
    }
    return diag;
}

// Host part of an admission: a slot with its row, entry and block ranges; the
// block table is staged in `blk` (room for 3 (n + 1) ints) and its upload
// enqueued. Returns the slot, or -1 when no slot or no contiguous room in the
// budget is free (nothing is held then).
int reserve(hgqp_pool* p, int64_t n, int64_t nnz, const int32_t* start, int32_t* blk) {
    if (p->free_slots.empty()) return -1;
    const int64_t rows_len = p->opt.slabs ? (n + 1 + 3) / 4 * 4 : n + 1;
    const int64_t r0 = p->rows.alloc(rows_len);
    if (r0 < 0) return -1;
    const int64_t e0 = p->ents.alloc(std::max<int64_t>(nnz, 1));
    if (e0 < 0) {
        p->rows.release(r0, rows_len);
        return -1;
    }
    const int s = p->free_slots.back();
    SlotHost& h = p->slots[s];
    h = SlotHost();
    h.n = n;
    h.nnz = nnz;
    h.row0 = r0;
    h.rows_len = rows_len;
    h.ent0 = e0;
    if (p->opt.slabs) {
        // slab blocks of at most kPoolSlabRows rows and about slab_block_entries entries
        const int64_t per = p->opt.slab_block_entries ? p->opt.slab_block_entries : 65536;
        const int target = int(std::max<int64_t>(1, (nnz + per - 1) / per));
        h.srb = slab_blocks(n, start, target, kPoolSlabRows);
        h.nsblk = int64_t(h.srb.size()) - 1;
        h.sblk0 = p->sblks.alloc(h.nsblk);
        if (h.sblk0 < 0) {
            p->rows.release(r0, rows_len);
            p->ents.release(e0, std::max<int64_t>(nnz, 1));
            h = SlotHost();
            return -1;
        }
    }
    if (!place_blocks(p, s, start, blk)) {
        p->rows.release(r0, rows_len);
        p->ents.release(e0, std::max<int64_t>(nnz, 1));
        if (h.nsblk) p->sblks.release(h.sblk0, h.nsblk);
        h = SlotHost();
        return -1;
    }
    h.start.assign(start, start + n + 1);
    h.used = true;
    h.gen = p->next_gen++;
    p->free_slots.pop_back();
    ++p->resident;
    return s;
}

// The device slot entry of a new instance: phase INIT (x0 and the iteration-0
// FP64 check run in the next period).
PoolInst slot_entry(const SlotHost& h, double lipschitz, double tol, uint64_t limit, int32_t maximize, double offset,
                    bool diag, bool zero_start) {
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
    t.r0 = int(h.row0);
    t.n = int(h.n);
    t.gen = h.gen;
    t.zero_start = zero_start ? 1 : 0;
    t.sb0 = int(h.sblk0);
    t.nsblk = int(h.nsblk);
    return t;
}

// Device part of an admission, enqueued on the pool stream: the uploads, the
// move to pool numbering, the FP32 values and the slot entry. All host memory
// named here must stay valid until the stream passes this point. index and
// value may also be device memory (a server request assembled on the device):
// their copies take the direction from the pointers.
void upload(hgqp_pool* p, int s, const int32_t* start, const int32_t* index, const double* value, const double* cost,
            const double* lower, const double* upper, const PoolInst* entry) {
    const SlotHost& h = p->slots[s];
    hipStream_t st = p->stream;
    int* dstart = const_cast<int*>(p->v.start) + h.row0;
    int* dcol = const_cast<int*>(p->v.col) + h.ent0;
    HIPCHECK(hipMemcpyAsync(dstart, start, (h.n + 1) * 4, hipMemcpyHostToDevice, st));
    if (h.nnz) {
        HIPCHECK(hipMemcpyAsync(dcol, index, h.nnz * 4, hipMemcpyDefault, st));
        HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.val) + h.ent0, value, h.nnz * 8, hipMemcpyDefault, st));
        launch_convert<float, double>(h.nnz, p->v.val + h.ent0, const_cast<float*>(p->v.val32) + h.ent0, st);
    }
    if (p->opt.slabs) {
        // the slab layout from the instance's own (local) CSR, before the move to pool numbering
        SlotHost& hs = p->slots[s];
        hs.slab = std::make_shared<SlabStore>();
        // the pool's row sums need not follow stored order: long runs are chained
        if (!build_slabs(h.n, h.nnz, h.srb, 1, dstart, dcol, p->v.val + h.ent0, *hs.slab, st, true))
            throw Error(HGQP_DEVICE_ERROR, "internal: slab layout of an admitted instance");
        const SlabMatrix& m = hs.slab->m;
        hs.slab_desc = PoolSlab{m.rb, m.tc, m.row, m.col, m.val, m.blocks, m.slabs, int(hs.slab->chunks - 1),
                                m.long_runs, m.chained};
        HIPCHECK(hipMemcpyAsync(p->slab_dev + s, &hs.slab_desc, sizeof(PoolSlab), hipMemcpyHostToDevice, st));
        std::vector<int32_t>& stage = hs.srb_stage;
        stage.resize(size_t(2 * h.nsblk));
        for (int64_t k = 0; k < h.nsblk; ++k) {
            stage[size_t(k)] = s;
            stage[size_t(h.nsblk + k)] = int32_t(k);
        }
        HIPCHECK(hipMemcpyAsync(p->sblk_inst + h.sblk0, stage.data(), h.nsblk * 4, hipMemcpyHostToDevice, st));
        HIPCHECK(hipMemcpyAsync(p->sblk_local + h.sblk0, stage.data() + h.nsblk, h.nsblk * 4, hipMemcpyHostToDevice,
                                st));
    }
    launch_pool_offset(dstart, h.n, int(h.ent0), dcol, h.nnz, int(h.row0), st);
    if (p->v.dstep) launch_row_abs(h.n, dstart, p->v.val, entry->lipschitz, nullptr, p->v.dstep + h.row0, st);
    HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.c) + h.row0, cost, h.n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.lo) + h.row0, lower, h.n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(const_cast<double*>(p->v.up) + h.row0, upper, h.n * 8, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(p->v.inst + s, entry, sizeof(PoolInst), hipMemcpyHostToDevice, st));
    HIPCHECK(hipGetLastError());
}

// The public admission (hgqp_pool_admit): validated, staged in host
// temporaries, synchronous.
int admit(hgqp_pool* p, int64_t n, int64_t nnz, const int32_t* start, const int32_t* index, const double* value,
          const double* cost, const double* lower, const double* upper, double lipschitz, int32_t maximize,
          double offset, double tol, uint64_t limit, int64_t tag, double* x_out) {
    bool zero = false;
    const bool diag = validate_instance(n, nnz, start, index, value, cost, lower, upper, x_out, lipschitz, tol,
                                        "hgqp_pool_admit", &zero);
    if (p->opt.slabs) {
        if (nnz < 1 || (n + kSlabWidth - 1) / kSlabWidth > kPoolSlabMax)
            throw Error(HGQP_UNSUPPORTED, "hgqp_pool_admit: a pool with slabs takes instances with stored entries "
                                          "and n <= 786432");
        for (int64_t i = 0; i < n; ++i)
            for (int32_t k = start[i] + 1; k < start[i + 1]; ++k)
                if (index[k] < index[k - 1])
                    throw Error(HGQP_UNSUPPORTED, "hgqp_pool_admit: a pool with slabs needs ascending columns in "
                                                  "every row");
    }
    if (!p->retired.empty()) {
        HIPCHECK(hipStreamSynchronize(p->stream));  // no queued launch reads a retired layout any more
        p->retired.clear();
    }
    std::vector<int32_t> blk(3 * (n + 1));
    const int s = reserve(p, n, nnz, start, blk.data());
    if (s < 0) return HGQP_BUSY;
    SlotHost& h = p->slots[s];
    h.tag = tag;
    h.x_out = x_out;
    const PoolInst entry = slot_entry(h, lipschitz, tol, limit, maximize, offset, diag, zero);
    upload(p, s, start, index, value, cost, lower, upper, &entry);
    HIPCHECK(hipStreamSynchronize(p->stream));  // blk and entry are host temporaries
    return HGQP_OK;
}

// Result record of a finished slot from its table entry.
hgqp_pool_result result_of(const SlotHost& h, const PoolInst& t, int s) {
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
    return r;
}

// Frees finished slots: their device entries go FREE in stream order (one
// small launch per 62 slots), their ranges and slots return to the host
// allocators at once (later uploads are ordered after the release on the
// stream).
void release_slots(hgqp_pool* p, const std::vector<int>& ids) {
    if (ids.empty()) return;
    launch_pool_release(p->v.inst, ids.data(), int(ids.size()), p->stream);
    HIPCHECK(hipGetLastError());
    for (int s : ids) {
        SlotHost& h = p->slots[s];
        p->rows.release(h.row0, h.rows_len);
        p->ents.release(h.ent0, std::max<int64_t>(h.nnz, 1));
        p->blks.release(h.blk0, h.nblk);
        if (h.nsblk) p->sblks.release(h.sblk0, h.nsblk);
        if (h.slab) p->retired.push_back(std::move(h.slab));  // queued launches may still read it
        h = SlotHost();
        p->free_slots.push_back(s);
        --p->resident;
    }
}

// Slots that a table snapshot reports finished and that still hold the
// instance the snapshot saw.
std::vector<int> finished_slots(const hgqp_pool* p, const PoolInst* table) {
    std::vector<int> done;
    for (int s = 0; s < p->opt.max_instances; ++s)
        if (p->slots[s].used && table[s].phase == kSlotDone && table[s].gen == p->slots[s].gen) done.push_back(s);
    return done;
}

struct HostCopy {
    void* dst;
    const void* src;
    size_t bytes;
};

// Host copies of collected iterates: pieces of at most 4 MiB dealt to up to
// eight threads once there are more than 8 MiB in all. The copy into a
// caller's array is also its first touch of those pages, and both the page
// faults and the pinned-to-pageable bandwidth scale with threads.
void host_copies(const std::vector<HostCopy>& jobs) {
    constexpr size_t kPiece = size_t(4) << 20;
    size_t total = 0;
    for (const HostCopy& j : jobs) total += j.bytes;
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const int threads = total <= 2 * kPiece ? 1 : int(std::min<size_t>(std::min<size_t>(8, hw), total / kPiece));
    if (threads <= 1) {
        for (const HostCopy& j : jobs) std::memcpy(j.dst, j.src, j.bytes);
        return;
    }
    std::vector<HostCopy> pieces;
    for (const HostCopy& j : jobs)
        for (size_t o = 0; o < j.bytes; o += kPiece)
            pieces.push_back({static_cast<char*>(j.dst) + o, static_cast<const char*>(j.src) + o,
                              std::min(kPiece, j.bytes - o)});
    auto work = [&](int t) {
        for (size_t i = size_t(t); i < pieces.size(); i += size_t(threads))
            std::memcpy(pieces[i].dst, pieces[i].src, pieces[i].bytes);
    };
    std::vector<std::thread> crew;
    int started = 1;
    try {
        for (int t = 1; t < threads; ++t) {
            crew.emplace_back(work, t);
            ++started;
        }
    } catch (...) {  // no more threads: this one takes the shares left over
    }
    for (int t = started; t < threads; ++t) work(t);
    work(0);
    for (std::thread& th : crew) th.join();
}

// hgqp_pool_collect: reads the slot table after the queued periods, copies x of
// finished slots to their callers and frees the slots. The iterates go to a
// pinned staging buffer first, one device-to-host copy per run of adjacent row
// ranges, then to the callers' arrays on several host threads.
void harvest(hgqp_pool* p) {
    if (!p->table_pending) {
        HIPCHECK(hipMemcpyAsync(p->table_host, p->v.inst, sizeof(PoolInst) * p->opt.max_instances,
                                hipMemcpyDeviceToHost, p->stream));
        HIPCHECK(hipEventRecord(p->table_event, p->stream));
    }
    HIPCHECK(hipEventSynchronize(p->table_event));
    p->table_pending = false;
    const std::vector<int> done = finished_slots(p, p->table_host);
    if (done.empty()) return;
    std::vector<int> by_row(done);
    std::sort(by_row.begin(), by_row.end(), [&](int a, int b) { return p->slots[a].row0 < p->slots[b].row0; });
    size_t rows = 0;
    for (int s : by_row) rows += size_t(p->slots[s].rows_len);
    if (p->x_stage_bytes < rows * 8) {
        HIPCHECK(hipStreamSynchronize(p->stream));
        if (p->x_stage) HIPCHECK(hipHostFree(p->x_stage));
        p->x_stage = nullptr;
        p->x_stage_bytes = 0;
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&p->x_stage), rows * 8, hipHostMallocDefault));
        p->x_stage_bytes = rows * 8;
    }
    std::vector<HostCopy> jobs;
    size_t at = 0;  // staging offset (rows) of the current run
    for (size_t k = 0; k < by_row.size();) {
        const int64_t r0 = p->slots[by_row[k]].row0;
        int64_t end = r0;
        size_t e = k;
        for (; e < by_row.size() && p->slots[by_row[e]].row0 == end; ++e) {
            const SlotHost& h = p->slots[by_row[e]];
            jobs.push_back({h.x_out, p->x_stage + at + size_t(h.row0 - r0), size_t(h.n) * 8});
            end += h.rows_len;
        }
        HIPCHECK(hipMemcpyAsync(p->x_stage + at, p->v.x + r0, size_t(end - r0) * 8, hipMemcpyDeviceToHost,
                                p->stream));
        at += size_t(end - r0);
        k = e;
    }
    for (int s : done) p->ready.push_back(result_of(p->slots[s], p->table_host[s], s));
    release_slots(p, done);
    HIPCHECK(hipStreamSynchronize(p->stream));
    host_copies(jobs);
}

}  // namespace

// ---- serving loop -------------------------------------------------------------------
namespace {

// Pinned host memory for staged transfers: one hipHostMalloc region handed out
// first-fit in 256-byte units; thread-safe (submit and the loop share it).
struct PinnedArena {
    char* base = nullptr;
    int64_t units = 0;
    RangeAlloc ranges;
    std::mutex mu;
    static constexpr int64_t kUnit = 256;
    void init(size_t bytes) {
        HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&base), bytes, hipHostMallocDefault));
        units = int64_t(bytes / kUnit);
        ranges.init(units);
    }
    // nullptr when the region has no room (the caller falls back to pageable memory)
    char* take(size_t bytes) {
        std::lock_guard<std::mutex> lk(mu);
        const int64_t u = ranges.alloc(int64_t((bytes + kUnit - 1) / kUnit));
        return u < 0 ? nullptr : base + u * kUnit;
    }
    void give(char* ptr, size_t bytes) {
        if (!ptr) return;
        std::lock_guard<std::mutex> lk(mu);
        ranges.release((ptr - base) / kUnit, int64_t((bytes + kUnit - 1) / kUnit));
    }
    ~PinnedArena() {
        if (base) (void)hipHostFree(base);
    }
};

constexpr size_t kStagingBytes = size_t(1) << 30;  // pinned staging of queued and uploading requests

size_t align16(size_t b) { return (b + 15) & ~size_t(15); }

}  // namespace

struct hgqp_server {
    hgqp_pool* pool = nullptr;
    hipStream_t copy_stream = nullptr;  // x read-back of finished slots, beside the next period
    hipEvent_t period_done[2] = {nullptr, nullptr};
    hipEvent_t copies_done = nullptr;
    PoolInst* tables[2] = {nullptr, nullptr};  // pinned slot-table snapshots, one per period in flight
    PinnedArena staging;
    std::thread thread;
    std::mutex mu;
    std::condition_variable cv;
    bool stopping = false;
    std::string error;

    // A submitted instance: its arrays in one host block (pinned staging, or
    // pageable when the staging region is full), laid out as start, index,
    // block table, value, cost, lower, upper, slot entry. A request assembled
    // on the device keeps its columns and values there (dev) and its block
    // leaves out index and value.
    struct Request {
        int64_t n = 0, nnz = 0;
        double lipschitz = 0, offset = 0, tol = 0;
        int32_t maximize = 0;
        bool diag = false, zero = false;
        uint64_t limit = 0;
        int64_t tag = 0;
        double* x_out = nullptr;
        char* block = nullptr;  // pinned
        size_t block_bytes = 0;
        std::vector<char> heap;  // pageable fallback
        int32_t *start = nullptr, *index = nullptr, *blk = nullptr;
        double *value = nullptr, *cost = nullptr, *lower = nullptr, *upper = nullptr;
        PoolInst* entry = nullptr;
        double* x_stage = nullptr;  // pinned x read-back while resident (nullptr: straight to x_out)
        int slot = -1;
        std::chrono::steady_clock::time_point submitted, admitted;
        std::unique_ptr<DeviceCsr> dev;  // device-resident columns and values, or null

        // entries: room for index and value in the block (host sources)
        static size_t bytes_for(int64_t n, int64_t nnz, bool entries = true) {
            const int64_t e = entries ? nnz : 0;
            return align16((n + 1) * 4) + align16(e * 4) + align16(3 * (n + 1) * 4) + align16(e * 8) +
                   3 * align16(n * 8) + align16(sizeof(PoolInst));
        }
        void carve(char* p, bool entries = true) {
            const int64_t e = entries ? nnz : 0;
            start = reinterpret_cast<int32_t*>(p);
            p += align16((n + 1) * 4);
            index = entries ? reinterpret_cast<int32_t*>(p) : nullptr;
            p += align16(e * 4);
            blk = reinterpret_cast<int32_t*>(p);
            p += align16(3 * (n + 1) * 4);
            value = entries ? reinterpret_cast<double*>(p) : nullptr;
            p += align16(e * 8);
            cost = reinterpret_cast<double*>(p);
            p += align16(n * 8);
            lower = reinterpret_cast<double*>(p);
            p += align16(n * 8);
            upper = reinterpret_cast<double*>(p);
            p += align16(n * 8);
            entry = reinterpret_cast<PoolInst*>(p);
        }
    };
    std::deque<std::unique_ptr<Request>> queue;           // submitted, not admitted
    std::map<int, std::unique_ptr<Request>> resident;     // by slot (touched by the loop thread only)
    std::deque<hgqp_served> done;
    int64_t outstanding = 0;                              // submitted, not polled
    // device CSRs of finished requests: freed by the next submission, on the
    // caller's thread (hipFree waits for the device), never by the loop
    std::vector<std::unique_ptr<DeviceCsr>> retired_dev;

    void give_block(Request& r) {
        staging.give(r.block, r.block_bytes);
        r.block = nullptr;
        r.heap.clear();
        r.heap.shrink_to_fit();
    }

    // Admits queued requests in arrival order while slots and budget allow;
    // returns the admitted ones (their uploads are enqueued, not finished).
    std::vector<Request*> admit_queued() {
        std::vector<std::unique_ptr<Request>> take;
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!queue.empty()) {
                take.push_back(std::move(queue.front()));
                queue.pop_front();
            }
        }
        std::vector<Request*> admitted;
        size_t k = 0;
        for (; k < take.size(); ++k) {
            Request& r = *take[k];
            const int s = reserve(pool, r.n, r.nnz, r.start, r.blk);
            if (s < 0) break;  // later requests wait behind this one (arrival order)
            SlotHost& h = pool->slots[s];
            h.tag = r.tag;
            h.x_out = r.x_out;
            *r.entry = slot_entry(h, r.lipschitz, r.tol, r.limit, r.maximize, r.offset, r.diag, r.zero);
            upload(pool, s, r.start, r.index, r.value, r.cost, r.lower, r.upper, r.entry);
            if (!r.block) HIPCHECK(hipStreamSynchronize(pool->stream));  // pageable source: finish now
            r.x_stage = reinterpret_cast<double*>(staging.take(size_t(r.n) * 8));
            r.slot = s;
            r.admitted = std::chrono::steady_clock::now();
            admitted.push_back(take[k].get());
            resident[s] = std::move(take[k]);
        }
        if (k < take.size()) {
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = take.size(); i-- > k;) queue.push_front(std::move(take[i]));
        }
        return admitted;
    }

    // Reads back x of the slots `table` reports finished (on the copy stream,
    // beside the period now running), hands the results to poll and frees the
    // slots in stream order.
    void deliver(const PoolInst* table, hipEvent_t table_ready) {
        const std::vector<int> fin = finished_slots(pool, table);
        if (fin.empty()) return;
        HIPCHECK(hipStreamWaitEvent(copy_stream, table_ready, 0));
        for (int s : fin) {
            const Request& r = *resident[s];
            double* dst = r.x_stage ? r.x_stage : r.x_out;
            HIPCHECK(hipMemcpyAsync(dst, pool->v.x + pool->slots[s].row0, r.n * 8, hipMemcpyDeviceToHost, copy_stream));
        }
        HIPCHECK(hipEventRecord(copies_done, copy_stream));
        HIPCHECK(hipEventSynchronize(copies_done));
        // later uploads into these ranges wait for the read-back on the device too
        HIPCHECK(hipStreamWaitEvent(pool->stream, copies_done, 0));
        const auto now = std::chrono::steady_clock::now();
        std::vector<hgqp_served> out;
        std::vector<std::unique_ptr<DeviceCsr>> dead;
        for (int s : fin) {
            std::unique_ptr<Request> r = std::move(resident[s]);
            resident.erase(s);
            if (r->dev) dead.push_back(std::move(r->dev));
            if (r->x_stage) std::memcpy(r->x_out, r->x_stage, size_t(r->n) * 8);
            staging.give(reinterpret_cast<char*>(r->x_stage), size_t(r->n) * 8);
            hgqp_served sv;
            sv.result = result_of(pool->slots[s], table[s], s);
            sv.queued_seconds = std::chrono::duration<double>(r->admitted - r->submitted).count();
            sv.latency_seconds = std::chrono::duration<double>(now - r->submitted).count();
            out.push_back(sv);
        }
        release_slots(pool, fin);
        std::lock_guard<std::mutex> lk(mu);
        for (const auto& sv : out) done.push_back(sv);
        for (auto& d : dead) retired_dev.push_back(std::move(d));
    }

    // Queues a filled request for the loop (the caller's thread).
    void enqueue(std::unique_ptr<Request> r) {
        std::vector<std::unique_ptr<DeviceCsr>> dead;
        {
            std::lock_guard<std::mutex> lk(mu);
            dead.swap(retired_dev);
            if (!error.empty()) {
                give_block(*r);
                throw Error(HGQP_DEVICE_ERROR, "hgqp_server: " + error);
            }
            if (stopping) {
                give_block(*r);
                throw Error(HGQP_BUSY, "hgqp_server: stopping");
            }
            queue.push_back(std::move(r));
            ++outstanding;
        }
        cv.notify_all();
    }

    // Each turn enqueues period i (after the admissions of this turn), then
    // waits for period i - 1 and delivers what it finished while period i runs.
    void loop() {
        try {
            int cur = 0;
            bool prev = false;                    // a period is in flight from the last turn
            std::vector<Request*> prev_uploads;   // admitted last turn: staging freed once that period ends
            for (;;) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stopping || !queue.empty() || pool->resident > 0 || prev; });
                    if (stopping && queue.empty() && pool->resident == 0 && !prev) return;
                }
                std::vector<Request*> uploads = admit_queued();
                const bool period = pool->resident > 0;
                if (period) {
                    if (!pool->tuned) autotune(pool);
                    // new instances take their first check now and step in this period
                    if (!uploads.empty())
                        launch_pool_checks(pool->v, pool->loads, int(std::min<int64_t>(pool->max_blocks, 1 << 30)),
                                           pool->stream);
                    launch_periods(pool, 1);
                    HIPCHECK(hipMemcpyAsync(tables[cur], pool->v.inst, sizeof(PoolInst) * pool->opt.max_instances,
                                            hipMemcpyDeviceToHost, pool->stream));
                    HIPCHECK(hipEventRecord(period_done[cur], pool->stream));
                }
                if (prev) {
                    HIPCHECK(hipEventSynchronize(period_done[cur ^ 1]));
                    for (Request* r : prev_uploads) give_block(*r);
                    deliver(tables[cur ^ 1], period_done[cur ^ 1]);
                }
                prev_uploads = std::move(uploads);
                prev = period;
                cur ^= 1;
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
        if (pool && pool->stream) (void)hipStreamSynchronize(pool->stream);
        if (copy_stream) {
            (void)hipStreamSynchronize(copy_stream);
            (void)hipStreamDestroy(copy_stream);
        }
        for (auto& e : period_done)
            if (e) (void)hipEventDestroy(e);
        if (copies_done) (void)hipEventDestroy(copies_done);
        for (auto& t : tables)
            if (t) (void)hipHostFree(t);
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
    o->step_metric = HGQP_STEP_JACOBI;
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

int hgqp_pool_time_step(hgqp_pool* pool, int32_t reps, double* avg_us, double* bytes) {
    return guarded([&] {
        check_pool(pool);
        if (reps < 1 || !avg_us || !bytes) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_time_step: need reps >= 1 and outputs");
        if (!pool->resident) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_pool_time_step: no resident instance");
        double b = 0;
        const double vectors = pool->opt.step_metric == HGQP_STEP_JACOBI ? 8 : 7;
        for (const SlotHost& sh : pool->slots)
            if (sh.used) b += 8.0 * double(sh.nnz) + 4.0 * double(sh.n + 1) + vectors * 4.0 * double(sh.n);
        *bytes = b;
        *avg_us = time_all_active(pool, reps);
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
        if (o.slabs)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_start: the server's pool runs without slabs (slabs = 0)");
        auto s = std::make_unique<hgqp_server>();
        s->pool = create_pool(o);
        HIPCHECK(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
        for (auto& e : s->period_done) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&s->copies_done, hipEventDisableTiming));
        for (auto& t : s->tables) {
            HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&t), sizeof(PoolInst) * o.max_instances,
                                   hipHostMallocDefault));
            std::memset(t, 0, sizeof(PoolInst) * o.max_instances);
        }
        s->staging.init(kStagingBytes);
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
        // checked here, on the caller's thread: the loop never sees an invalid instance
        bool zero = false;
        const bool diag = validate_instance(n, nnz, start, index, value, cost, lower, upper, x_out, lipschitz,
                                            tolerance, "hgqp_server_submit", &zero);
        if (n + 1 > server->pool->opt.max_rows || nnz > server->pool->opt.max_nnz)
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit: instance larger than the pool budget");
        auto r = std::make_unique<hgqp_server::Request>();
        r->n = n;
        r->nnz = nnz;
        r->block_bytes = hgqp_server::Request::bytes_for(n, nnz);
        r->block = server->staging.take(r->block_bytes);
        if (!r->block) r->heap.resize(r->block_bytes);
        r->carve(r->block ? r->block : r->heap.data());
        std::memcpy(r->start, start, size_t(n + 1) * 4);
        if (nnz) {
            std::memcpy(r->index, index, size_t(nnz) * 4);
            std::memcpy(r->value, value, size_t(nnz) * 8);
        }
        std::memcpy(r->cost, cost, size_t(n) * 8);
        std::memcpy(r->lower, lower, size_t(n) * 8);
        std::memcpy(r->upper, upper, size_t(n) * 8);
        r->lipschitz = lipschitz;
        r->maximize = maximize;
        r->offset = offset;
        r->tol = tolerance;
        r->diag = diag;
        r->zero = zero;
        r->limit = iteration_limit;
        r->tag = tag;
        r->x_out = x_out;
        r->submitted = std::chrono::steady_clock::now();
        server->enqueue(std::move(r));
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

namespace hgqp {

void server_submit_device(hgqp_server* server, std::unique_ptr<DeviceCsr> csr, const double* cost,
                          const double* lower, const double* upper, double lipschitz, int32_t maximize,
                          double offset, double tolerance, uint64_t iteration_limit, int64_t tag, double* x_out,
                          bool diag) {
    if (!server || !csr) throw Error(HGQP_INVALID_ARGUMENT, "null hgqp_server handle");
    const int64_t n = csr->n, nnz = csr->nnz;
    const int32_t* start = csr->host_start.data();
    // the ingest produced the CSR (columns in range, finite nonzero values);
    // the rest is checked as hgqp_server_submit checks it
    if (n <= 0 || int64_t(csr->host_start.size()) != n + 1 || !cost || !lower || !upper || !x_out)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit_problem: null array or bad size");
    if (!(lipschitz > 0) || !std::isfinite(lipschitz))
        throw Error(HGQP_UNSUPPORTED, "hgqp_server_submit_problem: needs a positive finite Lipschitz bound");
    if (!(tolerance > 0) || !std::isfinite(tolerance))
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit_problem: bad tolerance");
    if (start[0] != 0 || start[n] != nnz) throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit_problem: bad CSR");
    bool zero = true;
    for (int64_t i = 0; i < n; ++i) {
        if (std::isnan(lower[i]) || std::isnan(upper[i]) || lower[i] > upper[i])
            throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit_problem: bounds must satisfy lower <= upper");
        zero = zero && lower[i] <= 0 && 0 <= upper[i];
    }
    if (n + 1 > server->pool->opt.max_rows || nnz > server->pool->opt.max_nnz)
        throw Error(HGQP_INVALID_ARGUMENT, "hgqp_server_submit_problem: instance larger than the pool budget");
    auto r = std::make_unique<hgqp_server::Request>();
    r->n = n;
    r->nnz = nnz;
    r->block_bytes = hgqp_server::Request::bytes_for(n, nnz, false);
    r->block = server->staging.take(r->block_bytes);
    if (!r->block) r->heap.resize(r->block_bytes);
    r->carve(r->block ? r->block : r->heap.data(), false);
    std::memcpy(r->start, start, size_t(n + 1) * 4);
    std::memcpy(r->cost, cost, size_t(n) * 8);
    std::memcpy(r->lower, lower, size_t(n) * 8);
    std::memcpy(r->upper, upper, size_t(n) * 8);
    r->index = csr->col;
    r->value = csr->val;
    r->dev = std::move(csr);
    r->lipschitz = lipschitz;
    r->maximize = maximize;
    r->offset = offset;
    r->tol = tolerance;
    r->diag = diag;
    r->zero = zero;
    r->limit = iteration_limit;
    r->tag = tag;
    r->x_out = x_out;
    r->submitted = std::chrono::steady_clock::now();
    server->enqueue(std::move(r));
}

}  // namespace hgqp
